# AvSVD.jl -- Julia glue that keeps the reference's stage-1 entry points and
# calls the MI355X engine (libavsvd.so) through ccall.
#
# Drop-in for (paths relative to the reference repository root):
#   KAbasedSVD/src/TiledMatrix.jl:41-67      TiledMatrix(A, bm, bn; tiledim2d)
#   KAbasedSVD/src/tiledalgos.jl:6-34        tiled_QR!(A; kend), Blockbidiag!(A; kend)
#   KAbasedSVD/src/tiledalgos.jl:232-267     OOC_QR!/OOC_Bidiag!/OOC_SVD! (streamed), compile_kernels
#   KAbasedSVD/src/bulgechasing.jl:33-53     bidiagonalize, my_CU_svdval, diag_lapack
# A caller replaces
#   include("QRkernels.jl"); include("QRkernelscompile.jl"); include("TiledMatrix.jl");
#   include("tiledalgos.jl"); include("bulgechasing.jl")
# by `include("AvSVD.jl"); using .AvSVD` and keeps its code.  Host matrices
# (Matrix{Float64} / Matrix{Float32}) are passed; device memory lives in the
# library (tile-major, HBM).  Error behaviour: the reference's
# error("Not supported") for bad shapes is kept; other failures raise an
# ErrorException carrying avs_last_error().
module AvSVD

export TiledMatrix, tiled_QR!, myormqr!, Blockbidiag!, OOC_QR!, OOC_Bidiag!, OOC_SVD!, compile_kernels,
       bidiagonalize, my_CU_svdval, diag_lapack, svdvals_gpu, tune_set!

const libavsvd = get(ENV, "AVSVD_LIB", joinpath(@__DIR__, "..", "libavsvd.so"))

# mirrors avs_opts in include/avsvd.h
struct AvsOpts
    stream::Ptr{Cvoid}
    sched::Int32
    tree::Int32
    domain::Int32
    update_mode::Int32
    workers::Int32
    wait_ms::Int32
    phase_timing::Int32
    single_stream::Int32
    reserve_cus::Int32
    graph::Int32
    masked_modes::Int32
    reserved::NTuple{3,Int32}
end

function default_opts()
    o = Ref{AvsOpts}()
    ccall((:avs_default_opts, libavsvd), Cvoid, (Ref{AvsOpts},), o)
    return o[]
end

function check(st::Integer)
    if st != 0
        msg = unsafe_string(ccall((:avs_last_error, libavsvd), Cstring, ()))
        # the reference's shape errors keep their text
        occursin("Not supported", msg) ? error("Not supported") : error("avsvd status $st: $msg")
    end
    return nothing
end

dtype(::Type{Float64}) = Int32(0)
dtype(::Type{Float32}) = Int32(1)

const TILE = Ref(32)

# compile_kernels(; tilesize, tilefactor, tiledim2d)  (tiledalgos.jl:262-267):
# kernels are built ahead of time for b in (32, 64, 128); record the default.
function compile_kernels(; tilesize::Int=32, tilefactor::Int=1, tiledim2d::Bool=false)
    tilesize in (32, 64, 128) || error("Not supported")
    TILE[] = tilesize
    return nothing
end

# tune_set!(T, n_upto, b): one entry of the library's tile-size table (the
# autotuner's measured choice for served requests of order <= n_upto).
tune_set!(::Type{T}, n_upto::Integer, b::Integer) where {T<:Union{Float32,Float64}} =
    check(ccall((:avs_tune_set, libavsvd), Int32, (Int32, Int64, Int32), dtype(T), n_upto, b))

mutable struct TiledMatrix{T<:Union{Float32,Float64},AT<:AbstractMatrix{T}}
    handle::Ptr{Cvoid}
    A::AT          # the caller's array, host Matrix or device array; results go back into it
    device::Bool   # A lives in GPU memory (pointer(A) is a device address)
    M::Int
    N::Int
    m::Int
    n::Int
    no_tiles::Int
end

# TiledMatrix(A, blocksize_m, blocksize_n; tiledim2d)  (TiledMatrix.jl:41-67).
# A host Matrix is uploaded with avs_set_matrix_host.  Any other dense
# column-major array whose pointer(A) is a GPU address (a ROCArray{T,2} from
# AMDGPU.jl) is read in place on the device with avs_set_matrix_device, and the
# algorithms write their result back into that same device array, as the
# reference's TiledMatrix aliases its CuArray.  The copy runs on the handle's
# stream: work still queued on the array (another stream) must be finished
# first (AMDGPU.synchronize()), or opts.stream set to the caller's stream.
function TiledMatrix(A::AbstractMatrix{T}, blocksize_m::Int, blocksize_n::Int; tiledim2d::Int=0,
                     opts::AvsOpts=default_opts()) where {T<:Union{Float32,Float64}}
    if (blocksize_m != blocksize_n) || size(A, 1) != size(A, 2) || mod(size(A, 1), blocksize_n) != 0
        error("Not supported")
    end
    stride(A, 1) == 1 || error("A must be column-major with unit row stride")
    device = !(A isa Array)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:avs_create, libavsvd), Int32, (Ref{Ptr{Cvoid}}, Int32, Int64, Int32, Ref{AvsOpts}),
                h, dtype(T), size(A, 1), blocksize_m, opts))
    sym = device ? :avs_set_matrix_device : :avs_set_matrix_host
    st = ccall((sym, libavsvd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Int64), h[], Ptr{Cvoid}(pointer(A)), stride(A, 2))
    if st != 0
        ccall((:avs_destroy, libavsvd), Int32, (Ptr{Cvoid},), h[])
        check(st)
    end
    M = TiledMatrix{T,typeof(A)}(h[], A, device, size(A, 1), size(A, 2), blocksize_m, blocksize_n,
                                 div(size(A, 1), blocksize_n))
    finalizer(x -> ccall((:avs_destroy, libavsvd), Int32, (Ptr{Cvoid},), x.handle), M)
    return M
end

# the result back into the caller's array (host: avs_get_matrix_host; device:
# avs_get_matrix_device on the handle's stream, then a wait for it)
function sync_to_host!(A::TiledMatrix{T}) where {T}
    if A.device
        check(ccall((:avs_get_matrix_device, libavsvd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                    A.handle, Ptr{Cvoid}(pointer(A.A)), stride(A.A, 2)))
        check(ccall((:avs_synchronize, libavsvd), Int32, (Ptr{Cvoid},), A.handle))
    else
        check(ccall((:avs_get_matrix_host, libavsvd), Int32, (Ptr{Cvoid}, Ptr{T}, Int64), A.handle, A.A,
                    stride(A.A, 2)))
    end
    return A
end

# tiled_QR!(A; kend)  (tiledalgos.jl:6-15)
function tiled_QR!(A::TiledMatrix; kend=A.no_tiles)
    check(ccall((:avs_tiled_qr, libavsvd), Int32, (Ptr{Cvoid}, Int32), A.handle, kend))
    return sync_to_host!(A)
end

# myormqr!(C, A)  (CAQR/src/qr_kernels.jl:256-265): C <- Q' C (applyt) or Q C for the Q of
# A's last tiled_QR!, on a device array C (column-major, n x ncols, ncols a multiple of b)
function myormqr!(C::AbstractMatrix{T}, A::TiledMatrix{T}; applyt::Bool=true) where {T}
    C isa Array && error("myormqr!: C must be a device array")
    check(ccall((:avs_apply_q, libavsvd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int32),
                A.handle, Ptr{Cvoid}(pointer(C)), stride(C, 2), size(C, 2), Int32(applyt)))
    check(ccall((:avs_synchronize, libavsvd), Int32, (Ptr{Cvoid},), A.handle))
    return C
end

# Blockbidiag!(A; kend)  (tiledalgos.jl:17-34)
function Blockbidiag!(A::TiledMatrix; kend=A.no_tiles)
    check(ccall((:avs_blockbidiag, libavsvd), Int32, (Ptr{Cvoid}, Int32), A.handle, kend))
    return sync_to_host!(A)
end

# OOC_alg!  (tiledalgos.jl:232-251) on a host Matrix, in place: the first
# n/b - kswitch steps stream A through the GPU one tile row or tile column at a
# time (LargeTiledMatrix, TiledMatrix.jl:121-186, tiledalgos.jl:75-225), the
# trailing kswitch tiles run in core; n <= kswitch * tilesize runs in core.
# mydims2d / tilefactor / backend are accepted for the signature.
function OOC_alg!(A::Matrix{T}, sym::Symbol; kswitch=10, mydims2d=false, tilesize=32, tilefactor=1,
                  backend=nothing, opts::AvsOpts=default_opts()) where {T<:Union{Float32,Float64}}
    size(A, 1) == size(A, 2) || error("Not supported")
    check(ccall((sym, libavsvd), Int32, (Int32, Int64, Int32, Ptr{T}, Int64, Int32, Ref{AvsOpts}),
                dtype(T), size(A, 1), tilesize, A, stride(A, 2), kswitch, opts))
    return nothing
end
OOC_QR!(A::Matrix; kw...) = OOC_alg!(A, :avs_ooc_tiled_qr; kw...)
OOC_Bidiag!(A::Matrix; kw...) = OOC_alg!(A, :avs_ooc_blockbidiag; kw...)

# OOC_SVD!  (tiledalgos.jl:256-260): returns the bidiagonal matrix like the reference
function OOC_SVD!(A::Matrix; kswitch=10, mydims2d=false, tilesize=32, tilefactor=1, backend=nothing)
    OOC_Bidiag!(A; kswitch=kswitch, mydims2d=mydims2d, tilesize=tilesize, tilefactor=tilefactor)
    return bidiagonalize(A, tilesize)
end

function dense_to_band(A::AbstractMatrix, bw::Int)
    n = size(A, 1)
    band = zeros(Float64, bw + 1, n)
    for j in 1:n, i in max(1, j - bw):j
        band[bw + 1 + i - j, j] = A[i, j]
    end
    return band
end

# bidiagonalize(A, bandwidth)  (bulgechasing.jl:33-42), host FP64 in the library
function bidiagonalize(A::AbstractMatrix, bandwidth::Int)
    n = size(A, 1)
    band = dense_to_band(A, bandwidth)
    d = zeros(Float64, n)
    e = zeros(Float64, max(n - 1, 1))
    check(ccall((:avs_bidiagonalize_band, libavsvd), Int32,
                (Int64, Int32, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}), n, bandwidth, band, bandwidth + 1, d, e))
    A .= 0
    for i in 1:n
        A[i, i] = d[i]
        i < n && (A[i, i + 1] = e[i])
    end
    return A
end

# diag_lapack(A)  (bulgechasing.jl:50-53): dbdsdc('U','N')
function diag_lapack(A::AbstractMatrix)
    n = size(A, 1)
    d = Float64.(diag_of(A, 0))
    e = n > 1 ? Float64.(diag_of(A, 1)) : zeros(1)
    s = zeros(Float64, n)
    check(ccall((:avs_bdsdc_values, libavsvd), Int32, (Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), n, d, e, s))
    return s
end
diag_of(A, k) = [A[i, i + k] for i in 1:(size(A, 1) - k)]

# my_CU_svdval(A::Matrix, block_size)  (bulgechasing.jl:45-48)
my_CU_svdval(A::Matrix, block_size::Int) = diag_lapack(bidiagonalize(A, block_size))

# whole path: band reduction on the GPU, band-only D2H, host stage
function svdvals_gpu(A::Matrix{T}; tilesize::Int=TILE[]) where {T<:Union{Float32,Float64}}
    M = TiledMatrix(copy(A), tilesize, tilesize)
    s = zeros(Float64, size(A, 1))
    check(ccall((:avs_svdvals, libavsvd), Int32, (Ptr{Cvoid}, Ptr{Float64}), M.handle, s))
    return s
end

end # module
