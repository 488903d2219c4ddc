"""avsvd -- MI355X-native tiled SVD band reduction (host side).

Python mirror of the reference's Julia entry points for the stage-1 path
(paths relative to the reference repository root); Julia callers bind the
same C ABI with ccall (avocado-sandwich_amd/julia/AvSVD.jl):

  TiledMatrix(A, bm, bn; tiledim2d)      KAbasedSVD/src/TiledMatrix.jl:41-67
  tiled_QR!(A; kend)                     KAbasedSVD/src/tiledalgos.jl:6-15    -> tiled_QR_
  myormqr!(C, A, Tau)                    CAQR/src/qr_kernels.jl:256-265       -> apply_q_
  Blockbidiag!(A; kend)                  KAbasedSVD/src/tiledalgos.jl:17-34   -> Blockbidiag_
  OOC_QR! / OOC_Bidiag! / OOC_SVD!       KAbasedSVD/src/tiledalgos.jl:253-260 -> OOC_QR_ / OOC_Bidiag_ / OOC_SVD_
  compile_kernels(; tilesize, ...)       KAbasedSVD/src/tiledalgos.jl:262-267
  bidiagonalize / my_CU_svdval / diag_lapack
                                         KAbasedSVD/src/bulgechasing.jl:33-53
  tile kernels (QR_unsafe_kernel!, ...)  KAbasedSVD/src/QRkernels.jl:3-197  -> geqrt_, tsqrt_, unmqr_, tsmqr_

Julia's `!` suffix becomes a trailing underscore; every such function
mutates its argument in place like the reference.
"""
from ._lib import (AVS_F32, AVS_F64, SCHED_HOST, SCHED_PERSISTENT, SCHED_STREAMS, TREE_FLAT, TREE_HIER, UPD_NATIVE,
                   UPD_SPLIT_BF16, AvsError, EXPORTS, LIB_PATH, lib)
from .batch import Batch, batch_svdvals
from .server import Server
from .tiled import (TiledMatrix, Blockbidiag_, tiled_QR_, apply_q_, OOC_QR_, OOC_Bidiag_, OOC_SVD_, compile_kernels,
                    bidiagonalize, bidiagonalize_band, diag_lapack, bdsdc_values, my_CU_svdval, svdvals,
                    gaussian_matrix, geqrt_, tsqrt_, unmqr_, tsmqr_, schedule_stats, schedule_dump, phase_dump, mfma_layout)

__all__ = [
    "Batch", "batch_svdvals", "Server",
    "TiledMatrix", "Blockbidiag_", "tiled_QR_", "apply_q_", "OOC_QR_", "OOC_Bidiag_", "OOC_SVD_", "compile_kernels",
    "bidiagonalize", "bidiagonalize_band", "diag_lapack", "bdsdc_values", "my_CU_svdval", "svdvals",
    "gaussian_matrix", "geqrt_", "tsqrt_", "unmqr_", "tsmqr_", "schedule_stats", "schedule_dump", "phase_dump", "mfma_layout",
    "AvsError", "EXPORTS", "LIB_PATH", "lib", "AVS_F32", "AVS_F64", "SCHED_HOST", "SCHED_PERSISTENT", "SCHED_STREAMS",
    "TREE_FLAT", "TREE_HIER", "UPD_NATIVE", "UPD_SPLIT_BF16",
]
