"""Tile operators on the GPU through the C ABI vs LAPACK fixtures and the oracle.

Mirrors KAbasedSVD/test/test_QR_kernels.jl (single tile QR, double tile QR,
block QR over 5 tiles) for Float64 and Float32 and b in {32, 64, 128}.
"""
import os

import numpy as np
import pytest

import avsvd
import oracle as O

pytestmark = pytest.mark.gpu

DT = {np.float64: avsvd.AVS_F64, np.float32: avsvd.AVS_F32}


def tol(dt, b):
    return (50 * b * np.finfo(np.float64).eps) if dt == np.float64 else 2e-4


def dev(a, dt, gpu):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dt).T)).to(gpu)


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().T.astype(np.float64)


def ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def test_mfma_lane_layout(gpu):
    """Pins the 16x16x4 MFMA lane layout the update kernels assume:
    A[i][k] from lane i + 16k, B[k][j] from lane j + 16k; accumulator element m
    of lane l = D[4*(l//16) + m][l % 16] for f32 and D[l//16 + 4*m][l % 16] for
    f64.  Variants 1 and 2 (all-ones operands) isolate the B and A lane maps."""
    import json
    probes = {}
    for v in (0, 1, 2):
        o64, o32 = avsvd.mfma_layout(v)
        probes[v] = (o64, o32)
        print("MFMA_PROBE", v, json.dumps({"f64": o64.tolist(), "f32": o32.astype(float).tolist()}), flush=True)
    a = np.arange(1, 65, dtype=np.float64)
    bb = np.array([((i * 7) % 13) - 6 for i in range(64)], dtype=np.float64)
    A = np.array([[a[i + 16 * k] for k in range(4)] for i in range(16)])
    B = np.array([[bb[j + 16 * k] for j in range(16)] for k in range(4)])
    D = A @ B
    e32 = np.array([[D[4 * (l // 16) + m, l % 16] for m in range(4)] for l in range(64)])
    e64 = np.array([[D[l // 16 + 4 * m, l % 16] for m in range(4)] for l in range(64)])
    np.testing.assert_array_equal(probes[0][1].astype(np.float64), e32)
    np.testing.assert_array_equal(probes[0][0], e64)
    # variant 1 (A = 1, B = lane): every output in column j is sum_k (j + 16k)
    for v, (o64, o32) in ((1, probes[1]),):
        col = np.array([[4 * (l % 16) + 96 for _ in range(4)] for l in range(64)], dtype=np.float64)
        np.testing.assert_array_equal(o64, col)
        np.testing.assert_array_equal(o32.astype(np.float64), col)
    # variant 2 (A = lane, B = 1): output row i is sum_k (i + 16k) = 4i + 96
    o64, o32 = probes[2]
    np.testing.assert_array_equal(o64, np.array([[4 * (l // 16 + 4 * m) + 96 for m in range(4)] for l in range(64)]))
    np.testing.assert_array_equal(o32.astype(np.float64),
                                  np.array([[4 * (4 * (l // 16) + m) + 96 for m in range(4)] for l in range(64)]))


def test_lane_exchanges(gpu):
    """The panel reductions' VALU lane exchanges (DPP row_half_mirror +
    quad_perm, row_ror:8, v_permlane16_swap, v_permlane32_swap) move lane
    l ^ m to lane l, and the column sums are exact in every lane."""
    o64, o32 = avsvd.mfma_layout(3)
    lanes = np.arange(64)
    for k, m in enumerate((4, 8, 16, 32)):
        np.testing.assert_array_equal(o64[:, k], (lanes ^ m).astype(np.float64), err_msg="xor %d" % m)
        want = np.array([lanes[lanes % (4 << k) == l % (4 << k)].sum() for l in lanes], dtype=np.float64)
        np.testing.assert_array_equal(o32[:, k].astype(np.float64), want, err_msg="sum mod %d" % (4 << k))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("b", [32, 64])
def test_geqrt_vs_lapack(gpu, golden_dir, dt, b):
    """panel kernel alone (no matrix cores)"""
    import torch
    F = np.load(os.path.join(golden_dir, "kernels_b%d.npz" % b))
    A = dev(F["a"], dt, gpu)
    tau = torch.zeros(b, dtype=torch.float64, device=gpu)
    avsvd.geqrt_(DT[dt], b, ptr(A), b, ptr(tau))
    np.testing.assert_allclose(host(A), F["geqrt_factors"], atol=tol(dt, b) * 10)
    np.testing.assert_allclose(tau.cpu().numpy()[: b - 1], F["geqrt_tau"][: b - 1], atol=tol(dt, b))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("b", [32, 64])
def test_geqrt_and_unmqr_vs_lapack(gpu, golden_dir, dt, b):
    import torch
    F = np.load(os.path.join(golden_dir, "kernels_b%d.npz" % b))
    A = dev(F["a"], dt, gpu)
    tau = torch.zeros(b, dtype=torch.float64, device=gpu)
    avsvd.geqrt_(DT[dt], b, ptr(A), b, ptr(tau))
    f = host(A)
    np.testing.assert_allclose(f, F["geqrt_factors"], atol=tol(dt, b) * 10)
    np.testing.assert_allclose(tau.cpu().numpy()[: b - 1], F["geqrt_tau"][: b - 1], atol=tol(dt, b))
    ncol = F["c"].shape[1]
    for applyt, key in ((True, "unmqr_QtC"), (False, "unmqr_QC")):
        C = dev(F["c"], dt, gpu)
        avsvd.unmqr_(DT[dt], b, ptr(A), b, ptr(tau), ptr(C), b, ncol, applyt=applyt)
        np.testing.assert_allclose(host(C), F[key], atol=tol(dt, b) * 20, err_msg=key)
    # adjoint-view application (test_QR_kernels.jl:28-33): C' <- Q^T C'  ==  C <- C Q
    for applyt, key in ((True, "unmqr_CQ"), (False, "unmqr_CQt")):
        C = dev(F["cr"], dt, gpu)
        avsvd.unmqr_(DT[dt], b, ptr(A), b, ptr(tau), ptr(C), b, b, applyt=applyt, ctrans=True)
        np.testing.assert_allclose(host(C), F[key], atol=tol(dt, b) * 20, err_msg=key)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("b", [32, 64])
def test_tsqrt_and_tsmqr_vs_lapack(gpu, golden_dir, dt, b):
    import torch
    F = np.load(os.path.join(golden_dir, "kernels_b%d.npz" % b))
    A1, A2 = dev(F["a1"], dt, gpu), dev(F["a2"], dt, gpu)
    tau = torch.zeros(b, dtype=torch.float64, device=gpu)
    avsvd.tsqrt_(DT[dt], b, ptr(A1), b, ptr(A2), b, ptr(tau))
    r1, v2 = host(A1), host(A2)
    np.testing.assert_allclose(np.triu(r1), np.triu(F["tsqrt_factors"][:b]), atol=tol(dt, b) * 10)
    np.testing.assert_allclose(v2, F["tsqrt_factors"][b:], atol=tol(dt, b) * 10)
    np.testing.assert_allclose(tau.cpu().numpy(), F["tsqrt_tau"], atol=tol(dt, b))
    np.testing.assert_array_equal(np.tril(r1, -1), np.tril(np.asarray(F["a1"], dt).astype(np.float64), -1))
    c12, r12 = F["c12"], F["r12"]
    ncol = c12.shape[1]
    for applyt, key in ((True, "tsmqr_QtC"), (False, "tsmqr_QC")):
        C1, C2 = dev(c12[:b], dt, gpu), dev(c12[b:], dt, gpu)
        avsvd.tsmqr_(DT[dt], b, ptr(C1), b, ptr(C2), b, ptr(A2), b, ptr(tau), ncol, applyt=applyt)
        np.testing.assert_allclose(np.vstack([host(C1), host(C2)]), F[key], atol=tol(dt, b) * 20, err_msg=key)
    for applyt, key in ((True, "tsmqr_CQ"), (False, "tsmqr_CQt")):
        C1, C2 = dev(r12[:, :b], dt, gpu), dev(r12[:, b:], dt, gpu)
        avsvd.tsmqr_(DT[dt], b, ptr(C1), b, ptr(C2), b, ptr(A2), b, ptr(tau), b, applyt=applyt, ctrans=True)
        np.testing.assert_allclose(np.hstack([host(C1), host(C2)]), F[key], atol=tol(dt, b) * 20, err_msg=key)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("b", [32, 64, 128])
@pytest.mark.parametrize("trans", [False, True])
def test_tile_ops_vs_oracle_panel(gpu, dt, b, trans):
    """block QR over 5 tiles (test_QR_kernels.jl:66-87), both orientations."""
    import torch
    rng = np.random.default_rng(b + 2 * trans)
    a = rng.standard_normal((b, b)).astype(dt).astype(np.float64)
    a1 = rng.standard_normal((b, b)).astype(dt).astype(np.float64)
    a2 = rng.standard_normal((b, b)).astype(dt).astype(np.float64)
    w = 5 * b
    c = rng.standard_normal((w, b) if trans else (b, w)).astype(dt).astype(np.float64)
    c1 = rng.standard_normal((w, b) if trans else (b, w)).astype(dt).astype(np.float64)
    c2 = rng.standard_normal((w, b) if trans else (b, w)).astype(dt).astype(np.float64)
    # oracle: reference arithmetic on the same inputs
    of, otau = O.geqrt(a, trans=trans)
    oc = O.unmqr(of, otau, c, vtrans=trans, ctrans=trans)
    or1, ov2, otau2 = O.tsqrt(a1, a2, trans=trans)
    oc1, oc2 = O.tsmqr(c1, c2, ov2, otau2, ctrans=trans, vtrans=trans)
    # device
    A = dev(a, dt, gpu)
    tau = torch.zeros(b, dtype=torch.float64, device=gpu)
    avsvd.geqrt_(DT[dt], b, ptr(A), b, ptr(tau), trans=trans)
    C = dev(c, dt, gpu)
    ldc = c.shape[0]
    avsvd.unmqr_(DT[dt], b, ptr(A), b, ptr(tau), ptr(C), ldc, w, vtrans=trans, ctrans=trans)
    A1, A2 = dev(a1, dt, gpu), dev(a2, dt, gpu)
    tau2 = torch.zeros(b, dtype=torch.float64, device=gpu)
    avsvd.tsqrt_(DT[dt], b, ptr(A1), b, ptr(A2), b, ptr(tau2), trans=trans)
    C1, C2 = dev(c1, dt, gpu), dev(c2, dt, gpu)
    avsvd.tsmqr_(DT[dt], b, ptr(C1), ldc, ptr(C2), ldc, ptr(A2), b, ptr(tau2), w, vtrans=trans, ctrans=trans)
    s = 1e-12 if dt == np.float64 else 3e-4
    np.testing.assert_allclose(host(A), of, atol=s * b)
    np.testing.assert_allclose(host(C), oc, atol=s * b)
    np.testing.assert_allclose(host(A1), or1, atol=s * b)
    np.testing.assert_allclose(host(A2), ov2, atol=s * b)
    np.testing.assert_allclose(host(C1), oc1, atol=s * b)
    np.testing.assert_allclose(host(C2), oc2, atol=s * b)
