"""The oracle pinned against LAPACK fixtures, invariants and brute force (CPU).

Mirrors the reference's own kernel tests (KAbasedSVD/test/test_QR_kernels.jl)
and algorithm tests (KAbasedSVD/test/test_tiledalgo.jl), with committed
LAPACK fixtures instead of live CUDA runs.  Each check is chosen so that a
plausible mistake in the oracle (dropped term, wrong sign or index,
transposed operand) fails it.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

BASE = 20261203


def fix(golden_dir, b):
    return np.load(os.path.join(golden_dir, "kernels_b%d.npz" % b))


@pytest.mark.parametrize("b", [32, 64])
def test_geqrt_matches_dgeqrf(golden_dir, b):
    # test_QR_kernels.jl:16  Array(acopy) ≈ qr(Array(a)).factors
    F = fix(golden_dir, b)
    f, tau = O.geqrt(F["a"])
    np.testing.assert_allclose(f, F["geqrt_factors"], rtol=0, atol=50 * b * np.finfo(float).eps)
    np.testing.assert_allclose(tau[: b - 1], F["geqrt_tau"][: b - 1], rtol=0, atol=1e-13)


@pytest.mark.parametrize("b", [32, 64])
def test_geqrt_lq_twin_is_adjoint(golden_dir, b):
    # LQ1! = QR1! on the adjoint view (TiledMatrix.jl:85,98)
    F = fix(golden_dir, b)
    f, tau = O.geqrt(F["a"].T.copy(), trans=True)
    np.testing.assert_allclose(f.T, F["geqrt_factors"], rtol=0, atol=50 * b * np.finfo(float).eps)


@pytest.mark.parametrize("b", [32, 64])
def test_unmqr_four_ways(golden_dir, b):
    # test_QR_kernels.jl:20-33: Q'B, QB, BQ, BQ'
    F = fix(golden_dir, b)
    f, tau = O.geqrt(F["a"])
    tol = 1e-12
    np.testing.assert_allclose(O.unmqr(f, tau, F["c"], applyt=True), F["unmqr_QtC"], atol=tol)
    np.testing.assert_allclose(O.unmqr(f, tau, F["c"], applyt=False), F["unmqr_QC"], atol=tol)
    # applyQorQt(bcopy', ...) on the adjoint: true -> B*Q, false -> B*Q'
    np.testing.assert_allclose(O.unmqr(f, tau, F["cr"], applyt=True, ctrans=True), F["unmqr_CQ"], atol=tol)
    np.testing.assert_allclose(O.unmqr(f, tau, F["cr"], applyt=False, ctrans=True), F["unmqr_CQt"], atol=tol)


@pytest.mark.parametrize("b", [32, 64])
def test_tsqrt_matches_stacked_dgeqrf(golden_dir, b):
    # test_QR_kernels.jl:42  [triu(R); V2] ≈ qr([triu(a1); a2]).factors
    F = fix(golden_dir, b)
    r1, v2, tau = O.tsqrt(F["a1"], F["a2"])
    np.testing.assert_allclose(np.triu(r1), np.triu(F["tsqrt_factors"][:b]), atol=1e-12)
    np.testing.assert_allclose(v2, F["tsqrt_factors"][b:], atol=1e-12)
    np.testing.assert_allclose(tau, F["tsqrt_tau"], atol=1e-13)
    # the strict lower triangle of the top tile is preserved (test_QR_kernels.jl:41)
    np.testing.assert_array_equal(np.tril(r1, -1), np.tril(F["a1"], -1))


@pytest.mark.parametrize("b", [32, 64])
def test_tsmqr_four_ways(golden_dir, b):
    # test_QR_kernels.jl:44-62
    F = fix(golden_dir, b)
    _, v2, tau = O.tsqrt(F["a1"], F["a2"])
    c12, r12 = F["c12"], F["r12"]
    tol = 1e-12
    c1, c2 = O.tsmqr(c12[:b], c12[b:], v2, tau, applyt=True)
    np.testing.assert_allclose(np.vstack([c1, c2]), F["tsmqr_QtC"], atol=tol)
    c1, c2 = O.tsmqr(c12[:b], c12[b:], v2, tau, applyt=False)
    np.testing.assert_allclose(np.vstack([c1, c2]), F["tsmqr_QC"], atol=tol)
    c1, c2 = O.tsmqr(r12[:, :b], r12[:, b:], v2, tau, applyt=True, ctrans=True)
    np.testing.assert_allclose(np.hstack([c1, c2]), F["tsmqr_CQ"], atol=tol)
    c1, c2 = O.tsmqr(r12[:, :b], r12[:, b:], v2, tau, applyt=False, ctrans=True)
    np.testing.assert_allclose(np.hstack([c1, c2]), F["tsmqr_CQt"], atol=tol)


def test_zero_pivot_gives_identity_reflector():
    # all-zero column: tau = 0 (LAPACK convention) instead of the reference's NaN
    a = np.eye(8)
    a[:, 0] = 0.0
    f, tau = O.geqrt(a)
    assert tau[0] == 0.0 and np.all(np.isfinite(f))
    a = np.random.default_rng(0).standard_normal((8, 8))
    a[0, 0] = 0.0  # alpha = 0, x != 0: sign(0) = +1
    f, tau = O.geqrt(a)
    ref = sl.qr(a, mode="r")[0]
    np.testing.assert_allclose(np.abs(np.diag(f)), np.abs(np.diag(ref)), atol=1e-13)


@pytest.mark.parametrize("b,nt", [(4, 4), (4, 6), (16, 4), (16, 6)])
def test_blockbidiag_sigma_and_structure(golden_dir, b, nt):
    # test_tiledalgo.jl:32  svdvals(band) ≈ svdvals(A), plus exact zeros off the band
    n = b * nt
    a = O.gaussian_matrix(n, BASE + n)
    ref = np.load(os.path.join(golden_dir, "sigma_small.npz"))["gauss_n%d" % n]
    band = O.blockbidiag(a, b)
    assert O.sigma_err(sl.svdvals(band), ref) < 10 * n * np.finfo(float).eps
    np.testing.assert_array_equal(band, np.triu(np.tril(band, b)))
    np.testing.assert_allclose(np.linalg.norm(band), np.linalg.norm(a), rtol=1e-13)


def test_blockbidiag_tiny_brute_force():
    # n=8, b=4: the band equals U^T A W for orthogonal U, W; check via A^T A spectra
    rng = np.random.default_rng(3)
    a = rng.standard_normal((8, 8))
    band = O.blockbidiag(a, 4)
    np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(band.T @ band)),
                               np.sort(np.linalg.eigvalsh(a.T @ a)), atol=1e-12)
    # one QR step only (kend=1): first tile column is R over zeros
    part = O.blockbidiag(a, 4, kend=1)
    np.testing.assert_array_equal(part[4:, :4], 0.0)


@pytest.mark.parametrize("b,nt", [(4, 4), (16, 6)])
def test_tiled_qr_R_matches_lapack(b, nt):
    # test_tiledalgo.jl:26  triu(A) .* sign(diag) ≈ R .* sign(diag R)
    n = b * nt
    a = O.gaussian_matrix(n, BASE + 7 * n)
    out = O.tiled_qr(a, b)
    r = sl.qr(a, mode="r")[0]
    np.testing.assert_allclose(np.triu(out) * np.sign(np.diag(out))[:, None], r * np.sign(np.diag(r))[:, None],
                               atol=1e-11)


@pytest.mark.parametrize("n,bw", [(64, 8), (128, 32), (192, 64), (256, 64)])
def test_bidiagonalize_preserves_sigma(n, bw):
    # bulgechasing.jl:33-53 (no reference test pins it: checked against dgesdd of the band)
    a = O.gaussian_matrix(n, BASE + 11 * n + bw)
    band = np.triu(np.tril(a, bw))
    sig = O.svdvals_host_stage(band, bw)
    assert O.sigma_err(sig, sl.svdvals(band)) < 10 * n * np.finfo(float).eps
    bi = O.bidiagonalize(band, bw)
    off = bi - np.diag(np.diag(bi)) - np.diag(np.diag(bi, 1), 1)
    # only the bidiagonal is read by bdsdc; the rest must be rounding-level
    assert np.abs(off).max() < 1e-12 * np.abs(a).max()


def test_n1024_pipeline_against_gesdd(golden_dir):
    # the CPU plain-path workload: n=1024, b=64 Gaussian through band + host stage
    n, b = 1024, 64
    a = O.gaussian_matrix(n, BASE + 1)
    ref = np.load(os.path.join(golden_dir, "sigma_small.npz"))["gauss_n1024_seed1"]
    band = O.blockbidiag(a, b)
    sig = O.svdvals_host_stage(band, b)
    assert O.sigma_err(sig, ref) < 10 * n * np.finfo(float).eps


def test_philox_known_answer():
    # Philox4x32-10 known-answer vector (Random123 kat_vectors: ctr=0, key=0)
    # -> 0x6627e8d5 0xe169c58d 0xbc57ac4c 0x9b00dbd8; our first normal pair is a
    # deterministic function of that block, so recompute it here from the KAT.
    w = [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    u1 = (((w[0] << 32) | w[1]) >> 11) + 0.5
    u2 = (((w[2] << 32) | w[3]) >> 11) + 0.5
    u1 /= 2.0 ** 53
    u2 /= 2.0 ** 53
    r = np.sqrt(-2 * np.log(u1))
    exp = [r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)]
    np.testing.assert_allclose(O.philox_normal(0, 2), exp, rtol=1e-15)


# --------------------------------------------------------------------------
# the reference's CPU plain path (CUDAbasedSVD/SVD_GPU.jl:67-214)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n,b", [(128, 32), (256, 32), (512, 64)])
def test_cpu_plain_path_matches_tile_oracle(n, b):
    """BandBidiagonal! with LAPACK QR of each stacked 2b x b block gives the
    same band as the tile kernels' flat-tree Blockbidiag! (QR of [R; A] by
    dgeqrt and by TSQRT are the same Householder factorization), a band with
    exact zeros outside it, and sigma(A)."""
    a = O.gaussian_matrix(n, 20261203 + 5 * n + b)
    r = O.band_bidiagonal_cpu(a, b, 1)
    ref = O.blockbidiag(a, b)
    assert np.abs(r - ref).max() <= 1e-12 * np.sqrt(n) * np.abs(ref).max()
    assert np.array_equal(np.tril(r, -1), np.zeros_like(r))
    assert np.array_equal(np.triu(r, b + 1), np.zeros_like(r))
    assert O.sigma_err(O.svdvals_ref(r), O.svdvals_ref(a)) <= 10 * n * np.finfo(np.float64).eps


def test_cpu_plain_path_thread_count_independent():
    a = O.gaussian_matrix(384, 20261203 + 9)
    np.testing.assert_array_equal(O.band_bidiagonal_cpu(a, 32, 1), O.band_bidiagonal_cpu(a, 32, 5))
