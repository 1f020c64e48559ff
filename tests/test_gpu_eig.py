"""GPU checks of the Gram eigensolver through the tridiagonal form
(``eig.GramTridiag``: batched tridiagonalisation, bisection, twisted inverse
iteration, MFMA back-transformation, FP64 orthogonality re-check) against
LAPACK (numpy.linalg.eigh in float64 on the host), and of the batched GEMM
it runs on.

Tolerances (relative to the largest eigenvalue / ||G||): eigenvalues 1e-12;
eigenvector residual ||G u - lambda u|| <= 1e-11 ||G||; orthogonality of
the computed vectors within the re-check tolerance; the kept subspace
against LAPACK's, ||U U^H - U_ref U_ref^H|| <= 1e-9 where the cut has a
relative gap above 1e-4.  The measured values are printed.
"""
import numpy as np
import pytest
import torch

from atlas_q import kernels as K
from atlas_q._ffi import check, lib
from atlas_q.eig import ORTH_TOL, GramTridiag

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _gram(rng, n, cols, decay=None):
    X = rng.standard_normal((n, cols)) + 1j * rng.standard_normal((n, cols))
    if decay is not None:
        X = X * np.exp(-np.arange(cols) / decay)[None, :]
    return X @ X.conj().T


def _kron_gram(rng, n):
    a = rng.standard_normal((8, 3)) + 1j * rng.standard_normal((8, 3))
    q = n // 8
    b = rng.standard_normal((q, max(1, q // 2))) + 1j * rng.standard_normal((q, max(1, q // 2)))
    X = np.kron(a, b)
    G = np.zeros((n, n), complex)
    G[:X.shape[0], :X.shape[0]] = X @ X.conj().T
    return G


def _solve(Gs, ks):
    Gb = torch.from_numpy(np.stack(Gs)).to(DEV).contiguous()
    tri = GramTridiag(Gb)
    W = tri.eigenvalues().cpu().numpy()
    Et, err = tri.vectors(ks)
    torch.cuda.synchronize()
    return W, Et.cpu().numpy(), err.cpu().numpy()


def _check(Gs, ks, label, expect_ok=True):
    W, Et, err = _solve(Gs, ks)
    for b, (G, k) in enumerate(zip(Gs, ks)):
        ref_w, ref_v = np.linalg.eigh(G)
        ref_w, ref_v = ref_w[::-1], ref_v[:, ::-1]
        scale = max(abs(ref_w).max(), 1e-300)
        ew = np.abs(W[b] - ref_w).max() / scale
        U = Et[b, :k].conj().T                        # columns: eigenvectors
        res = np.linalg.norm(G @ U - U * ref_w[None, :k], axis=0).max() / np.linalg.norm(G, 2)
        orth = np.abs(U.conj().T @ U - np.eye(k)).max()
        gap = (ref_w[k - 1] - ref_w[k]) / scale if k < len(ref_w) else 1.0
        P = U @ U.conj().T
        Pr = ref_v[:, :k] @ ref_v[:, :k].conj().T
        span = np.linalg.norm(P - Pr, 2)
        print(f"  {label}[{b}] n={G.shape[0]} k={k}: eig {ew:.1e}, residual {res:.1e}, "
              f"orth {orth:.1e} (re-check {err[b]:.1e}), cut gap {gap:.1e}, span {span:.1e}",
              flush=True)
        assert ew <= 1e-12, (label, b, ew)
        if expect_ok:
            assert err[b] <= ORTH_TOL and orth <= ORTH_TOL, (label, b, err[b], orth)
            assert res <= 1e-11, (label, b, res)
            if gap > 1e-4:
                assert span <= 1e-9, (label, b, span)


@pytest.mark.parametrize("n", [65, 128, 512])
def test_tridiag_random_grams(n):
    rng = np.random.default_rng(n)
    Gs = [_gram(rng, n, n), _gram(rng, n, n // 2), _gram(rng, n, n, decay=n / 8)]
    _check(Gs, [n // 2, n // 2, n // 3], f"random n={n}")


def test_tridiag_chi1024_gram_shape():
    """Order 2048 (the chi = 1024 update's Gram core), kept rank 1024."""
    rng = np.random.default_rng(2048)
    Gs = [_gram(rng, 2048, 2048), _gram(rng, 2048, 2048)]
    _check(Gs, [1024, 1024], "n=2048")


def test_tridiag_degenerate_kronecker_grams():
    """Rank-deficient Gram cores with exactly repeated eigenvalues (the
    batched form's CZ-ladder states): eigenvalues exact, and the kept
    vectors of every repeated eigenvalue orthonormal (the cluster pass
    re-orthogonalises copies that share a block) with residuals at FP64
    level, so the re-check passes."""
    rng = np.random.default_rng(7)
    for n, k in ((128, 12), (256, 24)):
        Gs = [_kron_gram(rng, n), _kron_gram(rng, n)]
        _kron_case(Gs, [k, k], require=True)


@pytest.mark.parametrize("n", [96, 300])
def test_tridiag_repeated_eigenvalues_full_rank(n):
    """Full-rank Hermitian matrices whose spectra repeat each value 1-9
    times (exact multiplicities, random eigenbasis): every kept vector
    passes the re-check, residuals at FP64 level, also where the cut falls
    inside a repeated group."""
    rng = np.random.default_rng(n)
    vals = []
    while len(vals) < n:
        vals += [float(np.exp(-0.05 * len(vals)))] * int(rng.integers(1, 10))
    lam = np.array(vals[:n])
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    G = (Q * lam[None, :]) @ Q.conj().T
    G = (G + G.conj().T) / 2
    _kron_case([G, G.copy()], [n // 3, n // 2], require=True)


@pytest.mark.parametrize("n,k", [(1024, 1024), (2048, 1024)])
def test_tridiag_large_tail_cluster(n, k):
    """Spectra over twenty decades (a Haar circuit's growth phase, where the
    kept rank reaches the Gram order): several hundred eigenvalues chain
    within the cluster tolerance of ||T|| and form one cluster, which the
    blocked Gram-Schmidt orthonormalises.  Eigenvalues, residuals, the
    re-check and the kept subspace as in the other cases; the call's time
    is printed."""
    import time
    rng = np.random.default_rng(n + k)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.exp(-np.arange(n) * (48.0 / n))
    G = (Q * lam[None, :]) @ Q.conj().T
    G = (G + G.conj().T) / 2
    members = int((lam[:k] < 1e-5).sum())
    Gb = torch.from_numpy(np.stack([G])).to(DEV).contiguous()
    tri = GramTridiag(Gb.clone())
    tri.eigenvalues()
    tri.vectors([k])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tri = GramTridiag(Gb.clone())
    tri.eigenvalues()
    tri.vectors([k])
    torch.cuda.synchronize()
    print(f"  n={n} k={k}: {members} kept eigenvalues below 1e-5 ||G||, "
          f"solve {1e3 * (time.perf_counter() - t0):.1f} ms", flush=True)
    _check([G], [k], f"tail cluster n={n}")


def _kron_case(Gs, ks, require=False):
    W, Et, err = _solve(Gs, ks)
    for b, (G, k) in enumerate(zip(Gs, ks)):
        ref_w = np.linalg.eigvalsh(G)[::-1]
        assert np.abs(W[b] - ref_w).max() <= 1e-12 * ref_w.max()
        U = Et[b, :k].conj().T
        orth = np.abs(U.conj().T @ U - np.eye(k)).max()
        res = np.linalg.norm(G @ U - U * ref_w[None, :k], axis=0).max() / np.linalg.norm(G, 2)
        print(f"  kron[{b}]: orth {orth:.1e} (re-check {err[b]:.1e}), residual {res:.1e}",
              flush=True)
        if require:
            assert err[b] <= ORTH_TOL, (b, err[b])
        if err[b] <= ORTH_TOL:
            assert orth <= 10 * ORTH_TOL and res <= 1e-11


def test_bgemm_matches_float64_products():
    rng = np.random.default_rng(3)
    Bn, M, N, Kd = 3, 70, 150, 90
    A = rng.standard_normal((Bn, M, Kd)) + 1j * rng.standard_normal((Bn, M, Kd))
    Bm = rng.standard_normal((Bn, N, Kd)) + 1j * rng.standard_normal((Bn, N, Kd))   # used as B^T
    C0 = rng.standard_normal((Bn, M, N)) + 1j * rng.standard_normal((Bn, M, N))
    tA, tB = torch.from_numpy(A).to(DEV), torch.from_numpy(Bm).to(DEV)
    tC = torch.from_numpy(C0).to(DEV).contiguous()
    # C = -0.5 A conj(Bm)^T + 2 C0 (B given transposed and conjugated)
    check(lib().aq_bgemm(1, M, N, Kd, -0.5, K._ptr(tA), Kd, 1, M * Kd, 0, K._ptr(tB), 1, Kd,
                         N * Kd, 1, 2.0, K._ptr(tC), N, 1, M * N, Bn, K._stream(tC)), "aq_bgemm")
    want = -0.5 * A @ np.conj(np.transpose(Bm, (0, 2, 1))) + 2.0 * C0
    got = tC.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("n", [2, 33, 64, 65, 130, 144, 203, 256, 257, 512])
def test_own_tridiagonal_reduction_matches_lapack(n):
    """The library's blocked reduction against LAPACK's zhetrd (scipy, host,
    float64) on the column-major view conj(G) with uplo = "U": the same
    algorithm and conventions, so on full-rank Gram matrices D, E, tau and the
    Householder vectors agree to rounding (1e-12 relative to max |G|).  (On a
    rank-deficient matrix the entries of T are ill-conditioned functions of
    G and two valid reductions part ways; there the next test compares what
    is well defined.)"""
    import scipy.linalg.lapack as LA
    rng = np.random.default_rng(100 + n)
    Gs = np.stack([_gram(rng, n, n), _gram(rng, n, 2 * n)])
    Gb = torch.from_numpy(Gs).to(DEV).contiguous()
    tri = GramTridiag(Gb, reduction="own")
    torch.cuda.synchronize()
    D, E, tau, A = (tri.D.cpu().numpy(), tri.E.cpu().numpy(), tri.tau.cpu().numpy(),
                    Gb.cpu().numpy())
    for b in range(2):
        c, d, e, t, info = LA.zhetrd(np.asfortranarray(np.conj(Gs[b])), lower=0)
        assert info == 0
        scale = np.abs(Gs[b]).max()
        dD = np.abs(D[b] - d).max() / scale
        dE = np.abs(E[b][:n - 1] - e).max() / scale
        dT = np.abs(tau[b][:n - 1] - t).max() if n > 1 else 0.0
        dV = max([np.abs(A[b, i + 1, :i] - c[:i, i + 1]).max() for i in range(1, n - 1)],
                 default=0.0)
        print(f"  n={n}[{b}]: |dD| {dD:.1e}, |dE| {dE:.1e}, |dtau| {dT:.1e}, |dV| {dV:.1e}",
              flush=True)
        assert dD < 1e-12 and dE < 1e-12 and dT < 1e-10 and dV < 1e-10


@pytest.mark.parametrize("n", [130, 512])
def test_own_tridiagonal_reduction_rank_deficient(n):
    """Rank-deficient Gram matrices (rank n/3, and a Kronecker-structured one
    with repeated eigenvalues): the eigenvalues of the computed T equal
    LAPACK's eigenvalues of G (1e-13 relative), and Q T Q^H reproduces conj(G)
    through the library's own back-transformation (checked in the eigenpair
    tests above)."""
    rng = np.random.default_rng(200 + n)
    Gs = np.stack([_gram(rng, n, max(1, n // 3)), _kron_gram(rng, n)])
    Gb = torch.from_numpy(Gs).to(DEV).contiguous()
    tri = GramTridiag(Gb, reduction="own")
    D, E = tri.D.cpu().numpy(), tri.E.cpu().numpy()
    for b in range(2):
        T = np.diag(D[b]) + np.diag(E[b][:n - 1], 1) + np.diag(E[b][:n - 1], -1)
        ev = np.linalg.eigvalsh(T)
        ref = np.linalg.eigvalsh(Gs[b])
        err = np.abs(ev - ref).max() / np.abs(ref).max()
        print(f"  rank-deficient n={n}[{b}]: eig(T) vs eig(G) {err:.1e}", flush=True)
        assert err < 1e-13


def test_own_reduction_time_against_rocsolver():
    """Timing of both reductions on 16 Gram matrices of order 2048 (printed)."""
    rng = np.random.default_rng(5)
    G1 = _gram(rng, 2048, 2048)
    Gs = torch.from_numpy(np.stack([G1] * 16)).to(DEV)
    for red in ("own", "rocsolver", "own"):
        Gb = Gs.clone()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        GramTridiag(Gb, reduction=red)
        e.record()
        torch.cuda.synchronize()
        print(f"  {red}: {s.elapsed_time(e) / 16:.2f} ms per matrix (batch 16, n 2048)", flush=True)


def test_tridiag_kept_rank_near_order():
    """Kept ranks whose 64-row padding exceeds the order (k = 96 and 100 of
    100; 500 of 506, as in a growing circuit's Gram cores): the padded rows
    stay zero and the kept vectors are exact."""
    rng = np.random.default_rng(12)
    _check([_gram(rng, 100, 100), _gram(rng, 100, 100)], [96, 100], "n=100 near-full")
    _check([_gram(rng, 506, 506)], [500], "n=506 near-full")
