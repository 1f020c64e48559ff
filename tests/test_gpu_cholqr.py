"""GPU checks of the randomized SVD's own factorization kernels.

* ``aq_cholqr`` / ``rsvd.orth`` (shifted Cholesky QR, three passes):
  Q^H Q = I to 1e-13 and Q Q^H Y = Y to 1e-12 (relative to max |Y|) on
  sketch shapes m = 2048, l = 1 .. 300 (block edges 64, 65, 128, 130),
  including a sketch whose columns span nine decades (condition ~1e9), and
  the pivot code of a pass without shift on a matrix with a zero column.
  The triangular factor is not compared with LAPACK's: the three passes
  multiply several of them, so only the invariant properties pin the result.
* ``TopEig`` (eigenvalues by bisection, then the vectors of the k largest):
  eigenvalues against numpy's eigvalsh to 1e-12 relative to the largest,
  residuals |G v - lambda v| to 1e-11, orthonormal columns, at orders 16
  (Jacobi kernel), 65, 100, 200 and on a rank-deficient Gram matrix.
* ``aq_gram_svd_small`` (one-sided Jacobi SVD of the Cholesky factor of a
  Gram core, orders <= 64, wavefront reductions): eigenvalues against
  numpy's eigvalsh to 1e-12 of the largest, residuals 1e-11, the in-kernel
  FP64 re-check <= 1e-10 and equal to the host's |U^H U - I|; a
  rank-deficient core returns zero columns past its numerical rank (and
  TopEig then takes the two-sided kernel for vectors there).
* The Gram method's eigensolver-failure branch takes the plain method's SVD
  cascade with its complex128 promotion (forced by a failing eigensolver).
"""
import ctypes

import numpy as np
import pytest
import torch

from atlas_q import kernels as K

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _cz(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("l", [1, 5, 64, 65, 128, 130, 300])
def test_orth_sketch_orthonormal_and_same_span(l):
    from atlas_q.rsvd import orth
    rng = np.random.default_rng(l)
    m = 2048
    Y = _cz(rng, (m, l))
    Yd = torch.from_numpy(Y).to(DEV)
    infos = []
    Q = orth(Yd, infos).cpu().numpy()
    assert int(infos[0].abs().sum()) == 0
    orth_err = np.abs(Q.conj().T @ Q - np.eye(l)).max()
    span_err = np.abs(Q @ (Q.conj().T @ Y) - Y).max() / np.abs(Y).max()
    print(f"l={l}: |Q^H Q - I| {orth_err:.1e}, span {span_err:.1e}")
    assert orth_err < 1e-13
    assert span_err < 1e-12


def test_orth_ill_conditioned_sketch():
    """Columns scaled over nine decades: the shifted first pass keeps the
    factorization defined; after the two plain passes Q is orthonormal and
    spans Y's columns."""
    from atlas_q.rsvd import orth
    rng = np.random.default_rng(5)
    m, l = 2048, 96
    Y = _cz(rng, (m, l)) * np.logspace(0, -9, l)[None, :]
    infos = []
    Q = orth(torch.from_numpy(Y).to(DEV), infos).cpu().numpy()
    assert int(infos[0].abs().sum()) == 0
    orth_err = np.abs(Q.conj().T @ Q - np.eye(l)).max()
    # each column of Y lies in span(Q) to FP64 accuracy relative to its own norm
    res = np.linalg.norm(Q @ (Q.conj().T @ Y) - Y, axis=0) / np.linalg.norm(Y, axis=0)
    print(f"ill-conditioned: |Q^H Q - I| {orth_err:.1e}, worst column residual {res.max():.1e}")
    assert orth_err < 1e-13
    assert res.max() < 1e-6


def test_cholqr_reports_non_positive_pivot():
    """One unshifted pass on a sketch with a zero column: the pivot code is
    1 + that column's index and the output stays finite."""
    from atlas_q._ffi import check, lib
    rng = np.random.default_rng(9)
    m, l, z = 512, 70, 66
    Y = _cz(rng, (m, l))
    Y[:, z] = 0.0
    Yd = torch.from_numpy(Y).to(DEV)
    Qd = torch.empty_like(Yd)
    L = lib()
    nws = int(L.aq_cholqr_workspace_elems(m, l))
    ws = torch.empty(nws, dtype=torch.complex128, device=DEV)
    info = torch.full((1,), -7, dtype=torch.int32, device=DEV)
    check(L.aq_cholqr(m, l, K._ptr(Yd), l, K._ptr(Qd), l, 0.0, K._ptr(ws), nws,
                      ctypes.c_void_p(info.data_ptr()), K._stream(Yd)), "aq_cholqr")
    assert int(info.item()) == z + 1
    assert bool(torch.isfinite(torch.view_as_real(Qd)).all())


def _hermitian_with_spectrum(rng, lam):
    n = len(lam)
    Qm, _ = np.linalg.qr(_cz(rng, (n, n)))
    return (Qm * lam[None, :]) @ Qm.conj().T


@pytest.mark.parametrize("n", [16, 65, 100, 200])
def test_topeig_values_then_kept_vectors(n):
    from atlas_q.eig import TopEig
    rng = np.random.default_rng(n)
    lam = np.sort(np.exp(-np.linspace(0, 12, n)) * (1 + 0.1 * rng.random(n)))[::-1]
    G = _hermitian_with_spectrum(rng, lam)
    te = TopEig(torch.from_numpy(G).to(DEV))
    vals = te.values.cpu().numpy()
    ref = np.sort(np.linalg.eigvalsh(G))[::-1]
    assert np.abs(vals - ref).max() / ref[0] < 1e-12
    k = max(1, n // 3)
    V = te.vectors(k).cpu().numpy()
    assert V.shape == (n, k)
    res = np.abs(G @ V - V * vals[None, :k]).max() / ref[0]
    orth_err = np.abs(V.conj().T @ V - np.eye(k)).max()
    print(f"n={n}: residual {res:.1e}, orthogonality {orth_err:.1e}")
    assert res < 1e-11
    assert orth_err < 1e-10


def test_topeig_rank_deficient_gram():
    """Gram matrix of a rank-40 theta at order 130: eigenvalues (40 non-zero,
    90 at zero) and the kept 40 vectors."""
    from atlas_q.eig import TopEig
    rng = np.random.default_rng(3)
    X = _cz(rng, (130, 40)) @ _cz(rng, (40, 300))
    G = X @ X.conj().T
    te = TopEig(torch.from_numpy(G).to(DEV))
    vals = te.values.cpu().numpy()
    ref = np.sort(np.linalg.eigvalsh(G))[::-1]
    assert np.abs(vals - ref).max() / ref[0] < 1e-12
    V = te.vectors(40).cpu().numpy()
    res = np.abs(G @ V - V * vals[None, :40]).max() / ref[0]
    assert res < 1e-11
    assert np.abs(V.conj().T @ V - np.eye(40)).max() < 1e-10


def test_gram_eigensolver_failure_takes_plain_cascade_with_promotion(monkeypatch):
    """A failing eigensolver sends the update to the plain method's SVD
    cascade, which recomputes in complex128 when condS exceeds the policy's
    threshold: same kept rank, storage dtype and amplitudes as the plain
    method on the same gate."""
    import atlas_q.fastsvd as FS
    from atlas_q.adaptive_mps import AdaptiveMPS, DTypePolicy
    from atlas_q.circuits import brickwork

    class Failing:
        def __init__(self, G):
            raise RuntimeError("forced eigensolver failure")
    n, chi = 8, 16
    ops = brickwork(n, 6, 41)
    pol = dict(default=torch.complex64, promote_if_cond_gt=10.0)
    ref = AdaptiveMPS(n, bond_dim=1, eps_bond=1e-6, chi_max_per_bond=chi, device=DEV,
                      dtype_policy=DTypePolicy(**pol), svd_method="plain")
    for b, U in ops:
        ref.apply_two_site_gate(b, torch.from_numpy(U))
    monkeypatch.setattr(FS, "TopEig", Failing)
    got = AdaptiveMPS(n, bond_dim=1, eps_bond=1e-6, chi_max_per_bond=chi, device=DEV,
                      dtype_policy=DTypePolicy(**pol), svd_method="gram")
    for b, U in ops:
        got.apply_two_site_gate(b, torch.from_numpy(U))
    assert got.factor_stats.counts["plain"] == len(ops)
    assert got.statistics.logs["k_star"] == ref.statistics.logs["k_star"]
    assert [t.dtype for t in got.tensors] == [t.dtype for t in ref.tensors]
    assert any(t.dtype == torch.complex128 for t in got.tensors)
    for bstr in (0, 5, 2 ** n - 1):
        assert abs(got.get_amplitude(bstr) - ref.get_amplitude(bstr)) < 1e-6


@pytest.mark.parametrize("n", [1, 2, 17, 33, 64])
def test_gram_svd_small_matches_lapack(n):
    from atlas_q.eig import ORTH_TOL, gram_svd_small
    rng = np.random.default_rng(100 + n)
    Gs = []
    for decay in (None, 8.0):
        X = _cz(rng, (n, 3 * n + 1))
        if decay is not None:
            X = X * np.exp(-np.arange(n) / decay)[:, None]
        Gs.append(X @ X.conj().T)
    W, U, err = gram_svd_small(torch.from_numpy(np.stack(Gs)).to(DEV))
    W, U, err = W.cpu().numpy(), U.cpu().numpy(), err.cpu().numpy()
    for b, G in enumerate(Gs):
        ref = np.sort(np.linalg.eigvalsh(G))[::-1]
        ew = np.abs(W[b] - ref).max() / ref[0]
        res = np.abs(G @ U[b] - U[b] * W[b][None, :]).max() / ref[0]
        orth = np.abs(U[b].conj().T @ U[b] - np.eye(n)).max()
        print(f"n={n}[{b}]: eig {ew:.1e}, residual {res:.1e}, orth {orth:.1e} (re-check {err[b]:.1e})")
        assert ew < 1e-12 and res < 1e-11
        assert err[b] <= ORTH_TOL and orth <= ORTH_TOL
        assert abs(orth - err[b]) < 1e-13


def test_gram_svd_small_rank_deficient_core():
    from atlas_q.eig import TopEig, gram_svd_small
    rng = np.random.default_rng(11)
    n, r = 48, 20
    X = _cz(rng, (n, r)) @ _cz(rng, (r, 100))
    G = X @ X.conj().T
    W, U, err = gram_svd_small(torch.from_numpy(G).to(DEV))
    W, U = W.cpu().numpy(), U.cpu().numpy()
    ref = np.sort(np.linalg.eigvalsh(G))[::-1]
    assert np.abs(W - ref).max() < 1e-12 * ref[0]
    assert float(err) <= 1e-10
    Ur = U[:, :r]
    assert np.abs(G @ Ur - Ur * W[None, :r]).max() < 1e-11 * ref[0]
    assert np.abs(Ur.conj().T @ Ur - np.eye(r)).max() < 1e-10
    # vectors past the rank: TopEig takes the two-sided kernel
    before = TopEig.fallbacks
    te = TopEig(torch.from_numpy(G).to(DEV), small="svd")
    V = te.vectors(n).cpu().numpy()
    assert TopEig.fallbacks == before + 1
    assert np.abs(V.conj().T @ V - np.eye(n)).max() < 1e-10
