"""GPU checks of the library's blocked Householder QR (kernels.qr: aq_geqrf +
aq_ungqr) and of its callers.

Against the oracle's robust_qr (numpy LAPACK on the host), gauge-invariantly:
Q^H Q = I and Q R = A to 1e-13 (complex128) / 1e-6 (complex64) relative to
max |A|, R upper triangular.  Since the algorithm is LAPACK's (zgeqrf with
zlarfg reflectors), R itself equals LAPACK's R on full-rank input to 1e-11
relative after one phase per row is divided out: the phases are 1 except
where LAPACK's blocked code leaves the length-1 reflector of the last row
of a wide (or large square) matrix out and keeps that diagonal entry
complex, while the library, as zlarfg, makes every diagonal entry real.
Rank-deficient input is checked on the invariant properties only.
"""
import numpy as np
import pytest
import torch

from atlas_q import kernels as K
from oracle import ref_mps as R

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _rand(rng, shape, rank=None):
    A = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    if rank is not None:
        m, n = shape[-2:]
        B = rng.standard_normal((m, rank)) + 1j * rng.standard_normal((m, rank))
        C = rng.standard_normal((rank, n)) + 1j * rng.standard_normal((rank, n))
        A = B @ C
    return A


def _check(A, Q, Rm, tol, exact_r):
    m, n = A.shape
    k = min(m, n)
    scale = np.abs(A).max()
    orth = np.abs(Q.conj().T @ Q - np.eye(k)).max()
    rec = np.abs(Q @ Rm - A).max() / scale
    assert Q.shape == (m, k) and Rm.shape == (k, n)
    assert np.all(np.tril(Rm, -1) == 0)
    msg = f"  {m}x{n}: |Q^H Q - I| {orth:.1e}, |QR - A| {rec:.1e}"
    dr = 0.0
    if exact_r:
        Qo, Ro, _ = R.robust_qr(torch.from_numpy(A))
        Ro = Ro.numpy()
        ph = np.diag(Rm) / np.diag(Ro)
        assert np.abs(np.abs(ph) - 1.0).max() < 1e-10
        moved = int((np.abs(ph - 1.0) > 1e-10).sum())
        dr = np.abs(Rm - ph[:, None] * Ro).max() / scale
        msg += f", |R - phase R_lapack| {dr:.1e} ({moved} rows rephased)"
    print(msg, flush=True)
    assert orth < tol and rec < tol, (orth, rec)
    assert dr < 100 * tol, dr


@pytest.mark.parametrize("shape", [(1, 1), (5, 3), (3, 5), (33, 33), (100, 37), (37, 100),
                                   (64, 128), (130, 129), (200, 600), (600, 300), (2048, 1024),
                                   (1024, 2048), (4096, 64)])
def test_qr_complex128_matches_lapack(shape):
    rng = np.random.default_rng(sum(shape))
    A = _rand(rng, shape)
    Q, Rm = K.qr(torch.from_numpy(A).to(DEV))
    _check(A, Q.cpu().numpy(), Rm.cpu().numpy(), 1e-13, exact_r=True)


@pytest.mark.parametrize("shape", [(64, 64), (512, 200), (2048, 1024)])
def test_qr_complex64(shape):
    rng = np.random.default_rng(7 + sum(shape))
    A = _rand(rng, shape).astype(np.complex64)
    Q, Rm = K.qr(torch.from_numpy(A).to(DEV))
    assert Q.dtype == torch.complex64 and Rm.dtype == torch.complex64
    _check(A.astype(np.complex128), Q.cpu().numpy().astype(np.complex128),
           Rm.cpu().numpy().astype(np.complex128), 1e-6, exact_r=False)


@pytest.mark.parametrize("shape,rank", [((200, 80), 10), ((64, 64), 1), ((300, 300), 150)])
def test_qr_rank_deficient(shape, rank):
    rng = np.random.default_rng(rank)
    A = _rand(rng, shape, rank=rank)
    A[:, 3] = 0.0  # an exactly zero column as well
    Q, Rm = K.qr(torch.from_numpy(A).to(DEV))
    _check(A, Q.cpu().numpy(), Rm.cpu().numpy(), 1e-13, exact_r=False)


def test_qr_batched():
    rng = np.random.default_rng(11)
    A = _rand(rng, (3, 70, 50))
    Q, Rm = K.qr(torch.from_numpy(A).to(DEV))
    for b in range(3):
        _check(A[b], Q[b].cpu().numpy(), Rm[b].cpu().numpy(), 1e-13, exact_r=True)


def test_canonical_forms_use_the_library_qr():
    """to_left_canonical / to_mixed_canonical on a random MPS: the state is
    unchanged (1e-12) and the cores left of the centre are left-orthonormal,
    right of it right-orthonormal (1e-12)."""
    from atlas_q.adaptive_mps import AdaptiveMPS, DTypePolicy
    rng = np.random.default_rng(5)
    n = 8
    m = AdaptiveMPS(n, bond_dim=1, eps_bond=1e-10, chi_max_per_bond=16, device=DEV,
                    dtype_policy=DTypePolicy(default=torch.complex128))
    dims = [1, 2, 4, 8, 16, 8, 4, 2, 1]
    m.tensors = [torch.from_numpy(_rand(rng, (dims[i], 2, dims[i + 1]))).to(DEV)
                 for i in range(n)]
    psi0 = m.to_statevector().cpu().numpy()
    m.to_mixed_canonical(3)
    psi1 = m.to_statevector().cpu().numpy()
    assert np.abs(psi1 - psi0).max() < 1e-12 * np.abs(psi0).max()
    for i, T in enumerate(m.tensors):
        T = T.cpu().numpy()
        cl, p, cr = T.shape
        if i < 3:
            M = T.reshape(cl * p, cr)
            assert np.abs(M.conj().T @ M - np.eye(cr)).max() < 1e-12
        elif i > 3:
            M = T.reshape(cl, p * cr)
            assert np.abs(M @ M.conj().T - np.eye(cl)).max() < 1e-12
    m.to_left_canonical()
    psi2 = m.to_statevector().cpu().numpy()
    assert np.abs(psi2 - psi0).max() < 1e-12 * np.abs(psi0).max()
