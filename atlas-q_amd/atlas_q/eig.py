"""Hermitian eigen-decomposition dispatch for the truncated-SVD cores.

Orders up to 64 run in the library's batched FP64 Jacobi kernel
(``aq_heev_small``: one work-group, matrix in LDS, no launch-latency chain);
larger orders use rocSOLVER's divide-and-conquer, either through
``torch.linalg.eigh`` or, for a batch of cores (all Gram matrices of a
brickwork sub-layer), through the library's strided-batched
``aq_heevd_batched``, which pays rocSOLVER's kernel chain once per batch.
Both return eigenvalues in descending order with matching eigenvector columns.

rocSOLVER can report non-convergence on strongly clustered spectra (many
eigenvalues at zero, as in the Gram matrix of a rank-deficient theta), and
on such batches it has also returned non-finite eigenvectors with a zero
info code; both count as a failure here.  As the reference's SVD cascade
retries with a 1e-12 jitter of the input (linalg_robust.py:31-39),
``eigh_desc`` retries on G + E, E a random Hermitian matrix of norm ~1e-13,
then ~1e-11, times ||G||_max, which separates the cluster and moves every
eigenvalue by at most ||E||_2; ``heevd_batched`` reports the failed matrices
to its caller.
"""
from __future__ import annotations

import torch

JACOBI_MAX = 64
JITTERS = (1e-13, 1e-11)   # relative sizes of the retries' Hermitian perturbation


def heevd_batched(G: torch.Tensor, inplace: bool = False, return_ok: bool = False):
    """Eigenpairs of Hermitian complex128 matrices G [batch, n, n] (n > 64) in
    one rocSOLVER strided-batched call: (eigenvalues desc [batch, n],
    eigenvectors as columns [batch, n, n]).

    G is left unmodified unless ``inplace`` (then G must be a contiguous
    complex128 tensor and is overwritten by the eigenvectors).  Without
    ``return_ok`` a matrix that did not converge raises; with it the call
    returns a third value, a host list of per-matrix convergence flags, and
    the caller handles the failed ones; ``return_ok="device"`` returns those
    flags as a device bool tensor instead (no synchronisation)."""
    from . import kernels as K
    from ._ffi import check, lib
    if inplace:
        if G.dtype != torch.complex128 or not G.is_contiguous():
            raise ValueError("heevd_batched(inplace=True) needs a contiguous complex128 batch")
        A = G
    else:
        A = G.to(torch.complex128).contiguous().clone()
    bsz, n, _ = A.shape
    W = torch.empty((bsz, n), dtype=torch.float64, device=A.device)
    E = torch.empty((bsz, n), dtype=torch.float64, device=A.device)
    info = torch.zeros(bsz, dtype=torch.int32, device=A.device)
    check(lib().aq_heevd_batched(n, K._ptr(A), n * n, K._ptr(W), n, K._ptr(E), K._ptr(info),
                                 bsz, K._stream(A)), "aq_heevd_batched")
    good = (info == 0) & finite_rows(W, A)
    # row j of A holds conj(v_j): V = A^H, columns ascending -> flip to descending
    lam, V = W.flip(-1), A.transpose(-2, -1).conj().flip(-1)
    if return_ok == "device":
        return lam, V, good
    ok = good.cpu().tolist()
    if not return_ok and not all(ok):
        raise RuntimeError(f"aq_heevd_batched: not converged or non-finite (info {info.tolist()})")
    return (lam, V, ok) if return_ok else (lam, V)


def finite_rows(W: torch.Tensor, V: torch.Tensor) -> torch.Tensor:
    """Per-matrix flag [batch]: eigenvalues W [batch, n] and eigenvectors
    V [batch, n, n] are all finite."""
    bsz = W.shape[0]
    return (torch.isfinite(W).all(1)
            & torch.isfinite(torch.view_as_real(V.resolve_conj())).reshape(bsz, -1).all(1))


def _eigh_checked(G: torch.Tensor):
    lam, V = torch.linalg.eigh(G)
    if not bool(torch.isfinite(lam).all()) or not bool(
            torch.isfinite(torch.view_as_real(V)).all()):
        raise RuntimeError("linalg.eigh: non-finite eigenpairs")
    return lam, V


def eigh_desc(G: torch.Tensor):
    n = G.shape[-1]
    if n <= JACOBI_MAX:
        from .rank_mlp import heev_small
        return heev_small(G)
    try:
        lam, V = _eigh_checked(G)
        return lam.flip(-1), V.flip(-1)
    except RuntimeError as e:
        err = e
    dmax = float(torch.diagonal(G, dim1=-2, dim2=-1).real.abs().max())
    R = torch.randn(G.shape, dtype=G.dtype, device=G.device)
    R = (R + R.transpose(-2, -1).conj()) * (0.5 / n ** 0.5)
    # 1e-13, then 1e-11 of max|G_ii|: a large zero cluster (a rank-deficient
    # projected core) can need the second to come out finite
    for rel in JITTERS:
        try:
            lam, V = _eigh_checked(G + (rel * dmax) * R)
            return lam.flip(-1), V.flip(-1)
        except RuntimeError as e:
            err = e
    raise err
