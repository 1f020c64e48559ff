"""Hermitian eigen-decomposition of the truncated-SVD cores.

The fast paths run on the library's kernels: orders up to 64 in the batched
FP64 Jacobi kernel (``aq_heev_small``: one work-group, matrix in LDS);
larger orders through the tridiagonal form (``GramTridiag``: the library's
blocked reduction, all eigenvalues by bisection, eigenvectors of the kept
ones only by inverse iteration, back-transformed on MFMA and re-checked for
orthogonality in FP64).  ``TopEig`` serves one matrix that way: eigenvalues
first, the rank rule on them, then the vectors of the kept rank.

rocSOLVER's divide-and-conquer remains as the comparison point of the
batched form (``heevd_batched``) and as the fallback of a matrix whose
vectors fail the orthogonality re-check (``eigh_desc``).  It can report
non-convergence on strongly clustered spectra (many eigenvalues at zero, as
in the Gram matrix of a rank-deficient theta), and on such batches it has
also returned non-finite eigenvectors with a zero info code; both count as a
failure here.  As the reference's SVD cascade retries with a 1e-12 jitter of
the input (linalg_robust.py:31-39), ``eigh_desc`` retries on G + E, E a
random Hermitian matrix of norm ~1e-13, then ~1e-11, times ||G||_max, which
separates the cluster and moves every eigenvalue by at most ||E||_2;
``heevd_batched`` reports the failed matrices to its caller.
"""
from __future__ import annotations

import ctypes

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


# ---------------------------------------------------------------- tridiagonal path
NB_BACK = 128      # Householder vectors per block of the MFMA back-transformation
NB_BACK_SMALL = 64   # ... below order NB_BACK_MIN_ORDER (fewer, shorter blocks pay less
NB_BACK_MIN_ORDER = 1024   # for the wider factor join than they gain)
HETRD_NB = 32      # panel width of the library's tridiagonal reduction
REDUCTION = "own"  # "own": the library's kernels; "rocsolver": rocSOLVER's zhetrd
ORTH_TILE = 64     # rows per diagonal tile of the orthogonality re-check
ORTH_TOL = 1e-10   # largest |U_k^H U_k - I| entry accepted (else the matrix falls back)


def reduction_bytes(n: int, nb: int = HETRD_NB) -> float:
    """Algorithmic HBM bytes of the one-stage blocked reduction of one order-n
    matrix: every column's Hermitian matrix-vector product reads the stored
    triangle of the leading ii x ii block once (ii(ii+1)/2 complex128), and
    every panel's rank-2nb update reads and writes the lower triangle of the
    leading p0 x p0 block."""
    hemv = sum(ii * (ii + 1) // 2 for ii in range(1, n)) * 16.0
    upd = sum(p0 * (p0 + 1) // 2 for p0 in range(n - nb, 0, -nb)) * 32.0
    return hemv + upd


class GramTridiag:
    """Leading eigenpairs of a batch of Hermitian complex128 Gram cores
    [B, n, n] (row-major, lower triangle read, overwritten) through the
    tridiagonal form: ``eigenvalues`` gives the spectra by bisection, and
    ``vectors(k)`` the eigenvectors of the k[b] largest eigenvalues of matrix
    b by twisted inverse iteration, back-transformed on MFMA, with the FP64
    orthogonality re-check of the result.  ``groups`` (own reduction only):
    None follows the process-wide stream-group setting; 1 keeps the whole
    reduction on the current stream, so reductions started from one host
    thread on different streams overlap."""

    def __init__(self, Gb: torch.Tensor, reduction: str = None, groups: int = None):
        from . import kernels as K
        from ._ffi import check, lib
        if Gb.dtype != torch.complex128 or not Gb.is_contiguous() or Gb.dim() != 3:
            raise ValueError("GramTridiag needs a contiguous complex128 batch [B, n, n]")
        self.G = Gb
        self.B, self.n = int(Gb.shape[0]), int(Gb.shape[1])
        B, n, dev = self.B, self.n, Gb.device
        self.D = torch.empty((B, n), dtype=torch.float64, device=dev)
        self.E = torch.empty((B, max(n - 1, 1)), dtype=torch.float64, device=dev)
        self.tau = torch.empty((B, max(n - 1, 1)), dtype=torch.complex128, device=dev)
        reduction = reduction or REDUCTION
        if reduction == "own":
            L = lib()
            nws = int(L.aq_hetrd_workspace_elems(n, HETRD_NB, B))
            ws = torch.empty(nws, dtype=torch.complex128, device=dev)
            if groups is None:
                check(L.aq_hetrd_own(n, HETRD_NB, K._ptr(Gb), n * n, K._ptr(self.D),
                                     K._ptr(self.E), K._ptr(self.tau), K._ptr(ws), nws, B,
                                     K._stream(Gb)), "aq_hetrd_own")
            else:
                check(L.aq_hetrd_own_groups(n, HETRD_NB, K._ptr(Gb), n * n, K._ptr(self.D),
                                            K._ptr(self.E), K._ptr(self.tau), K._ptr(ws), nws,
                                            B, int(groups), K._stream(Gb)),
                      "aq_hetrd_own_groups")
            self._ws = ws
        elif reduction == "rocsolver":
            check(lib().aq_hetrd_batched(n, K._ptr(Gb), n * n, K._ptr(self.D), K._ptr(self.E),
                                         K._ptr(self.tau), B, K._stream(Gb)), "aq_hetrd_batched")
        else:
            raise ValueError(f"unknown tridiagonal reduction {reduction!r}")
        self.W = None

    def record_stream(self, stream) -> None:
        """Mark the solver's device buffers as used on ``stream`` (for a solver
        built on one stream and continued on another)."""
        for t in (self.G, self.D, self.E, self.tau, getattr(self, "_ws", None), self.W):
            if t is not None:
                t.record_stream(stream)

    def eigenvalues(self) -> torch.Tensor:
        """All eigenvalues, descending [B, n] (float64)."""
        from . import kernels as K
        from ._ffi import check, lib
        B, n = self.B, self.n
        self.W = torch.empty((B, n), dtype=torch.float64, device=self.G.device)
        check(lib().aq_tridiag_eigvals(n, K._ptr(self.D), n, K._ptr(self.E), n - 1,
                                       K._ptr(self.W), n, n, B, K._stream(self.G)),
              "aq_tridiag_eigvals")
        return self.W

    @staticmethod
    def vectors_buffer(B: int, n: int, kpad: int, device) -> torch.Tensor:
        """Zeroed storage for ``vectors(out=...)``: Et [B, kpad, n] plus the
        tail the orthogonality re-check's last tile reads."""
        return torch.zeros(B * kpad * n + ORTH_TILE * n, dtype=torch.complex128, device=device)

    def vectors(self, ks=None, kvec: torch.Tensor = None, kpad: int = None,
                out: torch.Tensor = None):
        """(Et [B, kpad, n] complex128, err [B] float64 on the device): row j <
        ks[b] of Et[b] is the conjugate transpose of the eigenvector of the
        j-th largest eigenvalue (rows >= ks[b] are zero); err[b] the largest
        |Et Et^H - I| entry over the diagonal tiles of the first ks[b] rows.
        The kept ranks come from the host list ``ks`` or, with no host round
        trip, from ``kvec`` (int32 [B] on the device) with the row count
        ``kpad`` (a multiple of 64 and at least every kvec[b]); ``out`` is
        optional storage from ``vectors_buffer``."""
        from . import kernels as K
        from ._ffi import check, lib
        B, n, dev = self.B, self.n, self.G.device
        st = K._stream(self.G)
        L = lib()
        if kvec is None:
            kmax = max(1, max(int(k) for k in ks))
            kpad = -(-kmax // ORTH_TILE) * ORTH_TILE
            kv = torch.tensor([int(k) for k in ks], dtype=torch.int32).to(dev)
        else:
            if kpad is None or kpad % ORTH_TILE or kvec.dtype != torch.int32 or kvec.numel() != B:
                raise ValueError("vectors: kvec needs int32 [B] and kpad a multiple of 64")
            kv = kvec
        if out is None:
            flat = self.vectors_buffer(B, n, kpad, dev)
        else:
            if out.numel() != B * kpad * n + ORTH_TILE * n or out.dtype != torch.complex128:
                raise ValueError("vectors: out must come from vectors_buffer(B, n, kpad)")
            flat = out
        Et = flat[:B * kpad * n].view(B, kpad, n)
        nsc = int(L.aq_tridiag_eigvecs_scratch_doubles(n, kpad, B))
        sc = torch.empty(nsc, dtype=torch.float64, device=dev)
        check(L.aq_tridiag_eigvecs(n, K._ptr(self.D), n, K._ptr(self.E), n - 1, K._ptr(self.W),
                                   n, K._ptr(kv), kpad, K._ptr(Et), kpad * n, K._ptr(sc), nsc,
                                   B, st), "aq_tridiag_eigvecs")
        del sc
        nb = NB_BACK if n >= NB_BACK_MIN_ORDER else NB_BACK_SMALL
        nblk = -(-(n - 1) // nb)
        V = torch.empty((B, nblk, nb, n), dtype=torch.complex128, device=dev)
        check(L.aq_hetrd_vblocks(n, nb, K._ptr(self.G), n * n, K._ptr(V), B, st),
              "aq_hetrd_vblocks")
        Gv = torch.empty((B, nblk, nb, nb), dtype=torch.complex128, device=dev)
        bg = L.aq_bgemm
        check(bg(1, nb, nb, n, 1.0, K._ptr(V), n, 1, nb * n, 1, K._ptr(V), 1, n, nb * n, 0, 0.0,
                 K._ptr(Gv), nb, 1, nb * nb, B * nblk, st), "aq_bgemm (block Gram)")
        Tf = torch.empty_like(Gv)
        check(L.aq_hetrd_tfactors(n, nb, K._ptr(self.tau), n - 1, K._ptr(Gv), K._ptr(Tf), B, st),
              "aq_hetrd_tfactors")
        Vt2 = torch.empty_like(V)
        check(bg(1, nb, n, nb, 1.0, K._ptr(Tf), nb, 1, nb * nb, 1, K._ptr(V), n, 1, nb * n, 0,
                 0.0, K._ptr(Vt2), n, 1, nb * n, B * nblk, st), "aq_bgemm (conj(T') V)")
        Wt = torch.empty((B, nb, kpad), dtype=torch.complex128, device=dev)
        vb, eb = nblk * nb * n, kpad * n
        for p in range(nblk):
            Kp = min(p * nb + nb - 1, n - 2) + 1            # rows the block's vectors touch
            off = p * nb * n * 16                          # bytes into V / Vt2 of matrix 0
            vp = ctypes.c_void_p(V.data_ptr() + off)
            tp = ctypes.c_void_p(Vt2.data_ptr() + off)
            # Wt = conj(V_p) Et^T   [nb, kpad]
            check(bg(1, nb, kpad, Kp, 1.0, vp, n, 1, vb, 1, K._ptr(Et), 1, n, eb, 0, 0.0,
                     K._ptr(Wt), kpad, 1, nb * kpad, B, st), "aq_bgemm (back 1)")
            # Et[:, :, :Kp] -= Wt^T (conj(T') V_p)
            check(bg(1, kpad, Kp, nb, -1.0, K._ptr(Wt), 1, kpad, nb * kpad, 0, tp, n, 1, vb, 0,
                     1.0, K._ptr(Et), n, 1, eb, B, st), "aq_bgemm (back 2)")
        del V, Vt2, Gv, Tf, Wt
        ntile = kpad // ORTH_TILE
        Q = torch.empty((B, ntile, ORTH_TILE, 2 * ORTH_TILE), dtype=torch.complex128, device=dev)
        check(bg(1, ORTH_TILE, 2 * ORTH_TILE, n, 1.0, K._ptr(Et), n, 1, ORTH_TILE * n, 0,
                 K._ptr(Et), 1, n, ORTH_TILE * n, 1, 0.0, K._ptr(Q), 2 * ORTH_TILE, 1,
                 ORTH_TILE * 2 * ORTH_TILE, B * ntile, st), "aq_bgemm (orthogonality)")
        err = torch.empty(B, dtype=torch.float64, device=dev)
        check(L.aq_orth_band_reduce(K._ptr(Q), ntile, K._ptr(kv), kpad, K._ptr(err), B, st),
              "aq_orth_band_reduce")
        self._keep = (flat, kv, Q)
        return Et, err


def gram_svd_small(G: torch.Tensor, max_sweeps: int = 30, tol: float = 1e-15):
    """One-sided Jacobi SVD of the Cholesky factor of Hermitian positive
    semidefinite complex128 matrices G [batch, n, n] or [n, n] (n <= 64):
    (eigenvalues desc, eigenvectors as columns, FP64 re-check err [batch])."""
    from . import kernels as K
    from ._ffi import check, lib
    single = G.dim() == 2
    if single:
        G = G.unsqueeze(0)
    G = G.to(torch.complex128).contiguous()
    bsz, n, _ = G.shape
    W = torch.empty((bsz, n), dtype=torch.float64, device=G.device)
    U = torch.empty((bsz, n, n), dtype=torch.complex128, device=G.device)
    err = torch.empty(bsz, dtype=torch.float64, device=G.device)
    check(lib().aq_gram_svd_small(n, K._ptr(G), n, n * n, K._ptr(W), K._ptr(U), n, n * n, bsz,
                                  max_sweeps, tol, K._ptr(err), K._stream(G)),
          "aq_gram_svd_small")
    return (W[0], U[0], err[0]) if single else (W, U, err)


class TopEig:
    """Eigenvalues (descending, ``values``) of one Hermitian complex128 matrix
    G [n, n] and, on request, the eigenvectors of its k largest
    (``vectors(k)``: columns [n, k]).  Orders <= 64: the two-sided Jacobi
    kernel, or with ``small="svd"`` (a positive semidefinite Gram core) the
    one-sided Jacobi SVD of its Cholesky factor.  Larger: the tridiagonal
    path (reduction and bisection now, inverse iteration and
    back-transformation for the k vectors later).  Vectors that fail the
    FP64 orthogonality re-check (or lie past the one-sided factor's
    numerical rank) are taken from the two-sided Jacobi kernel / ``eigh_desc``
    instead (``fallbacks`` counts them)."""

    fallbacks = 0

    def __init__(self, G: torch.Tensor, small: str = "jacobi"):
        self.G = G
        n = G.shape[-1]
        self._err = None
        if n <= JACOBI_MAX and small == "svd":
            self.values, self._V, self._err = gram_svd_small(G)
            self._tri = None
        elif n <= JACOBI_MAX:
            from .rank_mlp import heev_small
            self.values, self._V = heev_small(G)
            self._tri = None
        else:
            Gb = G.to(torch.complex128).contiguous().clone().unsqueeze(0)
            self._tri = GramTridiag(Gb)
            self.values = self._tri.eigenvalues()[0]
            self._V = None

    def vectors(self, k: int) -> torch.Tensor:
        k = int(k)
        if self._err is not None:
            host = torch.stack([self._err, self.values[k - 1]]).cpu().tolist()
            self._err = None
            if not (host[0] <= ORTH_TOL and host[1] > 0.0):
                from .rank_mlp import heev_small
                TopEig.fallbacks += 1
                _, self._V = heev_small(self.G)
        if self._V is not None:
            return self._V[:, :k]
        Et, err = self._tri.vectors([k])
        e = float(err[0])
        if not (e <= ORTH_TOL):
            TopEig.fallbacks += 1
            _, V = eigh_desc(self.G)
            self._V = V
            return V[:, :k]
        return Et[0, :k, :].transpose(0, 1).conj()
