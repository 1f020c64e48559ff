"""Many independent MPS advanced together by grouped persistent launches.

Storage: every site tensor of every MPS lives in one complex64 arena in HBM,
in a slot sized for the largest bond dimensions that site can reach
(min(chi_max, 2^q, 2^(n-q)) on the left, likewise on the right), so updates
happen in place and nothing is reallocated per gate; the current bond
dimensions are tracked on the host.  A circuit layer is executed as

  single-qubit gates  : 1 grouped launch for all sites of all MPS (in place)
  two-qubit sub-layer : grouped theta (MFMA, gate in the epilogue)
                        -> FP64 Gram of the smaller side (grouped GEMM)
                        -> eigen-decomposition per size bucket (Jacobi kernel
                           for orders <= 64; above, the library's tridiagonal
                           eigensolver: reduction, all eigenvalues by
                           bisection)
                        -> batched rank rule (one launch, one 48-byte-per-gate
                           device->host copy)
                        -> eigenvectors of the kept ranks only (tridiagonal
                           buckets: inverse iteration, MFMA back-transformation,
                           FP64 orthogonality re-check: a second small copy)
                        -> grouped projection GEMM + grouped rebuild into the
                           slots

Each gate follows AdaptiveMPS.apply_two_site_gate (adaptive_mps.py:182-308):
"U" gate convention, the rank rule with per-bond caps, S absorbed left.
Single-qubit gates follow adaptive_mps.py:177 (U^T).  The sampler runs the
sweep of mps_pytorch.py:247-355 for all MPS with two grouped launches per
site index and can be captured into a HIP graph.
"""
from __future__ import annotations

import contextlib
import ctypes
import gc
import math
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _ffi
from ._ffi import (COPY_DESC, GEMM_DESC, SAMPLE_DESC, SITE1Q_DESC, RankResult, check, lib)
from .autotune import Autotuner
from .eig import JACOBI_MAX, ORTH_TOL, GramTridiag, finite_rows, heevd_batched

BATCHED_JACOBI_MAX = 32
from .linalg_robust import robust_svd
from .rank_mlp import heev_small

C64 = torch.complex64
C128 = torch.complex128


def _vp(x) -> ctypes.c_void_p:
    return ctypes.c_void_p(x if isinstance(x, int) else x.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _h2d(arr: np.ndarray, device) -> torch.Tensor:
    """Host array -> device tensor, queued on the current stream without
    waiting for it: the values are staged in page-locked memory now (so the
    caller may reuse ``arr``) and copied asynchronously; torch's pinned-memory
    cache keeps the staging block until the copy has run.  (A copy from
    pageable memory would synchronise the stream, draining the GPU's queue at
    every upload.)"""
    src = torch.from_numpy(np.ascontiguousarray(arr))
    if torch.device(device).type == "cpu":
        return src.clone()
    staged = torch.empty(src.shape, dtype=src.dtype, pin_memory=True)
    staged.copy_(src)
    return staged.to(device, non_blocking=True)


def _upload(arr: np.ndarray, device) -> torch.Tensor:
    """A structured array's bytes on the device."""
    return _h2d(np.ascontiguousarray(arr).view(np.uint8), device)


def _prefix(counts: np.ndarray) -> Tuple[np.ndarray, int]:
    pre = np.zeros(len(counts), np.int64)
    if len(counts) > 1:
        pre[1:] = np.cumsum(counts[:-1])
    return pre, int(counts.sum())


def _cdiv(a, b):
    return -(-a // b)


def _gather_gates(us, d: int) -> np.ndarray:
    """Stack d x d gate matrices into a complex128 array [len(us), d, d]: their
    bytes joined when the gates are complex128 arrays of d * d entries (the
    common case, ~3x faster than numpy.array on the list; checked on the first
    gate's dtype and on the total byte count), else converted."""
    if not us:
        return np.zeros((0, d, d), np.complex128)
    if getattr(us[0], "dtype", None) == np.complex128:
        try:
            buf = b"".join([u.tobytes() for u in us])
        except AttributeError:
            buf = b""
        if len(buf) == 16 * d * d * len(us):
            return np.frombuffer(buf, np.complex128).reshape(len(us), d, d).copy()
    return np.array(us, np.complex128).reshape(len(us), d, d)


def _merge_per_site(jj: np.ndarray, qq: np.ndarray, U: np.ndarray, nmax: int):
    """Gates U[g] on sites (jj[g], qq[g]) in op order, merged per site into one
    matrix per site: U_first @ ... @ U_last.  Returns (jj, qq, U) per site."""
    key = jj * (nmax + 1) + qq
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.ones(len(ks), bool)
    first[1:] = ks[1:] != ks[:-1]
    gid = np.cumsum(first) - 1
    heads = np.flatnonzero(first)
    rank = np.arange(len(ks)) - heads[gid]
    Us = U[order]
    M = Us[heads].copy()
    for r in range(1, int(rank.max()) + 1):
        sel = rank == r
        M[gid[sel]] = M[gid[sel]] @ Us[sel]
    return jj[order[heads]], qq[order[heads]], M


def _ragged_arange(lens: np.ndarray) -> np.ndarray:
    """concatenate([arange(L) for L in lens]) without a Python loop."""
    lens = np.asarray(lens, np.int64)
    total = int(lens.sum())
    if total == 0:
        return np.zeros(0, np.int64)
    starts = np.zeros(len(lens), np.int64)
    starts[1:] = np.cumsum(lens[:-1])
    return np.arange(total, dtype=np.int64) - np.repeat(starts, lens)


def _condition_bucket(T: torch.Tensor, orders: np.ndarray, jitter: torch.Tensor):
    """Prepare a size bucket for rocSOLVER's divide-and-conquer.

    Matrix b holds its Gram matrix in the leading orders[b] rows/columns and
    zeros elsewhere; d_b is its largest diagonal entry (1 when it is zero).
    * The padding diagonal gets the distinct negative values
      -d_b (1 + i / size): the padding eigenvalues sort below every Gram
      eigenvalue (which is >= 0) instead of forming one large zero cluster
      with the Gram's own null space.
    * 1e-13 d_b ``jitter`` (a fixed Hermitian matrix of spectral norm ~1) is
      added: on the exactly degenerate spectra of structured states
      (CZ-ladder circuits) the strided-batched zheevd returns non-finite
      eigenvectors without an error code; the perturbation moves every
      eigenvalue by at most ~1e-13 d_b, as the reference's 1e-12 SVD jitter
      retry does (linalg_robust.py:31-39)."""
    B, sz, _ = T.shape
    dg = T.diagonal(dim1=1, dim2=2)
    d = dg.real.amax(1, keepdim=True)
    d = torch.where(d > 0, d, torch.ones_like(d))
    if int(orders.min()) < sz:
        i = torch.arange(sz, device=T.device, dtype=torch.float64)
        pad = _h2d(np.arange(sz)[None, :] >= orders[:, None], T.device)
        dg.copy_(torch.where(pad, (-d * (1.0 + i / sz)).to(T.dtype), dg))
    if jitter is not None:
        T.add_(jitter.unsqueeze(0) * (1e-13 * d).unsqueeze(-1).to(T.dtype))


def _pad_bucket(T: torch.Tensor, orders: np.ndarray):
    """Prepare a size bucket for the tridiagonal eigensolver: the padding
    diagonal of matrix b (rows/columns >= orders[b], zero elsewhere) gets the
    distinct negative values -d_b (1 + i / size), d_b the Gram's largest
    diagonal entry.  The padded matrix is the Gram block plus a diagonal
    block, so the reduction leaves the padding columns alone (their
    reflectors are the identity), the tridiagonal form splits exactly at the
    block boundary, and every padding eigenvalue sorts below the Gram's (which
    are >= 0) and apart from its null space.  No perturbation is added: the
    solver treats exactly repeated eigenvalues by splitting the tridiagonal."""
    _condition_bucket(T, orders, None)


@contextlib.contextmanager
def _gc_paused():
    """Pause Python's cyclic garbage collector (restored on exit).  The
    circuit loops allocate many short-lived arrays, which trigger collections
    whose cost grows with every live object of the calling program (a
    population run inside a large program was ~15 % slower); their own
    temporaries are freed by reference counting."""
    was_enabled = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was_enabled:
            gc.enable()


def _jitter_matrix(sz: int, device) -> torch.Tensor:
    """Fixed Hermitian matrix of order sz with spectral norm ~1 (seeded by sz)."""
    g = torch.Generator(device=device).manual_seed(1000003 + sz)
    R = torch.randn((sz, sz), dtype=torch.complex128, device=device, generator=g)
    return (R + R.conj().T) * (0.5 / math.sqrt(2.0 * sz))


class BatchedMPS:
    """Population of MPS (n_j qubits, cap chi_j) starting in |0...0>."""

    def __init__(self, specs: Sequence[Tuple[int, int]], eps_bond: float = 1e-8,
                 device="cuda", autotuner: Autotuner = None, eig_workers: int = 3,
                 eig_solver: str = "tridiag", jacobi_max: int = None):
        self.dev = torch.device(device)
        # Gram orders solved by the Jacobi kernel (each matrix at its own
        # order); larger ones by the eig_solver path on padded buckets.  At
        # orders 33-64 the single-work-group reduction, bisection and kept
        # vectors beat the order-64 Jacobi kernel (population: 205k -> 235k
        # updates/s; limit 16: 229k)
        self.jacobi_max = BATCHED_JACOBI_MAX if jacobi_max is None else int(jacobi_max)
        if not 2 <= self.jacobi_max <= JACOBI_MAX:
            raise ValueError(f"jacobi_max must be in [2, {JACOBI_MAX}]")
        # host threads (one HIP stream each) for the size buckets above the
        # Jacobi limit: each bucket's solve is a chain of small per-column
        # launches, so buckets solved concurrently overlap their latencies
        self.eig_workers = int(eig_workers)
        # orders above the Jacobi limit: "tridiag" (the library's reduction,
        # bisection and top-k inverse iteration) or "heevd" (rocSOLVER's
        # divide-and-conquer, kept as the comparison point)
        if eig_solver not in ("tridiag", "heevd"):
            raise ValueError(f"unknown eig_solver {eig_solver!r}")
        self.eig_solver = eig_solver
        if self.dev.type != "cuda":
            raise RuntimeError("BatchedMPS runs on a HIP device only")
        _ffi.check_layouts()
        self.n = np.array([s[0] for s in specs], np.int64)
        self.cap = np.array([s[1] for s in specs], np.int64)
        self.eps = float(eps_bond)
        self.M = len(specs)
        self.nmax = int(self.n.max())
        # bond capacities: bcap[j, q] for bond q (left of site q), q = 0..n_j
        self.bcap = np.zeros((self.M, self.nmax + 1), np.int64)
        for j in range(self.M):
            q = np.arange(self.n[j] + 1)
            e = np.minimum(q, self.n[j] - q)
            self.bcap[j, : self.n[j] + 1] = np.minimum(self.cap[j], 2.0 ** np.minimum(e, 62))
        self.slot_off = np.zeros((self.M, self.nmax), np.int64)   # in complex elements
        off = 0
        for j in range(self.M):
            for q in range(self.n[j]):
                self.slot_off[j, q] = off
                off += 2 * self.bcap[j, q] * self.bcap[j, q + 1]
        self.arena = torch.zeros(off, dtype=C64, device=self.dev)
        self.dims = np.ones((self.M, self.nmax + 1), np.int64)
        # |0...0>: every site [1, 2, 1] = (1, 0)
        first = torch.from_numpy(np.concatenate([self.slot_off[j, : self.n[j]]
                                                 for j in range(self.M)])).to(self.dev)
        self.arena[first] = 1.0
        self.tuner = autotuner if autotuner is not None else Autotuner.default()
        self.logs = {"mps": [], "bond": [], "k": [], "eps_local": [], "entropy": []}
        self.eig_fallbacks = 0     # matrices whose batched eigensolve failed
        self._jitter = {}          # order -> fixed Hermitian perturbation
        self._base = self.arena.data_ptr()

    # ------------------------------------------------------------ views
    def site(self, j: int, q: int) -> torch.Tensor:
        dl, dr = int(self.dims[j, q]), int(self.dims[j, q + 1])
        o = int(self.slot_off[j, q])
        return self.arena[o: o + dl * 2 * dr].view(dl, 2, dr)

    def tensors(self, j: int) -> List[torch.Tensor]:
        return [self.site(j, q) for q in range(int(self.n[j]))]

    def bond_dims(self, j: int) -> List[int]:
        return [int(x) for x in self.dims[j, 1: self.n[j]]]

    def _ptr(self, j, q) -> int:
        return self._base + 8 * int(self.slot_off[j, q])

    def _ptrs(self, jj: np.ndarray, qq: np.ndarray) -> np.ndarray:
        """Slot addresses of sites (jj[g], qq[g]) (vectorised ``_ptr``)."""
        return self._base + 8 * self.slot_off[jj, qq]

    # ------------------------------------------------------------ 1q layer
    def apply_1q_layer(self, ops: Sequence[Sequence[Tuple[int, np.ndarray]]]):
        """ops[j] = [(q, U2), ...]; at most one gate per site per call."""
        rows = [(j, q, U) for j, lst in enumerate(ops) for q, U in lst]
        if not rows:
            return
        self.apply_1q_arrays(np.array([r[0] for r in rows], np.int64),
                             np.array([r[1] for r in rows], np.int64),
                             np.stack([np.asarray(r[2], np.complex128).reshape(2, 2) for r in rows]))

    def apply_1q_arrays(self, jj: np.ndarray, qq: np.ndarray, U: np.ndarray):
        """Gate U[g] (2 x 2, U^T convention) on site qq[g] of MPS jj[g]; at most
        one gate per site per call."""
        if len(jj) == 0:
            return
        d = np.zeros(len(jj), SITE1Q_DESC)
        p = self._ptrs(jj, qq)
        d["T_in"] = p
        d["T_out"] = p
        d["chiL"] = self.dims[jj, qq]
        d["chiR"] = self.dims[jj, qq + 1]
        d["gate"] = np.ascontiguousarray(U, np.complex128).reshape(len(jj), 4).view(np.float64)
        cnt = d["chiL"] * d["chiR"]
        pre, total = _prefix(cnt)
        dd, pd = _upload(d, self.dev), _h2d(pre, self.dev)
        grid = self.tuner.grid("elem", total, None)   # in place: no timing trials
        check(lib().aq_grouped_1q(_vp(dd), _vp(pd), len(d), total, grid, _stream()),
              "aq_grouped_1q")

    # ------------------------------------------------------------ 2q layer
    def _grouped_gemm(self, dtype_code, descs: np.ndarray, layout: int, key: str):
        self._launch_grouped(self._prep_grouped(dtype_code, descs, layout, key))

    def _prep_grouped(self, dtype_code, descs: np.ndarray, layout: int, key: str):
        """Descriptors and tile offsets of one grouped GEMM, uploaded (queued
        on the stream); ``_launch_grouped`` launches it."""
        if len(descs) == 0:
            return None
        paired = descs["kind"] == _ffi.AQ_GROUPED_THETA
        tn = np.where(paired, _cdiv(descs["N"] // 2, 64), _cdiv(descs["N"], 128))
        dd = _upload(descs, self.dev)
        shapes = {}
        for rows in (64, 32):              # the grouped GEMM's two tile heights
            pre, total = _prefix(_cdiv(descs["M"], rows) * tn)
            shapes[rows] = (_h2d(pre, self.dev), total)
        return dtype_code, dd, shapes, len(descs), layout, key

    def _launch_grouped(self, prep):
        if prep is None:
            return
        dtype_code, dd, shapes, count, layout, key = prep

        def run(rows, g):
            pd, total = shapes[rows]
            lay = layout | (_ffi.AQ_GROUPED_LAYOUT_ROWS32 if rows == 32 else 0)
            check(lib().aq_grouped_gemm(dtype_code, _vp(dd), _vp(pd), count, total, lay,
                                        g, _stream()), "aq_grouped_gemm")
        run(*self.tuner.shape(key, {r: t for r, (_, t) in shapes.items()}, run))

    def apply_2q_layer(self, ops: Sequence[Sequence[Tuple[int, np.ndarray]]]):
        """ops[j] = [(bond, U4), ...] with disjoint bonds within each MPS."""
        gates = [(j, b, U) for j, lst in enumerate(ops) for b, U in lst]
        if not gates:
            return
        self.apply_2q_arrays(np.array([g[0] for g in gates], np.int64),
                             np.array([g[1] for g in gates], np.int64),
                             np.stack([np.asarray(g[2], np.complex128).reshape(4, 4)
                                       for g in gates]))

    def apply_2q_arrays(self, jj: np.ndarray, bb: np.ndarray, U4: np.ndarray):
        """Gate U4[g] (4 x 4, "U" convention) on bond bb[g] of MPS jj[g]; the
        bonds of one MPS must be disjoint."""
        G = len(jj)
        if G == 0:
            return
        dl = self.dims[jj, bb]
        dm = self.dims[jj, bb + 1]
        dr = self.dims[jj, bb + 2]
        m, n = 2 * dl, 2 * dr
        # ---- theta into the X arena (complex64)
        xoff, xtot = _prefix(m * n)
        X = torch.empty(xtot, dtype=C64, device=self.dev)
        th = np.zeros(G, GEMM_DESC)
        th["A"] = self._ptrs(jj, bb)
        th["B"] = self._ptrs(jj, bb + 1)
        th["C"] = X.data_ptr() + 8 * xoff
        th["M"], th["N"], th["K"] = m, n, dm
        th["a_s0"], th["a_s1"] = dm, 1
        th["b_s0"], th["b_s1"] = n, 1
        th["c_s0"], th["c_s1"] = n, 1
        th["kind"] = _ffi.AQ_GROUPED_THETA
        th["gate"] = np.ascontiguousarray(U4, np.complex128).reshape(G, 16).view(np.float64)
        self._grouped_gemm(_ffi.AQ_C64, th, 3, "theta")
        X64 = X.to(C128)
        x64 = X64.data_ptr() + 16 * xoff
        # ---- FP64 Gram of the smaller side into size buckets
        s = np.minimum(m, n)
        left = m <= n
        bsize = np.maximum(2, 2 ** np.ceil(np.log2(s)).astype(np.int64))
        buckets = {}
        for sz in np.unique(bsize):
            idx = np.nonzero(bsize == sz)[0]
            buckets[int(sz)] = (idx, torch.zeros((len(idx), int(sz), int(sz)), dtype=C128,
                                                 device=self.dev))
        gram = np.zeros(G, GEMM_DESC)
        gptr = np.zeros(G, np.int64)
        gld = np.zeros(G, np.int64)
        for sz, (idx, T) in buckets.items():
            gptr[idx] = T.data_ptr() + 16 * sz * sz * np.arange(len(idx))
            gld[idx] = sz
        gram["A"] = x64
        gram["B"] = x64
        gram["C"] = gptr
        gram["M"] = s
        gram["N"] = s
        gram["K"] = np.where(left, n, m)
        # left: X X^H  (A = X: (i,k) at i*n + k; B = X^H: (k,j) = conj X[j,k] at j*n + k)
        # right: X^H X (A = X^H: (i,k) = conj X[k,i] at k*n + i; B = X: (k,j) at k*n + j)
        gram["a_s0"] = np.where(left, n, 1)
        gram["a_s1"] = np.where(left, 1, n)
        gram["conj_a"] = np.where(left, 0, 1)
        gram["b_s0"] = np.where(left, 1, n)
        gram["b_s1"] = np.where(left, n, 1)
        gram["conj_b"] = np.where(left, 1, 0)
        gram["c_s0"], gram["c_s1"] = gld, 1
        gram["kind"] = _ffi.AQ_GROUPED_GEMM
        self._grouped_gemm(_ffi.AQ_C128, gram[left], 1 | 4, "gram")
        self._grouped_gemm(_ffi.AQ_C128, gram[~left], 2 | 4, "gram")
        # ---- eigen-decompositions per bucket (descending); success flags stay
        # on the device and come back with the rank results.  Buckets above
        # the Jacobi limit run on worker threads with their own streams while
        # the Jacobi buckets (one launch each) run on this one.  The
        # tridiagonal solver gives the eigenvalues first; the vectors of the
        # kept ranks follow the rank rule.
        big = [sz for sz in buckets if sz > self.jacobi_max]
        tri_mode = self.eig_solver == "tridiag"
        if not tri_mode:
            for sz in big:
                if sz not in self._jitter:
                    self._jitter[sz] = _jitter_matrix(sz, self.dev)

        def solve_big(sz):
            idx, T = buckets[sz]
            if tri_mode:
                _pad_bucket(T, s[idx])
                tri = GramTridiag(T)
                W = tri.eigenvalues()
                return W, tri, torch.ones(len(idx), dtype=torch.bool, device=self.dev)
            _condition_bucket(T, s[idx], self._jitter[sz])
            return heevd_batched(T, inplace=True, return_ok="device")

        main = torch.cuda.current_stream(self.dev)
        res = {}
        for sz, (idx, T) in buckets.items():
            if sz <= self.jacobi_max:
                W, V = heev_small(T, orders=_h2d(s[idx].astype(np.int64), self.dev))
                res[sz] = (W, V, finite_rows(W, V))
        res.update(self._on_workers(big, solve_big, main))
        W_all, V_all, oks, tris = {}, {}, [], {}
        for sz in buckets:
            W, V, ok = res[sz]
            W_all[sz] = W.contiguous()
            if isinstance(V, GramTridiag):
                tris[sz] = V
            else:
                V_all[sz] = V.contiguous()
            oks.append(ok)
        soff, _ = _prefix(s)
        rr, okh = self._rank_all(buckets, W_all, s, soff, jj, torch.cat(oks))
        if not okh.all():
            # rare: refill the failed rows from their theta's SVD, rank again
            at = 0
            for sz, (idx, T) in buckets.items():
                bad = np.flatnonzero(~okh[at: at + len(idx)]).tolist()
                at += len(idx)
                if bad:
                    self.eig_fallbacks += len(bad)
                    self._svd_fill(W_all[sz], V_all[sz], bad, idx, X64, xoff, m, n, left)
            rr, _ = self._rank_all(buckets, W_all, s, soff, jj, None)
        Sf64 = self._last_sv
        k = rr["k"].astype(np.int64)
        # eigenvector accessor per gate: V[r, c] = (conj) base[r vs0 + c vs1]
        # (component r of the eigenvector of the c-th largest eigenvalue)
        vptr = np.zeros(G, np.int64)
        vs0 = np.zeros(G, np.int64)
        vs1 = np.zeros(G, np.int64)
        vcj = np.zeros(G, np.int64)
        for sz, (idx, T) in buckets.items():
            if sz in V_all:            # columns of a row-major sz x sz block
                vptr[idx] = V_all[sz].data_ptr() + 16 * sz * sz * np.arange(len(idx))
                vs0[idx], vs1[idx] = sz, 1
        ok_host = ok_ev = None
        if tris:
            # on this stream: the solvers' buffers are marked as used here.
            # The re-check flags come back through pinned memory; the
            # projection and copy descriptors are built while the vectors
            # are computed, and the flags are waited for just before launch.
            vec = {sz: tris[sz].vectors(k[buckets[sz][0]].tolist()) for sz in tris}
            errs = torch.cat([vec[sz][1] for sz in tris])
            okd = (torch.isfinite(errs) & (errs <= ORTH_TOL)).to(torch.uint8)
            ok_host = torch.empty(okd.shape, dtype=torch.uint8, pin_memory=True)
            ok_host.copy_(okd, non_blocking=True)
            ok_ev = torch.cuda.current_stream(self.dev).record_event()
            for sz in tris:
                idx = buckets[sz][0]
                Et = vec[sz][0]
                kpad = Et.shape[1]
                V_all[sz] = Et
                # row j of Et[b] is the conjugate transpose of eigenvector j
                vptr[idx] = Et.data_ptr() + 16 * kpad * sz * np.arange(len(idx))
                vs0[idx], vs1[idx], vcj[idx] = 1, sz, 1
        # ---- projections (FP64) and rebuild into the slots
        sptr = np.zeros(G, np.int64)
        for sz, (idx, T) in buckets.items():
            sptr[idx] = Sf64[sz].data_ptr() + 8 * sz * np.arange(len(idx))
        inv = {sz: torch.where(v > 0, 1.0 / v, torch.zeros_like(v)) for sz, v in Sf64.items()}
        iptr = np.zeros(G, np.int64)
        for sz, (idx, T) in buckets.items():
            iptr[idx] = inv[sz].data_ptr() + 8 * sz * np.arange(len(idx))
        pcount = np.where(left, k * n, m * k)
        poff, ptot = _prefix(pcount)
        P = torch.empty(max(ptot, 1), dtype=C128, device=self.dev)
        pj = np.zeros(G, GEMM_DESC)
        pj["kind"] = _ffi.AQ_GROUPED_GEMM
        pj["C"] = P.data_ptr() + 16 * poff
        # left: Vh_k = U_k^H X : A(i,r) = conj V[r,i]
        pj["A"] = np.where(left, vptr, x64)
        pj["M"] = np.where(left, k, m)
        pj["N"] = np.where(left, n, k)
        pj["K"] = np.where(left, m, n)
        pj["a_s0"] = np.where(left, vs1, n)
        pj["a_s1"] = np.where(left, vs0, 1)
        pj["conj_a"] = np.where(left, 1 - vcj, 0)
        # right: U_k S_k = X V_k : B(r,c) = V[r,c]
        pj["B"] = np.where(left, x64, vptr)
        pj["b_s0"] = np.where(left, n, vs0)
        pj["b_s1"] = np.where(left, 1, vs1)
        pj["conj_b"] = np.where(left, 0, vcj)
        pj["c_s0"] = np.where(left, n, k)
        pj["c_s1"] = 1
        proj = (self._prep_grouped(_ffi.AQ_C128, pj[left], 2 | 4, "proj"),
                self._prep_grouped(_ffi.AQ_C128, pj[~left], 3 | 4, "proj"))
        # new shapes: site b -> [dl, 2, k], site b+1 -> [k, 2, dr]
        cp = np.zeros(2 * G, COPY_DESC)
        A_, B_ = cp[:G], cp[G:]
        A_["dst"] = self._ptrs(jj, bb)
        B_["dst"] = self._ptrs(jj, bb + 1)
        # left: A' = V[:, :k] * S (col scale), B' = P * (1/S) (row scale)
        A_["src"] = np.where(left, vptr, P.data_ptr() + 16 * poff)
        A_["rows"], A_["cols"] = m, k
        A_["s0"] = np.where(left, vs0, k)
        A_["s1"] = np.where(left, vs1, 1)
        A_["conj"] = np.where(left, vcj, 0)
        A_["col_scale"] = np.where(left, sptr, 0)
        B_["src"] = np.where(left, P.data_ptr() + 16 * poff, vptr)
        B_["rows"], B_["cols"] = k, n
        # right: B' = V_k^H : (r,c) = conj V[c, r]
        B_["s0"] = np.where(left, n, vs1)
        B_["s1"] = np.where(left, 1, vs0)
        B_["conj"] = np.where(left, 0, 1 - vcj)
        B_["row_scale"] = np.where(left, iptr, 0)
        cnt = cp["rows"] * cp["cols"]
        pre, total = _prefix(cnt)
        dd, pd = _upload(cp, self.dev), _h2d(pre, self.dev)
        if tris:
            ok_ev.synchronize()
            good = ok_host.numpy().astype(bool)
            at = 0
            for sz in tris:
                idx = buckets[sz][0]
                bad = np.flatnonzero(~good[at: at + len(idx)]).tolist()
                at += len(idx)
                if bad:
                    # vectors that failed the re-check: from their theta's SVD
                    self.eig_fallbacks += len(bad)
                    self._svd_fill_rows(V_all[sz], bad, idx, k, X64, xoff, m, n, left)
        for prep in proj:
            self._launch_grouped(prep)

        def runc(g):
            check(lib().aq_grouped_copy(_vp(dd), _vp(pd), len(cp), total, g, _stream()),
                  "aq_grouped_copy")
        runc(self.tuner.grid("elem", total, None))
        self.dims[jj, bb + 1] = k
        self.logs["mps"].extend(jj.tolist())
        self.logs["bond"].extend(bb.tolist())
        self.logs["k"].extend(k.tolist())
        self.logs["eps_local"].extend(rr["eps_local"].tolist())
        self.logs["entropy"].extend(rr["entropy"].tolist())
        # keep the temporaries alive until the stream has consumed them
        self._keep = (X, X64, buckets, W_all, V_all, tris, self._keep_rank, Sf64, inv, P,
                      dd, pd)

    def _rank_all(self, buckets, W_all, s, soff, jj, ok_dev):
        """Batched rank rule on the float32 singular values of every bucket;
        returns the host results and, when ``ok_dev`` is given, the host copy
        of those per-matrix flags (fetched in the same transfer)."""
        G = len(jj)
        Sflat = torch.empty(int(s.sum()), dtype=torch.float32, device=self.dev)
        Sf64 = {}
        for sz, (idx, T) in buckets.items():
            sv = torch.sqrt(torch.clamp(W_all[sz], min=0.0))       # [B, sz] float64
            Sf64[sz] = sv
            # scatter the first s_g values of each row into Sflat
            lens = s[idx]
            rows = np.repeat(np.arange(len(idx)), lens)
            cols = _ragged_arange(lens)
            dst = np.repeat(soff[idx], lens) + cols
            Sflat[_h2d(dst, self.dev)] = sv[_h2d(rows, self.dev), _h2d(cols, self.dev)].float()
        sp = _h2d((Sflat.data_ptr() + 4 * soff).astype(np.int64), self.dev)
        md = _h2d(s.astype(np.int64), self.dev)
        cd = _h2d(self.cap[jj].astype(np.int64), self.dev)
        nres = G * ctypes.sizeof(RankResult)
        nok = 0 if ok_dev is None else int(ok_dev.numel())
        buf = torch.zeros(nres + nok, dtype=torch.uint8, device=self.dev)
        check(lib().aq_rank_select_batched(_ffi.AQ_C64, _vp(sp), _vp(md), _vp(cd), self.eps, G,
                                           _vp(buf), _stream()), "aq_rank_select_batched")
        if nok:
            buf[nres:] = ok_dev.to(torch.uint8)
        host = buf.cpu().numpy()
        rr = np.frombuffer(host[:nres].tobytes(), dtype=np.dtype(
            [("k", "i4"), ("flags", "i4"), ("eps_local", "f8"), ("entropy", "f8"),
             ("condS", "f8"), ("total", "f8"), ("kept", "f8")]))
        self._last_sv = Sf64
        self._keep_rank = (Sflat, sp, md, cd, buf)
        return rr, (host[nres:] != 0 if nok else None)

    def _on_workers(self, sizes, fn, main):
        """{sz: fn(sz)} for the buckets ``sizes``, each on a worker thread's own
        stream (waiting for ``main``'s work so far) when there are several,
        ``main`` waiting for all of them; else on ``main``."""
        if not sizes:
            return {}
        if self.eig_workers <= 1 or len(sizes) == 1:
            return {sz: fn(sz) for sz in sizes}
        from .adaptive_mps import _thread_stream, _worker_pool
        ready = main.record_event()

        def job(sz):
            st = _thread_stream(self.dev)
            with torch.cuda.stream(st):
                st.wait_event(ready)
                out = fn(sz)
                done = st.record_event()
            return out, done

        pool = _worker_pool(self.eig_workers)
        futs = {sz: pool.submit(job, sz) for sz in sizes}
        out = {}
        for sz, f in futs.items():
            r, done = f.result()
            main.wait_event(done)
            for t in r:                   # produced on a worker stream, read on this one
                if isinstance(t, (torch.Tensor, GramTridiag)):
                    t.record_stream(main)
            out[sz] = r
        return out

    @staticmethod
    def _svd_fill_rows(Et, bad, idx, k, X64, xoff, m, n, left):
        """Rows of a tridiagonal bucket whose vectors failed the orthogonality
        re-check, from the SVD of each one's own theta: Et[b, j, :] is the
        conjugate transpose of the j-th left (X X^H) or right (X^H X)
        singular vector, j < k; the rest of the row block is zero."""
        for b in bad:
            g = int(idx[b])
            mg, ng, kg = int(m[g]), int(n[g]), int(k[g])
            X = X64[int(xoff[g]): int(xoff[g]) + mg * ng].view(mg, ng)
            U, S, Vh, _ = robust_svd(X)
            F = U if left[g] else Vh.conj().transpose(0, 1)
            Et[b].zero_()
            Et[b, :kg, :F.shape[0]] = F[:, :kg].conj().transpose(0, 1)

    @staticmethod
    def _svd_fill(W, V, bad, idx, X64, xoff, m, n, left):
        """Eigenpairs of the bucket rows ``bad`` whose eigensolve failed, from
        the SVD of each one's own theta: X = U S Vh gives (S^2, U) for X X^H
        and (S^2, Vh^H) for X^H X; the rest of the row is zero."""
        for b in bad:
            g = int(idx[b])
            mg, ng = int(m[g]), int(n[g])
            X = X64[int(xoff[g]): int(xoff[g]) + mg * ng].view(mg, ng)
            U, S, Vh, _ = robust_svd(X)
            F = U if left[g] else Vh.conj().transpose(0, 1)
            r = S.numel()
            W[b].zero_()
            V[b].zero_()
            W[b, :r] = S * S
            V[b, :F.shape[0], :r] = F

    # ------------------------------------------------------------ circuits
    def run_hea(self, circuits):
        """circuits[j]: op list of atlas_q.circuits.hea (same layer structure for
        all MPS: 2n 1q ops then CZ on even bonds then odd bonds per layer)."""
        with _gc_paused():
            self._run_steps(self.compile_circuits(circuits))

    @staticmethod
    def compile_circuits(circuits):
        """Split each op list into steps (a block of consecutive 1q ops, then 2q
        ops up to the first one whose bond touches an earlier bond of the
        block), as arrays: per circuit a list of (q [L], U [L, 2, 2], bonds
        [G], U4 [G, 4, 4]).  Runs of 1q ops are sliced from per-circuit
        arrays; only the 2q runs are walked, for the bond-conflict rule."""
        out = []
        z1 = (np.zeros(0, np.int64), np.zeros((0, 2, 2), np.complex128))
        z2 = (np.zeros(0, np.int64), np.zeros((0, 4, 4), np.complex128))
        for ops in circuits:
            L = len(ops)
            if L == 0:
                out.append([])
                continue
            kinds, wh, us = zip(*ops)
            two = np.fromiter(map("2q".__eq__, kinds), bool, L)
            where = np.array(wh, np.int64)
            i1 = np.flatnonzero(~two)
            i2 = np.flatnonzero(two)
            U1 = _gather_gates([us[i] for i in i1], 2)
            U2 = _gather_gates([us[i] for i in i2], 4)
            r1 = np.full(L, -1, np.int64)      # op index -> row in U1 / U2
            r1[i1] = np.arange(len(i1))
            r2 = np.full(L, -1, np.int64)
            r2[i2] = np.arange(len(i2))
            # runs of equal kind: [start, end)
            cut = np.flatnonzero(np.diff(two.astype(np.int8))) + 1
            starts = np.concatenate([[0], cut])
            ends = np.concatenate([cut, [L]])
            steps, pending = [], None          # pending: the 1q block of the next step
            for a, b in zip(starts.tolist(), ends.tolist()):
                if not two[a]:
                    if pending is not None:    # two 1q runs cannot be adjacent
                        steps.append(pending + z2)
                    pending = (where[a:b], U1[r1[a]:r1[a] + (b - a)])
                    continue
                bonds = where[a:b].tolist()
                blk, used = a, set()
                for t, bnd in enumerate(bonds):
                    if bnd in used or bnd - 1 in used or bnd + 1 in used:
                        q1 = pending if pending is not None else z1
                        steps.append(q1 + (where[blk:a + t], U2[r2[blk]:r2[a + t - 1] + 1]))
                        pending, blk, used = None, a + t, set()
                    used.add(bnd)
                q1 = pending if pending is not None else z1
                steps.append(q1 + (where[blk:b], U2[r2[blk]:r2[b - 1] + 1]))
                pending = None
            if pending is not None:
                steps.append(pending + z2)
            out.append(steps)
        return out

    def run_compiled(self, compiled):
        """Apply ``compile_circuits`` output: per step, one grouped 1q launch
        (the block's gates merged per site: applying U1 then U2 equals applying
        U1 @ U2 in the U^T convention) and one two-qubit sub-layer."""
        with _gc_paused():
            self._run_steps(compiled)

    def _run_steps(self, compiled):
        nsteps = max((len(c) for c in compiled), default=0)
        for t in range(nsteps):
            live = [(j, c[t]) for j, c in enumerate(compiled) if t < len(c)]
            j1 = np.concatenate([np.full(len(st[0]), j, np.int64) for j, st in live])
            if len(j1):
                q1 = np.concatenate([st[0] for _, st in live])
                u1 = np.concatenate([st[1] for _, st in live])
                self.apply_1q_arrays(*_merge_per_site(j1, q1, u1, self.nmax))
            j2 = np.concatenate([np.full(len(st[2]), j, np.int64) for j, st in live])
            if len(j2):
                self.apply_2q_arrays(j2, np.concatenate([st[2] for _, st in live]),
                                     np.concatenate([st[3] for _, st in live]))

    # ------------------------------------------------------------ sampling
    def prepare_sampler(self, shots: int):
        """Allocate sweep buffers and the per-site descriptor lists."""
        S = int(shots)
        self.shots = S
        words = (self.n + 63) // 64
        self.u_off, ut = _prefix(S * self.n)
        self.uniforms = torch.empty(max(ut, 1), dtype=torch.float64, device=self.dev)
        self.b_off, bt = _prefix(S * words)
        self.bits = torch.empty(max(bt, 1), dtype=torch.int64, device=self.dev)
        cmax = self.dims.max(axis=1)
        tb = S * 2 * cmax
        self.t_off, tt = _prefix(2 * tb)
        self.tbuf = torch.empty(max(tt, 1), dtype=C64, device=self.dev)
        pc = S * _cdiv(cmax, 32) * 2
        self.p_off, pt = _prefix(pc)
        self.part = torch.empty(max(pt, 1), dtype=torch.float32, device=self.dev)
        self.o_off, ot = _prefix(np.full(self.M, S))
        self.offs = torch.empty(max(ot, 1), dtype=torch.int32, device=self.dev)
        self._site_lists = []
        for i in range(self.nmax):
            js = np.nonzero(self.n > i)[0]
            d = np.zeros(len(js), SAMPLE_DESC)
            for r, j in enumerate(js):
                cl, cr = int(self.dims[j, i]), int(self.dims[j, i + 1])
                cur = self.tbuf.data_ptr() + 8 * (int(self.t_off[j]) + (i & 1) * int(tb[j]))
                prev = self.tbuf.data_ptr() + 8 * (int(self.t_off[j]) + ((i + 1) & 1) * int(tb[j]))
                d[r]["L"] = 0 if i == 0 else prev
                d[r]["off"] = 0 if i == 0 else self.offs.data_ptr() + 4 * int(self.o_off[j])
                d[r]["site"] = self._ptr(j, i)
                d[r]["Tout"] = cur
                d[r]["part"] = self.part.data_ptr() + 4 * int(self.p_off[j])
                d[r]["uniforms"] = self.uniforms.data_ptr() + 8 * int(self.u_off[j])
                d[r]["bits"] = self.bits.data_ptr() + 8 * int(self.b_off[j])
                d[r]["off_next"] = self.offs.data_ptr() + 4 * int(self.o_off[j])
                d[r]["ldL"] = 2 * cl
                d[r]["S"], d[r]["chiL"], d[r]["chiR"] = S, cl, cr
                d[r]["n"], d[r]["site_index"], d[r]["words"] = self.n[j], i, words[j]
            tiles = _cdiv(np.full(len(js), S), 64) * _cdiv(d["chiR"], 64)
            tpre, ttot = _prefix(tiles)
            spre, stot = _prefix(np.full(len(js), S))
            self._site_lists.append((_upload(d, self.dev), _h2d(tpre, self.dev),
                                     _h2d(spre, self.dev), len(js), ttot, stot))
        self._grid_sample = self.tuner.grid("sample", self._site_lists[0][4], None)
        # the descriptors are on the device when this returns, so the sweep may
        # be launched (or captured) on any stream
        torch.cuda.current_stream(self.dev).synchronize()

    def sweep(self):
        """The full sweep for all MPS (launches only; uniforms must be filled)."""
        check(lib().aq_zero_u64(_vp(self.bits), self.bits.numel(), _stream()), "aq_zero_u64")
        for dd, tp, sp, cnt, ttot, stot in self._site_lists:
            check(lib().aq_sample_grouped_site(_vp(dd), _vp(tp), _vp(sp), cnt, ttot, stot,
                                               self._grid_sample, _stream()),
                  "aq_sample_grouped_site")

    def sample_bits(self, j: int) -> torch.Tensor:
        words = (int(self.n[j]) + 63) // 64
        o = int(self.b_off[j])
        return self.bits[o: o + self.shots * words].view(self.shots, words)

    def uniforms_of(self, j: int) -> torch.Tensor:
        o = int(self.u_off[j])
        return self.uniforms[o: o + self.shots * int(self.n[j])].view(self.shots, int(self.n[j]))
