"""Truncated-SVD methods for the two-site update, selected by
``AdaptiveMPS(svd_method=...)``.

All methods produce the reference's outputs (adaptive_mps.py:249-290): the
rank rule of truncation.py on the singular values, S absorbed into the left
core, the right core with orthonormal rows, statistics.  They differ in how
the leading singular triplets of X [m = 2 chiL, n = 2 chiR] are found:

* "gram"    exact to FP64: G = X X^H (or X^H X, the smaller side) by the FP64
            MFMA GEMM, FP64 Hermitian eigensolver, Vh_k = S_k^-1 U_k^H X in FP64.
* "rsvd"    randomized range finder of width l = k_guess + p with q power
            steps (complex GEMMs on MFMA, BF16x3 optional), orthonormal basis
            Q, projected core B = Q^H X, FP64 core factorization; certified by
            the FP64 tail check  ||X||_F^2 - ||B||_F^2 <= tail_tol * eps^2 *
            ||X||_F^2, else the exact "gram" method runs (the fallback).
* "learned" "rsvd" whose width comes from the rank-predictor MLP (features of
            a cheap sketch); wide predictions go straight to "gram".

For complex128 policies the plain SVD is used (the Gram form squares the
condition number, which the complex128 contract of 1e-12 per amplitude
cannot absorb for singular values below ~1e-8 sigma_max).
"""
from __future__ import annotations

import math
from typing import Callable, Optional

import numpy as np
import torch

from . import kernels as K
from . import roofline as RL
from .eig import TopEig
from .linalg_robust import robust_svd
from .truncation import choose_rank_from_sigma


class FactorStats:
    """Counters of which factorization served each update (thread-safe)."""

    def __init__(self):
        import threading
        self.counts = {"gram": 0, "rsvd": 0, "rsvd_fallback": 0, "plain": 0}
        self.predicted = []   # (predicted width, kept rank)
        self.rejects = {}     # why the randomized factorization fell back
        self.tails = []       # uncaptured weight / ||X||^2 per randomized attempt
        self.orth = []        # FP64 orthogonality re-check of the tridiagonal path
        self.lock = threading.Lock()

    def bump(self, key):
        with self.lock:
            self.counts[key] = self.counts.get(key, 0) + 1

    def reject(self, reason):
        with self.lock:
            self.rejects[reason] = self.rejects.get(reason, 0) + 1


RSVD_ATTEMPTS = 2      # a rejected sketch is retried once at twice the width
GROUPED_GRAM = True    # factor a sub-layer's Gram cores with grouped launches
RETRY_REASONS = ("tail", "uncertified", "rank_at_width", "orth")

_DEFAULT = {}
_INIT_LOCK = __import__("threading").Lock()


def _factor_stats(mps):
    with _INIT_LOCK:
        st = getattr(mps, "factor_stats", None)
        if st is None:
            st = mps.factor_stats = FactorStats()
        return st


def default_predictor(device):
    """The shipped trained rank predictor (rank_mlp_weights.npz), loaded once per device."""
    key = str(device)
    if key not in _DEFAULT:
        from .rank_mlp import DEFAULT_WEIGHTS, RankPredictor
        _DEFAULT[key] = RankPredictor.load(DEFAULT_WEIGHTS, device) if DEFAULT_WEIGHTS.exists() \
            else None
    return _DEFAULT[key]


def _gram_matrix(X: torch.Tensor, out: torch.Tensor = None):
    """FP64 Gram matrix of the smaller side of X: (G, left, X64) with left
    True for G = X X^H [m, m], False for G = X^H X [n, n]; written into
    ``out`` (contiguous complex128 of that order) when given."""
    m, n = X.shape
    X64 = X.to(torch.complex128)
    if m <= n:
        with RL.timed("gram_fp64_cgemm", 8.0 * m * m * n, "fp64_mfma"):
            return K.cgemm(X64, X64.T, conj_b=True, out=out), True, X64
    with RL.timed("gram_fp64_cgemm", 8.0 * n * n * m, "fp64_mfma"):
        return K.cgemm(X64.T, X64, conj_a=True, out=out), False, X64


def _finish_from_gram(mps, X, vec, left, S, X64, k_dtype, cap, budget_ok, cl, cr, eps):
    """Rank rule on S (in the policy's real dtype) and the two cores; vec(k)
    gives the Gram matrix's eigenvectors of the k largest eigenvalues as
    columns (left singular vectors of X when ``left``, else right ones)."""
    S_rule = S.to(torch.float32) if k_dtype == torch.complex64 else S
    k, eps_loc, entropy, condS = choose_rank_from_sigma(S_rule, eps, cap, budget_ok)
    local = k_dtype
    if math.isfinite(condS) and condS > mps.dtype_policy.promote_if_cond_gt:
        k, eps_loc, entropy, condS = choose_rank_from_sigma(S, eps, cap, budget_ok)
        local = torch.complex128
    Sk = S[:k]
    inv = torch.where(Sk > 0, 1.0 / Sk, torch.zeros_like(Sk))
    if left:
        Uk = vec(k)
        with RL.timed("proj_fp64_cgemm", 8.0 * k * X64.shape[0] * X64.shape[1], "fp64_mfma"):
            Vh = K.cgemm(Uk.T, X64, conj_a=True)             # U_k^H X  [k, n]
        Vh = Vh * inv[:, None]
    else:
        Vk = vec(k)
        Uk = K.cgemm(X64, Vk) * inv[None, :]                   # X V_k S^-1 [m, k]
        Vh = Vk.T.conj()
    Uk = Uk.to(local)
    Vh = Vh.to(local).contiguous()
    A_new, B_new = K.rebuild(Uk, Sk.to(torch.float32 if local == torch.complex64 else torch.float64),
                             Vh, k, cl, cr)
    return A_new, B_new, k, eps_loc, entropy, condS, local


def plain_factor(mps, X, cap, budget_ok, local, cl, cr):
    """The plain method's factorization of theta X (adaptive_mps.py:249-290):
    SVD cascade, rank rule, and the complex128 recompute when condS exceeds
    the policy's promotion threshold.  Returns the update's result tuple."""
    U, S, Vh, driver = robust_svd(X)
    k, e, s, c = choose_rank_from_sigma(S, mps.eps_bond, cap, budget_ok)
    if math.isfinite(c) and c > mps.dtype_policy.promote_if_cond_gt:
        U, S, Vh, driver = robust_svd(X.to(torch.complex128))
        k, e, s, c = choose_rank_from_sigma(S, mps.eps_bond, cap, budget_ok)
        local = torch.complex128
    A_new, B_new = K.rebuild(U, S, Vh, k, cl, cr)
    return A_new, B_new, k, e, s, c, driver, local


def two_site_update(mps, i: int, A: torch.Tensor, B: torch.Tensor, gate: np.ndarray, cap: int,
                    budget_ok: Callable[[int], bool], local: torch.dtype):
    """Returns (A_new, B_new, k, eps_local, entropy, condS, driver, dtype)."""
    r = prepare_update(mps, i, A, B, gate, cap, budget_ok, local)
    return r if not isinstance(r, GramJob) else finish_gram(mps, [r])[0]


class GramJob:
    """An update waiting for the exact Gram factorization."""

    def __init__(self, X, cap, budget_ok, local, cl, cr, feats):
        self.X, self.cap, self.budget_ok, self.local = X, cap, budget_ok, local
        self.cl, self.cr, self.feats = cl, cr, feats


def prepare_update(mps, i: int, A: torch.Tensor, B: torch.Tensor, gate: np.ndarray, cap: int,
                   budget_ok: Callable[[int], bool], local: torch.dtype):
    """Theta and, for the randomized methods, the certified randomized SVD.
    Returns the update's result tuple, or a GramJob for ``finish_gram``."""
    method = mps.svd_method
    cl, cr = A.shape[0], B.shape[2]
    st = _factor_stats(mps)
    prec = mps.theta_precision if local == torch.complex64 else "native"
    cm = A.shape[2]
    # operands and output prepared outside the timed launch (host work inside
    # the event pair would be counted; with several worker streams the pair
    # still spans the other streams' concurrent kernels)
    A, B = A.contiguous(), B.contiguous()
    X = torch.empty((2 * cl, 2 * cr), dtype=A.dtype, device=A.device)
    with RL.timed(f"theta_{prec}", 32.0 * cl * cm * cr + 128.0 * cl * cr,
                  "bf16_mfma" if prec == "bf16x3" else "fp32_mfma"):
        K.theta(A, B, gate, prec=prec, out=X)
    if local == torch.complex128 or method == "plain_in_fast":
        st.bump("plain")
        return plain_factor(mps, X, cap, budget_ok, local, cl, cr)
    feats = None
    if (method == "learned" and getattr(mps, "harvest", None) is None and cm >= cap
            and getattr(mps, "_at_cap", {}).get(i, False) and 2 * cap >= min(X.shape)):
        # the last update of this bond kept the cap (its rule wanted at least
        # that many values) and a randomized sketch as wide as the cap would
        # cover half the spectrum or more: the exact Gram method serves the
        # update whatever the network predicts, so the sketch and the
        # prediction are not paid for
        st.reject("at_cap")
        return GramJob(X, cap, budget_ok, local, cl, cr, None)
    pred = getattr(mps, "rank_predictor", None)
    if pred is None and method == "learned":
        with _INIT_LOCK:
            pred = mps.rank_predictor = default_predictor(X.device)
    if pred is not None and (method == "learned" or getattr(mps, "harvest", None) is not None):
        feats = pred.features(mps, i, X, gate, cap)
    if method in ("rsvd", "learned"):
        from . import rsvd
        if method == "learned" and pred is not None and pred.trained:
            k_guess = pred.predict_width(feats, cap, min(X.shape))
        else:
            k_guess = min(cap, 2 * cm)          # operator-Schmidt growth bound, no model
        out = None
        for attempt in range(RSVD_ATTEMPTS):
            try:
                out, why = rsvd.try_rsvd(mps, i, X, cap, budget_ok, local, cl, cr, st, k_guess)
            except RuntimeError:        # core eigensolver did not converge
                st.reject("eigensolver")
                break
            if out is not None or why not in RETRY_REASONS:
                break
            k_guess *= 2                # a wider sketch, fresh random columns
        if out is not None:
            _harvest(mps, feats, out[2])
            return out
        st.bump("rsvd_fallback")
    elif method != "gram":
        raise ValueError(f"unknown svd_method {method!r}")
    return GramJob(X, cap, budget_ok, local, cl, cr, feats)


def _harvest(mps, feats, k):
    if feats is not None and getattr(mps, "harvest", None) is not None:
        mps.harvest.append((feats.cpu().numpy(), k))


def _rank_inputs(S32: torch.Tensor, caps):
    """Device inputs of the batched rank rule over the rows of S32 [B, m]
    (row pointers, orders, caps) and its output buffer; the uploads
    synchronise the current stream, so they are made before long work is
    queued on it."""
    from ._ffi import RankResult
    import ctypes
    B, m = S32.shape
    dev = S32.device
    ptrs = torch.tensor([S32[b].data_ptr() for b in range(B)], dtype=torch.int64).to(dev)
    mm = torch.full((B,), m, dtype=torch.int64).to(dev)
    cp = torch.tensor([int(c) for c in caps], dtype=torch.int64).to(dev)
    out = torch.empty(B * ctypes.sizeof(RankResult), dtype=torch.uint8, device=dev)
    return S32, ptrs, mm, cp, out


def _rank_launch(rin, eps: float):
    """Queue the rank rule (aq_rank_select semantics) on every row and the copy
    of its results to pinned host memory; returns (host buffer, event)."""
    from ._ffi import check, lib
    S32, ptrs, mm, cp, out = rin
    check(lib().aq_rank_select_batched(0, K._ptr(ptrs), K._ptr(mm), K._ptr(cp), float(eps),
                                       S32.shape[0], K._ptr(out), K._stream(S32)),
          "aq_rank_select_batched")
    host = torch.empty(out.numel(), dtype=torch.uint8, pin_memory=True)
    host.copy_(out, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record()
    return host, ev


def _rank_read(host: torch.Tensor, ev, B: int):
    from ._ffi import RankResult
    import ctypes
    ev.synchronize()
    raw = host.numpy().tobytes()
    sz = ctypes.sizeof(RankResult)
    return [RankResult.from_buffer_copy(raw[b * sz:(b + 1) * sz]) for b in range(B)]


def _rank_batched(S32: torch.Tensor, caps, eps: float):
    """Rank rule (aq_rank_select semantics) on every row of S32 [B, m] in one
    launch; one device->host copy of the B results."""
    host, ev = _rank_launch(_rank_inputs(S32, caps), eps)
    return _rank_read(host, ev, S32.shape[0])


def _gram_group(mps, jobs, st):
    """Exact Gram factorization of a group of complex64 updates with m <= n
    and one order m, every step one launch for the whole group: the FP64
    Gram matrices (lower-triangle tiles only: the eigensolver reads that
    triangle) in one grouped GEMM, one batched rocSOLVER eigensolve, the rank
    rule for all, the projections Vh_k = U_k^H X in one grouped GEMM that
    reads the eigensolver's output rows directly, and the cores (A' = U_k
    S_k, B' = S_k^-1 Vh_k, rounded to complex64) in one grouped copy.
    Returns per-job results, or None for a job that needs the per-job path
    (eigensolver failure, precision promotion)."""
    import ctypes
    from ._ffi import AQ_GROUPED_HERK, COPY_DESC, check, lib
    from .eig import heevd_batched
    c128 = torch.complex128
    B = len(jobs)
    m = jobs[0].X.shape[0]
    dev = jobs[0].X.device
    keep = []
    X64 = [j.X.to(c128) for j in jobs]
    Gb = torch.zeros((B, m, m), dtype=c128, device=dev)
    gd = np.stack([K.gemm_desc(X64[b], X64[b].T, Gb[b], conj_b=True, kind=AQ_GROUPED_HERK)
                   for b in range(B)])
    work = float(sum(4.0 * m * m * x.shape[1] for x in X64))
    with RL.timed("gram_fp64_herk_grouped", work, "fp64_mfma"):
        K.grouped_gemm_descs(gd, c128, dev, keep)
    # algorithmic flops of a complex Hermitian eigen-decomposition with all
    # vectors: tridiagonalisation 16/3 n^3, tridiagonal solve ~4/3 n^3,
    # back-transformation 8 n^3
    with RL.timed("eig_zheevd_batched", B * (16.0 / 3.0 + 4.0 / 3.0 + 8.0) * m ** 3, "fp64_mfma"):
        lam, _, ok = heevd_batched(Gb, inplace=True, return_ok=True)
    S64 = torch.sqrt(torch.clamp(lam, min=0.0)).contiguous()          # [B, m] descending
    ranks = _rank_batched(S64.to(torch.float32).contiguous(), [j.cap for j in jobs],
                          mps.eps_bond)
    out = [None] * B
    good = []
    for b, (j, r) in enumerate(zip(jobs, ranks)):
        if not ok[b]:
            continue
        if r.flags & 1:
            k, e, sv, c = 1, 0.0, 0.0, float("inf")
        else:
            k, e, sv, c = int(r.k), float(r.eps_local), float(r.entropy), float(r.condS)
        if math.isfinite(c) and c > mps.dtype_policy.promote_if_cond_gt:
            continue                                   # per-job path promotes
        good.append((b, k, e, sv, c))
    if not good:
        return out
    inv = torch.where(S64 > 0, 1.0 / S64, torch.zeros_like(S64))
    ks = np.array([g[1] for g in good], np.int64)
    ns = np.array([X64[g[0]].shape[1] for g in good], np.int64)
    off = np.zeros(len(good), np.int64)
    off[1:] = np.cumsum((ks * ns)[:-1])
    P = torch.empty(int((ks * ns).sum()), dtype=c128, device=dev)
    pd = []
    for (b, k, _, _, _), o, n in zip(good, off, ns):
        Vh = P[o:o + k * n].view(k, n)
        # U_k^H row i = conj(v_i) = row m-1-i of the eigensolver output
        pd.append(K.gemm_desc(Gb[b], X64[b], Vh, M=k, K=m, N=n,
                              a_ptr=Gb[b, m - 1].data_ptr(), a_strides=(-m, 1)))
    work = float(sum(8.0 * k * m * n for k, n in zip(ks, ns)))
    with RL.timed("proj_fp64_grouped", work, "fp64_mfma"):
        K.grouped_gemm_descs(np.stack(pd), c128, dev, keep)
    cps = np.zeros(2 * len(good), COPY_DESC)
    res = []
    for q, ((b, k, e, sv, c), o, n) in enumerate(zip(good, off, ns)):
        j = jobs[b]
        A_new = torch.empty((j.cl, 2, k), dtype=torch.complex64, device=dev)
        B_new = torch.empty((k, 2, j.cr), dtype=torch.complex64, device=dev)
        a, bb = cps[2 * q], cps[2 * q + 1]
        # A'[r, c] = U[r, c] S[c], U[r, c] = conj(output[m-1-c, r])
        a["src"], a["dst"] = Gb[b, m - 1].data_ptr(), A_new.data_ptr()
        a["col_scale"] = S64[b].data_ptr()
        a["rows"], a["cols"], a["s0"], a["s1"], a["conj"] = m, k, 1, -m, 1
        # B'[r, c] = Vh[r, c] / S[r]
        bb["src"], bb["dst"] = P[o:].data_ptr(), B_new.data_ptr()
        bb["row_scale"] = inv[b].data_ptr()
        bb["rows"], bb["cols"], bb["s0"], bb["s1"] = k, n, n, 1
        res.append((b, (A_new, B_new, k, e, sv, c, "gram_fp64", j.local)))
    cnt = cps["rows"] * cps["cols"]
    pre = np.zeros(len(cnt), np.int64)
    pre[1:] = np.cumsum(cnt[:-1])
    dd = torch.from_numpy(cps.view(np.uint8).copy()).to(dev)
    pp = torch.from_numpy(pre).to(dev)
    total = int(cnt.sum())
    check(lib().aq_grouped_copy(K._ptr(dd), K._ptr(pp), len(cps), total,
                                int(min(-(-total // 256), 65536)), K._stream(Gb)),
          "aq_grouped_copy")
    keep += [dd, pp, P, inv, S64, Gb, X64]
    # the temporaries are released in stream order (one stream), so the
    # caching allocator cannot hand them out before the launches above ran
    for b, r in res:
        out[b] = r
    return out


EIG_SOLVER = "tridiag"   # "tridiag": the library's tridiagonal top-k solver; "heevd": rocSOLVER


# Shares of a sub-layer's Gram batch factored as separate chunks on their own
# streams, smallest first, so that a finished chunk's vectors, projection and
# cores (FP64 MFMA and latency-bound work) could overlap the reductions still
# running (HBM-bound).  Off by default: measured at chi = 1024 (64 matrices of
# order 2048) every split was slower than one batch whose reduction runs as
# two interleaved stream groups (123.8 updates/s; 0.375/0.625: 111.6;
# 0.25/0.3125/0.4375: 108.9; four chunks: 103.1), with 4, 8 or 16 hardware
# queues alike: a chunk's reduction on one stream loses the two-group
# interleave of its per-column launch chain, and the chunks' reductions end
# together, so little post-processing overlaps them.  Batches below
# GRAM_CHUNK_MIN are always one chunk.
GRAM_CHUNKS = ()
GRAM_CHUNK_MIN = 24
_CHUNK_STREAMS = {}


def _chunk_stream(dev, i: int):
    """The i-th chunk stream of a device (default priority, created once)."""
    d = torch.device(dev)
    key = (d.index if d.index is not None else torch.cuda.current_device(), i)
    s = _CHUNK_STREAMS.get(key)
    if s is None:
        s = _CHUNK_STREAMS[key] = torch.cuda.Stream(device=d)
    return s


def _chunk_sizes(B: int):
    if B < GRAM_CHUNK_MIN or len(GRAM_CHUNKS) < 2:
        return [B]
    sizes = [max(1, int(round(B * f))) for f in GRAM_CHUNKS[:-1]]
    sizes.append(B - sum(sizes))
    return [c for c in sizes if c > 0]


def _drain(gen):
    try:
        while True:
            next(gen)
    except StopIteration as stop:
        return stop.value


def _gram_group_tridiag(mps, jobs, st):
    """``_gram_group`` with the eigen-decomposition through the tridiagonal
    form: one batched tridiagonalisation, all eigenvalues by bisection, the
    rank rule on them, then eigenvectors only for the kept ranks (twisted
    inverse iteration, MFMA back-transformation, FP64 orthogonality
    re-check); a matrix whose vectors fail the re-check takes the per-job
    path.  The projection reads the back-transformed rows, which are the rows
    of U_k^H.  A batch of GRAM_CHUNK_MIN or more is split into chunks
    (GRAM_CHUNKS) on chunk streams: every chunk's Gram matrices, reduction,
    bisection and rank rule are queued first, then the chunks are finished in
    order of size, so the host waits for one chunk's ranks while the others
    still reduce.  Results are those of one batch: every matrix goes through
    the same kernels either way."""
    B = len(jobs)
    dev = jobs[0].X.device
    sizes = _chunk_sizes(B)
    red = RL.Span("eig_zhetrd_batched", "fp64_mfma")
    if len(sizes) == 1:
        out = _drain(_tridiag_chunk(mps, jobs, st, red, None))
        red.close()
        return out
    main = torch.cuda.current_stream(dev)
    parts = []
    o = 0
    for c, size in enumerate(sizes):
        s = _chunk_stream(dev, c)
        s.wait_stream(main)
        parts.append((s, o, size, _tridiag_chunk(mps, jobs[o:o + size], st, red, 1)))
        o += size
    for s, _, _, gen in parts:
        with torch.cuda.stream(s):
            next(gen)
    out = [None] * B
    for s, o, size, gen in parts:
        with torch.cuda.stream(s):
            out[o:o + size] = _drain(gen)
    for s, _, _, _ in parts:
        main.wait_stream(s)
    for r in out:
        if r is not None:            # produced on a chunk stream, used on the caller's
            r[0].record_stream(main)
            r[1].record_stream(main)
    red.close()
    return out


def _tridiag_chunk(mps, jobs, st, red, groups):
    """One chunk of ``_gram_group_tridiag`` on the current stream, as a
    generator: it queues the Gram matrices, the reduction (timed as a part of
    ``red``), bisection and the rank rule, yields; reads the ranks, queues the
    kept vectors and their re-check, yields; reads the re-check and queues the
    projection and the cores; returns the per-job results (None: per-job
    path)."""
    from ._ffi import AQ_GROUPED_HERK, COPY_DESC, check, lib
    from .eig import ORTH_TOL, GramTridiag, reduction_bytes
    c128 = torch.complex128
    B = len(jobs)
    m = jobs[0].X.shape[0]
    dev = jobs[0].X.device
    keep = []
    X64 = [j.X.to(c128) for j in jobs]
    Gb = torch.zeros((B, m, m), dtype=c128, device=dev)
    S32 = torch.empty((B, m), dtype=torch.float32, device=dev)
    rin = _rank_inputs(S32, [j.cap for j in jobs])
    gd = np.stack([K.gemm_desc(X64[b], X64[b].T, Gb[b], conj_b=True, kind=AQ_GROUPED_HERK)
                   for b in range(B)])
    work = float(sum(4.0 * m * m * x.shape[1] for x in X64))
    with RL.timed("gram_fp64_herk_grouped", work, "fp64_mfma"):
        K.grouped_gemm_descs(gd, c128, dev, keep)
    # algorithmic flops: complex Hermitian tridiagonalisation 16/3 n^3 (and the
    # one-stage method's bytes, which bound it); the Sturm counts of
    # bisection ~53 steps of 4 flops per entry and eigenvalue
    with red.part(B * 16.0 / 3.0 * m ** 3, nbytes=B * reduction_bytes(m)):
        tri = GramTridiag(Gb, groups=groups)
    with RL.timed("eig_bisection", B * 53.0 * 4.0 * m * m, "fp64_valu"):
        lam = tri.eigenvalues()
    S64 = torch.sqrt(torch.clamp(lam, min=0.0)).contiguous()          # [B, m] descending
    S32.copy_(S64)
    rank_host, rank_ev = _rank_launch(rin, mps.eps_bond)
    yield
    ranks = _rank_read(rank_host, rank_ev, B)
    out = [None] * B
    good = []
    for b, (j, r) in enumerate(zip(jobs, ranks)):
        if r.flags & 1:
            k, e, sv, c = 1, 0.0, 0.0, float("inf")
        else:
            k, e, sv, c = int(r.k), float(r.eps_local), float(r.entropy), float(r.condS)
        if math.isfinite(c) and c > mps.dtype_policy.promote_if_cond_gt:
            continue                                   # per-job path promotes
        good.append((b, k, e, sv, c))
    if not good:
        return out
    kv = [0] * B
    for b, k, _, _, _ in good:
        kv[b] = k
    kmax = max(kv)
    with RL.timed("eig_vectors_backtransform", B * 8.0 * m * m * kmax, "fp64_mfma"):
        Et, err = tri.vectors(kv)
    err_host = torch.empty(B, dtype=torch.float64, pin_memory=True)
    err_host.copy_(err, non_blocking=True)
    err_ev = torch.cuda.Event()
    err_ev.record()
    yield
    err_ev.synchronize()
    errs = err_host.tolist()
    st.orth.extend(errs[g[0]] for g in good)
    good = [g for g in good if errs[g[0]] <= ORTH_TOL]
    if not good:
        return out
    inv = torch.where(S64 > 0, 1.0 / S64, torch.zeros_like(S64))
    ks = np.array([g[1] for g in good], np.int64)
    ns = np.array([X64[g[0]].shape[1] for g in good], np.int64)
    off = np.zeros(len(good), np.int64)
    off[1:] = np.cumsum((ks * ns)[:-1])
    P = torch.empty(int((ks * ns).sum()), dtype=c128, device=dev)
    pd = []
    for (b, k, _, _, _), o, n in zip(good, off, ns):
        Vh = P[o:o + k * n].view(k, n)
        pd.append(K.gemm_desc(Et[b], X64[b], Vh, M=k, K=m, N=n,
                              a_ptr=Et[b].data_ptr(), a_strides=(m, 1)))
    work = float(sum(8.0 * k * m * n for k, n in zip(ks, ns)))
    with RL.timed("proj_fp64_grouped", work, "fp64_mfma"):
        K.grouped_gemm_descs(np.stack(pd), c128, dev, keep)
    cps = np.zeros(2 * len(good), COPY_DESC)
    res = []
    for q, ((b, k, e, sv, c), o, n) in enumerate(zip(good, off, ns)):
        j = jobs[b]
        A_new = torch.empty((j.cl, 2, k), dtype=torch.complex64, device=dev)
        B_new = torch.empty((k, 2, j.cr), dtype=torch.complex64, device=dev)
        a, bb = cps[2 * q], cps[2 * q + 1]
        # A'[r, c] = U[r, c] S[c], U[r, c] = conj(Et[c, r])
        a["src"], a["dst"] = Et[b].data_ptr(), A_new.data_ptr()
        a["col_scale"] = S64[b].data_ptr()
        a["rows"], a["cols"], a["s0"], a["s1"], a["conj"] = m, k, 1, m, 1
        # B'[r, c] = Vh[r, c] / S[r]
        bb["src"], bb["dst"] = P[o:].data_ptr(), B_new.data_ptr()
        bb["row_scale"] = inv[b].data_ptr()
        bb["rows"], bb["cols"], bb["s0"], bb["s1"] = k, n, n, 1
        res.append((b, (A_new, B_new, k, e, sv, c, "gram_fp64", j.local)))
    cnt = cps["rows"] * cps["cols"]
    pre = np.zeros(len(cnt), np.int64)
    pre[1:] = np.cumsum(cnt[:-1])
    dd = torch.from_numpy(cps.view(np.uint8).copy()).to(dev)
    pp = torch.from_numpy(pre).to(dev)
    total = int(cnt.sum())
    check(lib().aq_grouped_copy(K._ptr(dd), K._ptr(pp), len(cps), total,
                                int(min(-(-total // 256), 65536)), K._stream(Gb)),
          "aq_grouped_copy")
    # the temporaries (held in keep until here) are released in the order of
    # this chunk's stream, so the caching allocator cannot hand them out before
    # the launches above ran
    keep += [dd, pp, P, inv, S64, Gb, X64, Et, tri, rin]
    for b, r in res:
        out[b] = r
    return out


def finish_gram(mps, jobs, workers: int = 1):
    """Exact Gram factorization of several updates.  Complex64 updates of one
    order above 64 with m <= n are factored together (``_gram_group_tridiag``:
    one launch per step for the whole group); every other update gets its own
    Gram matrix and ``TopEig`` (the Jacobi kernel up to order 64, else the
    tridiagonal path with the kept vectors only).  If the eigensolver fails
    (vectors failing the orthogonality re-check and rocSOLVER's jittered
    retries after them), the update takes the plain method's SVD cascade,
    precision promotion included.  With ``workers`` > 1 the updates factored
    one by one run concurrently on that many host threads, each with its own
    HIP stream (an order above 64 is a chain of small per-column launches,
    which overlap across streams)."""
    from .eig import JACOBI_MAX
    st = _factor_stats(mps)
    orders = {}
    for idx, j in enumerate(jobs):
        orders.setdefault(min(j.X.shape), []).append(idx)
    done = {}
    for n, idxs in orders.items():
        grouped = [i for i in idxs if jobs[i].X.shape[0] <= jobs[i].X.shape[1]
                   and jobs[i].X.dtype == torch.complex64]
        if (n > JACOBI_MAX and len(grouped) > 1 and mps.budget_global_mb is None
                and GROUPED_GRAM):
            group = _gram_group_tridiag if EIG_SOLVER == "tridiag" else _gram_group
            res = group(mps, [jobs[i] for i in grouped], st)
            for i, r in zip(grouped, res):
                if r is not None:
                    done[i] = r
    def one(j):
        try:
            G, left, X64 = _gram_matrix(j.X)
            te = TopEig(G)
            S = torch.sqrt(torch.clamp(te.values, min=0.0))
            A_new, B_new, k, e, s, c, dt = _finish_from_gram(
                mps, j.X, te.vectors, left, S, X64, j.local, j.cap, j.budget_ok, j.cl, j.cr,
                mps.eps_bond)
        except RuntimeError:
            # the FP64 eigensolver failed: the plain method's SVD cascade
            res = plain_factor(mps, j.X, j.cap, j.budget_ok, j.local, j.cl, j.cr)
            st.bump("plain")
            return res
        st.bump("gram")
        return (A_new, B_new, k, e, s, c, "gram_fp64", dt)

    rest = [idx for idx in range(len(jobs)) if idx not in done]
    if workers > 1 and len(rest) > 1:
        from .adaptive_mps import _thread_stream, _worker_pool
        main = torch.cuda.current_stream(jobs[0].X.device)
        ready = main.record_event()
        dev = jobs[0].X.device

        def job(idx):
            stw = _thread_stream(dev)
            with torch.cuda.stream(stw):
                stw.wait_event(ready)
                res = one(jobs[idx])
                fin = stw.record_event()
            return res, fin

        pool = _worker_pool(workers)
        futs = {idx: pool.submit(job, idx) for idx in rest}
        for idx, f in futs.items():
            res, fin = f.result()
            main.wait_event(fin)
            for t in res[0:2]:          # produced on a worker stream, used on this one
                t.record_stream(main)
            done[idx] = res
    else:
        for idx in rest:
            done[idx] = one(jobs[idx])
    outs = []
    for idx, j in enumerate(jobs):
        if idx not in rest:             # factored by the grouped path
            st.bump("gram")
        outs.append(done[idx])
        _harvest(mps, j.feats, done[idx][2])
    return outs
