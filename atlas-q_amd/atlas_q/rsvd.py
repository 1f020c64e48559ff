"""Randomized truncated SVD of the theta matrix, certified in FP64.

For X [m, n] (complex64) and a width l = k_guess + p, p = max(16, k_guess):
  Y = X Omega                      range finder (MFMA GEMM, BF16x3 optional)
  q x { Z = X^H Q ; Y = X Z' }     optional power steps (q = POWER_STEPS),
                                   re-orthonormalised
  Q = orth(Y)                      shifted Cholesky-QR, twice, in FP64
                                   (the library's aq_cholqr)
  B = Q^H X        [l, n]          FP64 MFMA GEMM
  B B^H = U_B diag(lam) U_B^H      FP64 Hermitian core (up to order 64 the
                                   one-sided Jacobi SVD of its Cholesky
                                   factor, else the tridiagonal path; only the
                                   kept k vectors are formed)
  tail = ||X||_F^2 - sum(lam)      FP64: weight outside range(Q)
The captured eigenvalues bound the true squared singular values:
sigma_i^2 >= lam_i (Cauchy interlacing for the projection Q^H X) and
sum_{i<=j} sigma_i^2 <= E_j + tail with E_j = sum_{i<=j} lam_i, while
||X||_F^2 is exact.  The result is accepted only when
(i)  tail <= FID_TOL * ||X||_F^2: the kept weight of the rank-k result is then
     within FID_TOL of the exact truncation's (sum_{i<k} (sigma_i^2 - lam_i)
     <= tail), i.e. the truncation fidelity agrees to 1e-7 per update; and
(ii) the rank is certified (``certify_rank``): the reference's float32 rule,
     evaluated in its own arithmetic, returns the same rank for every
     spectrum consistent with the bounds, and the rank kernel returned it on
     the captured spectrum.
Otherwise the caller falls back to the exact Gram factorization.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import kernels as K
from . import roofline as RL
from .eig import TopEig
from .truncation import choose_rank_from_sigma

FID_TOL = 1e-7         # admissible uncaptured weight, of ||X||_F^2
POWER_STEPS = 0        # power iterations; 100% oversampling captures decaying spectra
WIDE_FRACTION = 0.5    # widths beyond this share of min(m, n) go to the Gram path


def rule_rank_f32(S: np.ndarray, eps: float) -> int:
    """k_tol of the reference's rule (truncation.py:52-61) in its float32
    arithmetic: float32 squares, cumulative sums accumulated in double and
    rounded to float32, threshold fl32(fl32(1 - eps^2) * total), left search."""
    S = np.asarray(S, np.float32)
    E = np.cumsum((S * S).astype(np.float64)).astype(np.float32)
    thresh = np.float32(np.float32(1.0 - eps * eps) * E[-1])
    return min(int(np.searchsorted(E, thresh, side="left")) + 1, len(S))


def _rule_weights(x: np.ndarray) -> np.ndarray:
    """Per-value weights as the rule forms them: fl32(fl32(sqrt(x))^2)."""
    s = np.sqrt(np.maximum(x, 0.0)).astype(np.float32)
    return (s * s).astype(np.float64)


def _threshold(total_double: float, eps: float) -> np.float32:
    return np.float32(np.float32(1.0 - eps * eps) * np.float32(total_double))


def certify_rank(lam: np.ndarray, tail: float, eps: float, cap: int, total: float = None):
    """Rank the float32 rule returns for every spectrum consistent with the
    captured bounds, or None if it is not pinned.

    The true squared singular values satisfy sigma_i^2 >= lam_i and
    sum_i (sigma_i^2 - lam_i) + sum_{i >= l} sigma_i^2 = tail (up to the FP64
    error of lam, added to both sides).  In the rule's arithmetic a value whose
    interval [lam_i - d, lam_i + tail + d] rounds to one float32 weight has
    that weight exactly; any other value's weight exceeds the lower one by at
    most its share of the tail plus three float32 ulps.  This bounds every
    float32 partial sum; the float32 total is the exact weight of the stable
    values plus the remaining exact weight (||X||^2 minus theirs, known to
    the tail) within two float32 roundings, which bounds the threshold; the rank
    is certified when the earliest and the latest admissible crossings agree
    after the cap.  The rule is monotone in the partial sums and in the
    threshold, so every admissible spectrum then yields that rank."""
    lam = np.asarray(lam, np.float64)
    if total is None:
        total = float(lam.sum() + tail)
    d = 1e-13 * total                                  # FP64 error of lam and of sums
    tail_e = tail + d
    w_lo = _rule_weights(lam - d)
    w_hi = _rule_weights(lam + tail_e)
    unstable = w_hi != w_lo
    slack = np.where(unstable, 3.0 * np.spacing(w_hi.astype(np.float32)).astype(np.float64), 0.0)
    L = np.cumsum(w_lo)
    U = L + tail_e + np.cumsum(slack)
    # float32 total: stable values contribute their weight exactly; the rest of
    # the exact total (unstable values and everything beyond the captured
    # width) contributes its weight within two float32 roundings
    stable = ~unstable
    base = float(w_lo[stable].sum())
    rest = total - float(lam[stable].sum())
    rest_lo = max(rest - (tail_e if stable.any() else 0.0), 0.0)
    rest_hi = rest + d * len(lam)
    thr_lo = _threshold(base + rest_lo * (1.0 - 2.0 ** -22) - d, eps)
    thr_hi = _threshold(base + rest_hi * (1.0 + 2.0 ** -22) + d, eps)
    early = np.nonzero(U.astype(np.float32) >= thr_lo)[0]
    k_min = int(early[0]) + 1 if early.size else len(lam) + 1
    if k_min > cap:
        return cap                  # no admissible spectrum reaches the threshold by cap
    late = np.nonzero(L.astype(np.float32) >= thr_hi)[0]
    if not late.size:
        return None
    k_min, k_max = min(k_min, cap), min(int(late[0]) + 1, cap)
    return k_min if k_min == k_max else None


def orth(Y: torch.Tensor, infos: list) -> torch.Tensor:
    """Orthonormal basis of the columns of Y [m, l] (complex128) by shifted
    Cholesky-QR followed by Cholesky-QR2 (shiftedCholeskyQR3: Fukaya,
    Kannan, Nakatsukasa, Yamamoto, Yanagisawa, SIAM J. Sci. Comput. 42 (2020)),
    with their shift 11 (m l + l (l + 1)) u ||Y||^2 (||Y||_F^2 as the bound),
    each pass one ``aq_cholqr`` call on the library's kernels (Gram matrix and
    triangular solves on MFMA GEMMs, 64 x 64 diagonal blocks factored and
    inverted in LDS).  The first pass keeps the span of Y and makes the next
    two well posed.  No host round trip: the passes' pivot codes (int32 [3]
    on the device) are appended to `infos` and checked once by the caller.
    (The library's Householder QR, ``kernels.qr``, would need no shift, but
    its 32-column panel factorisation is a chain of column steps: on the
    sketches here, m = 2048 and l ~ 100-250, it halved the low-entanglement
    rate; the small l x l Cholesky factors cost far less.)"""
    from ._ffi import check, lib
    m, l = Y.shape
    u = 2.0 ** -53
    Y = Y.contiguous()
    L = lib()
    nws = int(L.aq_cholqr_workspace_elems(m, l))
    ws = torch.empty(nws, dtype=torch.complex128, device=Y.device)
    info = torch.zeros(3, dtype=torch.int32, device=Y.device)
    bufs = (torch.empty_like(Y), torch.empty_like(Y))
    src = Y
    for it in range(3):
        dst = bufs[it % 2]
        scale = 11.0 * (m * l + l * (l + 1)) * u if it == 0 else 0.0
        check(L.aq_cholqr(m, l, K._ptr(src), l, K._ptr(dst), l, scale, K._ptr(ws), nws,
                          ctypes.c_void_p(info.data_ptr() + 4 * it), K._stream(Y)), "aq_cholqr")
        src = dst
    infos.append(info)
    return src


_TLS = __import__("threading").local()


def _generator(mps, device):
    """Sketch generator, one per host thread (updates may run concurrently)."""
    gens = getattr(_TLS, "gens", None)
    if gens is None:
        gens = _TLS.gens = {}
    key = (id(mps), str(device))
    if key not in gens:
        g = torch.Generator(device=device)
        g.manual_seed(1234567)
        gens[key] = g
    return gens[key]


def width_for(k_guess: int, m: int, n: int, cap: int):
    """Range-finder width for a rank guess: k_guess plus oversampling."""
    r = min(m, n)
    k_guess = max(1, min(k_guess, cap, r))
    p = max(16, k_guess)
    return k_guess, min(r, k_guess + p)


def try_rsvd(mps, i, X, cap, budget_ok, local, cl, cr, st, k_guess):
    """One certified randomized factorization at width k_guess + p.
    Returns (result tuple, None) or (None, reason for the fallback)."""
    m, n = X.shape
    k_guess, l = width_for(k_guess, m, n, cap)
    if l >= WIDE_FRACTION * min(m, n):
        st.reject("wide")
        return None, "wide"
    prec = mps.theta_precision
    gen = _generator(mps, X.device)
    Om = torch.randn(n, l, dtype=torch.complex64, device=X.device, generator=gen)
    with RL.timed("rsvd_rangefinder", 8.0 * m * n * l,
                  "bf16_mfma" if prec == "bf16x3" else "fp32_mfma"):
        Y = K.cgemm(X, Om, prec=prec)
    infos = []
    Q = orth(Y.to(torch.complex128), infos)
    for _ in range(POWER_STEPS):
        Z = K.cgemm(X.T, Q.to(torch.complex64), conj_a=True, prec=prec)
        Zq = orth(Z.to(torch.complex128), infos)
        Y = K.cgemm(X, Zq.to(torch.complex64), prec=prec)
        Q = orth(Y.to(torch.complex128), infos)
    X64 = X.to(torch.complex128)
    with RL.timed("rsvd_project_fp64", 8.0 * l * m * n, "fp64_mfma"):
        Bc = K.cgemm(Q.T, X64, conj_a=True)                       # [l, n]
    G = K.cgemm(Bc, Bc.T, conj_b=True)                           # [l, l]
    # eigenvalues now, the kept vectors after the rank rule; orders <= 64 by
    # the one-sided Jacobi SVD of the core's Cholesky factor
    core = TopEig(G, small="svd")
    lam = torch.clamp(core.values, min=0.0)
    # one device-to-host transfer: captured spectrum, ||X||^2, factorization codes
    host = torch.cat([lam, torch.vdot(X64.reshape(-1), X64.reshape(-1)).real.view(1),
                      torch.stack(infos).to(torch.float64).abs().sum().view(1)]).cpu().numpy()
    lam_h, total, bad = host[:-2], float(host[-2]), host[-1] != 0
    if bad or not np.isfinite(host).all():
        st.reject("orth")
        return None, "orth"
    tail = max(total - float(lam_h.sum()), 0.0)
    S = torch.sqrt(lam)
    S_aug = torch.cat([S, torch.tensor([math.sqrt(tail)], dtype=S.dtype, device=S.device)])
    # keep the augmented tail last: the rule needs a descending spectrum
    S_aug, _ = torch.sort(S_aug, descending=True)
    eps = mps.eps_bond
    S_rule = S_aug.to(torch.float32) if local == torch.complex64 else S_aug
    k, eps_loc, entropy, condS = choose_rank_from_sigma(S_rule, eps, cap, budget_ok)
    with st.lock:
        st.predicted.append((k_guess, k))
        st.tails.append(tail / total if total > 0 else 0.0)
    if k >= l:
        st.reject("rank_at_width")
        return None, "rank_at_width"
    if tail > FID_TOL * total:
        st.reject("tail")
        return None, "tail"
    k_cert = certify_rank(lam_h, tail, eps, cap, total)
    # a rank below the certified one comes from the memory budget, which
    # depends on k only and binds identically on the exact spectrum
    if k_cert is None or k > k_cert or (k < k_cert and mps.budget_global_mb is None):
        st.reject("uncertified")
        return None, "uncertified"
    if math.isfinite(condS) and condS > mps.dtype_policy.promote_if_cond_gt:
        st.reject("condition")
        return None, "condition"  # promotion semantics: let the exact path decide in FP64
    Sk = S[:k]
    inv = torch.where(Sk > 0, 1.0 / Sk, torch.zeros_like(Sk))
    UBk = core.vectors(k).contiguous()                            # [l, k]
    Uk = K.cgemm(Q, UBk)                                          # [m, k]
    Vh = K.cgemm(UBk.T, Bc, conj_a=True) * inv[:, None]           # [k, n]
    A_new, B_new = K.rebuild(Uk.to(local), Sk.to(torch.float32), Vh.to(local).contiguous(),
                             k, cl, cr)
    st.bump("rsvd")
    return (A_new, B_new, k, eps_loc, entropy, condS, "rsvd_fp64_core", local), None
