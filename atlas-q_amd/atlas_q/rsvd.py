"""Randomized truncated SVD of the theta matrix, certified in FP64.

For X [m, n] (complex64) and a width l = k_guess + p, p = max(16, k_guess/2):
  Y = X Omega                      range finder (MFMA GEMM, BF16x3 optional)
  q x { Z = X^H Q ; Y = X Z' }     optional power steps (q = POWER_STEPS),
                                   re-orthonormalised
  Q = orth(Y)                      shifted Cholesky-QR, twice, in FP64
  B = Q^H X        [l, n]          FP64 MFMA GEMM
  B B^H = U_B diag(lam) U_B^H      FP64 Hermitian core
  tail = ||X||_F^2 - sum(lam)      FP64: weight outside range(Q)
The captured eigenvalues bound the true squared singular values:
sigma_i^2 >= lam_i (Cauchy interlacing for the projection Q^H X) and
sum_{i<=j} sigma_i^2 <= E_j + tail with E_j = sum_{i<=j} lam_i, while
||X||_F^2 is exact.  The result is accepted only when
(i)  tail <= FID_TOL * ||X||_F^2: the kept weight of the rank-k result is then
     within FID_TOL of the exact truncation's (sum_{i<k} (sigma_i^2 - lam_i)
     <= tail), i.e. the truncation fidelity agrees to 1e-7 per update; and
(ii) the rank is certified (``certify_rank``): every spectrum consistent with
     the bounds crosses the rule's threshold (1 - eps^2) ||X||^2, widened by
     one float32 ulp of ||X||^2 (the resolution of the rule's own float32
     arithmetic), within MAX_SPREAD consecutive ranks, and the rank the rule
     returns on the captured spectrum is one of them.  Where the exact partial
     sums sit farther than that resolution from the threshold this is the
     exact rule's rank; inside it the reference's float32 rule itself cannot
     separate the candidates, whose weights all lie within the band (so the
     kept weight, and the fidelity, differ by at most 2 ulp + tail).
Otherwise the caller falls back to the exact Gram factorization.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import kernels as K
from . import roofline as RL
from .eig import eigh_desc
from .truncation import choose_rank_from_sigma

FID_TOL = 1e-7         # admissible uncaptured weight, of ||X||_F^2
POWER_STEPS = 0        # power iterations; oversampling k/2 captures decaying spectra
WIDE_FRACTION = 0.5    # widths beyond this share of min(m, n) go to the Gram path
RULE_BAND = 2.0 ** -23 # float32 resolution of the rank rule's partial sums, of ||X||_F^2
MAX_SPREAD = 2         # certified ranks may differ by this much (all within RULE_BAND)


def certify_rank(lam: np.ndarray, tail: float, total: float, eps: float, cap: int):
    """Ranks the rule (truncation.py:52-63) can return for any spectrum with
    sigma_i^2 >= lam_i and partial sums <= E_j + tail, total fixed.

    Returns (k_lo, k_hi) with k_hi - k_lo <= MAX_SPREAD (after the cap), or
    None when the crossing is not pinned that closely within the captured
    range.  FP64 throughout; lam descending."""
    E = np.cumsum(np.asarray(lam, np.float64))
    thresh = (1.0 - eps * eps) * total
    band = RULE_BAND * total
    early = np.nonzero(E + tail >= thresh - band)[0]     # largest admissible sums
    k_early = int(early[0]) + 1 if early.size else len(E) + 1
    if k_early > cap:
        return (cap, cap)            # no admissible spectrum reaches the threshold by cap
    late = np.nonzero(E >= thresh + band)[0]             # smallest admissible sums
    if not late.size:
        return None
    k_late = int(late[0]) + 1
    lo, hi = min(k_early, cap), min(k_late, cap)
    return (lo, hi) if hi - lo <= MAX_SPREAD else None


def orth(Y: torch.Tensor, infos: list) -> torch.Tensor:
    """Orthonormal basis of the columns of Y (complex128) by shifted Cholesky-QR
    x2, without a host round trip: the factorizations' info codes are appended
    to `infos` and checked once by the caller."""
    for it in range(2):
        W = K.cgemm(Y.T, Y, conj_a=True)
        if it == 0:
            shift = 1e-13 * torch.diagonal(W).real.sum()
            W = W + shift * torch.eye(W.shape[0], dtype=W.dtype, device=W.device)
        L, info = torch.linalg.cholesky_ex(W)
        infos.append(info)
        Y = torch.linalg.solve_triangular(L.conj().T, Y, upper=True, left=False)
    return Y


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
    p = max(16, k_guess // 2)
    return k_guess, min(r, k_guess + p)


def try_rsvd(mps, i, X, cap, budget_ok, local, cl, cr, st, k_guess):
    m, n = X.shape
    k_guess, l = width_for(k_guess, m, n, cap)
    st.predicted.append((k_guess, None))
    if l >= WIDE_FRACTION * min(m, n):
        st.reject("wide")
        return None
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
    lam, UB = eigh_desc(G)
    lam = torch.clamp(lam, min=0.0)
    # one device-to-host transfer: captured spectrum, ||X||^2, factorization codes
    host = torch.cat([lam, torch.vdot(X64.reshape(-1), X64.reshape(-1)).real.view(1),
                      torch.stack(infos).to(torch.float64).abs().sum().view(1)]).cpu().numpy()
    lam_h, total, bad = host[:-2], float(host[-2]), host[-1] != 0
    if bad or not np.isfinite(host).all():
        st.reject("orth")
        return None
    tail = max(total - float(lam_h.sum()), 0.0)
    S = torch.sqrt(lam)
    S_aug = torch.cat([S, torch.tensor([math.sqrt(tail)], dtype=S.dtype, device=S.device)])
    # keep the augmented tail last: the rule needs a descending spectrum
    S_aug, _ = torch.sort(S_aug, descending=True)
    eps = mps.eps_bond
    S_rule = S_aug.to(torch.float32) if local == torch.complex64 else S_aug
    k, eps_loc, entropy, condS = choose_rank_from_sigma(S_rule, eps, cap, budget_ok)
    st.predicted[-1] = (k_guess, k)
    st.tails.append(tail / total if total > 0 else 0.0)
    if k >= l:
        st.reject("rank_at_width")
        return None
    if tail > FID_TOL * total:
        st.reject("tail")
        return None
    rng = certify_rank(lam_h, tail, total, eps, cap)
    # a rank below the certified pair comes from the memory budget, which
    # depends on k only and binds identically on the exact spectrum
    if rng is None or k > rng[1] or (k < rng[0] and mps.budget_global_mb is None):
        st.reject("uncertified")
        return None
    if math.isfinite(condS) and condS > mps.dtype_policy.promote_if_cond_gt:
        st.reject("condition")
        return None  # promotion semantics: let the exact path decide in FP64
    Sk = S[:k]
    inv = torch.where(Sk > 0, 1.0 / Sk, torch.zeros_like(Sk))
    Uk = K.cgemm(Q, UB[:, :k].contiguous())                       # [m, k]
    Vh = K.cgemm(UB[:, :k].T, Bc, conj_a=True) * inv[:, None]     # [k, n]
    A_new, B_new = K.rebuild(Uk.to(local), Sk.to(torch.float32), Vh.to(local).contiguous(),
                             k, cl, cr)
    st.bump("rsvd")
    return A_new, B_new, k, eps_loc, entropy, condS, "rsvd_fp64_core", local
