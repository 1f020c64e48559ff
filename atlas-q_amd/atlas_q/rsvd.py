"""Randomized truncated SVD of the theta matrix, certified by an FP64 tail check.

For X [m, n] (complex64) and a width l = k_guess + p:
  Y = X Omega                      range finder (MFMA GEMM, BF16x3 optional)
  q x { Z = X^H Q ; Y = X Z' }     power steps, re-orthonormalised each time
  Q = orth(Y)                      shifted Cholesky-QR, twice, in FP64
  B = Q^H X        [l, n]          FP64 MFMA GEMM
  B B^H = U_B diag(lam) U_B^H      FP64 Hermitian core
  tail = ||X||_F^2 - sum(lam)      FP64: weight outside range(Q)
The result is accepted only when (i) the rank decision is certified: with
captured energies E_i (descending, lower bounds of the true partial sums by
Ky Fan) and the exact total ||X||_F^2, the first index k with E_{k-1} >=
thresh + margin satisfies E_{k-2} + tail < thresh - margin (the true partial
sums are at most E_i + tail), so neither the uncaptured weight nor the
rule's rounding in the storage precision (margin = 1e-6 ||X||^2) can move k,
and the reference's rounding of the rule (run on [sqrt(lam), sqrt(tail)])
agrees;
and (ii) the tail is at most TAIL_TOL * max(eps^2, 2^-24) * ||X||_F^2, which
bounds the kept-weight (fidelity) difference from the exact truncation.
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

TAIL_TOL = 1e-2
POWER_STEPS = 2
WIDE_FRACTION = 0.75
RULE_MARGIN = 1e-6     # of ||X||_F^2, around the rank-rule threshold  # predicted widths beyond this share of min(m, n) go to the Gram path


def orth(Y: torch.Tensor) -> torch.Tensor:
    """Orthonormal basis of the columns of Y (complex128) by shifted Cholesky-QR x2."""
    for it in range(2):
        W = K.cgemm(Y.T, Y, conj_a=True)
        if it == 0:
            shift = 1e-13 * float(torch.diagonal(W).real.sum())
            W = W + shift * torch.eye(W.shape[0], dtype=W.dtype, device=W.device)
        L, info = torch.linalg.cholesky_ex(W)
        if int(info) != 0:
            Q, _ = torch.linalg.qr(Y)
            return Q
        Y = torch.linalg.solve_triangular(L.conj().T, Y, upper=True, left=False)
    return Y


def width_for(k_guess: int, m: int, n: int, cap: int):
    """Range-finder width for a rank guess: k_guess plus oversampling."""
    r = min(m, n)
    k_guess = max(1, min(k_guess, cap, r))
    p = max(10, k_guess // 4)
    return k_guess, min(r, k_guess + p)


def try_rsvd(mps, i, X, cap, budget_ok, local, cl, cr, st, k_guess):
    m, n = X.shape
    k_guess, l = width_for(k_guess, m, n, cap)
    st.predicted.append((k_guess, None))
    if l >= WIDE_FRACTION * min(m, n):
        return None
    prec = mps.theta_precision
    gen = getattr(mps, "_rsvd_gen", None)
    if gen is None:
        gen = mps._rsvd_gen = torch.Generator(device=X.device)
        gen.manual_seed(1234567)
    Om = torch.randn(n, l, dtype=torch.complex64, device=X.device, generator=gen)
    with RL.timed("rsvd_rangefinder", 8.0 * m * n * l,
                  "bf16_mfma" if prec == "bf16x3" else "fp32_mfma"):
        Y = K.cgemm(X, Om, prec=prec)
    Q = orth(Y.to(torch.complex128))
    for _ in range(POWER_STEPS):
        Z = K.cgemm(X.T, Q.to(torch.complex64), conj_a=True, prec=prec)
        Zq = orth(Z.to(torch.complex128))
        Y = K.cgemm(X, Zq.to(torch.complex64), prec=prec)
        Q = orth(Y.to(torch.complex128))
    X64 = X.to(torch.complex128)
    with RL.timed("rsvd_project_fp64", 8.0 * l * m * n, "fp64_mfma"):
        Bc = K.cgemm(Q.T, X64, conj_a=True)                       # [l, n]
    G = K.cgemm(Bc, Bc.T, conj_b=True)                           # [l, l]
    lam, UB = eigh_desc(G)
    lam = torch.clamp(lam, min=0.0)
    total = float(torch.vdot(X64.reshape(-1), X64.reshape(-1)).real)
    captured = float(lam.sum())
    tail = max(total - captured, 0.0)
    S = torch.sqrt(lam)
    S_aug = torch.cat([S, torch.tensor([math.sqrt(tail)], dtype=S.dtype, device=S.device)])
    # keep the augmented tail last: the rule needs a descending spectrum
    S_aug, _ = torch.sort(S_aug, descending=True)
    eps = mps.eps_bond
    S_rule = S_aug.to(torch.float32) if local == torch.complex64 else S_aug
    k, eps_loc, entropy, condS = choose_rank_from_sigma(S_rule, eps, cap, budget_ok)
    ok = (k < l) and tail <= TAIL_TOL * max(eps * eps, 2.0 ** -24) * total
    if ok:
        # decide in FP64 with a margin that covers the rule's rounding in the
        # storage precision (float32 energies round at ~3e-8 of the total)
        E = torch.cumsum(lam, 0).cpu().numpy()
        thresh = (1.0 - eps * eps) * total
        margin = RULE_MARGIN * total
        hit = np.nonzero(E >= thresh + margin)[0]
        k_c = int(hit[0]) + 1 if hit.size else None
        if k_c is None:
            ok = k == cap and k < l and E[k - 1] + tail < thresh - margin
        else:
            ok = (k == min(k_c, cap)) and (k_c == 1 or E[k_c - 2] + tail < thresh - margin)
    st.predicted[-1] = (k_guess, k)
    if not ok:
        return None
    if math.isfinite(condS) and condS > mps.dtype_policy.promote_if_cond_gt:
        return None  # promotion semantics: let the exact path decide in FP64
    Sk = S[:k]
    inv = torch.where(Sk > 0, 1.0 / Sk, torch.zeros_like(Sk))
    Uk = K.cgemm(Q, UB[:, :k].contiguous())                       # [m, k]
    Vh = K.cgemm(UB[:, :k].T, Bc, conj_a=True) * inv[:, None]     # [k, n]
    A_new, B_new = K.rebuild(Uk.to(local), Sk.to(torch.float32), Vh.to(local).contiguous(),
                             k, cl, cr)
    st.bump("rsvd")
    return A_new, B_new, k, eps_loc, entropy, condS, "rsvd_fp64_core", local
