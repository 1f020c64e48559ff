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
from .eig import eigh_desc
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
        self.lock = threading.Lock()

    def bump(self, key):
        with self.lock:
            self.counts[key] = self.counts.get(key, 0) + 1

    def reject(self, reason):
        with self.lock:
            self.rejects[reason] = self.rejects.get(reason, 0) + 1


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


def _gram_factors(X: torch.Tensor):
    """Exact leading factors of X via the FP64 Gram matrix of the smaller side.
    Returns (U [m, r] c128, S [r] f64 descending, Vh [r, n] c128, X64)."""
    m, n = X.shape
    X64 = X.to(torch.complex128)
    if m <= n:
        with RL.timed("gram_fp64_cgemm", 8.0 * m * m * n, "fp64_mfma"):
            G = K.cgemm(X64, X64.T, conj_b=True)         # X X^H  [m, m]
        lam, U = eigh_desc(G)
        S = torch.sqrt(torch.clamp(lam, min=0.0))
        return U, S, None, X64
    with RL.timed("gram_fp64_cgemm", 8.0 * n * n * m, "fp64_mfma"):
        G = K.cgemm(X64.T, X64, conj_a=True)             # X^H X  [n, n]
    lam, Vr = eigh_desc(G)
    S = torch.sqrt(torch.clamp(lam, min=0.0))
    return None, S, Vr, X64


def _finish_from_gram(mps, X, U, S, Vr, X64, k_dtype, cap, budget_ok, cl, cr, eps):
    """Rank rule on S (in the policy's real dtype) and the two cores."""
    S_rule = S.to(torch.float32) if k_dtype == torch.complex64 else S
    k, eps_loc, entropy, condS = choose_rank_from_sigma(S_rule, eps, cap, budget_ok)
    local = k_dtype
    if math.isfinite(condS) and condS > mps.dtype_policy.promote_if_cond_gt:
        k, eps_loc, entropy, condS = choose_rank_from_sigma(S, eps, cap, budget_ok)
        local = torch.complex128
    Sk = S[:k]
    inv = torch.where(Sk > 0, 1.0 / Sk, torch.zeros_like(Sk))
    if U is not None:
        Uk = U[:, :k]
        with RL.timed("proj_fp64_cgemm", 8.0 * k * X64.shape[0] * X64.shape[1], "fp64_mfma"):
            Vh = K.cgemm(Uk.T, X64, conj_a=True)             # U_k^H X  [k, n]
        Vh = Vh * inv[:, None]
    else:
        Vk = Vr[:, :k]
        Uk = K.cgemm(X64, Vk) * inv[None, :]                   # X V_k S^-1 [m, k]
        Vh = Vk.T.conj()
    Uk = Uk.to(local)
    Vh = Vh.to(local).contiguous()
    A_new, B_new = K.rebuild(Uk, Sk.to(torch.float32 if local == torch.complex64 else torch.float64),
                             Vh, k, cl, cr)
    return A_new, B_new, k, eps_loc, entropy, condS, local


def two_site_update(mps, i: int, A: torch.Tensor, B: torch.Tensor, gate: np.ndarray, cap: int,
                    budget_ok: Callable[[int], bool], local: torch.dtype):
    """Returns (A_new, B_new, k, eps_local, entropy, condS, driver, dtype)."""
    method = mps.svd_method
    cl, cr = A.shape[0], B.shape[2]
    st = _factor_stats(mps)
    prec = mps.theta_precision if local == torch.complex64 else "native"
    cm = A.shape[2]
    with RL.timed(f"theta_{prec}", 32.0 * cl * cm * cr + 128.0 * cl * cr,
                  "bf16_mfma" if prec == "bf16x3" else "fp32_mfma"):
        X = K.theta(A, B, gate, prec=prec)
    if local == torch.complex128 or method == "plain_in_fast":
        U, S, Vh, driver = robust_svd(X)
        k, e, s, c = choose_rank_from_sigma(S, mps.eps_bond, cap, budget_ok)
        A_new, B_new = K.rebuild(U, S, Vh, k, cl, cr)
        st.bump("plain")
        return A_new, B_new, k, e, s, c, driver, local
    feats = None
    pred = getattr(mps, "rank_predictor", None)
    if pred is None and method == "learned":
        with _INIT_LOCK:
            pred = mps.rank_predictor = default_predictor(X.device)
    if pred is not None and (method == "learned" or getattr(mps, "harvest", None) is not None):
        feats = pred.features(mps, i, X, gate, cap)
    out = None
    if method in ("rsvd", "learned"):
        from . import rsvd
        if method == "learned" and pred is not None and pred.trained:
            k_guess = pred.predict_width(feats, cap, min(X.shape))
        else:
            k_guess = min(cap, 2 * cm)          # operator-Schmidt growth bound, no model
        out = rsvd.try_rsvd(mps, i, X, cap, budget_ok, local, cl, cr, st, k_guess)
        if out is None:
            st.bump("rsvd_fallback")
    elif method != "gram":
        raise ValueError(f"unknown svd_method {method!r}")
    if out is None:
        U, S, Vr, X64 = _gram_factors(X)
        A_new, B_new, k, e, s, c, dt = _finish_from_gram(mps, X, U, S, Vr, X64, local, cap,
                                                        budget_ok, cl, cr, mps.eps_bond)
        st.bump("gram")
        out = (A_new, B_new, k, e, s, c, "gram_fp64", dt)
    if feats is not None and getattr(mps, "harvest", None) is not None:
        mps.harvest.append((feats.cpu().numpy(), out[2]))
    return out
