"""torch-CPU restatement of the reference's tools_qih/tn_core.py two-site
update and its learned rank predictors (the oracle).

TEST INFRASTRUCTURE ONLY -- imported by ``tests/`` and never by the shipped
package.  Every function cites the reference file:line it restates (paths
relative to the reference checkout).

SVD control flow: on the reference's target device the requested driver
succeeds on the first try (tn_core.py:44-51, no jitter).  On a CPU tensor
torch rejects the ``driver`` argument, so a CPU run of the reference ends in
its last resort, a jittered CPU SVD (tn_core.py:53-56).  The oracle restates
the device control flow; the golden vectors reproduce it in the reference by
dropping the ``driver`` argument (tests/golden/make_golden_tn_core.py).

Parity pinning
--------------
* ``theta_t``: closed-form index loop (U^T convention) in tests.
* ``adaptive_keep`` / ``truncate``: hand-computed spectra, policy arithmetic
  with Python's half-to-even rounding, eps filter.
* ``mps_apply_2q``: golden runs written by the reference (chi and truncation
  error per gate, amplitudes), the sqrt(S) split identity.
* ``features_ft135`` / ``features_log1p64``: golden feature vectors written by
  the reference's own feature code.
* ``dense_forward``: torch.nn Linear / LayerNorm / ReLU / Sigmoid modules (the
  library routine) and the reference's predictor outputs on seeded weights.
* ``features_spec32``: this build's own predictor inputs (no reference
  counterpart); pinned by the closed forms of its definition in tests.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np
import torch

POLICIES = ("baseline", "guardrail", "blend", "clamp", "clamp+blend")


def theta_t(Ai: torch.Tensor, Aj: torch.Tensor, U: torch.Tensor) -> torch.Tensor:
    """tn_core.py:224-230: T'[l, u, r] = sum_d T[l, d, r] U[d, u], reshaped to
    [2 li, 2 rj]."""
    li, rj = Ai.shape[0], Aj.shape[2]
    T = torch.einsum("abc,cde->abde", Ai, Aj).reshape(li, 4, rj)
    T = torch.einsum("ldr,du->lur", T, U).reshape(li, 2, 2, rj)
    return T.reshape(li * 2, 2 * rj)


def svd_robust(T: torch.Tensor):
    """tn_core.py:27-51, first-driver branch: SVD of T / ||T||, S rescaled."""
    scale = T.norm()
    Xn = T / scale if float(scale) > 0 else T
    U, S, Vh = torch.linalg.svd(Xn, full_matrices=False)
    if float(scale) > 0:
        S = S * scale
    return U, S, Vh


def adaptive_keep(S: torch.Tensor, tol: float, chi_max: Optional[int]) -> int:
    """tn_core.py:237-246 / 300-309: #(cumsum(S^2) <= (1 - tol) total) + 1, at
    most len(S), 1 when none; capped by chi_max."""
    S2 = S ** 2
    total = S2.sum()
    keep_mask = torch.cumsum(S2, dim=0) <= (1 - tol) * total
    if keep_mask.any():
        k = min(keep_mask.sum().item() + 1, S.numel())
    else:
        k = 1
    if chi_max is not None:
        k = min(k, chi_max)
    return int(k)


def policy_keep(ai_keep: int, adapt: int, policy: str, alpha: float) -> int:
    """tn_core.py:311-329."""
    if policy in ("baseline", "clamp"):
        return max(1, ai_keep)
    if policy in ("blend", "clamp+blend"):
        blend = int(round(alpha * ai_keep + (1 - alpha) * adapt))
        return max(1, min(blend, adapt))
    return max(1, min(ai_keep, adapt))


def mps_apply_2q(cores: List[torch.Tensor], i: int, j: int, U: torch.Tensor, chi_max=None,
                 eps: float = 1e-10, adaptive: bool = False, tol: float = 1e-4,
                 ai_predictor=None, policy: str = "guardrail", alpha: float = 0.7,
                 events: Optional[list] = None):
    """tn_core.py:186-384 (policy / alpha passed instead of read from the
    environment; the SVD event appended to ``events`` when given)."""
    assert j == i + 1
    Ai, Aj = cores[i], cores[j]
    li, _, ri = Ai.shape
    _, _, rj = Aj.shape
    T = theta_t(Ai, Aj, U)
    U1, S, Vh = svd_robust(T)
    S_squared = S ** 2
    total_weight = S_squared.sum()
    adapt_log = S.numel()
    if adaptive and tol > 0:
        adapt_log = adaptive_keep(S, tol, chi_max)
    if events is not None:
        top = min(128, S.numel())
        Sv = S[:top].detach().abs().double().tolist() + [0.0] * (128 - top)
        events.append(dict(kind="svd_spectrum", i=int(i), j=int(j), pre_rank=int(ri),
                           chi_max=int(chi_max) if chi_max is not None else None,
                           tol=float(tol), adaptive=bool(adaptive),
                           total_weight=float(total_weight), S_top128=Sv,
                           adaptive_keep=int(adapt_log)))
    if ai_predictor is not None and S.numel() > 10:
        ai_keep = int(ai_predictor.predict(S))
        if chi_max is not None:
            ai_keep = min(ai_keep, chi_max)
        adapt = S.numel()
        if adaptive and tol > 0:
            adapt = adaptive_keep(S, tol, chi_max)
        keep_count = policy_keep(ai_keep, adapt, policy, alpha)
    elif adaptive and tol > 0:
        keep_count = adaptive_keep(S, tol, chi_max)
    else:
        keep_count = chi_max if (chi_max is not None and S.numel() > chi_max) else S.numel()
    S_t, U_t, V_t = S[:keep_count], U1[:, :keep_count], Vh[:keep_count, :]
    kept_weight = (S_t ** 2).sum()
    trunc = 1.0 - (kept_weight / total_weight).item() if total_weight > 0 else 0.0
    keep = S_t > eps
    S_f, U_f, V_f = S_t[keep], U_t[:, keep], V_t[keep, :]
    chi = S_f.numel()
    U_f = U_f.reshape(li, 2, chi)
    V_f = V_f.reshape(chi, 2, rj)
    Ssqrt = torch.sqrt(S_f)
    cores[i] = torch.einsum("ldr,r->ldr", U_f, Ssqrt)
    cores[j] = torch.einsum("r,rde->rde", Ssqrt, V_f)
    return cores, chi, trunc


def mps_init_plus(n: int, dtype=torch.complex64) -> List[torch.Tensor]:
    """tn_core.py:171-174."""
    v = torch.tensor([1 / np.sqrt(2), 1 / np.sqrt(2)], dtype=dtype)
    return [v.view(1, 2, 1).clone() for _ in range(n)]


def mps_apply_1q(cores, i, U):
    """tn_core.py:177-181."""
    cores[i] = torch.einsum("ldr,du->lur", cores[i], U)
    return cores


def contract(cores: Sequence[torch.Tensor]) -> np.ndarray:
    """State vector of open-boundary cores (qubit 0 most significant)."""
    psi = cores[0].to(torch.complex128).reshape(2, -1)
    for c in cores[1:]:
        c = c.to(torch.complex128)
        psi = (psi @ c.reshape(c.shape[0], -1)).reshape(-1, c.shape[2])
    return psi.reshape(-1).numpy()


# ---------------------------------------------------------------- predictors
def features_log1p64(S: np.ndarray, max_rank: int) -> np.ndarray:
    """ai_rank_predictor.py:60-69: log1p(S[:n] / S[0]), n = min(max_rank, 64),
    in S's precision, stored as float32."""
    S = np.asarray(S)
    x = np.zeros(64, np.float32)
    n = min(int(max_rank), 64, S.size)
    if n > 0 and S[0] > 0:
        x[:n] = np.log1p(S[:n] / S[0]).astype(np.float32)
    return x


def features_ft135(S: np.ndarray) -> np.ndarray:
    """finetuned_predictor.py:51-75 (float32 throughout)."""
    sv = np.asarray(S, np.float32)
    sv = np.pad(sv, (0, 128 - sv.size)) if sv.size < 128 else sv[:128]
    sv2 = sv * sv
    sw = float(sv2.sum() + 1e-12)
    mass = sv2 / sw
    nz = mass > 0
    H = float(-(mass[nz] * np.log(mass[nz] + 1e-12)).sum())
    logs = np.log(sv + 1e-12)
    logs -= logs.max()
    cm = [float(mass[:k].sum()) for k in (8, 16, 32, 64, 96, 128)]
    return np.concatenate([logs, np.array([H], np.float32), np.array(cm, np.float32)])


def n_sv(S: np.ndarray) -> int:
    return int((np.asarray(S) > np.asarray(1e-10, dtype=np.asarray(S).dtype)).sum())


def features_spec32(S_top: np.ndarray, total_weight: float) -> np.ndarray:
    """This build's spectral predictor inputs from the leading (<= 128)
    singular values and the total weight sum S^2:
      0-23  log10 max(S_i^2 / total, 1e-12)   (-12 past the spectrum)
      24-29 log10 max(1 - sum_{i<c} S_i^2 / total, 1e-12), c = 8,16,32,64,96,128
      30    entropy -sum p ln p of p_i = S_i^2 / total over the leading values
      31    log2 max(1, #(S_i > 1e-10) among the leading values)."""
    s = np.asarray(S_top, np.float64)[:128]
    tot = float(total_weight)
    f = np.full(32, -12.0, np.float64)
    if tot <= 0:
        f[24:30] = 0.0
        f[30] = 0.0
        f[31] = 0.0
        return f.astype(np.float32)
    p = s * s / tot
    n = min(24, s.size)
    f[:n] = np.log10(np.maximum(p[:n], 1e-12))
    c = np.cumsum(p)
    for q, cut in enumerate((8, 16, 32, 64, 96, 128)):
        mass = c[min(cut, s.size) - 1] if s.size else 0.0
        f[24 + q] = math.log10(max(1.0 - mass, 1e-12))
    pp = p[p > 0]
    f[30] = float(-(pp * np.log(pp)).sum())
    f[31] = math.log2(max(1, int((s > 1e-10).sum())))
    return f.astype(np.float32)


def dense_forward(layers: Sequence[dict], x: np.ndarray) -> np.ndarray:
    """Linear (+LayerNorm, biased variance) (+ReLU / GELU / Sigmoid) stack in
    float64 on float32 parameters; layers: dicts of numpy arrays W, b and
    optional ln_g, ln_b, ln_eps, and act."""
    h = np.asarray(x, np.float64)
    for L in layers:
        h = h @ np.asarray(L["W"], np.float64).T
        if L.get("b") is not None:
            h = h + np.asarray(L["b"], np.float64)
        if L.get("ln_g") is not None:
            mu = h.mean(-1, keepdims=True)
            var = ((h - mu) ** 2).mean(-1, keepdims=True)
            h = (h - mu) / np.sqrt(var + L.get("ln_eps", 1e-5))
            h = h * np.asarray(L["ln_g"], np.float64) + np.asarray(L["ln_b"], np.float64)
        act = L.get("act", "none")
        if act == "relu":
            h = np.maximum(h, 0.0)
        elif act == "sigmoid":
            h = 1.0 / (1.0 + np.exp(-h))
        elif act == "gelu":
            from scipy.special import erf
            h = 0.5 * h * (1.0 + erf(h / math.sqrt(2.0)))
    return h


def layers_from_sequential(net) -> List[dict]:
    """Oracle layer list from a torch nn.Sequential of Linear / LayerNorm /
    activation / Dropout modules."""
    import torch.nn as nn
    out = []
    for mod in net:
        if isinstance(mod, nn.Linear):
            out.append({"W": mod.weight.detach().numpy(),
                        "b": None if mod.bias is None else mod.bias.detach().numpy(),
                        "act": "none"})
        elif isinstance(mod, nn.LayerNorm):
            out[-1].update(ln_g=mod.weight.detach().numpy(), ln_b=mod.bias.detach().numpy(),
                           ln_eps=mod.eps)
        elif isinstance(mod, nn.ReLU):
            out[-1]["act"] = "relu"
        elif isinstance(mod, nn.Sigmoid):
            out[-1]["act"] = "sigmoid"
        elif isinstance(mod, nn.GELU):
            out[-1]["act"] = "gelu"
    return out


def rank_from_fraction(frac: float, nsv: int, zero_is_one: bool = True) -> int:
    """ai_rank_predictor.py:80-88 / finetuned_predictor.py:77-81."""
    if nsv == 0:
        return 1
    return max(1, min(int(frac * nsv), nsv))


class SeqPredictor:
    """The reference's predictors (ai_rank_predictor.py:71-89,
    finetuned_predictor.py:77-81) evaluated by the oracle: features in NumPy,
    forward by ``dense_forward`` on the weights of a torch nn.Sequential."""

    def __init__(self, net, kind: str):
        self.layers = layers_from_sequential(net)
        self.kind = kind

    def fraction(self, S: torch.Tensor) -> float:
        s = S.detach().cpu().numpy()
        nsv = n_sv(s)
        x = features_log1p64(s, nsv) if self.kind == "log1p64" else features_ft135(s)
        return float(dense_forward(self.layers, x[None, :])[0, 0])

    def predict(self, S: torch.Tensor) -> int:
        s = S.detach().cpu().numpy()
        nsv = n_sv(s)
        if self.kind == "log1p64" and nsv == 0:
            return 1
        return rank_from_fraction(self.fraction(S), nsv if self.kind == "log1p64" else (nsv or 1))
