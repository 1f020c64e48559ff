"""Learned rank prediction for the two-site update.

The predictor sizes the randomized SVD *before* any factorization of the
theta matrix X: from 32 features it predicts log2 of the rank the plain
method's rule keeps, and the range finder is built with that width plus
oversampling (``atlas_q.rsvd``).  The network is the 32 -> 128 -> 128 -> 1
GELU MLP with BF16 weights of ``aq_mlp_*``; training runs on the GPU with the
library's forward / backward / AdamW kernels on (features, kept rank) pairs
harvested from the plain method's own updates.

Features (index: meaning)
  0-15  log10 of the normalised eigenvalues of Y^H Y, Y = X Omega (16 Gaussian
        columns), i.e. the spectrum of a 16-column sketch of X (Jacobi kernel)
  16-18 log2 chi_left, chi_mid, chi_right     19 log2 cap
  20    log2 min(m, n)                        21 bond position i / (n - 1)
  22    -log10 eps_bond
  23-25 last entropy (nats) at bonds i, i-1, i+1 (0 before any update)
  26-29 operator-Schmidt coefficients of the 4x4 gate (unit 2-norm)
  30    log10 of the last discarded weight eps_local^2 at bond i
  31    chi_mid / cap
The reference's own learned predictors (tools_qih/ai_rank_predictor.py:14-89,
finetuned_predictor.py:8-81) consume the singular values of a completed SVD;
their data-harvest fields (tn_core.py:256-272) and training recipe
(ai_rank_predictor.py:117-195: AdamW, MSE) carry over.
"""
from __future__ import annotations

import ctypes
import math
from pathlib import Path
from typing import Optional

import numpy as np
import torch

from . import _ffi
from . import kernels as K
from ._ffi import check, lib

NFEAT = 32
SKETCH = 16
DEFAULT_WEIGHTS = Path(__file__).with_name("rank_mlp_weights.npz")


def _p(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _s(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def heev_small(G: torch.Tensor, max_sweeps: int = 30, tol: float = 1e-15,
               orders: torch.Tensor = None):
    """Eigen-decomposition of Hermitian complex128 matrices [batch, n, n] or
    [n, n] (n <= 64) with the Jacobi kernel: (eigenvalues desc, eigenvectors).

    ``orders`` (int64 [batch] on the device, each <= n): matrix b is its
    leading orders[b] block (the rest is padding, not read); its eigenvalues
    fill W[b, :orders[b]] in descending order, W[b, orders[b]:] = 0, and the
    padding block of V[b] is the identity."""
    single = G.dim() == 2
    if single:
        G = G.unsqueeze(0)
    G = G.to(torch.complex128).contiguous()
    bsz, n, _ = G.shape
    W = torch.empty((bsz, n), dtype=torch.float64, device=G.device)
    V = torch.empty((bsz, n, n), dtype=torch.complex128, device=G.device)
    if orders is None:
        check(lib().aq_heev_small(n, _p(G), n, n * n, _p(W), _p(V), n, n * n, bsz, max_sweeps,
                                  tol, None, _s(G)), "aq_heev_small")
    else:
        if orders.dtype != torch.int64 or orders.device != G.device or orders.numel() != bsz:
            raise ValueError("heev_small: orders must be int64 [batch] on the matrices' device")
        orders = orders.contiguous()
        check(lib().aq_heev_small_orders(n, _p(orders), _p(G), n, n * n, _p(W), _p(V), n, n * n,
                                         bsz, max_sweeps, tol, None, _s(G)),
              "aq_heev_small_orders")
    return (W[0], V[0]) if single else (W, V)


def gate_schmidt(gate64: np.ndarray) -> np.ndarray:
    """Operator-Schmidt coefficients of U[(p,q),(p1,p2)] across the bond."""
    U = gate64.view(np.complex128).reshape(2, 2, 2, 2)       # [p, q, p1, p2]
    M = U.transpose(0, 2, 1, 3).reshape(4, 4)                # [(p,p1), (q,p2)]
    s = np.linalg.svd(M, compute_uv=False)
    return s / max(np.linalg.norm(s), 1e-300)


CENSOR_BAND = 0.25     # log2 units


class RankPredictor:
    """Device-resident MLP with feature standardisation."""

    def __init__(self, device="cuda", seed: int = 0):
        self.device = torch.device(device)
        self.nparam = int(lib().aq_mlp_num_params())
        self.master = torch.from_numpy(self._init(seed)).to(self.device)
        self.wb = torch.empty(self.nparam, dtype=torch.int16, device=self.device)
        self.mean = torch.zeros(NFEAT, dtype=torch.float32, device=self.device)
        self.std = torch.ones(NFEAT, dtype=torch.float32, device=self.device)
        self.margin = 1.15          # width multiplier on the predicted rank
        # largest log2(kept rank) of the training targets: a prediction within
        # CENSOR_BAND of it is censored (the rank may be larger than anything
        # seen in training) and read as "at least the cap"
        self.log2_max = 20.0
        self._sync()
        self._gen = torch.Generator(device=self.device)
        self._gen.manual_seed(777)
        self.trained = False

    def _thread_gen(self):
        import threading
        tls = self.__dict__.setdefault("_tls", threading.local())
        g = getattr(tls, "gen", None)
        if g is None:
            g = tls.gen = torch.Generator(device=self.device)
            g.manual_seed(777)
        return g

    # ------------------------------------------------------------ parameters
    def _init(self, seed):
        rng = np.random.default_rng(seed)
        p = np.zeros(self.nparam, dtype=np.float32)
        i = 0
        for rows, cols, fan in ((128, 32, 32), (128, 128, 128)):
            p[i:i + rows * cols] = rng.normal(0, math.sqrt(2.0 / fan), rows * cols)
            i += rows * cols + rows
        p[i:i + 128] = rng.normal(0, math.sqrt(1.0 / 128), 128)
        return p

    def _sync(self):
        check(lib().aq_mlp_sync_bf16(_p(self.master), _p(self.wb), _s(self.master)),
              "aq_mlp_sync_bf16")

    def save(self, path):
        np.savez(path, master=self.master.cpu().numpy(), mean=self.mean.cpu().numpy(),
                 std=self.std.cpu().numpy(), margin=np.float64(self.margin),
                 log2_max=np.float64(self.log2_max))

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, device="cuda"):
        d = np.load(path, allow_pickle=False)
        r = cls(device)
        r.master = torch.from_numpy(d["master"].astype(np.float32)).to(r.device)
        r.mean = torch.from_numpy(d["mean"].astype(np.float32)).to(r.device)
        r.std = torch.from_numpy(d["std"].astype(np.float32)).to(r.device)
        r.margin = float(d["margin"])
        # weights saved before the field existed were trained on kept ranks <= 1024
        r.log2_max = float(d["log2_max"]) if "log2_max" in d.files else 10.0
        r._sync()
        r.trained = True
        return r

    # ------------------------------------------------------------ inference
    def forward(self, F: torch.Tensor) -> torch.Tensor:
        """Raw network output for standardised features F [B, 32] (float32)."""
        F = F.contiguous()
        y = torch.empty(F.shape[0], dtype=torch.float32, device=F.device)
        check(lib().aq_mlp_forward(_p(F), F.shape[0], _p(self.wb), _p(self.master), _p(y),
                                   _s(F)), "aq_mlp_forward")
        return y

    def predict_log2(self, feats: torch.Tensor) -> torch.Tensor:
        return self.forward((feats - self.mean) / self.std)

    def features(self, mps, i: int, X: torch.Tensor, gate64: np.ndarray, cap: int) -> torch.Tensor:
        m, n = X.shape
        Om = torch.randn(n, SKETCH, dtype=X.dtype, device=X.device, generator=self._thread_gen())
        Y = K.cgemm(X, Om).to(torch.complex128)
        G = K.cgemm(Y.T, Y, conj_a=True)
        mu, _ = heev_small(G)
        mu = torch.clamp(mu, min=0.0)
        sk = torch.log10(mu / torch.clamp(mu.sum(), min=1e-300) + 1e-12).to(torch.float32)
        T = mps.tensors
        cl, cm, cr = T[i].shape[0], T[i].shape[2], T[i + 1].shape[2]
        ent = getattr(mps, "_last_entropy", {})
        epsd = getattr(mps, "_last_eps", {})
        host = np.zeros(NFEAT - SKETCH, dtype=np.float32)
        host[0:3] = np.log2([cl, cm, cr])
        host[3] = math.log2(cap)
        host[4] = math.log2(min(m, n))
        host[5] = i / max(1, mps.num_qubits - 2)
        host[6] = -math.log10(max(mps.eps_bond, 1e-300))
        host[7] = ent.get(i, 0.0)
        host[8] = ent.get(i - 1, 0.0)
        host[9] = ent.get(i + 1, 0.0)
        host[10:14] = gate_schmidt(gate64)
        host[14] = math.log10(epsd.get(i, 0.0) ** 2 + 1e-30)
        host[15] = cm / cap
        return torch.cat([sk, torch.from_numpy(host).to(X.device)])

    def predict_width(self, feats: torch.Tensor, cap: int, rmax: int) -> int:
        y = float(self.predict_log2(feats.unsqueeze(0))[0].item())
        if y >= self.log2_max - CENSOR_BAND:
            return max(1, min(cap, rmax))      # at the edge of the training range
        k = int(math.ceil(self.margin * 2.0 ** max(0.0, min(y, 20.0))))
        return max(1, min(k, cap, rmax))

    # ------------------------------------------------------------ training
    def fit(self, F: np.ndarray, k: np.ndarray, steps: int = 3000, batch: int = 4096,
            lr: float = 2e-3, wd: float = 1e-5, seed: int = 0, log_every: int = 500):
        """AdamW on MSE(log2 k) with the library's kernels; returns the loss log."""
        F = np.asarray(F, np.float32)
        t = np.log2(np.asarray(k, np.float64)).astype(np.float32)
        self.log2_max = float(t.max())
        mean = F.mean(0)
        std = F.std(0) + 1e-3
        self.mean = torch.from_numpy(mean).to(self.device)
        self.std = torch.from_numpy(std).to(self.device)
        Xd = torch.from_numpy((F - mean) / std).to(self.device)
        td = torch.from_numpy(t).to(self.device)
        N = Xd.shape[0]
        bsz = min(batch, N)
        grad = torch.zeros(self.nparam, dtype=torch.float32, device=self.device)
        m1 = torch.zeros_like(grad)
        m2 = torch.zeros_like(grad)
        ws_n = int(lib().aq_mlp_workspace_floats(bsz))
        ws = torch.empty(ws_n, dtype=torch.float32, device=self.device)
        g = torch.Generator(device=self.device)
        g.manual_seed(seed)
        log = []
        for step in range(1, steps + 1):
            idx = torch.randint(0, N, (bsz,), device=self.device, generator=g)
            xb = Xd[idx].contiguous()
            tb = td[idx]
            y = self.forward(xb)
            dy = (2.0 / bsz) * (y - tb)
            check(lib().aq_mlp_backward(_p(xb), _p(dy), bsz, _p(self.wb), _p(self.master),
                                        _p(grad), _p(ws), ws_n, _s(xb)), "aq_mlp_backward")
            gn = float(grad.norm())  # clip to 1.0 (ai_rank_predictor.py recipe)
            scale = 1.0 if gn <= 1.0 else 1.0 / gn
            cur_lr = lr * 0.5 * (1 + math.cos(math.pi * step / steps))
            check(lib().aq_mlp_adamw(_p(self.master), _p(grad), _p(m1), _p(m2), _p(self.wb),
                                     cur_lr, 0.9, 0.999, 1e-8, wd, scale, step, _s(xb)),
                  "aq_mlp_adamw")
            if step % log_every == 0 or step == steps:
                log.append((step, float(((y - tb) ** 2).mean())))
        self.trained = True
        return log

    def evaluate(self, F: np.ndarray, k: np.ndarray):
        feats = torch.from_numpy(np.asarray(F, np.float32)).to(self.device)
        y = self.predict_log2(feats).cpu().numpy()
        kh = np.ceil(self.margin * 2.0 ** np.clip(y, 0, 20))
        k = np.asarray(k)
        return {"rmse_log2": float(np.sqrt(np.mean((y - np.log2(k)) ** 2))),
                "under_predicted": float(np.mean(kh < k)),
                "mean_overshoot": float(np.mean(kh / k))}

    # ------------------------------------------------------------ hook used by rsvd
    def predict_for_update(self, mps, i, X, gate64, cap):
        feats = self.features(mps, i, X, gate64, cap)
        return self.predict_width(feats, cap, min(X.shape))
