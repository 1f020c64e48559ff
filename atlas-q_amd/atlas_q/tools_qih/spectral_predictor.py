"""This library's learned rank predictor for the reference's ``predict(S)``
seam (tools_qih/tn_core.py:294).

The reference's predictors see the singular values of a finished SVD and
predict the fraction of the spectrum to keep.  ``SpectralRankPredictor`` plugs
into the same seam with this library's network -- the 32 -> 128 -> 128 -> 1
GELU MLP with BF16 weights of ``aq_mlp_*``, forward and training on the
device -- over 32 spectral features (``aq_spectrum_features`` kind
AQ_FEAT_SPEC32: leading weights, tail weights beyond 8 ... 128 values,
entropy, non-zero count), predicting log2 of the kept rank.

Training data are the reference's SVD events (``S_top128``, ``total_weight``,
``adaptive_keep``; ``svd_logger`` format), so logs of either implementation
train it; ``spectral_features_host`` computes the same 32 features from an
event.  With ``CalibratedPolicyAdapter`` or the ``QIH_AI_POLICY`` policies of
``tn_core.mps_apply_2q`` it behaves as the reference's predictors do.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .. import kernels as K
from ..rank_mlp import RankPredictor as _DeviceMLP

CUTS = (8, 16, 32, 64, 96, 128)


def spectral_features_host(S_top, total_weight: float) -> np.ndarray:
    """The AQ_FEAT_SPEC32 features from the leading (<= 128) singular values
    and the total weight sum S^2 (e.g. an SVD event's fields)."""
    s = np.asarray(S_top, np.float64)[:128]
    tot = float(total_weight)
    f = np.full(32, -12.0)
    if tot <= 0:
        f[24:] = 0.0
        return f.astype(np.float32)
    p = s * s / tot
    f[:min(24, s.size)] = np.log10(np.maximum(p[:24], 1e-12))
    c = np.cumsum(p)
    for q, cut in enumerate(CUTS):
        f[24 + q] = math.log10(max(1.0 - (c[min(cut, s.size) - 1] if s.size else 0.0), 1e-12))
    pp = p[p > 0]
    f[30] = float(-(pp * np.log(pp)).sum())
    f[31] = math.log2(max(1, int((s > 1e-10).sum())))
    return f.astype(np.float32)


class SpectralRankPredictor:
    """predict(S) -> rank on the device; fit from SVD events."""

    def __init__(self, device="cuda", seed: int = 0):
        self.mlp = _DeviceMLP(device, seed=seed)
        self.device = self.mlp.device

    @property
    def trained(self) -> bool:
        return self.mlp.trained

    def _log2_and_count(self, S: torch.Tensor):
        f, nsv = K.spectrum_features(S.to(self.device), "spec32")
        y = self.mlp.predict_log2(f)
        host = torch.cat([y.reshape(-1).double(), nsv.double()]).cpu()
        return float(host[0]), int(host[1])

    def predict(self, S: torch.Tensor) -> int:
        y, nsv = self._log2_and_count(S)
        k = int(round(2.0 ** max(0.0, min(y, 30.0))))
        return max(1, min(k, max(nsv, 1)))

    def predict_fraction(self, S: torch.Tensor) -> float:
        y, nsv = self._log2_and_count(S)
        return min(1.0, 2.0 ** max(0.0, min(y, 30.0)) / max(nsv, 1))

    @staticmethod
    def event_arrays(events):
        """(features [N, 32], kept ranks [N]) of svd_spectrum events."""
        F = np.stack([spectral_features_host(e["S_top128"], e["total_weight"]) for e in events])
        k = np.array([max(1, int(e["adaptive_keep"])) for e in events])
        return F, k

    def fit_events(self, events, steps: int = 2000, **kw):
        F, k = self.event_arrays(events)
        return self.mlp.fit(F, k, steps=steps, **kw)

    def save(self, path):
        self.mlp.save(path)

    @classmethod
    def load(cls, path, device="cuda"):
        r = cls.__new__(cls)
        r.mlp = _DeviceMLP.load(path, device)
        r.device = r.mlp.device
        return r
