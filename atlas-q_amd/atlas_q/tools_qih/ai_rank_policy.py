"""Calibrated rank policy for the ``predict(S)`` seam of tn_core.

Interface of the reference's adapter (tools_qih/ai_rank_policy.py:11-55):
``CalibratedPolicyAdapter(base_predictor, PolicyConfig(...), chi_max)`` is
itself a predictor whose ``predict(sigmas)`` returns a rank.  The rank is a
pure function of the base predictor's kept fraction f and the non-zero count
n (#(S > 1e-10), at least 1):

    f -> calibration table (piecewise linear) or [0, 1] -> [f_min, f_max]
      -> max(1, floor(f * n + 1e-6)) -> min(chi_max, chi_budget) -> >= 1

The fraction is the base's ``predict_fraction`` when it has one, otherwise
its rank divided by n.  The rank arithmetic is ``policy_rank`` below, shared
with anything else that turns a fraction into a rank under a policy.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

NONZERO = 1e-10


@dataclass
class PolicyConfig:
    """Clamp interval of the kept fraction, optional bond budget and
    optional monotone calibration points (cal_x ascending)."""
    f_min: float = 0.70
    f_max: float = 0.98
    chi_budget: Optional[int] = None
    cal_x: Optional[np.ndarray] = None
    cal_y: Optional[np.ndarray] = None


def policy_rank(fraction: float, n_sv: int, cfg: PolicyConfig, chi_max: int) -> int:
    """The adapter's rank for a raw kept fraction over n_sv non-zero values."""
    if cfg.cal_x is not None and cfg.cal_y is not None:
        g = float(np.interp(fraction, cfg.cal_x, cfg.cal_y))
    else:
        g = min(max(float(fraction), 0.0), 1.0)
    g = min(max(g, cfg.f_min), cfg.f_max)
    ceiling = int(chi_max) if cfg.chi_budget is None else min(int(chi_max), int(cfg.chi_budget))
    rank = max(1, math.floor(g * n_sv + 1e-6))
    return max(1, min(rank, ceiling))


class CalibratedPolicyAdapter:
    """A base rank predictor behind the calibrated clamp policy."""

    def __init__(self, base_predictor, cfg: PolicyConfig, chi_max: int):
        self.base = base_predictor
        self.cfg = cfg
        self.chi_max = int(chi_max)

    def fraction(self, sigmas: torch.Tensor, n_sv: int) -> float:
        """The base predictor's kept fraction of the n_sv non-zero values."""
        frac_fn = getattr(self.base, "predict_fraction", None)
        if frac_fn is not None:
            return float(frac_fn(sigmas))
        return int(self.base.predict(sigmas)) / float(n_sv)

    @torch.no_grad()
    def predict(self, sigmas: torch.Tensor) -> int:
        n_sv = max(1, int(torch.count_nonzero(sigmas > NONZERO)))
        return policy_rank(self.fraction(sigmas, n_sv), n_sv, self.cfg, self.chi_max)
