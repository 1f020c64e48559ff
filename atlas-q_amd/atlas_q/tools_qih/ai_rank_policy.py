"""Calibrated rank-policy adapter (tools_qih/ai_rank_policy.py:11-55).

Wraps a base predictor and exposes ``predict(sigmas) -> rank``: the base
fraction (``predict_fraction`` when the base has it, else ``predict`` divided
by the non-zero count), an optional monotone calibration by linear
interpolation, a clamp to [f_min, f_max], ``floor(f * n_sv + 1e-6)`` at least
1, then the chi_max and chi_budget caps.  Host logic on the predictor's
outputs; the predictors themselves run on the device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch


@dataclass
class PolicyConfig:
    f_min: float = 0.70
    f_max: float = 0.98
    chi_budget: Optional[int] = None
    cal_x: Optional[np.ndarray] = None
    cal_y: Optional[np.ndarray] = None


class CalibratedPolicyAdapter:
    def __init__(self, base_predictor, cfg: PolicyConfig, chi_max: int):
        self.base = base_predictor
        self.cfg = cfg
        self.chi_max = int(chi_max)

    def _calibrate(self, f: float) -> float:
        if self.cfg.cal_x is None or self.cfg.cal_y is None:
            return float(np.clip(f, 0.0, 1.0))
        return float(np.interp(f, self.cfg.cal_x, self.cfg.cal_y))

    @staticmethod
    def _fraction_to_rank(f: float, n_sv: int) -> int:
        return max(1, int(math.floor(f * n_sv + 1e-6)))

    def predict(self, sigmas: torch.Tensor) -> int:
        with torch.no_grad():
            n_sv = int((sigmas > 1e-10).sum().item()) or 1
            if hasattr(self.base, "predict_fraction"):
                f = float(self.base.predict_fraction(sigmas))
            else:
                f = int(self.base.predict(sigmas)) / float(n_sv)
            f = float(np.clip(self._calibrate(f), self.cfg.f_min, self.cfg.f_max))
            r = min(self._fraction_to_rank(f, n_sv), self.chi_max)
            if self.cfg.chi_budget is not None:
                r = min(r, int(self.cfg.chi_budget))
            return max(1, int(r))
