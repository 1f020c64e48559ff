"""Per-operation statistics (API of the reference's ``src/atlas_q/diagnostics.py``)."""
from __future__ import annotations

import math
from typing import Dict, List

import numpy as np
import torch


def bond_entropy_from_S(S: torch.Tensor) -> float:
    """diagnostics.py:17-35: entanglement entropy in bits."""
    if len(S) == 0:
        return 0.0
    p = (S * S) / (S * S).sum()
    return float((-(p * torch.log2(torch.clamp(p, min=1e-30))).sum()).item())


def effective_rank(S: torch.Tensor, threshold: float = 0.99) -> int:
    """diagnostics.py:38-55."""
    if len(S) == 0:
        return 0
    E = (S * S).cumsum(0)
    k = int(torch.searchsorted(E, threshold * E[-1]).item()) + 1
    return min(k, len(S))


def spectral_gap(S: torch.Tensor, k: int) -> float:
    """diagnostics.py:58-75: sigma_{k-1} / sigma_k."""
    if k >= len(S) or k < 1 or S[k] == 0:
        return float("inf")
    return float((S[k - 1] / S[k]).item())


class MPSStatistics:
    """diagnostics.py:78-152: per-operation log and aggregates."""

    KEYS = ("step", "bond", "k_star", "chi_before", "chi_after", "eps_local", "entropy",
            "svd_driver", "dtype", "ms_elapsed", "condS")

    def __init__(self):
        self.logs: Dict[str, List] = {k: [] for k in self.KEYS}

    def record(self, **kwargs):
        for key, value in kwargs.items():
            if key in self.logs:
                self.logs[key].append(value)

    def summary(self) -> Dict[str, float]:
        def agg(key, fn):
            if not self.logs[key]:
                return 0.0
            try:
                return float(fn(np.array(self.logs[key])))
            except Exception:
                return 0.0

        return {
            "total_operations": len(self.logs["step"]),
            "max_chi": agg("chi_after", np.max),
            "mean_chi": agg("chi_after", np.mean),
            "sum_eps2": agg("eps_local", lambda x: (x ** 2).sum()),
            "max_eps": agg("eps_local", np.max),
            "mean_entropy": agg("entropy", np.mean),
            "p95_entropy": agg("entropy", lambda x: np.percentile(x, 95)),
            "total_time_ms": agg("ms_elapsed", np.sum),
            "cuda_svd_pct": self._driver_percentage("torch_cuda"),
            "cpu_fallback_pct": self._driver_percentage("torch_cpu"),
        }

    def _driver_percentage(self, name: str) -> float:
        d = self.logs["svd_driver"]
        return 100.0 * sum(1 for x in d if x == name) / len(d) if d else 0.0

    def global_error_bound(self) -> float:
        e = self.logs["eps_local"]
        return math.sqrt(sum(x ** 2 for x in e)) if e else 0.0

    def reset(self):
        for key in self.logs:
            self.logs[key].clear()
