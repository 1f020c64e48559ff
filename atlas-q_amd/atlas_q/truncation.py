"""Rank rule of the two-site update, on the HIP device.

API of the reference's ``src/atlas_q/truncation.py``.  ``choose_rank_from_sigma``
runs the rank rule in the HIP kernel ``aq_rank_select`` (one device->host copy
of the 48-byte result); the budget callback is evaluated on the host exactly as
the reference's loop does (truncation.py:66-67).
"""
from __future__ import annotations

import math
from typing import Callable, Tuple

import torch

from . import kernels as K


def choose_rank_from_sigma(S: torch.Tensor, eps_bond: float, chi_cap: int,
                           budget_ok: Callable[[int], bool] = lambda k: True
                           ) -> Tuple[int, float, float, float]:
    """truncation.py:21-83: (k, eps_local, entropy[nats], condS)."""
    if len(S) == 0:
        return 0, 0.0, 0.0, float("inf")
    r = K.rank_select(S, eps_bond, int(chi_cap))
    if r.flags & 1:
        return 1, 0.0, 0.0, float("inf")
    k = int(r.k)
    k0 = k
    while k > 1 and not budget_ok(k):
        k -= 1
    if k != k0:  # diagnostics of the reduced rank (k <= k_tol, so cap=k fixes it)
        r = K.rank_select(S, eps_bond, k)
    return k, float(r.eps_local), float(r.entropy), float(r.condS)


def compute_global_error_bound(local_errors: list) -> float:
    """truncation.py:86-101: sqrt(sum eps_local^2)."""
    return math.sqrt(sum(e ** 2 for e in local_errors))


def check_entropy_sanity(entropy: float, chi_left: int, chi_right: int) -> bool:
    """truncation.py:104-122: entropy <= log2(min(2 chiL, 2 chiR)) + 1e-6."""
    return entropy <= math.log2(min(chi_left * 2, 2 * chi_right)) + 1e-6
