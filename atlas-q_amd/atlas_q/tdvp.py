"""The truncating split of the reference's two-site TDVP step
(``TDVP2Site._evolve_two_site``, src/atlas_q/tdvp.py:530-562) on the HIP
device.  The Krylov time evolution of the two-site tensor with the MPO
environments (tdvp.py:480-528) is outside this build; ``two_site_split``
takes the evolved tensor Theta [chiL, 2, 2, chiR] and returns the new cores:

  U, S, Vh = SVD(Theta as [2 chiL, 2 chiR])                (robust_svd)
  k = min(len(S), chi_max, searchsorted(cumsum(S^2), (1 - eps) total) + 1)
  A = (U[:, :k] S[:k]) as [chiL, 2, k], B = Vh[:k] as [k, 2, chiR]

with the threshold (1 - eps) -- not (1 - eps^2) as in the gate update's
rule -- in the rank kernel's ``AQ_RANK_TDVP`` mode and the rebuild kernel.
"""
from __future__ import annotations

import torch

from . import _ffi
from . import kernels as K
from .linalg_robust import robust_svd


def two_site_split(theta: torch.Tensor, chi_max: int, eps_bond: float):
    """Returns (A [chiL, 2, k], B [k, 2, chiR], k)."""
    shape = theta.shape
    X = theta.reshape(shape[0] * shape[1], shape[2] * shape[3])
    U, S, Vh, _ = robust_svd(X)
    k = int(K.rank_select(S, eps_bond, int(chi_max), _ffi.AQ_RANK_TDVP).k)
    A, B = K.rebuild(U, S, Vh, k, shape[0], shape[3])
    return A, B, k


def evolve_split_into(mps, site: int, theta_evolved: torch.Tensor, chi_max: int,
                      eps_bond: float) -> int:
    """tdvp.py:546-562 applied to ``mps``: writes the split cores to sites
    (site, site + 1) and the bond dimension; returns k."""
    A, B, k = two_site_split(theta_evolved, chi_max, eps_bond)
    mps.tensors[site] = A
    mps.tensors[site + 1] = B
    if site < len(mps.bond_dims):
        mps.bond_dims[site] = k
    return k
