"""Hermitian eigen-decomposition dispatch for the truncated-SVD cores.

Orders up to 64 run in the library's batched FP64 Jacobi kernel
(``aq_heev_small``: one work-group, matrix in LDS, no launch-latency chain);
larger orders use rocSOLVER's divide-and-conquer through ``torch.linalg.eigh``.
Both return eigenvalues in descending order with matching eigenvector columns.
"""
from __future__ import annotations

import torch

JACOBI_MAX = 64


def eigh_desc(G: torch.Tensor):
    n = G.shape[-1]
    if n <= JACOBI_MAX:
        from .rank_mlp import heev_small
        return heev_small(G)
    lam, V = torch.linalg.eigh(G)
    return lam.flip(-1), V.flip(-1)
