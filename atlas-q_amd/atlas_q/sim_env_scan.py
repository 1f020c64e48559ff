"""Left / right environment scans (the reference's
``src/atlas_q/sim_env_scan.py:53-130``) on the HIP device.

L_{k+1} = L_k E_k (left to right, L_0 = I) and R_k = E_k R_{k+1} (right to
left, R_L = I) for a batch of B chains of transfer matrices E [B, L, chi,
chi]; every step is one grouped MFMA GEMM over the batch
(``aq_grouped_gemm``), and every ``tile_size`` steps the new environments are
divided by their Frobenius norm (at least 1e-12) in
``aq_fro_normalize_batched``, as the reference renormalises to keep the
products finite.
"""
from __future__ import annotations

import torch

from . import kernels as K


def build_left_envs(E: torch.Tensor, eye: torch.Tensor, tile_size: int = 64,
                    normalize: bool = True) -> torch.Tensor:
    B, L_len, chi, _ = E.shape
    E = E.contiguous()
    L_envs = torch.zeros(B, L_len + 1, chi, chi, dtype=E.dtype, device=E.device)
    L_envs[:, 0] = eye.to(E.dtype).unsqueeze(0).expand(B, -1, -1)
    for k in range(L_len):
        K.grouped_gemm([(L_envs[b, k], E[b, k], L_envs[b, k + 1], False, False)
                        for b in range(B)], E.dtype, E.device)
        if normalize and (k + 1) % tile_size == 0:
            blk = L_envs[:, k + 1].contiguous()
            K.fro_normalize_(blk)
            L_envs[:, k + 1] = blk
    return L_envs


def build_right_envs(E: torch.Tensor, eye: torch.Tensor, tile_size: int = 64,
                     normalize: bool = True) -> torch.Tensor:
    B, L_len, chi, _ = E.shape
    E = E.contiguous()
    R_envs = torch.zeros(B, L_len + 1, chi, chi, dtype=E.dtype, device=E.device)
    R_envs[:, L_len] = eye.to(E.dtype).unsqueeze(0).expand(B, -1, -1)
    for k in range(L_len - 1, -1, -1):
        K.grouped_gemm([(E[b, k], R_envs[b, k + 1], R_envs[b, k], False, False)
                        for b in range(B)], E.dtype, E.device)
        if normalize and k % tile_size == 0:
            blk = R_envs[:, k].contiguous()
            K.fro_normalize_(blk)
            R_envs[:, k] = blk
    return R_envs
