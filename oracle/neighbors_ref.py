"""torch-CPU restatement of the reference's callers next to the two-site path
(the oracle): MPO x MPS application, MPO expectation value, environment
scans and the TDVP two-site split.

TEST INFRASTRUCTURE ONLY -- imported by ``tests/`` and never by the shipped
package.  Each function cites the reference file:line it restates.

Parity pinning (tests/test_neighbors_oracle.py): ``apply_mpo_to_mps`` and
``expectation_value`` against the dense operator built from the MPO (state
vector O psi and psi^H O^T psi, the latter being what the reference's index
order computes), ``build_left_envs`` / ``build_right_envs`` against explicit
matrix products with the renormalisation applied by hand,
``tdvp_two_site_split`` against a hand-computed spectrum (threshold (1 - eps),
not (1 - eps^2)) and the U S Vh identity.
"""
from __future__ import annotations

from typing import List

import torch


def mpo_apply_sites(mpo_tensors: List[torch.Tensor], mps_tensors: List[torch.Tensor]):
    """mpo_ops.py:1055-1072: M[l a, s, r b] = sum_s' W[l,s,s',r] A[a,s',b]."""
    out = []
    for W, A in zip(mpo_tensors, mps_tensors):
        W = W.to(A.dtype)
        M = torch.einsum("lstr, atb -> lasrb", W, A)
        l, a, s, r, b = M.shape
        out.append(M.reshape(l * a, s, r * b))
    return out


def expectation_value(mpo_tensors: List[torch.Tensor], mps_tensors: List[torch.Tensor]) -> complex:
    """mpo_ops.py:1118-1150 (standard contraction path)."""
    dtype = mps_tensors[0].dtype
    E = torch.ones(1, 1, 1, dtype=dtype)
    for W, A in zip(mpo_tensors, mps_tensors):
        W = W.to(dtype)
        E = torch.einsum("Lab, atr, LstR, bsB -> RrB", E, A.conj(), W, A)
    num = E[0, 0, 0] if E.numel() > 1 else E
    I_t = torch.zeros(1, 2, 2, 1, dtype=dtype)
    I_t[0, 0, 0, 0] = 1.0
    I_t[0, 1, 1, 0] = 1.0
    En = torch.ones(1, 1, 1, dtype=dtype)
    for A in mps_tensors:
        En = torch.einsum("Lab, atr, LstR, bsB -> RrB", En, A.conj(), I_t, A)
    den = En[0, 0, 0] if En.numel() > 1 else En
    return complex((num / den).reshape(-1)[0].item())


def build_left_envs(E: torch.Tensor, eye: torch.Tensor, tile_size: int = 64,
                    normalize: bool = True) -> torch.Tensor:
    """sim_env_scan.py:53-90."""
    B, L_len, chi, _ = E.shape
    L_envs = torch.zeros(B, L_len + 1, chi, chi, dtype=E.dtype)
    L_envs[:, 0] = eye.unsqueeze(0).expand(B, -1, -1)
    for k in range(L_len):
        L_envs[:, k + 1] = torch.bmm(L_envs[:, k], E[:, k])
        if normalize and (k + 1) % tile_size == 0:
            norms = torch.norm(L_envs[:, k + 1], p="fro", dim=(-2, -1),
                               keepdim=True).clamp_min(1e-12)
            L_envs[:, k + 1] = L_envs[:, k + 1] / norms
    return L_envs


def build_right_envs(E: torch.Tensor, eye: torch.Tensor, tile_size: int = 64,
                     normalize: bool = True) -> torch.Tensor:
    """sim_env_scan.py:93-130."""
    B, L_len, chi, _ = E.shape
    R_envs = torch.zeros(B, L_len + 1, chi, chi, dtype=E.dtype)
    R_envs[:, L_len] = eye.unsqueeze(0).expand(B, -1, -1)
    for k in range(L_len - 1, -1, -1):
        R_envs[:, k] = torch.bmm(E[:, k], R_envs[:, k + 1])
        if normalize and k % tile_size == 0:
            norms = torch.norm(R_envs[:, k], p="fro", dim=(-2, -1),
                               keepdim=True).clamp_min(1e-12)
            R_envs[:, k] = R_envs[:, k] / norms
    return R_envs


def tdvp_two_site_split(theta: torch.Tensor, chi_max: int, eps_bond: float):
    """tdvp.py:546-562 (robust_svd's first attempt: LAPACK SVD)."""
    shape = theta.shape
    X = theta.reshape(shape[0] * shape[1], shape[2] * shape[3])
    U, S, Vh = torch.linalg.svd(X, full_matrices=False)
    k = min(len(S), chi_max)
    S2 = S ** 2
    cumsum = torch.cumsum(S2, dim=0)
    total = cumsum[-1]
    k_trunc = torch.searchsorted(cumsum, (1 - eps_bond) * total).item() + 1
    k = min(k, k_trunc)
    A = (U[:, :k] * S[:k]).reshape(shape[0], shape[1], k)
    B = Vh[:k, :].reshape(k, shape[2], shape[3])
    return A, B, k
