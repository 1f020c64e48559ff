"""MPO operations next to the two-site path (the reference's
``src/atlas_q/mpo_ops.py``), on the HIP device.

* ``MPO``: the container the functions below read -- ``tensors`` ([l, s, s',
  r] per site) and ``n_sites`` (plus an optional dense ``full_matrix``).
* ``apply_mpo_to_mps`` (mpo_ops.py:1034-1082): the site-wise contraction
  M[(l a), s, (r b)] = sum_s' W[l, s, s', r] A[a, s', b] (``aq_mpo_apply_site``)
  into a new ``AdaptiveMPS``, then its left-canonical QR sweep -- as in the
  reference, the sweep does not truncate.
* ``expectation_value`` (mpo_ops.py:1083-1150): <psi|O|psi> / <psi|psi> by the
  reference's transfer E' = einsum("Lab,atr,LstR,bsB->RrB", E, conj(A), W, A)
  per site, evaluated as Y_L = E_L [A_0 | A_1] and Z[L, t] = conj(A_t)^T Y_L
  on the MFMA GEMM and the weighted sum over (L, s, t) in
  ``aq_mpo_env_combine``; the dense ``full_matrix`` shortcut for n <= 12 as in
  the reference.
"""
from __future__ import annotations

from typing import List, Optional

import torch

from . import kernels as K


class MPO:
    """Matrix product operator: tensors[i] of shape [l, s, s', r]."""

    def __init__(self, tensors: List[torch.Tensor], n_sites: Optional[int] = None):
        self.tensors = list(tensors)
        self.n_sites = len(self.tensors) if n_sites is None else n_sites
        self.full_matrix = None


def apply_mpo_to_mps(mpo, mps, chi_max: int = 128, eps: float = 1e-8):
    """|psi'> = O |psi> (mpo_ops.py:1034-1082)."""
    from .adaptive_mps import AdaptiveMPS
    assert mpo.n_sites == mps.num_qubits, "MPO and MPS must have same number of sites"
    n = mpo.n_sites
    device = mps.tensors[0].device
    new_tensors = [K.mpo_apply_site(mpo.tensors[i], mps.tensors[i]) for i in range(n)]
    result = AdaptiveMPS(n, bond_dim=2, device=device)
    result.tensors = new_tensors
    result.to_left_canonical()
    return result


def _transfer(E: torch.Tensor, A: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """E'[R, r, B] = sum E[L,a,b] conj(A[a,t,r]) W[L,s,t,R] A[b,s,B]."""
    DL, ca, cb = E.shape
    chir, chiB = A.shape[2], A.shape[2]
    Am = A.reshape(A.shape[0], 2 * chiB)                   # [b, (s, B)]
    Y = torch.empty((DL, ca, 2 * chiB), dtype=A.dtype, device=A.device)
    for L in range(DL):
        K.cgemm(E[L], Am, out=Y[L])
    Z = torch.empty((DL, 2, chir, 2 * chiB), dtype=A.dtype, device=A.device)
    A3 = A.reshape(A.shape[0], 2, chir)
    for L in range(DL):
        for t in range(2):
            K.cgemm(A3[:, t, :].T, Y[L], conj_a=True, out=Z[L, t])
    return K.mpo_env_combine(Z, W)


def expectation_value(mpo, mps, use_gpu_optimized: bool = True) -> complex:
    """<psi|O|psi> / <psi|psi> (mpo_ops.py:1083-1150)."""
    n = mpo.n_sites
    assert n == mps.num_qubits
    dtype = mps.tensors[0].dtype
    device = mps.tensors[0].device
    if getattr(mpo, "full_matrix", None) is not None and n <= 12:
        psi = mps.to_statevector().to(device=device, dtype=dtype)
        nrm = torch.linalg.vector_norm(psi)
        if float(nrm) == 0:
            raise ValueError("Statevector has zero norm.")
        psi = (psi / nrm).reshape(-1, 1)
        H = mpo.full_matrix.to(device=device, dtype=dtype)
        Hpsi = K.cgemm(H, psi)
        e = K.cgemm(psi.T, Hpsi, conj_a=True)
        return complex(e.reshape(-1)[0].item())
    E = torch.ones((1, 1, 1), dtype=dtype, device=device)
    for i in range(n):
        W = mpo.tensors[i].to(device=device, dtype=dtype)
        E = _transfer(E, mps.tensors[i].to(dtype), W)
    I_t = torch.zeros((1, 2, 2, 1), dtype=dtype, device=device)
    I_t[0, 0, 0, 0] = 1.0
    I_t[0, 1, 1, 0] = 1.0
    En = torch.ones((1, 1, 1), dtype=dtype, device=device)
    for i in range(n):
        En = _transfer(En, mps.tensors[i].to(dtype), I_t)
    num = E.reshape(-1)[0]
    den = En.reshape(-1)[0]
    return complex((num / den).item())
