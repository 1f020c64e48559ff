"""The reference's tensor-network core (tools_qih/tn_core.py) on the HIP device.

``mps_apply_2q`` (tn_core.py:186-384) is the reference's second two-site
update, the one its learned rank predictors plug into:

* theta in the U^T convention T'[l, u, r] = sum_d T[l, d, r] U[d, u]
  (tn_core.py:224-230) -- the library's MFMA theta kernel with the gate
  transposed;
* ``svd_robust``: the SVD of theta normalised by its Frobenius norm, driver
  cascade requested -> gesvda -> gesvdj -> gesvd, each plain then with a
  1e-6 relative jitter (tn_core.py:18-58), factorised by rocSOLVER through
  torch; the reference's last resort, a CPU SVD, is not taken (a device
  failure raises);
* the SVD event (``S_top128``, ``adaptive_keep``, ...) to the reference's
  JSONL.gz log when ``QIH_SVD_LOG_DIR`` is set (tn_core.py:248-279);
* the truncation -- adaptive rule, AI policy (``QIH_AI_POLICY`` /
  ``QIH_AI_ALPHA``), fixed cap, S > eps filter, truncation error -- in the
  ``aq_tn_truncate`` kernel, with the predictor's ``predict(S)`` called on
  the device spectrum as the reference calls it (tn_core.py:282-372);
* the sqrt(S) split into both cores in the ``aq_rebuild_sqrt`` kernel
  (tn_core.py:374-384).

The gate builders and ``mps_init_plus`` / ``mps_apply_1q`` keep the
reference's definitions (tn_core.py:62-183).
"""
from __future__ import annotations

import os

import numpy as np
import torch

from .. import kernels as K
from . import svd_logger

SVD_EVENT_LOGGER = svd_logger.get_logger()


# ---------------------------------------------------------------- SVD
def svd_robust(T: torch.Tensor, driver: str = "gesvdj"):
    """tn_core.py:18-58 on the device (no CPU step)."""
    tried = []

    def _try(drv, jitter):
        X = T
        if jitter:
            eps = (T.norm() + 1e-12) * 1e-6
            noise = (torch.randn_like(T.real) + 1j * torch.randn_like(T.real)) * eps
            X = T + noise.to(T.dtype)
        scale = X.norm()
        pos = float(scale) > 0
        Xn = X / scale if pos else X
        try:
            U, S, Vh = torch.linalg.svd(Xn, full_matrices=False, driver=drv)
        except (RuntimeError, ValueError, TypeError):
            if drv is None:
                raise
            # torch's ROCm build may not expose the cuSOLVER driver names
            U, S, Vh = torch.linalg.svd(Xn, full_matrices=False)
        if pos:
            S = S * scale
        return U, S, Vh

    for drv in [driver, "gesvda", "gesvdj", "gesvd"]:
        if drv in tried:
            continue
        for jitter in (False, True):
            try:
                return _try(drv, jitter)
            except Exception as e:  # noqa: BLE001 -- the reference's cascade catches all
                tried.append(f"{drv}{'+j' if jitter else ''}:{type(e).__name__}")
    raise RuntimeError("All SVD attempts failed: " + " ; ".join(tried))


# ---------------------------------------------------------------- gates
def Rx(t, device, dtype):
    t = torch.as_tensor(t, device=device, dtype=torch.float32)
    c, s = torch.cos(t / 2), torch.sin(t / 2)
    M = torch.empty((2, 2), device=device, dtype=dtype)
    M[0, 0] = c
    M[0, 1] = -1j * s
    M[1, 0] = -1j * s
    M[1, 1] = c
    return M


def Rz(t, device, dtype):
    t = torch.as_tensor(t, device=device, dtype=torch.float32)
    c, s = torch.cos(t / 2), torch.sin(t / 2)
    M = torch.zeros((2, 2), device=device, dtype=dtype)
    M[0, 0] = (c - 1j * s).to(dtype)
    M[1, 1] = (c + 1j * s).to(dtype)
    return M


def Ry(t, device, dtype):
    t = torch.as_tensor(t, device=device, dtype=torch.float32)
    c, s = torch.cos(t / 2), torch.sin(t / 2)
    M = torch.empty((2, 2), device=device, dtype=dtype)
    M[0, 0] = c
    M[0, 1] = -s
    M[1, 0] = s
    M[1, 1] = c
    return M


def CZ_gate(device, dtype):
    return torch.diag(torch.tensor([1, 1, 1, -1], device=device, dtype=dtype))


def CNOT_gate(device, dtype):
    M = torch.zeros((4, 4), device=device, dtype=dtype)
    M[0, 0] = M[1, 1] = M[2, 3] = M[3, 2] = 1
    return M


def RZZ_gate(theta, device, dtype):
    """exp(-i theta/2 Z x Z)."""
    t = torch.as_tensor(theta, device=device, dtype=torch.float32)
    c, s = torch.cos(t / 2).to(dtype), torch.sin(t / 2).to(dtype)
    return torch.diag(torch.stack([c - 1j * s, c + 1j * s, c + 1j * s, c - 1j * s]))


def rand_su4_gate(device, dtype):
    """QR of a complex Gaussian, determinant phase removed (tn_core.py:111-119)."""
    A = torch.randn(4, 4, device=device) + 1j * torch.randn(4, 4, device=device)
    Q, _ = torch.linalg.qr(A)
    phase = torch.angle(torch.linalg.det(Q))
    return (Q * torch.exp(-1j * phase / 4)).to(dtype)


def build_entangler(entangler: str, theta: float, device, dtype):
    """U_right @ E @ U_left with Rz Rx Rz locals at angles theta*(k+1),
    k = 0..11 (tn_core.py:123-167); "none" is the identity, "su4" random."""
    if entangler == "none":
        return torch.eye(4, device=device, dtype=dtype)
    if entangler == "su4":
        return rand_su4_gate(device, dtype)
    a, b, c, d, e, f, g, h, i, j, k, l = [theta * (n + 1) for n in range(12)]
    U_left = torch.kron(Rz(a, device, dtype) @ Rx(b, device, dtype) @ Rz(c, device, dtype),
                        Rz(d, device, dtype) @ Rx(e, device, dtype) @ Rz(f, device, dtype))
    if entangler == "cz":
        E = CZ_gate(device, dtype)
    elif entangler == "cnot":
        E = CNOT_gate(device, dtype)
    elif entangler == "rzz":
        E = RZZ_gate(theta, device, dtype)
    else:
        raise ValueError(f"Unknown entangler: {entangler}")
    U_right = torch.kron(Rz(g, device, dtype) @ Rx(h, device, dtype) @ Rz(i, device, dtype),
                         Rz(j, device, dtype) @ Rx(k, device, dtype) @ Rz(l, device, dtype))
    return U_right @ E @ U_left


# ---------------------------------------------------------------- MPS
def mps_init_plus(n, device="cuda", dtype=torch.complex64):
    """|+>^n as n cores [1, 2, 1]."""
    v = torch.tensor([1 / np.sqrt(2), 1 / np.sqrt(2)], device=device, dtype=dtype)
    return [v.view(1, 2, 1).clone() for _ in range(n)]


def mps_apply_1q(cores, i, U):
    """cores[i][l, u, r] = sum_d cores[i][l, d, r] U[d, u] (HIP kernel)."""
    A = cores[i]
    dt = torch.promote_types(A.dtype, torch.as_tensor(U).dtype)
    cores[i] = K.apply_1q(A.to(dt), K.gate_to_host(U, (2, 2)))
    return cores


def _policy():
    policy = os.getenv("QIH_AI_POLICY", "guardrail").lower()
    try:
        alpha = float(os.getenv("QIH_AI_ALPHA", "0.7"))
    except Exception:  # noqa: BLE001
        alpha = 0.7
    return policy, max(0.0, min(1.0, alpha))


@torch.no_grad()
def mps_apply_2q(cores, i, j, U, chi_max=None, eps=1e-10, svd_driver="gesvdj", adaptive=False,
                 tol=1e-4, ai_predictor=None):
    """Apply the 4x4 gate U to adjacent sites i, j = i + 1 and truncate.

    Returns (cores, chi, truncation_error) as tn_core.py:186-384: chi the new
    bond dimension, truncation_error = 1 - kept/total spectral weight."""
    assert j == i + 1, "Sites must be adjacent"
    Ai, Aj = cores[i], cores[j]
    li, _, ri = Ai.shape
    r_i, _, rj = Aj.shape
    assert ri == r_i, "Bond dimensions must match"
    Ut = torch.as_tensor(U)
    dt = torch.promote_types(torch.promote_types(Ai.dtype, Aj.dtype), Ut.dtype)
    gT = K.gate_to_host(Ut.transpose(0, 1), (4, 4))           # U^T convention
    X = K.theta(Ai.to(dt), Aj.to(dt), gT)

    U1, S, Vh = svd_robust(X, driver=svd_driver)
    m = S.numel()

    ai_keep = -1
    policy, alpha = _policy()
    if ai_predictor is not None and m > 10:
        ai_keep = int(ai_predictor.predict(S))
        if chi_max is not None:
            ai_keep = min(ai_keep, chi_max)
    r = K.tn_truncate(S, tol, adaptive, chi_max, ai_keep, policy, alpha, eps)

    if SVD_EVENT_LOGGER.active:
        try:
            top = min(128, m)
            Sv = S[:top].detach().abs().double().tolist()
            Sv = Sv + [0.0] * (128 - top)
            SVD_EVENT_LOGGER.log(
                kind="svd_spectrum", i=int(i), j=int(j), pre_rank=int(ri),
                chi_max=int(chi_max) if chi_max is not None else None, tol=float(tol),
                adaptive=bool(adaptive), entangler=os.getenv("QIH_ENTANGLER", "unknown"),
                total_weight=float(r.total_weight), S_top128=Sv,
                adaptive_keep=int(r.adaptive_keep))
        except Exception:  # noqa: BLE001 -- logging never breaks the update
            pass

    chi = int(r.chi)
    if chi == 0:
        cores[i] = torch.zeros((li, 2, 0), dtype=X.dtype, device=X.device)
        cores[j] = torch.zeros((0, 2, rj), dtype=X.dtype, device=X.device)
        return cores, 0, float(r.truncation_error)
    if not r.prefix:
        # S > eps survivors not leading (only for an unsorted spectrum)
        keep = int(r.keep_count)
        sel = torch.nonzero(S[:keep] > eps).reshape(-1)
        U1, S, Vh = U1[:, sel], S[sel], Vh[sel, :]
    cores[i], cores[j] = K.rebuild_sqrt(U1, S, Vh, chi, li, rj)
    return cores, chi, float(r.truncation_error)
