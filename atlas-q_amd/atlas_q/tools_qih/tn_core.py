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

The gate builders follow the reference's definitions (tn_core.py:61-167),
written from their closed forms; ``mps_init_plus`` / ``mps_apply_1q`` keep
the reference's state and convention (tn_core.py:170-183).
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
# The gate family of tn_core.py:61-167 (names, angle conventions and the
# entangler's composition are the reference's interface).  Built here from
# closed forms: a rotation about Pauli axis P is cos(t/2) I - i sin(t/2) P,
# evaluated in float64 from the angle rounded to float32 (the reference
# evaluates its angles in float32, so the two agree to float32 rounding).
_PAULI = {
    "x": np.array([[0, 1], [1, 0]], np.complex128),
    "y": np.array([[0, -1j], [1j, 0]], np.complex128),
    "z": np.array([[1, 0], [0, -1]], np.complex128),
}
_ZZ = np.kron(_PAULI["z"], _PAULI["z"])
_CZ = np.diag([1.0, 1.0, 1.0, -1.0]).astype(np.complex128)
_CNOT = np.eye(4, dtype=np.complex128)[[0, 1, 3, 2]]     # |10> <-> |11>


def _half_angle(t):
    h = 0.5 * float(np.float32(t))
    return np.cos(h), np.sin(h)


def _pauli_rotation(axis: str, t) -> np.ndarray:
    c, s = _half_angle(t)
    return c * np.eye(2, dtype=np.complex128) - 1j * s * _PAULI[axis]


def _to_torch(M: np.ndarray, device, dtype) -> torch.Tensor:
    return torch.as_tensor(M, dtype=torch.complex128).to(device=device, dtype=dtype)


def Rx(t, device, dtype):
    return _to_torch(_pauli_rotation("x", t), device, dtype)


def Ry(t, device, dtype):
    return _to_torch(_pauli_rotation("y", t), device, dtype)


def Rz(t, device, dtype):
    return _to_torch(_pauli_rotation("z", t), device, dtype)


def CZ_gate(device, dtype):
    return _to_torch(_CZ, device, dtype)


def CNOT_gate(device, dtype):
    return _to_torch(_CNOT, device, dtype)


def RZZ_gate(theta, device, dtype):
    """exp(-i theta/2 Z x Z) = cos(theta/2) I - i sin(theta/2) Z x Z."""
    return _to_torch(_zz_rotation(theta), device, dtype)


def rand_su4_gate(device, dtype):
    """A random SU(4) gate drawn from torch's generator as the reference draws
    it (tn_core.py:111-119: real then imaginary Gaussian parts, the Q factor
    of their QR, determinant phase divided out), so a seed gives its gate."""
    re = torch.randn(4, 4, device=device)
    im = torch.randn(4, 4, device=device)
    Q = torch.linalg.qr(torch.complex(re, im))[0]
    phi = torch.angle(torch.linalg.det(Q))
    return (Q * torch.polar(torch.ones_like(phi), -phi / 4)).to(dtype)


def _euler_zxz(a, b, c) -> np.ndarray:
    return _pauli_rotation("z", a) @ _pauli_rotation("x", b) @ _pauli_rotation("z", c)


def _zz_rotation(theta) -> np.ndarray:
    c, s = _half_angle(theta)
    return c * np.eye(4, dtype=np.complex128) - 1j * s * _ZZ


_ENTANGLING = {"cz": lambda th: _CZ, "cnot": lambda th: _CNOT, "rzz": _zz_rotation}


def build_entangler(entangler: str, theta: float, device, dtype):
    """Two layers of Rz Rx Rz rotations on both qubits around an entangling
    gate, rotation angles theta * (1 .. 12) in order (tn_core.py:123-167);
    "none" gives the identity and "su4" a random gate."""
    if entangler == "none":
        return torch.eye(4, device=device, dtype=dtype)
    if entangler == "su4":
        return rand_su4_gate(device, dtype)
    if entangler not in _ENTANGLING:
        raise ValueError(f"Unknown entangler: {entangler}")
    ang = [theta * q for q in range(1, 13)]
    before = np.kron(_euler_zxz(*ang[0:3]), _euler_zxz(*ang[3:6]))
    after = np.kron(_euler_zxz(*ang[6:9]), _euler_zxz(*ang[9:12]))
    return _to_torch(after @ _ENTANGLING[entangler](theta) @ before, device, dtype)


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
