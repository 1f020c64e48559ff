"""SVD / QR with the reference's retry cascade (``src/atlas_q/linalg_robust.py``).

The SVD runs on the HIP device (rocSOLVER through torch: it is the plain
method by definition), the QR through the library's Householder kernels.
The reference's last resort -- a CPU SVD/QR (linalg_robust.py:49-52, 72-76) --
is replaced by a third device attempt with a relative jitter of 1e-9 max|X|
(rocSOLVER's SVD can fail to converge on exactly degenerate spectra, e.g.
product-state blocks of CZ-ladder circuits, where the reference's absolute
1e-12 jitter is not enough); a failure after it raises.
"""
from __future__ import annotations

from typing import Tuple

import torch


def _require_device(X: torch.Tensor) -> None:
    if X.device.type != "cuda":
        raise RuntimeError("atlas_q: linear algebra runs on the HIP device only "
                           f"(got a tensor on {X.device})")


def robust_svd(X: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor, str]:
    """linalg_robust.py:17-52 without the CPU step: (U, S, Vh, driver)."""
    _require_device(X)
    try:
        U, S, Vh = torch.linalg.svd(X, full_matrices=False)
        return U, S, Vh, "torch_cuda"
    except Exception:
        pass
    Xj = X + (1e-12 * torch.randn_like(X))
    try:
        U, S, Vh = torch.linalg.svd(Xj, full_matrices=False)
        return U, S, Vh, "torch_cuda_jitter"
    except Exception:
        pass
    scale = float(X.abs().max()) if X.numel() else 0.0
    Xj = X + ((1e-9 * scale) * torch.randn_like(X))
    try:
        U, S, Vh = torch.linalg.svd(Xj, full_matrices=False)
    except Exception as e:
        raise RuntimeError(f"atlas_q: device SVD failed after jitter retries: {e}") from e
    return U, S, Vh, "torch_cuda_jitter_rel"


def robust_qr(X: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor, str]:
    """linalg_robust.py:55-76: (Q, R, driver), by the library's blocked
    Householder QR on the device (``kernels.qr``; LAPACK's zgeqrf/zungqr
    algorithm, so the factors are the ones the reference's torch.linalg.qr
    returns up to rounding).  Householder QR does not fail on rank-deficient
    input, so the reference's retry cascade has nothing to retry."""
    _require_device(X)
    from . import kernels as K
    Q, R = K.qr(X)
    return Q, R, "atlas_householder"


def condition_number(S: torch.Tensor) -> float:
    """linalg_robust.py:79-91."""
    if len(S) == 0 or S[-1] == 0:
        return float("inf")
    return float((S[0] / S[-1]).item())
