"""Synthetic steady-state workloads for the two-site update.

A *slice* is a run of consecutive sites whose bonds (including the two open
ends) all have dimension chi, as in the middle of a long chain at its cap.
Site tensors are Lambda_{q-1} G_q with G_q a random right-isometry
(sum_s G_s G_s^H = I) and Lambda a diagonal Schmidt profile:

* "flat":  weights of a saturated random circuit (Marchenko-Pastur-like,
           sigma_i ~ sqrt(1 - i/chi)), where the rank rule keeps the cap;
* "decay": sigma_i = exp(-i / tau) + 1e-6 exp(-i / chi), a low-entanglement
           state where the kept rank is far below chi (the regime the learned
           width targets).
"""
from __future__ import annotations

import numpy as np
import torch


def schmidt_profile(chi: int, kind: str = "flat", tau: float = 24.0) -> np.ndarray:
    i = np.arange(chi, dtype=np.float64)
    if kind == "flat":
        s = np.sqrt(np.clip(1.0 - i / chi, 1e-3, None))
    else:
        # exponential decay onto a floor at 1e-6 sigma_0 (distinct values): a
        # truncated float32 state holds no Schmidt weight below ~1e-7 of the
        # largest, and an unfloored profile reaches float32 subnormals
        s = np.exp(-i / tau) + 1e-6 * np.exp(-i / chi)
    return s / np.linalg.norm(s)


def random_right_isometry(chi_l: int, chi_r: int, gen: torch.Generator) -> torch.Tensor:
    """G [chi_l, 2, chi_r] with orthonormal rows over (s, r) (needs chi_l <= 2 chi_r),
    drawn and factored on the generator's device."""
    dev = gen.device
    Z = torch.complex(torch.randn(2 * chi_r, chi_l, generator=gen, dtype=torch.float64, device=dev),
                      torch.randn(2 * chi_r, chi_l, generator=gen, dtype=torch.float64, device=dev))
    if dev.type == "cuda":
        from . import kernels as K
        Q, _ = K.qr(Z)                 # the library's Householder QR
    else:
        Q, _ = torch.linalg.qr(Z)      # host setup draws (LAPACK on the host)
    return Q.T.conj().reshape(chi_l, 2, chi_r)


def slice_tensors(sites: int, chi: int, kind: str = "flat", seed: int = 0, tau: float = 24.0,
                  device="cuda", dtype=torch.complex64, draw_on_device: bool = False):
    """``sites`` tensors Lambda G [chi, 2, chi].  The random isometries are
    drawn from a CPU generator (reproducible across machines) unless
    ``draw_on_device``, which draws and factors them on ``device`` (fast for
    long slices at large chi; reproducible on one device type)."""
    gdev = torch.device(device) if draw_on_device else torch.device("cpu")
    gen = torch.Generator(device=gdev).manual_seed(seed)
    lam = torch.from_numpy(schmidt_profile(chi, kind, tau)).to(torch.complex128).to(gdev)
    out = []
    for _ in range(sites):
        G = random_right_isometry(chi, chi, gen)
        out.append((lam[:, None, None] * G).to(dtype).to(device))
    return out


def weak_gate(rng: np.random.Generator, angle: float) -> np.ndarray:
    """exp(-i angle H) for a random Hermitian 4x4 H (GUE-like)."""
    H = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    H = (H + H.conj().T) / 2
    w, v = np.linalg.eigh(H)
    return (v * np.exp(-1j * angle * w)) @ v.conj().T


def gate_schedule(sites: int, count: int, seed: int, angle: float = None):
    """Brickwork order over the slice's interior bonds; Haar SU(4) gates, or
    weak gates exp(-i angle H) when angle is given."""
    from .circuits import haar_su4
    rng = np.random.default_rng(seed)
    out, layer = [], 0
    while len(out) < count:
        for b in range(layer % 2, sites - 1, 2):
            out.append((b, haar_su4(rng) if angle is None else weak_gate(rng, angle)))
        layer += 1
    return out[:count]
