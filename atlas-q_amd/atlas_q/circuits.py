"""Seeded synthetic circuits for the adaptive-MPS hot path.

* Haar SU(4): Z = (N(0,1) + i N(0,1))/sqrt(2) (4x4), Q,R = qr(Z),
  Q <- Q diag(R_ii/|R_ii|) (Mezzadri's phase fix, as the reference's
  ``tests/integration/test_mps_triton_kernels.py:20-30``), Q <- Q/det(Q)^(1/4).
* Brickwork: layer l acts on bonds b = l (mod 2), ascending; one gate per
  (layer, bond) drawn in that order from ``numpy.random.default_rng(seed)``.
* Hardware-efficient ansatz: per layer RY(theta) on every qubit, RZ(phi) on
  every qubit, then CZ on even bonds and on odd bonds (ordering of
  ``vqe_qaoa.py:72-85`` with an extra RZ layer), angles U[0, 2 pi).

Gate index convention: ``U[out, in]`` with pair index ``2*s_i + s_{i+1}``.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np


def haar_su4(rng: np.random.Generator) -> np.ndarray:
    z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    q = q * (d / np.abs(d))[None, :]
    q = q / np.linalg.det(q) ** 0.25
    return q.astype(np.complex128)


def brickwork(n: int, depth: int, seed: int) -> List[Tuple[int, np.ndarray]]:
    """[(bond, U4)] in application order."""
    rng = np.random.default_rng(seed)
    ops = []
    for layer in range(depth):
        for b in range(layer % 2, n - 1, 2):
            ops.append((b, haar_su4(rng)))
    return ops


def brickwork_count(n: int, depth: int) -> int:
    return sum(len(range(layer % 2, n - 1, 2)) for layer in range(depth))


def ry(theta: float) -> np.ndarray:
    c, s = np.cos(theta / 2.0), np.sin(theta / 2.0)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def rz(phi: float) -> np.ndarray:
    return np.array([[np.exp(-0.5j * phi), 0], [0, np.exp(0.5j * phi)]], dtype=np.complex128)


CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)


def hea(n: int, layers: int, seed: int):
    """[("1q", q, U2) | ("2q", bond, U4)] for the RY/RZ + CZ-ladder ansatz."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(layers):
        th = rng.uniform(0.0, 2 * np.pi, n)
        ph = rng.uniform(0.0, 2 * np.pi, n)
        for q in range(n):
            ops.append(("1q", q, ry(th[q])))
        for q in range(n):
            ops.append(("1q", q, rz(ph[q])))
        for b in range(0, n - 1, 2):
            ops.append(("2q", b, CZ))
        for b in range(1, n - 1, 2):
            ops.append(("2q", b, CZ))
    return ops


def batched_hea_population(count: int, seed: int, n_lo: int = 12, n_hi: int = 96,
                           chi_lo: int = 16, chi_hi: int = 256):
    """(n_j, chi_j) log-uniform, as the batched workload's population."""
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), count))).astype(int)
    chis = np.round(np.exp(rng.uniform(np.log(chi_lo), np.log(chi_hi), count))).astype(int)
    return [(int(a), int(b)) for a, b in zip(ns, chis)]
