"""Exact NumPy state-vector simulator (n <= 20) -- the mathematical check of
the MPS path.  TEST INFRASTRUCTURE ONLY.

Conventions (those of the reference MPS, adaptive_mps.py:177,
mps_complex.py:343-368, mps_pytorch.py:614):
* amplitude index: qubit 0 is the most significant bit;
* two-site gate ``U[4,4]`` acts as ``psi'[.., p, q, ..] = sum U[2p+q, 2p1+p2]
  psi[.., p1, p2, ..]`` (row = output pair, left qubit = high bit);
* ``AdaptiveMPS.apply_single_qubit_gate`` applies ``U^T``:
  ``psi'[.., t, ..] = sum_s U[s, t] psi[.., s, ..]``.

Parity pinning: GHZ/Bell closed forms and unitarity (norm) invariants in
``tests/test_oracle.py``.
"""
from __future__ import annotations

import numpy as np


def zero_state(n: int) -> np.ndarray:
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1.0
    return psi


def apply_1q_transposed(psi: np.ndarray, n: int, q: int, U: np.ndarray) -> np.ndarray:
    """psi'[..t..] = sum_s U[s,t] psi[..s..] (AdaptiveMPS convention)."""
    t = psi.reshape(2 ** q, 2, 2 ** (n - q - 1))
    return np.einsum("st,asb->atb", np.asarray(U, np.complex128), t).reshape(-1)


def apply_2q(psi: np.ndarray, n: int, i: int, U4: np.ndarray) -> np.ndarray:
    """Gate on qubits (i, i+1), U convention (row = output pair index)."""
    t = psi.reshape(2 ** i, 4, 2 ** (n - i - 2))
    return np.einsum("uv,avb->aub", np.asarray(U4, np.complex128).reshape(4, 4), t).reshape(-1)


def apply_2q_transposed(psi: np.ndarray, n: int, i: int, U4: np.ndarray) -> np.ndarray:
    """The reference's einsum fallback convention (applies U^T)."""
    return apply_2q(psi, n, i, np.asarray(U4).reshape(4, 4).T)


def run_circuit(n: int, ops, convention: str = "U") -> np.ndarray:
    """ops: sequence of ("1q", q, U2) / ("2q", i, U4)."""
    psi = zero_state(n)
    for op in ops:
        if op[0] == "1q":
            psi = apply_1q_transposed(psi, n, op[1], op[2])
        elif convention == "U":
            psi = apply_2q(psi, n, op[1], op[2])
        else:
            psi = apply_2q_transposed(psi, n, op[1], op[2])
    return psi


def marginal_p0(psi: np.ndarray, n: int, prefix_bits) -> float:
    """P(qubit len(prefix) = 0 | qubits 0..len(prefix)-1 = prefix), Born rule."""
    k = len(prefix_bits)
    idx = 0
    for b in prefix_bits:
        idx = (idx << 1) | int(b)
    p = np.abs(psi.reshape(2 ** k, 2, -1)[idx]) ** 2
    tot = p.sum()
    return float(p[0].sum() / tot) if tot > 0 else 0.5
