"""Run a gate list on ``AdaptiveMPS``, sample, count bitstrings: the MPS
branch of the reference's Qiskit adapter (``adapters/qiskit_adapter.py``:
``_run_mps`` :383-405, ``_apply_gate_to_mps`` :505-550, ``_samples_to_counts``
:729-744) without Qiskit.  Instructions are ``(name, qubits, params)``
tuples, the form the adapter's own gate extraction yields (:495-503).

Counts use the reference's bitstrings (``format(sample, '0{n}b')``: qubit 0
leftmost).  The reference builds them from int64 sample values, which turn
negative at 64 qubits and lose bits beyond; here they are built from the
sampler's packed words (``sample_bits``), so every qubit count works.
"""
from __future__ import annotations

from typing import Dict, Iterable, Optional, Sequence, Tuple

import numpy as np
import torch

SINGLE_QUBIT = ("h", "x", "y", "z", "s", "sdg", "t", "tdg", "rx", "ry", "rz")
TWO_QUBIT = ("cx", "cnot", "cz", "cy", "swap")


def apply_gate(mps, gate_name: str, qubits: Sequence[int], params: Sequence = ()) -> None:
    """qiskit_adapter.py:505-550: measure / barrier are skipped; unknown
    names raise ValueError with the reference's messages."""
    name = gate_name.lower()
    if name in ("measure", "barrier"):
        return
    if len(qubits) == 1:
        q = qubits[0]
        if name not in SINGLE_QUBIT:
            raise ValueError(f"Unknown single-qubit gate: {name}")
        if name in ("rx", "ry", "rz"):
            getattr(mps, name)(q, params[0])
        else:
            getattr(mps, name)(q)
    elif len(qubits) == 2:
        q0, q1 = qubits
        if name not in TWO_QUBIT:
            raise ValueError(f"Unknown two-qubit gate: {name}")
        {"cx": mps.cnot, "cnot": mps.cnot, "cz": mps.cz, "cy": mps.cy,
         "swap": mps.swap}[name](q0, q1)


def samples_to_counts(samples, n_qubits: Optional[int] = None) -> Dict[str, int]:
    """qiskit_adapter.py:729-744 on the reference's int samples (a dict
    passes through, an unknown type gives {}).  At 64 qubits an int64 sample
    with qubit 0 set is negative; it is read as the unsigned 64-bit value it
    encodes.  Beyond 64 qubits ints have lost bits: use ``bits_to_counts``."""
    if isinstance(samples, dict):
        return samples
    if not isinstance(samples, (np.ndarray, list, tuple)):
        return {}
    if n_qubits is not None and n_qubits > 64:
        raise ValueError("int samples hold only the low 64 bits; count from sample_bits()")
    counts: Dict[str, int] = {}
    for s in samples:
        v = int(s)
        if v < 0:
            v &= (1 << 64) - 1
        key = format(v, f"0{n_qubits}b") if n_qubits else bin(v)[2:]
        counts[key] = counts.get(key, 0) + 1
    return counts


def bits_to_counts(bits: torch.Tensor, n_qubits: int) -> Dict[str, int]:
    """Counts from packed words [shots, ceil(n/64)] (the bitstring's integer
    sum_q b_q 2^(n-1-q), 64-bit words least significant first): the
    reference's keys, qubit 0 leftmost, for any n."""
    w = np.ascontiguousarray(bits.detach().cpu().numpy()).view(np.uint64)
    if w.ndim != 2 or w.shape[1] != (n_qubits + 63) // 64:
        raise ValueError(f"bits_to_counts: expected [shots, {(n_qubits + 63) // 64}] words")
    rows, inv = np.unique(w, axis=0, return_inverse=True)
    freq = np.bincount(inv.reshape(-1), minlength=len(rows))
    counts: Dict[str, int] = {}
    for row, c in zip(rows, freq):
        v = 0
        for j in range(len(row) - 1, -1, -1):
            v = (v << 64) | int(row[j])
        counts[format(v, f"0{n_qubits}b")] = int(c)
    return counts


def run_mps(gates: Iterable[Tuple[str, Sequence[int], Sequence]], n_qubits: int, shots: int,
            seed: Optional[int] = None, max_bond_dim: int = 8, device="cuda",
            born_exact: bool = False, **mps_kwargs):
    """qiskit_adapter.py:383-405: an ``AdaptiveMPS(n_qubits,
    bond_dim=max_bond_dim)`` (further keyword arguments go to it), the gates in
    order, ``shots`` samples, their counts, and the state vector for
    n_qubits <= 20 (else None).  ``seed`` seeds the sampler's uniforms (the
    reference seeds NumPy, which its torch sampler does not read).

    Sampling follows the reference's ``sample()`` dispatch: at <= 10 qubits a
    non-canonical chain is canonicalised left to right before a sweep that
    assumes right-canonical sites, so the counts are not the Born
    distribution there (as in the reference).  ``born_exact=True``
    (extension) right-canonicalises the chain first instead: the sweep then
    samples the Born distribution at every size."""
    from .adaptive_mps import AdaptiveMPS
    mps = AdaptiveMPS(n_qubits, bond_dim=max_bond_dim, device=device, **mps_kwargs)
    for name, qubits, params in gates:
        apply_gate(mps, name, list(qubits), list(params))
    # the reference's sample() dispatch (mps_pytorch.py:154-175): a
    # non-canonical chain of <= 10 qubits is canonicalised left to right first,
    # one shot takes the single-shot path
    if born_exact:
        mps.canonicalize_right_to_left()
        mps.is_canonical = True
    elif not mps.is_canonical and n_qubits <= 10:
        mps.canonicalize_left_to_right()
    if int(shots) == 1 and n_qubits < 64 and seed is None:
        counts = samples_to_counts(mps.sample(1), n_qubits)
    else:
        u = None
        if seed is not None:
            g = torch.Generator(device=mps.device).manual_seed(int(seed))
            u = torch.rand((int(shots), n_qubits), dtype=torch.float64, device=mps.device,
                           generator=g)
        counts = bits_to_counts(mps.sample_bits(int(shots), uniforms=u), n_qubits)
    statevector = mps.to_statevector() if n_qubits <= 20 else None
    return counts, statevector
