"""CPU checks of the gate-list runner's host logic (qiskit_adapter.py
_apply_gate_to_mps :505-550, _samples_to_counts :729-744): gate-name
dispatch, the reference's count keys from int samples (qubit 0 leftmost),
the 64-qubit sign fix, and counts from packed words at any qubit count."""
import numpy as np
import pytest
import torch

from atlas_q.runner import apply_gate, bits_to_counts, samples_to_counts


class _Recorder:
    def __getattr__(self, name):
        calls = self.__dict__.setdefault("calls", [])
        return lambda *a: calls.append((name,) + a)


def test_apply_gate_dispatch_matches_reference_names():
    r = _Recorder()
    for name, qs, ps in (("H", [0], []), ("x", [1], []), ("sdg", [2], []), ("rx", [0], [0.3]),
                         ("rz", [3], [1.5]), ("cx", [0, 1], []), ("cnot", [1, 2], []),
                         ("cz", [2, 3], []), ("cy", [0, 1], []), ("swap", [3, 4], []),
                         ("measure", [0], []), ("barrier", [0, 1], [])):
        apply_gate(r, name, qs, ps)
    assert r.calls == [("h", 0), ("x", 1), ("sdg", 2), ("rx", 0, 0.3), ("rz", 3, 1.5),
                       ("cnot", 0, 1), ("cnot", 1, 2), ("cz", 2, 3), ("cy", 0, 1), ("swap", 3, 4)]
    with pytest.raises(ValueError, match="Unknown single-qubit gate: u3"):
        apply_gate(r, "u3", [0], [0.1, 0.2, 0.3])
    with pytest.raises(ValueError, match="Unknown two-qubit gate: iswap"):
        apply_gate(r, "iswap", [0, 1], [])


def test_samples_to_counts_reference_semantics():
    assert samples_to_counts([5, 5, 2], 3) == {"101": 2, "010": 1}
    assert samples_to_counts(np.array([1, 3]), None) == {"1": 1, "11": 1}
    assert samples_to_counts({"00": 4}, 2) == {"00": 4}
    assert samples_to_counts("junk", 2) == {}
    # 64 qubits: qubit 0 set makes the int64 negative; read as the unsigned value
    v = np.int64(-(1 << 63) + 5)
    assert samples_to_counts([v], 64) == {"1" + "0" * 60 + "101": 1}
    with pytest.raises(ValueError):
        samples_to_counts([0], 70)


def test_bits_to_counts_any_width():
    def pack(strs, n):
        w = (n + 63) // 64
        out = np.zeros((len(strs), w), dtype=np.uint64)
        for i, s in enumerate(strs):
            v = int(s, 2)
            for j in range(w):
                out[i, j] = (v >> (64 * j)) & ((1 << 64) - 1)
        return torch.from_numpy(out.view(np.int64))
    for n, strs in ((3, ["101", "101", "000"]),
                    (64, ["1" + "0" * 63, "0" * 63 + "1", "1" + "0" * 63]),
                    (70, ["1" * 70, "0" * 70, "1" + "0" * 68 + "1", "1" * 70])):
        want = {}
        for s in strs:
            want[s] = want.get(s, 0) + 1
        assert bits_to_counts(pack(strs, n), n) == want
    with pytest.raises(ValueError):
        bits_to_counts(torch.zeros((2, 1), dtype=torch.int64), 70)
