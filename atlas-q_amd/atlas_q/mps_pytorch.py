"""MPS base class on the HIP device (API of the reference's
``src/atlas_q/mps_pytorch.py``: ``MatrixProductStatePyTorch``).

Site tensors are torch tensors ``[chi_left, 2, chi_right]`` on a HIP device.
The sweep sampler, single-qubit gates, amplitude contractions and the
canonical-form absorptions run in the hand-written kernels of libatlasq.so;
the QR factorizations of the canonical forms use rocSOLVER through torch.
There is no CPU path: constructing on a machine without a HIP device raises.
"""
from __future__ import annotations

import random
from abc import ABC, abstractmethod
from typing import List, Optional

import numpy as np
import torch

from . import kernels as K
from .linalg_robust import robust_qr, robust_svd


def hip_device(device) -> torch.device:
    d = torch.device(device)
    if d.type != "cuda" or not torch.cuda.is_available():
        raise RuntimeError(
            "atlas_q (MI355X build) needs a HIP device: got device="
            f"{device!r}, torch.cuda.is_available()={torch.cuda.is_available()}; "
            "there is no CPU path")
    if d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


def _mat(T: torch.Tensor, rows: int, cols: int) -> torch.Tensor:
    return T.reshape(rows, cols)


class CompressedQuantumStatePyTorch(ABC):
    """mps_pytorch.py:26-42."""

    def __init__(self, num_qubits: int, device: str = "cuda"):
        self.num_qubits = num_qubits
        self.dim = 2 ** num_qubits
        self.device = hip_device(device)

    @abstractmethod
    def get_amplitude(self, basis_state: int) -> complex:
        ...

    def get_probability(self, basis_state: int) -> float:
        return abs(self.get_amplitude(basis_state)) ** 2


class MatrixProductStatePyTorch(CompressedQuantumStatePyTorch):
    """mps_pytorch.py:45-655 on the HIP device."""

    def __init__(self, num_qubits: int, bond_dim: int = 8, device: str = "cuda",
                 dtype: torch.dtype = torch.complex64):
        super().__init__(num_qubits, device)
        self.bond_dim = bond_dim
        self.dtype = dtype
        self.is_canonical = False
        real = torch.float32 if dtype == torch.complex64 else torch.float64

        def rnd(*shape):  # mps_pytorch.py:72-87: two randn draws per site
            re = torch.randn(*shape, device=self.device, dtype=real)
            im = torch.randn(*shape, device=self.device, dtype=real)
            return torch.complex(re, im)

        self.tensors: List[torch.Tensor] = [rnd(1, 2, bond_dim)]
        for _ in range(num_qubits - 2):
            self.tensors.append(rnd(bond_dim, 2, bond_dim))
        if num_qubits > 1:
            self.tensors.append(rnd(bond_dim, 2, 1))
        self._normalize()

    # ------------------------------------------------------------ canonical forms
    def _absorb_right(self, i: int, R: torch.Tensor) -> None:
        """tensors[i] <- R @ tensors[i] (over the left bond), HIP GEMM."""
        nxt = self.tensors[i].to(R.dtype)
        l, p, r = nxt.shape
        out = K.cgemm(R, _mat(nxt, l, p * r))
        self.tensors[i] = out.reshape(R.shape[0], p, r)

    def _absorb_left(self, i: int, Rt: torch.Tensor) -> None:
        """tensors[i] <- tensors[i] @ Rt (over the right bond), HIP GEMM."""
        prv = self.tensors[i].to(Rt.dtype)
        l, p, r = prv.shape
        out = K.cgemm(_mat(prv, l * p, r), Rt)
        self.tensors[i] = out.reshape(l, p, Rt.shape[1])

    def canonicalize_left_to_right(self):
        """mps_pytorch.py:91-121."""
        for i in range(self.num_qubits - 1):
            T = self.tensors[i]
            l, p, r = T.shape
            Q, R = K.qr(T.reshape(l * p, r))
            self.tensors[i] = Q.reshape(l, p, Q.shape[1])
            self._absorb_right(i + 1, R)
        self.is_canonical = True

    def canonicalize_right_to_left(self):
        """mps_pytorch.py:123-152 (plain transposes, as the reference)."""
        for i in range(self.num_qubits - 1, 0, -1):
            T = self.tensors[i]
            l, p, r = T.shape
            Q, R = K.qr(T.reshape(l, p * r).T)
            Q, R = Q.T, R.T
            self.tensors[i] = Q.reshape(Q.shape[0], p, r)
            self._absorb_left(i - 1, R)

    def _normalize(self):
        """mps_pytorch.py:595-604."""
        self.canonicalize_left_to_right()
        if self.num_qubits > 0:
            last = self.tensors[-1]
            nrm2 = torch.sum(torch.abs(last) ** 2)
            if nrm2 > 0:
                self.tensors[-1] = last / torch.sqrt(nrm2)

    # ------------------------------------------------------------ sampling
    def sweep_sample(self, num_shots: int = 1) -> List[int]:
        """mps_pytorch.py:154-175 dispatch (canonicalisation policy kept)."""
        if not self.is_canonical and self.num_qubits > 10:
            pass
        elif not self.is_canonical:
            self.canonicalize_left_to_right()
        if num_shots > 1:
            return self._batch_sweep_sample(num_shots)
        return self._single_shot_sample()

    def _batch_sweep_sample(self, num_shots: int, uniforms: Optional[torch.Tensor] = None,
                            prec: str = "native") -> List[int]:
        """mps_pytorch.py:247-355 in one fused HIP sweep; draws are
        ``outcome = (u >= p0)`` over device uniforms (float64)."""
        u = self._uniforms(num_shots, uniforms)
        bits, _ = K.sample_bits(self.tensors, u, prec=prec)
        return K.bits_to_ints(bits)

    def _single_shot_sample(self, u: Optional[torch.Tensor] = None) -> List[int]:
        """mps_pytorch.py:177-245: one shot, Python ``random.random()`` draws."""
        if u is None:
            u = torch.tensor([[random.random() for _ in range(self.num_qubits)]],
                             dtype=torch.float64)
        u = self._uniforms(1, u)
        bits, _ = K.sample_bits(self.tensors, u, single_shot_rules=True)
        return K.bits_to_ints(bits)

    def _uniforms(self, shots: int, u) -> torch.Tensor:
        if u is None:
            return torch.rand((shots, self.num_qubits), dtype=torch.float64, device=self.device)
        u = torch.as_tensor(u, dtype=torch.float64).to(self.device)
        if tuple(u.shape) != (shots, self.num_qubits):
            raise ValueError(f"uniforms must have shape {(shots, self.num_qubits)}")
        return u.contiguous()

    def sample_bits(self, num_shots: int, uniforms=None, prec: str = "native"):
        """Full bitstrings as packed uint64 words [shots, ceil(n/64)] (qubit 0 =
        most significant bit); no 64-qubit limit.  Same draws as ``sample``."""
        u = self._uniforms(num_shots, uniforms)
        return K.sample_bits(self.tensors, u, prec=prec)[0]

    def sample(self, num_shots: int = 1) -> List[int]:
        return self.measure(num_shots)

    def measure(self, num_shots: int = 1) -> List[int]:
        """mps_pytorch.py:373-385."""
        return self.sweep_sample(num_shots)

    # ------------------------------------------------------------ gates
    def apply_single_qubit_gate(self, qubit: int, gate: torch.Tensor):
        """mps_pytorch.py:387-412 (applies ``gate``: T'[a,t,b] = sum_s g[t,s] T[a,s,b])."""
        g = K.gate_to_host(gate, (2, 2))
        gT = g.reshape(2, 2, 2).transpose(1, 0, 2).reshape(-1).copy()
        self.tensors[qubit] = K.apply_1q(self.tensors[qubit], gT)

    def apply_two_qubit_gate(self, qubit1: int, qubit2: int, gate: torch.Tensor):
        """mps_pytorch.py:414-476: fixed bond_dim truncation, S absorbed right."""
        if abs(qubit1 - qubit2) != 1:
            raise NotImplementedError("Only adjacent qubit gates supported")
        q = min(qubit1, qubit2)
        A, B = self.tensors[q], self.tensors[q + 1]
        cl, cr = A.shape[0], B.shape[2]
        X = K.theta(A, B.to(A.dtype), K.gate_to_host(gate, (4, 4)))
        U, S, Vh, _ = robust_svd(X)
        keep = min(self.bond_dim, len(S))
        self.tensors[q] = U[:, :keep].reshape(cl, 2, keep)
        self.tensors[q + 1] = (S[:keep, None].to(Vh.dtype) * Vh[:keep]).reshape(keep, 2, cr)

    # ------------------------------------------------------------ contractions
    def get_amplitude(self, basis_state: int) -> complex:
        """mps_pytorch.py:606-627 (qubit 0 = most significant bit)."""
        n = self.num_qubits
        if n == 1:
            a = self.tensors[0][0, basis_state & 1, 0]
            return complex(a.real.item(), a.imag.item())
        bits = [(basis_state >> (n - 1 - i)) & 1 for i in range(n)]
        res = self.tensors[0][:, bits[0], :]
        for i in range(1, n):
            T = self.tensors[i].to(res.dtype)
            res = K.cgemm(res, T[:, bits[i], :])
        a = res[0, 0]
        return complex(a.real.item(), a.imag.item())

    def to_statevector(self):
        """mps_pytorch.py:629-647: NumPy vector of all amplitudes."""
        return self._contract_all(self.tensors[0].dtype).cpu().numpy().astype(complex)

    def _contract_all(self, dtype: torch.dtype) -> torch.Tensor:
        """Left-to-right contraction with the HIP GEMM: [2^i, chi_i] @ [chi_i, 2 chi_{i+1}]."""
        res = self.tensors[0].to(dtype).reshape(2, -1)
        for i in range(1, self.num_qubits):
            T = self.tensors[i].to(dtype)
            l, p, r = T.shape
            res = K.cgemm(res, T.reshape(l, p * r)).reshape(-1, r)
        return res.reshape(-1)

    def memory_usage(self) -> int:
        """mps_pytorch.py:649-655."""
        return sum(t.element_size() * t.nelement() for t in self.tensors)

    def to_numpy_mps(self):
        return {"tensors": [t.cpu().numpy() for t in self.tensors],
                "num_qubits": self.num_qubits, "bond_dim": self.bond_dim,
                "is_canonical": self.is_canonical}

    @staticmethod
    def from_numpy_mps(numpy_mps_dict, device: str = "cuda"):
        mps = MatrixProductStatePyTorch(numpy_mps_dict["num_qubits"], numpy_mps_dict["bond_dim"],
                                        device)
        mps.tensors = [torch.from_numpy(t).to(mps.device) for t in numpy_mps_dict["tensors"]]
        mps.is_canonical = numpy_mps_dict.get("is_canonical", False)
        return mps

    # ------------------------------------------------------------ named gates
    def h(self, qubit: int):
        self.apply_single_qubit_gate(qubit, _H)

    def x(self, qubit: int):
        self.apply_single_qubit_gate(qubit, _X)

    def y(self, qubit: int):
        self.apply_single_qubit_gate(qubit, _Y)

    def z(self, qubit: int):
        self.apply_single_qubit_gate(qubit, _Z)

    def s(self, qubit: int):
        self.apply_single_qubit_gate(qubit, _S)

    def sdg(self, qubit: int):
        self.apply_single_qubit_gate(qubit, _SDG)

    def t(self, qubit: int):
        self.apply_single_qubit_gate(qubit, _T)

    def tdg(self, qubit: int):
        self.apply_single_qubit_gate(qubit, _TDG)

    def rx(self, qubit: int, theta):
        self.apply_single_qubit_gate(qubit, rx_matrix(theta))

    def ry(self, qubit: int, theta):
        self.apply_single_qubit_gate(qubit, ry_matrix(theta))

    def rz(self, qubit: int, theta):
        self.apply_single_qubit_gate(qubit, rz_matrix(theta))

    def cnot(self, control: int, target: int):
        self.apply_two_qubit_gate(control, target, _CNOT)

    def cx(self, control: int, target: int):
        self.cnot(control, target)

    def cz(self, control: int, target: int):
        self.apply_two_qubit_gate(control, target, _CZ)

    def cy(self, control: int, target: int):
        self.apply_two_qubit_gate(control, target, _CY)

    def swap(self, qubit1: int, qubit2: int):
        self.apply_two_qubit_gate(qubit1, qubit2, _SWAP)


# gate matrices (complex128, host) ------------------------------------------
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2.0)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_S = np.array([[1, 0], [0, 1j]], dtype=np.complex128)
_SDG = np.array([[1, 0], [0, -1j]], dtype=np.complex128)
_T = np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]], dtype=np.complex128)
_TDG = np.array([[1, 0], [0, np.exp(-1j * np.pi / 4)]], dtype=np.complex128)
_CNOT = np.eye(4, dtype=np.complex128)[[0, 1, 3, 2]]
_CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
_CY = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]], dtype=np.complex128)
_SWAP = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]


def _angle(theta) -> float:
    return float(theta.item()) if isinstance(theta, torch.Tensor) else float(theta)


def rx_matrix(theta) -> np.ndarray:
    t = _angle(theta)
    c, s = np.cos(t / 2), np.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def ry_matrix(theta) -> np.ndarray:
    t = _angle(theta)
    c, s = np.cos(t / 2), np.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def rz_matrix(theta) -> np.ndarray:
    t = _angle(theta)
    return np.array([[np.exp(-0.5j * t), 0], [0, np.exp(0.5j * t)]], dtype=np.complex128)
