"""torch-CPU restatement of ATLAS-Q's adaptive-MPS hot path (the oracle).

TEST INFRASTRUCTURE ONLY -- imported by ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py``; never by the shipped package.

The reference is Python on top of torch, so the restatement is written with
the same torch CPU operations in the same order: the float rounding of the
rank rule (float32 threshold saturation, CPU cumsum accumulating in double),
the LAPACK SVD/QR calls and the einsum shapes all follow the reference.  Every
function cites the reference file:line it restates (paths relative to the
reference checkout).

Parity pinning
--------------
* ``choose_rank_from_sigma``: hand-computed known answers (tests/test_oracle.py)
  + the per-gate kept ranks of the reference's golden runs.
* ``theta``: closed-form einsum known answer (the reference's own
  ``tests/integration/test_mps_triton_kernels.py:34-57`` formula) + golden.
* ``OracleMPS`` two-site update, canonical forms, statevector: golden vectors
  from the reference (per-gate k/eps/entropy/condS sequences, amplitudes) and
  the exact NumPy state-vector simulator (``oracle.statevector``).
* ``OracleMPS.sample``: golden bitstrings from the reference with its
  multinomial replaced by the same inverse-CDF rule over recorded uniforms,
  plus Born-rule marginals from the exact state vector for n <= 12.
"""
from __future__ import annotations

import math
import random as _pyrandom
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch


# --------------------------------------------------------------------------
# Policy / statistics (adaptive_mps.py:43-48, diagnostics.py:78-152)
# --------------------------------------------------------------------------
@dataclass
class DTypePolicy:
    """adaptive_mps.py:43-48."""

    default: torch.dtype = torch.complex64
    promote_if_cond_gt: float = 1e6


class MPSStatistics:
    """diagnostics.py:78-152: per-operation log + aggregates."""

    KEYS = ("step", "bond", "k_star", "chi_before", "chi_after", "eps_local",
            "entropy", "svd_driver", "dtype", "ms_elapsed", "condS")

    def __init__(self):
        self.logs: Dict[str, list] = {k: [] for k in self.KEYS}

    def record(self, **kw):
        for key, val in kw.items():
            if key in self.logs:
                self.logs[key].append(val)

    def global_error_bound(self) -> float:
        e = self.logs["eps_local"]
        return math.sqrt(sum(x * x for x in e)) if e else 0.0


def bond_entropy_from_S(S: torch.Tensor) -> float:
    """diagnostics.py:17-35 (bits)."""
    if len(S) == 0:
        return 0.0
    p = (S * S) / (S * S).sum()
    return float((-(p * torch.log2(torch.clamp(p, min=1e-30))).sum()).item())


# --------------------------------------------------------------------------
# Rank rule (truncation.py:21-83, 104-122)
# --------------------------------------------------------------------------
def choose_rank_from_sigma(S: torch.Tensor, eps_bond: float, chi_cap: int,
                           budget_ok: Callable[[int], bool] = lambda k: True
                           ) -> Tuple[int, float, float, float]:
    """truncation.py:21-83.  All arithmetic in S's dtype, as the reference."""
    if len(S) == 0:
        return 0, 0.0, 0.0, float("inf")
    energy = (S * S).cumsum(0)                      # :52
    total = energy[-1]
    if total < 1e-30:                               # :55-56
        return 1, 0.0, 0.0, float("inf")
    thresh = (1.0 - eps_bond ** 2) * total          # :58 (0-dim tensor op)
    k_tol = int(torch.searchsorted(energy, thresh).item()) + 1   # :59 left
    k_tol = min(k_tol, len(S))
    k = min(k_tol, chi_cap)                         # :63
    while k > 1 and not budget_ok(k):               # :66-67
        k -= 1
    eps_local = torch.sqrt(torch.clamp(total - energy[k - 1], min=0.0))
    kept = S[:k]
    p = (kept * kept) / (kept * kept).sum()
    entropy = float(-(p * torch.log(torch.clamp(p, min=1e-30))).sum().item())
    if k > 1 and S[k - 1] > 0:
        cond = float((S[0] / S[k - 1]).item())
    else:
        cond = float("inf")
    return k, float(eps_local.item()), entropy, cond


def check_entropy_sanity(entropy: float, chi_left: int, chi_right: int) -> bool:
    """truncation.py:104-122 (log2 bound, nats compared as in the reference)."""
    return entropy <= math.log2(min(chi_left * 2, 2 * chi_right)) + 1e-6


# --------------------------------------------------------------------------
# Linear algebra wrappers (linalg_robust.py:17-76)
# --------------------------------------------------------------------------
def robust_svd(X: torch.Tensor):
    """linalg_robust.py:17-52; on CPU the first attempt is LAPACK."""
    try:
        U, S, Vh = torch.linalg.svd(X, full_matrices=False)
        return U, S, Vh, "torch_cuda"
    except Exception:
        pass
    try:
        Xj = X + (1e-12 * torch.randn_like(X))
        U, S, Vh = torch.linalg.svd(Xj, full_matrices=False)
        return U, S, Vh, "torch_cuda_jitter"
    except Exception:
        pass
    U, S, Vh = torch.linalg.svd(X.cpu(), full_matrices=False)
    return U, S, Vh, "torch_cpu"


def robust_qr(X: torch.Tensor):
    """linalg_robust.py:55-76."""
    Q, R = torch.linalg.qr(X)
    return Q, R, "torch_cuda"


# --------------------------------------------------------------------------
# Theta contraction (triton_kernels/mps_complex.py:343-368)
# --------------------------------------------------------------------------
def theta(A: torch.Tensor, B: torch.Tensor, U: torch.Tensor) -> torch.Tensor:
    """X[2l+p, q*chiR+j] = sum_{p1,p2} U[2p+q, 2p1+p2] (A[l,p1,:] @ B[:,p2,j]).

    Restates ``fused_two_qubit_gate_pytorch`` (mps_complex.py:343-368): the
    "U" convention (row index of U = output pair, left qubit = high bit).
    """
    li, rj = A.shape[0], B.shape[2]
    psi = torch.einsum("lpr,rqj->lpqj", A, B)
    psi = psi.permute(1, 2, 0, 3).reshape(4, li * rj)
    out = U @ psi
    return out.reshape(2, 2, li, rj).permute(2, 0, 1, 3).reshape(li * 2, 2 * rj)


def theta_transposed(A: torch.Tensor, B: torch.Tensor, U: torch.Tensor) -> torch.Tensor:
    """The reference's no-triton fallback (adaptive_mps.py:244-247): applies U^T."""
    li, rj = A.shape[0], B.shape[2]
    th = torch.einsum("asm,mtb->astb", A, B)
    th = torch.einsum("stuv,astb->auvb", U.view(2, 2, 2, 2), th)
    return th.reshape(li * 2, 2 * rj)


# --------------------------------------------------------------------------
# The MPS
# --------------------------------------------------------------------------
class OracleMPS:
    """Restatement of ``AdaptiveMPS`` (adaptive_mps.py:51-623) on top of the
    base ``MatrixProductStatePyTorch`` (mps_pytorch.py:45-655), CPU only."""

    def __init__(self, num_qubits: int, bond_dim: int = 8, *, eps_bond: float = 1e-6,
                 chi_max_per_bond=256, budget_global_mb: Optional[float] = None,
                 dtype_policy: DTypePolicy = DTypePolicy(), dtype=None,
                 gate_convention: str = "U"):
        if dtype is None:
            dtype = dtype_policy.default
        self.num_qubits = num_qubits
        self.bond_dim = bond_dim
        self.dtype = dtype
        self.device = torch.device("cpu")
        self.is_canonical = False
        self.gate_convention = gate_convention
        # mps_pytorch.py:65-87: random complex tensors from the global RNG
        real_dtype = torch.float32 if dtype == torch.complex64 else torch.float64
        self.tensors: List[torch.Tensor] = []

        def rnd(*shape):
            re = torch.randn(*shape, dtype=real_dtype)
            im = torch.randn(*shape, dtype=real_dtype)
            return torch.complex(re, im)

        self.tensors.append(rnd(1, 2, bond_dim))
        for _ in range(num_qubits - 2):
            self.tensors.append(rnd(bond_dim, 2, bond_dim))
        if num_qubits > 1:
            self.tensors.append(rnd(bond_dim, 2, 1))
        self._normalize()                           # mps_pytorch.py:89
        # adaptive_mps.py:105-126
        self.eps_bond = eps_bond
        self.dtype_policy = dtype_policy
        self.budget_global_mb = budget_global_mb
        if isinstance(chi_max_per_bond, int):
            self.chi_max_per_bond = [chi_max_per_bond] * (num_qubits - 1)
        else:
            assert len(chi_max_per_bond) == num_qubits - 1
            self.chi_max_per_bond = list(chi_max_per_bond)
        self.bond_dims = [bond_dim] * (num_qubits - 1)
        self.statistics = MPSStatistics()
        self._operation_counter = 0
        self._initialize_zero_state()

    # adaptive_mps.py:128-139
    def _initialize_zero_state(self):
        for T in self.tensors:
            T[:, 1, :] = 0
            T[:, 0, :] = torch.eye(T.shape[0], T.shape[2], dtype=T.dtype)
        self._normalize()

    # mps_pytorch.py:595-604
    def _normalize(self):
        self.to_left_canonical()
        if self.num_qubits > 0:
            last = self.tensors[-1]
            nrm2 = torch.sum(torch.abs(last) ** 2)
            if nrm2 > 0:
                self.tensors[-1] /= torch.sqrt(nrm2)

    # adaptive_mps.py:310-339
    def to_left_canonical(self):
        for i in range(self.num_qubits - 1):
            A = self.tensors[i]
            cl, p, cr = A.shape
            Q, R, _ = robust_qr(A.reshape(cl * p, cr))
            cm = Q.shape[1]
            self.tensors[i] = Q.reshape(cl, p, cm)
            self.tensors[i + 1] = torch.einsum("ij,jkl->ikl", R, self.tensors[i + 1])
            if hasattr(self, "bond_dims") and i < len(self.bond_dims):
                self.bond_dims[i] = cm
        self.is_canonical = True

    canonicalize_left_to_right = to_left_canonical   # adaptive_mps.py:342-344

    # adaptive_mps.py:346-385
    def to_mixed_canonical(self, center: int):
        assert 0 <= center < self.num_qubits
        for i in range(center):
            A = self.tensors[i]
            cl, p, cr = A.shape
            Q, R, _ = robust_qr(A.reshape(cl * p, cr))
            self.tensors[i] = Q.reshape(cl, p, Q.shape[1])
            self.tensors[i + 1] = torch.einsum("ij,jkl->ikl", R, self.tensors[i + 1])
        for i in range(self.num_qubits - 1, center, -1):
            B = self.tensors[i]
            cl, p, cr = B.shape
            Q, R, _ = robust_qr(B.permute(1, 2, 0).reshape(p * cr, cl))
            self.tensors[i] = Q.t().reshape(Q.shape[1], p, cr)
            if i > 0:
                self.tensors[i - 1] = torch.einsum("ijk,kl->ijl", self.tensors[i - 1], R.t())

    # adaptive_mps.py:159-180 (U^T convention)
    def apply_single_qubit_gate(self, q: int, U2: torch.Tensor):
        T = self.tensors[q]
        self.tensors[q] = torch.einsum("st,asb->atb", U2.to(dtype=T.dtype), T)
        self.is_canonical = False

    def memory_usage(self) -> int:
        """mps_pytorch.py:649-655."""
        return sum(t.element_size() * t.nelement() for t in self.tensors)

    # adaptive_mps.py:182-308
    def apply_two_site_gate(self, i: int, U4: torch.Tensor):
        assert 0 <= i < self.num_qubits - 1
        A, B = self.tensors[i], self.tensors[i + 1]
        cl, cm, cr = A.shape[0], A.shape[2], B.shape[2]
        local = self.dtype_policy.default
        A = A.to(dtype=local)
        B = B.to(dtype=local)
        if tuple(U4.shape) == (2, 2, 2, 2):
            Um = U4.reshape(4, 4)
        elif tuple(U4.shape) == (4, 4):
            Um = U4
        else:
            raise ValueError(f"U4 must be (2,2,2,2) or (4,4), got {tuple(U4.shape)}")
        Um = Um.to(dtype=local)
        if self.gate_convention == "U":
            X = theta(A, B, Um)
        else:
            X = theta_transposed(A, B, Um)
        U, S, Vh, driver = robust_svd(X)
        cap = self.chi_max_per_bond[i]

        def budget_ok(k: int) -> bool:            # adaptive_mps.py:255-264
            if self.budget_global_mb is None:
                return True
            bpe = torch.finfo(local).bits // 8
            before = (cl * cm + cm * cr) * 2 * bpe
            after = (cl * k + k * cr) * 2 * bpe
            delta = max(0, after - before)
            cur = self.memory_usage() / (1024 ** 2)
            return (cur + delta / (1024 ** 2)) < self.budget_global_mb

        k, eps_loc, entropy, cond = choose_rank_from_sigma(S, self.eps_bond, cap, budget_ok)
        if math.isfinite(cond) and cond > self.dtype_policy.promote_if_cond_gt:
            U, S, Vh, driver = robust_svd(X.to(torch.complex128))
            k, eps_loc, entropy, cond = choose_rank_from_sigma(S, self.eps_bond, cap, budget_ok)
            local = torch.complex128
        check_entropy_sanity(entropy, cl, cr)
        US = U[:, :k] * S[:k]
        self.tensors[i] = US.reshape(cl, 2, k)
        self.tensors[i + 1] = Vh[:k, :].reshape(k, 2, cr)
        self.bond_dims[i] = k
        self.statistics.record(step=self._operation_counter, bond=i, k_star=k,
                               chi_before=cm, chi_after=k, eps_local=eps_loc,
                               entropy=entropy, svd_driver=driver, dtype=str(local),
                               ms_elapsed=0.0, condS=cond)
        self._operation_counter += 1
        self.is_canonical = False

    # adaptive_mps.py:444-481
    def to_statevector(self) -> torch.Tensor:
        if self.num_qubits > 20:
            raise ValueError("statevector limited to n <= 20")
        res = self.tensors[0].to(dtype=self.dtype)
        for i in range(1, self.num_qubits):
            res = torch.einsum("...i,ijk->...jk", res, self.tensors[i].to(dtype=self.dtype))
        return res.squeeze(-1).reshape(-1)

    # mps_pytorch.py:606-627
    def get_amplitude(self, basis_state: int) -> complex:
        n = self.num_qubits
        if n == 1:
            a = self.tensors[0][0, basis_state & 1, 0]
            return complex(a.real.item(), a.imag.item())
        bits = [(basis_state >> (n - 1 - i)) & 1 for i in range(n)]
        res = self.tensors[0][:, bits[0], :]
        for i in range(1, n - 1):
            res = res @ self.tensors[i][:, bits[i], :]
        res = res @ self.tensors[-1][:, bits[-1], :]
        a = res[0, 0]
        return complex(a.real.item(), a.imag.item())

    # ----------------------------------------------------------------------
    # Sampling (mps_pytorch.py:154-355) with injected uniforms
    # ----------------------------------------------------------------------
    def sample(self, num_shots: int = 1, uniforms: Optional[np.ndarray] = None) -> List[int]:
        """``AdaptiveMPS.sample`` -> ``measure`` -> ``sweep_sample``
        (adaptive_mps.py:621-623; mps_pytorch.py:373-385,154-175).

        ``uniforms[shot, site]`` (float64) replace the reference's draws: the
        outcome is 0 iff ``u < p0`` where ``p0`` is the normalised probability
        of outcome 0 computed exactly as the reference computes it.
        """
        if not self.is_canonical and self.num_qubits > 10:
            pass
        elif not self.is_canonical:
            self.canonicalize_left_to_right()
        if uniforms is None:
            uniforms = np.random.default_rng().random((num_shots, self.num_qubits))
        uniforms = np.asarray(uniforms, dtype=np.float64)
        if num_shots > 1:
            return batch_sweep_sample(self.tensors, uniforms)
        return single_shot_sample(self.tensors, uniforms[0])


def batch_sweep_sample(tensors: Sequence[torch.Tensor], uniforms: np.ndarray,
                       return_p0: bool = False):
    """mps_pytorch.py:247-355 with multinomial replaced by ``u < p0``.

    Returns the reference's int64 sample words (qubit 0 = MSB, wrapping as
    int64 shifts wrap) as a Python list; with ``return_p0`` also the float
    ``p0`` used at every (shot, site), for margin analysis.
    """
    S, n = uniforms.shape
    tdt = tensors[0].dtype
    L = torch.ones((S, 1), dtype=tdt)
    samples = torch.zeros(S, dtype=torch.int64)
    p0_all = np.zeros((S, n))
    u = torch.from_numpy(uniforms)
    for i in range(n):
        T = tensors[i]
        if T.dtype != tdt:
            T = T.to(dtype=tdt)
        if i == 0:
            t0 = T[0, 0, :].unsqueeze(0)
            t1 = T[0, 1, :].unsqueeze(0)
            p0 = torch.sum(torch.abs(t0) ** 2, dim=-1).expand(S)
            p1 = torch.sum(torch.abs(t1) ** 2, dim=-1).expand(S)
        elif i == n - 1:
            t0 = torch.sum(L * T[:, 0, 0].unsqueeze(0), dim=-1)
            t1 = torch.sum(L * T[:, 1, 0].unsqueeze(0), dim=-1)
            p0 = torch.abs(t0) ** 2
            p1 = torch.abs(t1) ** 2
        else:
            t0 = L @ T[:, 0, :]
            t1 = L @ T[:, 1, :]
            p0 = torch.sum(torch.abs(t0) ** 2, dim=-1)
            p1 = torch.sum(torch.abs(t1) ** 2, dim=-1)
        probs = torch.stack([p0, p1], dim=-1)
        probs = torch.abs(probs.real) if torch.is_complex(probs) else torch.abs(probs)
        probs = probs + 1e-15
        probs = probs / torch.sum(probs, dim=-1, keepdim=True)
        p0n = probs[:, 0]
        p0_all[:, i] = p0n.double().numpy()
        out = (u[:, i] >= p0n.double()).to(torch.int64)
        samples = (samples << 1) | out
        if i == 0:
            L = T[0, out, :]
        elif i < n - 1:
            sel = T[:, out, :].permute(1, 0, 2)
            L = torch.bmm(L.unsqueeze(1), sel).squeeze(1)
    res = samples.tolist()
    return (res, p0_all) if return_p0 else res


def single_shot_sample(tensors: Sequence[torch.Tensor], u: np.ndarray) -> List[int]:
    """mps_pytorch.py:177-245 (one shot), ``random.random()`` replaced by u."""
    n = len(tensors)
    tdt = tensors[0].dtype
    left = torch.ones((1,), dtype=tdt)
    sample = 0
    for i in range(n):
        T = tensors[i]
        if T.dtype != left.dtype:
            T = T.to(dtype=left.dtype)
        if i == 0:
            p0 = torch.abs(torch.sum(T[0, 0, :])) ** 2
            p1 = torch.abs(torch.sum(T[0, 1, :])) ** 2
        elif i == n - 1:
            p0 = torch.abs(left @ T[:, 0, 0]) ** 2
            p1 = torch.abs(left @ T[:, 1, 0]) ** 2
        else:
            p0 = torch.sum(torch.abs(left @ T[:, 0, :]) ** 2)
            p1 = torch.sum(torch.abs(left @ T[:, 1, :]) ** 2)
        a, b = p0.item(), p1.item()
        tot = a + b
        if tot > 1e-15:
            a /= tot
        else:
            a = 0.5
        out = 0 if float(u[i]) < a else 1
        sample = (sample << 1) | out
        if i == 0:
            left = T[0, out, :]
        elif i < n - 1:
            left = left @ T[:, out, :]
    return [sample]


def mps_overlap(ta: Sequence[torch.Tensor], tb: Sequence[torch.Tensor]) -> complex:
    """<a|b> of two open-boundary MPS by left-to-right transfer contraction
    (complex128).  Test helper for fidelity, not a reference function."""
    E = torch.ones((1, 1), dtype=torch.complex128)
    for a, b in zip(ta, tb):
        a = a.to(torch.complex128)
        b = b.to(torch.complex128)
        E = torch.einsum("xy,xsa,ysb->ab", E, a.conj(), b)
    return complex(E.reshape(-1)[0].item())
