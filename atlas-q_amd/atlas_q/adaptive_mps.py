"""Adaptive MPS on MI355X (API of the reference's ``src/atlas_q/adaptive_mps.py``).

``AdaptiveMPS.apply_two_site_gate`` keeps the reference's semantics
(adaptive_mps.py:182-308): theta in the "U" convention, thin SVD, the rank
rule of truncation.py with the per-bond cap and the global memory budget,
promotion to complex128 when cond(S) exceeds the policy, S absorbed into the
left core, no renormalisation, statistics per update.  The arithmetic runs in
the HIP kernels of libatlasq.so: theta (MFMA), rank rule, core rebuild; the
plain method's SVD is rocSOLVER through ``torch.linalg.svd``.

Extensions (keyword-only, defaults reproduce the reference):
* ``svd_method``: "plain" (full thin SVD, the reference's path) or one of the
  truncated-SVD methods of ``atlas_q.fastsvd`` ("gram", "rsvd", "learned").
* ``theta_precision``: "native" (FP32 MFMA for complex64) or "bf16x3".
"""
from __future__ import annotations

import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import Dict, List, Optional, Union

import numpy as np
import torch

from . import kernels as K
from .diagnostics import MPSStatistics
from .linalg_robust import robust_qr, robust_svd
from .mps_pytorch import (MatrixProductStatePyTorch, _CNOT, _CY, _CZ, _H, _S, _SDG, _SWAP,
                          _T, _TDG, _X, _Y, _Z, rx_matrix, ry_matrix, rz_matrix)
from .truncation import check_entropy_sanity, choose_rank_from_sigma

# The gate convention of the reference's installed-triton path
# (mps_complex.py:363): row index of U = output pair, left qubit = high bit.
GATE_CONVENTION = "U"
fused_two_qubit_gate = K.theta


@dataclass
class DTypePolicy:
    """adaptive_mps.py:43-48."""

    default: torch.dtype = torch.complex64
    promote_if_cond_gt: float = 1e6


_POOLS = {}
_TLS = threading.local()


def _worker_pool(workers: int):
    """Process-wide host thread pools, one per worker count."""
    if workers not in _POOLS:
        _POOLS[workers] = ThreadPoolExecutor(max_workers=workers,
                                             thread_name_prefix="aq-update")
    return _POOLS[workers]


def _thread_stream(device) -> torch.cuda.Stream:
    """The calling thread's own HIP stream (default priority) on `device`."""
    streams = getattr(_TLS, "streams", None)
    if streams is None:
        streams = _TLS.streams = {}
    key = torch.device(device).index or 0
    if key not in streams:
        streams[key] = torch.cuda.Stream(device)
    return streams[key]


def _gate4(U4) -> np.ndarray:
    shape = tuple(U4.shape)
    if shape not in ((2, 2, 2, 2), (4, 4)):
        raise ValueError(f"U4 must be (2,2,2,2) or (4,4), got {U4.shape}")
    return K.gate_to_host(U4, (4, 4))


class AdaptiveMPS(MatrixProductStatePyTorch):
    """adaptive_mps.py:51-623 on the HIP device."""

    def __init__(self, num_qubits: int, bond_dim: int = 8, *, eps_bond: float = 1e-6,
                 chi_max_per_bond: Optional[Union[int, List[int]]] = 256,
                 budget_global_mb: Optional[float] = None,
                 dtype_policy: DTypePolicy = DTypePolicy(), device: str = "cuda",
                 dtype: Optional[torch.dtype] = None, svd_method: str = "plain",
                 theta_precision: str = "native", rank_predictor=None):
        if dtype is None:
            dtype = dtype_policy.default
        super().__init__(num_qubits, bond_dim, device, dtype)
        self.eps_bond = eps_bond
        self.dtype_policy = dtype_policy
        self.budget_global_mb = budget_global_mb
        if isinstance(chi_max_per_bond, int):
            self.chi_max_per_bond = [chi_max_per_bond] * (num_qubits - 1)
        else:
            assert len(chi_max_per_bond) == num_qubits - 1, \
                f"chi_max_per_bond must have length {num_qubits - 1}"
            self.chi_max_per_bond = list(chi_max_per_bond)
        self.bond_dims = [bond_dim] * (num_qubits - 1)
        self.statistics = MPSStatistics()
        self._operation_counter = 0
        self.svd_method = svd_method
        self.theta_precision = theta_precision
        self.rank_predictor = rank_predictor
        self.harvest = None          # list -> (features, kept rank) pairs are appended
        self._last_entropy = {}
        self._last_eps = {}
        self._at_cap = {}      # bond -> the last update there kept the per-bond cap
        self._initialize_zero_state()

    def _initialize_zero_state(self):
        """adaptive_mps.py:128-139."""
        for T in self.tensors:
            T[:, 1, :] = 0
            T[:, 0, :] = torch.eye(T.shape[0], T.shape[2], dtype=T.dtype, device=T.device)
        self._normalize()

    # ------------------------------------------------------------ statistics
    def _log_operation(self, **kwargs):
        self.statistics.record(step=self._operation_counter, **kwargs)
        self._operation_counter += 1

    def stats_summary(self) -> Dict[str, float]:
        return self.statistics.summary()

    def global_error_bound(self) -> float:
        return self.statistics.global_error_bound()

    def reset_stats(self):
        self.statistics.reset()
        self._operation_counter = 0

    # ------------------------------------------------------------ gates
    @torch.no_grad()
    def apply_single_qubit_gate(self, q: int, U2: torch.Tensor):
        """adaptive_mps.py:159-180: T'[a,t,b] = sum_s U[s,t] T[a,s,b]."""
        assert 0 <= q < self.num_qubits, f"Qubit index {q} out of range"
        assert tuple(U2.shape) == (2, 2), "U2 must be 2x2"
        self.tensors[q] = K.apply_1q(self.tensors[q], K.gate_to_host(U2, (2, 2)))
        self.is_canonical = False

    def _budget_ok_factory(self, cl: int, cm: int, cr: int, local: torch.dtype):
        def budget_ok(k: int) -> bool:  # adaptive_mps.py:255-264
            if self.budget_global_mb is None:
                return True
            bpe = torch.finfo(local).bits // 8
            before = (cl * cm + cm * cr) * 2 * bpe
            after = (cl * k + k * cr) * 2 * bpe
            delta = max(0, after - before)
            cur = self.memory_usage() / (1024 ** 2)
            return (cur + delta / (1024 ** 2)) < self.budget_global_mb
        return budget_ok

    @torch.no_grad()
    def apply_two_site_gate(self, i: int, U4: torch.Tensor):
        """adaptive_mps.py:182-308."""
        assert 0 <= i < self.num_qubits - 1, f"Bond index {i} out of range"
        start = time.time()
        self._commit(i, self._compute_update(i, U4), start)

    @torch.no_grad()
    def apply_two_site_gates(self, gates, workers: int = 4):
        """Apply two-site gates on pairwise disjoint bonds (one brickwork
        sub-layer).  The updates are independent: the truncated-SVD methods
        factor all their Gram cores in one batched eigensolver call; the plain
        method runs them concurrently on `workers` host threads, each
        launching on its own HIP stream.  Tensors, bond dimensions and
        statistics are committed in the given order, so the result equals
        applying the gates one by one."""
        gates = list(gates)
        used = set()
        for i, _ in gates:
            assert 0 <= i < self.num_qubits - 1, f"Bond index {i} out of range"
            if {i, i + 1} & used:
                raise ValueError("apply_two_site_gates needs gates on disjoint bonds")
            used |= {i, i + 1}
        if workers <= 1 or len(gates) <= 1 or self.budget_global_mb is not None:
            # the global budget reads the whole chain's memory: keep the order
            for i, U in gates:
                self.apply_two_site_gate(i, U)
            return
        if self.svd_method != "plain":
            # truncated-SVD methods: one pass over the sub-layer on the current
            # stream, the Gram cores of all its bonds factored in one batch
            start = time.time()
            for (i, _), out in zip(gates, self._compute_sublayer(gates, workers)):
                self._commit(i, out, start)
            return
        pool = _worker_pool(workers)
        start = time.time()
        ready = torch.cuda.current_stream(self.device)
        ready_ev = ready.record_event()

        def job(i, U):
            st = _thread_stream(self.device)
            with torch.cuda.stream(st):
                st.wait_event(ready_ev)
                out = self._compute_update(i, U)
                done = st.record_event()
            return out, done

        futs = [pool.submit(job, i, U) for i, U in gates]
        results = [f.result() for f in futs]
        for (i, _), (out, done) in zip(gates, results):
            ready.wait_event(done)
            for t in out[0:2]:
                t.record_stream(ready)
            self._commit(i, out, start)

    def _compute_sublayer(self, gates, workers: int = 1):
        """Results of the truncated-SVD methods for gates on disjoint bonds (no
        state change).  Theta and the randomized factorizations run per bond,
        concurrently on `workers` host threads with their own HIP streams; the
        exact Gram factorizations left over are batched on the current stream."""
        from . import fastsvd

        def prep(i, U4):
            A, B = self.tensors[i], self.tensors[i + 1]
            cl, cm, cr = A.shape[0], A.shape[2], B.shape[2]
            local = self.dtype_policy.default
            r = fastsvd.prepare_update(self, i, A.to(dtype=local), B.to(dtype=local), _gate4(U4),
                                       self.chi_max_per_bond[i],
                                       self._budget_ok_factory(cl, cm, cr, local), local)
            return [r, cm]

        if workers > 1 and len(gates) > 1:
            pool = _worker_pool(workers)
            ready = torch.cuda.current_stream(self.device)
            ready_ev = ready.record_event()

            def job(i, U4):
                st = _thread_stream(self.device)
                with torch.cuda.stream(st):
                    st.wait_event(ready_ev)
                    out = prep(i, U4)
                    done = st.record_event()
                return out, done

            outs = []
            for f in [pool.submit(job, i, U) for i, U in gates]:
                out, done = f.result()
                ready.wait_event(done)
                r = out[0]
                held = ([r.X] + ([r.feats] if r.feats is not None else [])
                        if isinstance(r, fastsvd.GramJob) else list(r[0:2]))
                for t in held:          # produced on a worker stream, consumed on this one
                    t.record_stream(ready)
                outs.append(out)
        else:
            outs = [prep(i, U) for i, U in gates]
        jobs = [j for j, (r, _) in enumerate(outs) if isinstance(r, fastsvd.GramJob)]
        for j, res in zip(jobs, fastsvd.finish_gram(self, [outs[j][0] for j in jobs], workers)):
            outs[j][0] = res
        return [tuple(r) + (cm,) for r, cm in outs]

    def _compute_update(self, i: int, U4):
        """Theta, factorization, rank rule and rebuild for bond i; no state change."""
        A, B = self.tensors[i], self.tensors[i + 1]
        cl, cm, cr = A.shape[0], A.shape[2], B.shape[2]
        local = self.dtype_policy.default
        A = A.to(dtype=local)
        B = B.to(dtype=local)
        gate = _gate4(U4)
        cap = self.chi_max_per_bond[i]
        budget_ok = self._budget_ok_factory(cl, cm, cr, local)
        if self.svd_method != "plain":
            from . import fastsvd
            return fastsvd.two_site_update(self, i, A, B, gate, cap, budget_ok, local) + (cm,)
        prec = self.theta_precision if local == torch.complex64 else "native"
        X = K.theta(A, B, gate, prec=prec)
        feats = None
        if self.harvest is not None and self.rank_predictor is not None:
            feats = self.rank_predictor.features(self, i, X, gate, cap)
        U, S, Vh, driver = robust_svd(X)
        k, eps_loc, entropy, condS = choose_rank_from_sigma(S, self.eps_bond, cap, budget_ok)
        if feats is not None:
            self.harvest.append((feats.cpu().numpy(), k))
        if math.isfinite(condS) and condS > self.dtype_policy.promote_if_cond_gt:
            U, S, Vh, driver = robust_svd(X.to(torch.complex128))
            k, eps_loc, entropy, condS = choose_rank_from_sigma(S, self.eps_bond, cap,
                                                                budget_ok)
            local = torch.complex128
        A_new, B_new = K.rebuild(U, S, Vh, k, cl, cr)
        return A_new, B_new, k, eps_loc, entropy, condS, driver, local, cm

    def _commit(self, i: int, out, start: float):
        A_new, B_new, k, eps_loc, entropy, condS, driver, local, cm = out
        cl, cr = A_new.shape[0], B_new.shape[2]
        self._last_entropy[i] = entropy
        self._last_eps[i] = eps_loc
        self._at_cap[i] = k >= self.chi_max_per_bond[i]
        if not check_entropy_sanity(entropy, cl, cr):
            print(f"Warning: Entropy {entropy:.3f} exceeds physical bound at bond {i}")
        self.tensors[i] = A_new
        self.tensors[i + 1] = B_new
        self.bond_dims[i] = k
        self._log_operation(bond=i, k_star=k, chi_before=cm, chi_after=k, eps_local=eps_loc,
                            entropy=entropy, svd_driver=driver, dtype=str(local),
                            ms_elapsed=(time.time() - start) * 1000, condS=condS)
        self.is_canonical = False

    # ------------------------------------------------------------ canonical forms
    @torch.no_grad()
    def to_left_canonical(self):
        """adaptive_mps.py:310-339."""
        for i in range(self.num_qubits - 1):
            A = self.tensors[i]
            cl, p, cr = A.shape
            Q, R, _ = robust_qr(A.reshape(cl * p, cr))
            cmid = Q.shape[1]
            self.tensors[i] = Q.reshape(cl, p, cmid)
            self._absorb_right(i + 1, R)
            if hasattr(self, "bond_dims") and i < len(self.bond_dims):
                self.bond_dims[i] = cmid
        self.is_canonical = True

    def canonicalize_left_to_right(self):
        """adaptive_mps.py:342-344."""
        self.to_left_canonical()

    @torch.no_grad()
    def to_mixed_canonical(self, center: int):
        """adaptive_mps.py:346-385 (does not touch is_canonical / bond_dims)."""
        assert 0 <= center < self.num_qubits, "Center out of range"
        for i in range(center):
            A = self.tensors[i]
            cl, p, cr = A.shape
            Q, R, _ = robust_qr(A.reshape(cl * p, cr))
            self.tensors[i] = Q.reshape(cl, p, Q.shape[1])
            self._absorb_right(i + 1, R)
        for i in range(self.num_qubits - 1, center, -1):
            B = self.tensors[i]
            cl, p, cr = B.shape
            Q, R, _ = robust_qr(B.permute(1, 2, 0).reshape(p * cr, cl))
            self.tensors[i] = Q.t().reshape(Q.shape[1], p, cr)
            if i > 0:
                self._absorb_left(i - 1, R.t())

    # ------------------------------------------------------------ checkpoint
    def snapshot(self, path: str):
        """adaptive_mps.py:387-404: the reference's dict (tensors, bond_dims,
        num_qubits, eps_bond, chi_max_per_bond, statistics) plus two extension
        keys the reference's loader ignores: ``dtype`` and ``dtype_policy``."""
        torch.save({"tensors": [t.cpu() for t in self.tensors], "bond_dims": self.bond_dims,
                    "num_qubits": self.num_qubits, "eps_bond": self.eps_bond,
                    "chi_max_per_bond": self.chi_max_per_bond,
                    "statistics": self.statistics.logs, "dtype": str(self.dtype),
                    "dtype_policy": {"default": str(self.dtype_policy.default),
                                     "promote_if_cond_gt":
                                         float(self.dtype_policy.promote_if_cond_gt)}}, path)

    @staticmethod
    def load_snapshot(path: str, device: str = "cuda") -> "AdaptiveMPS":
        """adaptive_mps.py:406-433 (read with weights_only=True).  Snapshots
        written by the reference carry no dtype policy and load with the
        default one, as there; the extension keys restore it."""
        data = torch.load(path, weights_only=True)
        kw = {}
        pol = data.get("dtype_policy")
        if isinstance(pol, dict) and "default" in pol:
            dt = {"torch.complex64": torch.complex64,
                  "torch.complex128": torch.complex128}.get(pol["default"])
            if dt is not None:
                kw["dtype_policy"] = DTypePolicy(default=dt, promote_if_cond_gt=float(
                    pol.get("promote_if_cond_gt", 1e6)))
        mps = AdaptiveMPS(num_qubits=data["num_qubits"], bond_dim=max(data["bond_dims"]),
                          eps_bond=data["eps_bond"], chi_max_per_bond=data["chi_max_per_bond"],
                          device=device, **kw)
        mps.tensors = [t.to(mps.device) for t in data["tensors"]]
        mps.bond_dims = data["bond_dims"]
        if "statistics" in data:
            mps.statistics.logs = data["statistics"]
        return mps

    def get_memory_usage(self) -> int:
        return self.memory_usage()

    @torch.no_grad()
    def to_statevector(self) -> torch.Tensor:
        """adaptive_mps.py:444-481 (device tensor of 2^n amplitudes)."""
        if self.num_qubits > 20:
            raise ValueError(
                f"Converting {self.num_qubits} qubits to statevector would require "
                f"{2 ** self.num_qubits} amplitudes. Use get_amplitude() for large systems.")
        return self._contract_all(self.dtype)

    # ------------------------------------------------------------ named gates
    def _g2(self, q: int, m: np.ndarray):
        self.apply_single_qubit_gate(q, torch.from_numpy(m))

    def h(self, qubit: int):
        self._g2(qubit, _H)

    def x(self, qubit: int):
        self._g2(qubit, _X)

    def y(self, qubit: int):
        self._g2(qubit, _Y)

    def z(self, qubit: int):
        self._g2(qubit, _Z)

    def s(self, qubit: int):
        self._g2(qubit, _S)

    def sdg(self, qubit: int):
        self._g2(qubit, _SDG)

    def t(self, qubit: int):
        self._g2(qubit, _T)

    def tdg(self, qubit: int):
        self._g2(qubit, _TDG)

    def rx(self, qubit: int, theta):
        self._g2(qubit, rx_matrix(theta))

    def ry(self, qubit: int, theta):
        self._g2(qubit, ry_matrix(theta))

    def rz(self, qubit: int, theta):
        self._g2(qubit, rz_matrix(theta))

    def cnot(self, control: int, target: int):
        """adaptive_mps.py:553-571 (adjacent, control < target only)."""
        if abs(control - target) != 1:
            raise NotImplementedError("Only adjacent qubit gates supported in MPS")
        if control > target:
            raise NotImplementedError(
                "Non-adjacent CNOT or reversed control/target not yet supported")
        self.apply_two_site_gate(control, torch.from_numpy(_CNOT))

    def cx(self, control: int, target: int):
        self.cnot(control, target)

    def cz(self, q0: int, q1: int):
        if abs(q0 - q1) != 1:
            raise NotImplementedError("Only adjacent qubit gates supported in MPS")
        self.apply_two_site_gate(min(q0, q1), torch.from_numpy(_CZ))

    def cy(self, control: int, target: int):
        if abs(control - target) != 1:
            raise NotImplementedError("Only adjacent qubit gates supported in MPS")
        if control > target:
            raise NotImplementedError("Reversed control/target not yet supported")
        self.apply_two_site_gate(control, torch.from_numpy(_CY))

    def swap(self, q0: int, q1: int):
        if abs(q0 - q1) != 1:
            raise NotImplementedError("Only adjacent qubit gates supported in MPS")
        self.apply_two_site_gate(min(q0, q1), torch.from_numpy(_SWAP))

    def sample(self, num_shots: int = 1):
        """adaptive_mps.py:621-623."""
        return self.measure(num_shots)
