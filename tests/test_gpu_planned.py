"""GPU checks of the device-planned batched circuit (PlannedCircuit: bond
dimensions on the device, every launch shape fixed from bounds, the whole
circuit replayable from one HIP graph).

* Without truncation the planned run equals the exact state vector
  (complex64: fidelity within 1e-5 of 1 per MPS), with no failed eigensolve,
  on a mixed population whose Gram matrices fill Jacobi buckets (orders
  <= 64) and tridiagonal buckets (orders 128-512).
* A graph captured from the same plan, replayed several times from the
  initial state (with eager work and synchronisations between the replays),
  gives bit for bit the eager planned run's states, bond dimensions and
  statistics every time, also where caps bind.
* Every bond of the run stays within its planned bound (no kept rank was
  lowered to a bound); with the operator-Schmidt-rank bound the plan is
  tighter and the state still the exact one.
* The planned run keeps the host-planned BatchedMPS's kept ranks except
  where the float32 rank rule is saturated (no resolvable weight discarded
  by either run).
"""
import numpy as np
import pytest
import torch

from oracle import statevector as SV
from atlas_q.autotune import Autotuner
from atlas_q.batched import BatchedMPS
from atlas_q.batched_graph import PlannedCircuit
from atlas_q.circuits import hea

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _tuner():
    return Autotuner(path="/tmp/atlasq_test_autotune.json", tune=True)


def _dense(ts):
    res = ts[0].reshape(2, -1).to(torch.complex128)
    for T in ts[1:]:
        T = T.to(torch.complex128)
        res = (res @ T.reshape(T.shape[0], -1)).reshape(-1, T.shape[2])
    return res.reshape(-1).cpu().numpy()


def _fidelities(bm, specs, circs):
    fids = []
    for j, (n, _) in enumerate(specs):
        psi = SV.run_circuit(n, circs[j])
        v = _dense(bm.tensors(j))
        fid = abs(np.vdot(psi, v)) ** 2 / np.vdot(v, v).real
        print(f"mps {j} ({n} qubits): fidelity {fid:.12f}, bond max {int(bm.dims[j].max())}")
        fids.append(fid)
    return fids


def test_planned_circuit_exact_against_statevector():
    specs = [(6, 64), (9, 64), (12, 64), (14, 128), (16, 256), (15, 256)]
    circs = [hea(n, 7, 700 + j) for j, (n, _) in enumerate(specs)]
    bm = BatchedMPS(specs, eps_bond=1e-10, device=DEV, autotuner=_tuner())
    plan = PlannedCircuit(bm, bm.compile_circuits(circs))
    plan.run()
    plan.finish()
    # the bounds min(cap, 2 chi_left, 2 chi_right) are the rank rule's own
    # limits: no rank is ever lowered, and every bond stays within its bound
    assert plan.clamped == 0
    assert (bm.dims <= plan.bound).all()
    assert max(int(bm.dims[j].max()) for j in range(len(specs))) > 64
    fids = _fidelities(bm, specs, circs)
    ref = BatchedMPS(specs, eps_bond=1e-10, device=DEV, autotuner=_tuner())
    ref.run_hea(circs)
    diff = [(i, a, b, e1, e2) for i, (a, b, e1, e2) in enumerate(zip(
        bm.logs["k"], ref.logs["k"], bm.logs["eps_local"], ref.logs["eps_local"])) if a != b]
    print(f"{len(diff)} of {len(bm.logs['k'])} kept ranks differ from the host-planned run:",
          diff[:12])
    assert all(abs(1 - f) < 1e-5 for f in fids)
    assert len(bm.logs["k"]) == len(ref.logs["k"])
    assert bm.logs["mps"] == ref.logs["mps"] and bm.logs["bond"] == ref.logs["bond"]
    # a kept rank may differ only where the float32 rule is saturated: both
    # runs discard no weight the storage precision resolves (the noise-level
    # eigenvalues past the cut differ between the Gram sides the two plans use)
    assert all(e1 <= 1e-7 and e2 <= 1e-7 for _, _, _, e1, e2 in diff)


def test_planned_circuit_graph_replays_bit_identically():
    specs = [(8, 8), (11, 16), (13, 24), (14, 100), (16, 200), (12, 40)]
    circs = [hea(n, 8, 800 + j) for j, (n, _) in enumerate(specs)]
    # eager planned run
    a = BatchedMPS(specs, eps_bond=1e-8, device=DEV, autotuner=_tuner())
    pa = PlannedCircuit(a, a.compile_circuits(circs))
    pa.run()
    pa.finish()
    # captured once, replayed three times from the initial state
    b = BatchedMPS(specs, eps_bond=1e-8, device=DEV, autotuner=_tuner())
    arena0 = b.arena.clone()
    pb = PlannedCircuit(b, b.compile_circuits(circs))
    dims0 = pb.dims.clone()
    g = pb.capture()
    for rep in range(3):
        b.arena.copy_(arena0)
        pb.dims.copy_(dims0)
        pb.fails.zero_()
        pb.overflow.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _ = (b.arena.abs().sum()).item()            # eager work between replays
        assert torch.equal(b.arena, a.arena), rep
        assert torch.equal(pb.dims, pa.dims), rep
        assert int(pb.fails.item()) == 0 and int(pb.overflow.item()) == 0
        for s_a, s_b in zip(pa.gate_order, pb.gate_order):
            assert torch.equal(s_a[0].log_k, s_b[0].log_k), rep
            assert torch.equal(s_a[0].log_eps, s_b[0].log_eps), rep
    print("bond max", [int(x) for x in pa.dims.max(1).values.cpu()])
    assert int(pa.dims.max()) > 64                   # tridiagonal buckets were used


def test_planned_circuit_schmidt_bound_exact_against_statevector():
    """With the operator-Schmidt-rank bound (CZ: rank 2) the plan lowers only
    ranks made of round-off singular values past theta's exact rank: the
    state stays the exact one, every bond within its (tighter) bound."""
    specs = [(9, 64), (12, 64), (14, 128), (16, 256)]
    circs = [hea(n, 6, 750 + j) for j, (n, _) in enumerate(specs)]
    bm = BatchedMPS(specs, eps_bond=1e-10, device=DEV, autotuner=_tuner())
    loose = PlannedCircuit(BatchedMPS(specs, eps_bond=1e-10, device=DEV, autotuner=_tuner()),
                           bm.compile_circuits(circs))
    plan = PlannedCircuit(bm, bm.compile_circuits(circs), schmidt_bound=True)
    assert (plan.bound <= loose.bound).all() and (plan.bound < loose.bound).any()
    assert plan.arena_bytes < loose.arena_bytes
    plan.run()
    plan.finish()
    print(f"ranks lowered to the bound: {plan.clamped}")
    assert (bm.dims <= plan.bound).all()
    fids = _fidelities(bm, specs, circs)
    assert all(abs(1 - f) < 1e-5 for f in fids)
