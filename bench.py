#!/usr/bin/env python
"""Two-site gate updates/s at chi = 1024 on one MI355X.

Workload: the steady state of the split-precision brickwork circuit (128
qubits, depth 60, Haar SU(4), chi_max = 1024, cutoff 1e-8, complex64): a chain
of SITES = 128 sites whose bonds -- including the two open ends -- are all at
the cap 1024 (sites Lambda G: random right-isometries G, Schmidt weights of a
saturated random circuit), so every update contracts a 2048 x 2048 theta and
keeps 1024 singular triplets, and a brickwork sub-layer holds 64 (63) gates as
in the circuit itself.  Gates are seeded Haar SU(4) in brickwork order over
the slice's interior bonds.  A step is one two-site update
(theta -> truncated SVD -> rank rule -> rebuild), tensors resident in HBM.
The gates of one brickwork sub-layer act on disjoint bonds; both paths run
them concurrently on WORKERS host threads with one HIP stream each (the
eigen/SVD solvers are launch-latency bound, so overlap raises throughput) and
commit in circuit order.  "sequential" repeats both paths one gate at a time.

value        : updates/s of the learned-rank + BF16x3 path (svd_method=learned,
               theta_precision=bf16x3)
plain        : updates/s of the reference-compatible plain path (FP32-MFMA
               theta + torch.linalg.svd) on the same slice and gates
fidelity     : per-update truncation fidelity of both paths against an FP64
               truncation of the same input (complex128 SVD), |dF| reported
roofline     : the recorded call with the largest share of the timed region,
               whoever wrote it -- the batched tridiagonal reduction of the
               sub-layer's Gram matrices (one call per sub-layer) -- against
               the bound its algorithmic work sets (HBM for the one-stage
               reduction: the bytes its matrix-vector products must stream),
               from HIP events; with the headline efficiency over all calls
cpu_baseline : the CPU oracle (torch-CPU restatement of the reference update)
               on the host cores, a bounded sample of updates at the same shape
batched      : the batched form (256 uneven MPS, RY/RZ + CZ ansatz, 8 layers,
               1000 shots each): updates/s of the grouped engine, shots/s of
               its graph-captured sweep, and the one-at-a-time plain rate
circuit      : the split-precision circuit itself from |0...0> (128 qubits,
               depth 60, chi_max 1024): updates/s per phase (growth, saturated),
               which factorization served the updates, and growth-phase
               truncation fidelities against FP64
sampler      : shots/s of the fused sweep sampler on the plain-GPU circuit's
               shape (64 qubits, chi_max 512, 10,000 shots, complex64, a random
               right-canonical MPS at the circuit's bond profile), its FP32-MFMA
               roofline and the CPU oracle's shots/s on a bounded sample
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "atlas-q_amd"))

METRIC = "2-site gate updates/s @chi=1024, learned-rank vs plain SVD, at equal fidelity"
_T0 = time.perf_counter()


def log(msg: str) -> None:
    """Progress on stderr (the JSON line is the only stdout output)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=127,
                    help="timed updates (127 = one odd and one even sub-layer of 128 sites)")
    ap.add_argument("--warmup", type=int, default=127,
                    help="untimed updates (127 = both sub-layers once: every bond of the "
                         "slice has had one update before the timed region)")
    ap.add_argument("--chi", type=int, default=1024)
    ap.add_argument("--sites", type=int, default=128)
    ap.add_argument("--low-sites", type=int, default=10,
                    help="slice length of the low-entanglement measurement")
    ap.add_argument("--workers", type=int, default=4,
                    help="host threads / HIP streams for the independent gates of a sub-layer")
    ap.add_argument("--method", default="learned")
    ap.add_argument("--theta-precision", default="bf16x3")
    ap.add_argument("--plain-steps", type=int, default=9)
    ap.add_argument("--seq-steps", type=int, default=4,
                    help="updates timed one at a time (workers=1) for both paths")
    ap.add_argument("--fidelity-samples", type=int, default=2)
    ap.add_argument("--cpu-updates", type=int, default=8)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-low", action="store_true",
                    help="skip the low-entanglement (learned-width) measurement")
    ap.add_argument("--sampler-qubits", type=int, default=64)
    ap.add_argument("--sampler-chi", type=int, default=512)
    ap.add_argument("--sampler-shots", type=int, default=10000)
    ap.add_argument("--sampler-reps", type=int, default=5)
    ap.add_argument("--sampler-cpu-shots", type=int, default=128)
    ap.add_argument("--no-sampler", action="store_true")
    ap.add_argument("--batched-mps", type=int, default=256)
    ap.add_argument("--batched-layers", type=int, default=8)
    ap.add_argument("--batched-shots", type=int, default=1000)
    ap.add_argument("--batched-ref", type=int, default=8,
                    help="MPS of the population run one by one (plain method) as the "
                         "unbatched reference rate")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--wide-chi", type=int, default=2048,
                    help="bond of the learned-rank form's slice (chi_max 2048)")
    ap.add_argument("--wide-sites", type=int, default=12)
    ap.add_argument("--wide-plain-steps", type=int, default=2)
    ap.add_argument("--wide-fidelity", type=int, default=1)
    ap.add_argument("--wide-cpu-updates", type=int, default=4)
    ap.add_argument("--batched-cpu-mps", type=int, default=2,
                    help="MPS of the batched population run by the oracle on the host cores")
    ap.add_argument("--no-wide", action="store_true")
    ap.add_argument("--circuit-qubits", type=int, default=128)
    ap.add_argument("--circuit-depth", type=int, default=60)
    ap.add_argument("--circuit-fidelity", type=int, default=6,
                    help="growth-phase layers whose middle gate's truncation fidelity is "
                         "checked against the FP64 truncation")
    ap.add_argument("--no-circuit", action="store_true",
                    help="skip the whole split-precision circuit from |0...0>")
    ap.add_argument("--circuit-only", action="store_true",
                    help="only run the whole-circuit measurement and print its JSON object")
    ap.add_argument("--no-peaks", action="store_true",
                    help="skip the copy / rocBLAS GEMM reference rates")
    ap.add_argument("--wide-only", action="store_true")
    ap.add_argument("--batched-only", action="store_true")
    ap.add_argument("--no-batched-graph", action="store_true",
                    help="skip the device-planned, graph-captured run of the batched population")
    ap.add_argument("--batched-eig", default="tridiag", choices=("tridiag", "heevd"),
                    help="eigensolver of the batched form's buckets above order 64")
    ap.add_argument("--mark", action="store_true",
                    help="launch a one-element zero-fill kernel right before and after the "
                         "timed fast-path region (dispatch markers for trace filtering)")
    ap.add_argument("--sampler-only", action="store_true",
                    help="only run the sampler measurement and print its JSON object")
    ap.add_argument("--theta-probe", type=int, default=0, metavar="N",
                    help="N BF16x3 theta calls at bond --chi (random operands), for profiling")
    ap.add_argument("--tile-rows", default=None, metavar="R,R",
                    help="grouped GEMM tile heights the autotuner may choose (default 64,32)")
    ap.add_argument("--gram-chunks", default=None, metavar="F,F,...",
                    help="shares of a sub-layer's Gram batch factored as overlapped chunks "
                         "(default: the library's setting; 1 = one batch)")
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--gemm-probe", action="store_true",
                    help="only run 5 FP64 Gram GEMMs X X^H of the bench's 2048 x 2048 theta "
                         "(the dominant hand-written kernel; for counter runs) and exit")
    ap.add_argument("--eig-probe", action="store_true",
                    help="only time rocSOLVER's strided-batched Hermitian eigensolvers on Gram "
                         "matrices (the variants, the batch-size scan at the bench's order and "
                         "concurrent calls) and exit")
    ap.add_argument("--hetrd-probe", type=int, default=0, metavar="B",
                    help="only run one batched tridiagonal reduction of B Gram matrices of "
                         "order 2 chi (the update's dominant call; for counter runs) and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters for the traffic figure) and exit")
    return ap.parse_args()


def slice_state(torch, sites, chi, seed, device):
    """Steady-state slice: Lambda G sites, G random right-isometries, Lambda the
    Schmidt weights of a saturated random circuit (all bonds at the cap)."""
    from atlas_q.workloads import slice_tensors
    return slice_tensors(sites, chi, "flat", seed=seed, device=device, draw_on_device=True)


def gate_schedule(sites, count, seed):
    import numpy as np
    from atlas_q.circuits import haar_su4
    rng = np.random.default_rng(seed)
    out, layer = [], 0
    while len(out) < count:
        for b in range(layer % 2, sites - 1, 2):
            out.append((b, haar_su4(rng)))
        layer += 1
    return out[:count]


def sublayers(gates):
    """Split a brickwork-ordered gate list into runs of gates on disjoint bonds."""
    runs, cur, used = [], [], set()
    for b, U in gates:
        if {b, b + 1} & used:
            runs.append(cur)
            cur, used = [], set()
        cur.append((b, U))
        used |= {b, b + 1}
    if cur:
        runs.append(cur)
    return runs


def run_gates(torch, mps, gates, workers):
    """Apply gates in order, timing only the updates; independent gates of a
    sub-layer run concurrently.  Between sub-layers (outside the timed region)
    the touched sites are rescaled to unit norm: the slice's open ends carry
    no normalisation, and without it the scale of theta drifts from update
    to update until the float32 SVD underflows."""
    total = 0.0
    for run in sublayers(gates):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mps.apply_two_site_gates(run, workers=workers)
        torch.cuda.synchronize()
        total += time.perf_counter() - t0
        for b, _ in run:
            for q in (b, b + 1):
                mps.tensors[q] = mps.tensors[q] / mps.tensors[q].norm()
    return total


def make_mps(torch, tensors, chi, method, prec):
    from atlas_q.adaptive_mps import AdaptiveMPS, DTypePolicy
    m = AdaptiveMPS(len(tensors), bond_dim=1, eps_bond=1e-8, chi_max_per_bond=chi,
                    dtype_policy=DTypePolicy(default=torch.complex64), device="cuda",
                    svd_method=method, theta_precision=prec)
    m.tensors = [t.clone() for t in tensors]
    m.bond_dims = [chi] * (len(tensors) - 1)
    return m


def fidelity_vs_fp64(torch, A, B, gate, chi, methods):
    """Truncation fidelity of each method's output against the FP64 truncation
    of the same theta at the smallest kept rank among the methods."""
    from atlas_q import kernels as K
    A64, B64 = A.to(torch.complex128), B.to(torch.complex128)
    X = K.theta(A64, B64, gate)
    U, S, Vh = torch.linalg.svd(X, full_matrices=False)
    k = min(chi, S.numel(), min(An.shape[2] for An, _ in methods.values()))
    P64 = (U[:, :k] * S[:k]) @ Vh[:k]
    out = {}
    for name, (An, Bn) in methods.items():
        kk = An.shape[2]
        P = (An.reshape(-1, kk).to(torch.complex128) @ Bn.reshape(kk, -1).to(torch.complex128))
        ov = torch.vdot(P64.reshape(-1), P.reshape(-1))
        F = float((ov.abs() ** 2 / (P64.norm() ** 2 * P.norm() ** 2)).item())
        kept = float((P.norm() ** 2 / X.norm() ** 2).item())
        out[name] = {"F": F, "kept_weight": kept}
    out["fp64_kept_weight"] = float((S[:k] ** 2).sum() / (S ** 2).sum())
    return out


def circuit_bench(torch, a, dev):
    """The split-precision circuit itself, from |0...0>: 128 qubits, depth 60,
    Haar SU(4) brickwork, chi_max 1024, cutoff 1e-8, complex64, learned +
    BF16x3, every layer's gates concurrently as in the slice measurement.
    Per layer: wall time, which factorization served each update, the largest
    kept rank; phases split where the largest kept rank first reaches the cap.
    Between layers (outside the timing) the touched sites are rescaled to unit
    norm, as in the slice measurement (the reference's update does not
    renormalise, and ~9 % discarded weight per saturated update would take the
    complex64 state's scale to underflow within a few hundred layers' worth of
    updates).  In the growth phase the middle gate of every few layers is
    checked against the FP64 truncation of the same theta."""
    from atlas_q.adaptive_mps import AdaptiveMPS, DTypePolicy, _gate4
    from atlas_q.circuits import brickwork
    n, depth, chi = a.circuit_qubits, a.circuit_depth, a.chi
    ops = brickwork(n, depth, a.seed + 31)
    m = AdaptiveMPS(n, bond_dim=1, eps_bond=1e-8, chi_max_per_bond=chi,
                    dtype_policy=DTypePolicy(default=torch.complex64), device="cuda",
                    svd_method=a.method, theta_precision=a.theta_precision)
    layers = sublayers(ops)
    fid_every = max(1, 12 // max(1, a.circuit_fidelity))
    per, fids = [], []
    counts0 = None
    for li, run in enumerate(layers):
        st = getattr(m, "factor_stats", None)
        before = dict(st.counts) if st is not None else {}
        sample = None
        if len(fids) < a.circuit_fidelity and li % fid_every == fid_every - 1:
            b, U = run[len(run) // 2]
            sample = (b, U, m.tensors[b].clone(), m.tensors[b + 1].clone())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.apply_two_site_gates(run, workers=a.workers)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = getattr(m, "factor_stats", None)
        after = dict(st.counts) if st is not None else {}
        delta = {k: after.get(k, 0) - before.get(k, 0) for k in after}
        kmax = max(int(m.tensors[b].shape[2]) for b, _ in run)
        per.append({"layer": li, "updates": len(run), "s": dt, "factor_counts": delta,
                    "max_kept_rank": kmax})
        if sample is not None:
            b, U, A0, B0 = sample
            f = fidelity_vs_fp64(torch, A0, B0, _gate4(U), chi,
                                 {"fast": (m.tensors[b], m.tensors[b + 1])})
            fids.append({"layer": li, "bond": b, "kept_rank": int(m.tensors[b].shape[2]),
                         "F": f["fast"]["F"], "fp64_kept_weight": f["fp64_kept_weight"]})
        for b, _ in run:
            for q in (b, b + 1):
                m.tensors[q] = m.tensors[q] / m.tensors[q].norm()
    sat = next((p["layer"] for p in per if p["max_kept_rank"] >= chi), len(per))

    def phase(rows, name):
        if not rows:
            return None
        cnt = {}
        for r in rows:
            for k, v in r["factor_counts"].items():
                cnt[k] = cnt.get(k, 0) + v
        ups = sum(r["updates"] for r in rows)
        t = sum(r["s"] for r in rows)
        return {"phase": name, "layers": f"{rows[0]['layer']}-{rows[-1]['layer']}",
                "updates": ups, "s": t, "updates_per_s": ups / t, "factor_counts": cnt}
    total_u = sum(r["updates"] for r in per)
    total_t = sum(r["s"] for r in per)
    return {"workload": f"{n}-qubit Haar SU(4) brickwork, depth {depth}, chi_max {chi}, cutoff "
                        f"1e-8, complex64, from |0...0>, {a.method} + {a.theta_precision}",
            "updates": total_u, "s": total_t, "updates_per_s": total_u / total_t,
            "phases": [p for p in (phase(per[:sat], "growth (largest kept rank below the cap)"),
                                   phase(per[sat:], "saturated")) if p],
            "per_layer_updates_per_s": [round(r["updates"] / r["s"], 1) for r in per],
            "per_layer_max_kept_rank": [r["max_kept_rank"] for r in per],
            "fidelity_growth_samples": fids,
            "min_fidelity": min((f["F"] for f in fids), default=None)}


def low_entanglement(torch, a, dev):
    """The learned width where it applies: the first brickwork layer of weak
    gates (exp(-0.1 i H)) on a fresh slice at bond CHI whose Schmidt weights
    decay (tau = 10), cutoff 1e-8: kept ranks ~100 of 1024.  All three
    methods on the same slice and gates, sub-layer concurrency as above."""
    from atlas_q.adaptive_mps import _gate4
    from atlas_q.workloads import gate_schedule as weak_schedule
    from atlas_q.workloads import slice_tensors
    chi, sites = a.chi, a.low_sites
    ts = slice_tensors(sites, chi, "decay", seed=a.seed + 7, tau=10.0, device=dev)
    gates = weak_schedule(sites, (sites - 1 + 1) // 2, seed=a.seed + 8, angle=0.1)
    out = {"workload": f"first layer of weak gates exp(-0.1iH) on a {sites}-site slice at bond "
                       f"{chi}, Schmidt weights exp(-2i/10), cutoff 1e-8, complex64",
           "updates": len(gates)}
    ranks = {}
    for name, meth, prec in (("learned", a.method, a.theta_precision), ("gram", "gram", a.theta_precision),
                             ("plain", "plain", "native")):
        run_gates(torch, make_mps(torch, ts, chi, meth, prec), gates, a.workers)   # warm-up
        m = make_mps(torch, ts, chi, meth, prec)
        t = run_gates(torch, m, gates, a.workers)
        out[f"{name}_updates_per_s"] = len(gates) / t
        ranks[name] = m.statistics.logs["k_star"]
        if hasattr(m, "factor_stats"):
            out[f"{name}_factor_counts"] = dict(m.factor_stats.counts)
            if m.factor_stats.rejects:
                out[f"{name}_rejects"] = dict(m.factor_stats.rejects)
    out["kept_ranks"] = ranks
    out["ranks_equal_to_plain"] = {k: v == ranks["plain"] for k, v in ranks.items()}
    out["speedup_learned_vs_plain"] = out["learned_updates_per_s"] / out["plain_updates_per_s"]
    b, U = gates[0]
    outs = {}
    for name, meth, prec in (("fast", a.method, a.theta_precision), ("plain", "plain", "native")):
        mm = make_mps(torch, ts, chi, meth, prec)
        mm.apply_two_site_gate(b, U)
        outs[name] = (mm.tensors[b], mm.tensors[b + 1])
    f = fidelity_vs_fp64(torch, ts[b], ts[b + 1], _gate4(U), chi, outs)
    out["fidelity"] = f
    out["abs_dF_learned_vs_plain"] = abs(f["fast"]["F"] - f["plain"]["F"])
    return out


def sampler_bench(torch, a, dev):
    """Fused sweep sampler (one aq_sample call: per site the MFMA left-state
    GEMM and the draw kernel) at the plain-GPU circuit's shape: n qubits,
    bonds min(chi_max, 2^(b+1), 2^(n-b-1)), complex64, a random
    right-canonical MPS (site 0 a unit vector, every other site a random
    right-isometry) so that the sweep samples the Born distribution exactly."""
    import numpy as np
    from atlas_q import kernels as K
    from atlas_q.workloads import random_right_isometry
    n, cmax, S = a.sampler_qubits, a.sampler_chi, a.sampler_shots
    chi = [1] + [min(cmax, 2 ** min(b + 1, n - b - 1)) for b in range(n - 1)] + [1]
    gen = torch.Generator(device=dev).manual_seed(a.seed + 11)
    ts = [random_right_isometry(chi[i], chi[i + 1], gen).to(torch.complex64).contiguous()
          for i in range(n)]
    u = torch.rand((S, n), dtype=torch.float64, device=dev, generator=gen)
    K.sample_bits(ts, u)                                            # warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(a.sampler_reps):
        bits, _ = K.sample_bits(ts, u)
    e1.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / a.sampler_reps
    ms = e0.elapsed_time(e1) / a.sampler_reps
    flops = float(S) * sum(16.0 * chi[i] * chi[i + 1] for i in range(n))
    # tensors once, per shot and site: left state read (8 chiL), both outcome
    # halves written (16 chiR), the uniform (8)
    nbytes = sum(16.0 * chi[i] * chi[i + 1] for i in range(n)) + \
        float(S) * sum(8.0 * chi[i] + 16.0 * chi[i + 1] + 8.0 for i in range(n))
    tf = flops / (ms * 1e-3) / 1e12
    gbs = nbytes / (ms * 1e-3) / 1e9
    out = {"workload": f"{n} qubits, bonds min({cmax}, 2^(b+1), 2^(n-b-1)), {S} shots, "
                       f"complex64, random right-canonical MPS",
           "shots_per_s": S / wall, "ms_per_call": 1e3 * wall, "kernel_ms_per_call": ms,
           "roofline": {"bound": "mfma", "achieved": tf, "peak": 157.3, "unit": "TFLOP/s",
                        "frac": tf / 157.3, "flops_per_call": flops,
                        "hbm_GBps": gbs, "hbm_frac": gbs / 8000.0}}
    out["roofline"]["traffic"], out["roofline"]["traffic_source"] = traffic_from_profile(
        "aq_sample", flops)
    if not a.no_cpu and a.sampler_cpu_shots > 0:
        from oracle import ref_mps as R
        cores = len(os.sched_getaffinity(0))
        torch.set_num_threads(cores)
        cs = min(a.sampler_cpu_shots, S)
        tc = [t.cpu() for t in ts]
        uc = u[:cs].cpu().numpy()
        t0 = time.perf_counter()
        exp, p0 = R.batch_sweep_sample(tc, uc, return_p0=True)
        t_cpu = time.perf_counter() - t0
        got = np.array(K.bits_to_ints(bits[:cs]))
        clear = (np.abs(uc - p0) > 1e-5).all(axis=1)
        out["cpu_baseline"] = {"value": cs / t_cpu, "unit": "shots/s", "cores": cores,
                               "kind": "port",
                               "sample": f"{cs} of the {S} shots on the same MPS, torch CPU "
                                         f"with {cores} threads"}
        out["bit_exact_vs_oracle"] = {"shots": int(clear.sum()),
                                      "equal": int((got[clear] == np.array(exp)[clear]).sum()),
                                      "margin": 1e-5}
    return out


def wide_bench(torch, a, dev):
    """The learned-rank form's bond (chi_max 2048, theta 4096 x 4096): a
    steady-state slice of WIDE_SITES sites at bond 2048 (the 256-qubit
    depth-80 brickwork's saturated bulk), one sub-layer of warm-up, then one
    odd and one even sub-layer timed; the plain GPU SVD path on a few of the
    same gates; truncation fidelity of both against the FP64 truncation."""
    from atlas_q.adaptive_mps import _gate4
    chi, sites = a.wide_chi, a.wide_sites
    base = slice_state(torch, sites, chi, a.seed + 31, dev)
    sched = gate_schedule(sites, 3 * sites, a.seed + 32)
    runs = sublayers(sched)
    warm, timed = runs[0], runs[1] + runs[2]
    rest = [g for r in runs[3:] for g in r]
    fast = make_mps(torch, base, chi, a.method, a.theta_precision)
    run_gates(torch, fast, warm, a.workers)
    t = run_gates(torch, fast, timed, a.workers)
    out = {"workload": f"steady-state slice of the 256-qubit depth-80 Haar brickwork at "
                       f"chi_max={chi}: {sites} sites, all bonds {chi}, cutoff 1e-8, complex64; "
                       f"theta {2 * chi}x{2 * chi}",
           "updates": len(timed), "updates_per_s": len(timed) / t,
           "kept_ranks": sorted(set(fast.statistics.logs["k_star"][len(warm):])),
           "factor_counts": dict(fast.factor_stats.counts)}
    plain = None
    if not a.no_plain and a.wide_plain_steps > 0:
        pm = make_mps(torch, fast.tensors, chi, "plain", "native")
        run_gates(torch, pm, rest[:1], 1)                                   # library warm-up
        plain = a.wide_plain_steps / run_gates(torch, pm, rest[1: 1 + a.wide_plain_steps],
                                               a.workers)
    out["plain_updates_per_s"] = plain
    out["speedup_vs_plain"] = out["updates_per_s"] / plain if plain else None
    if not a.no_cpu and a.wide_cpu_updates > 0:
        out["cpu_baseline"] = cpu_slice_baseline(torch, base, sched, chi, a.wide_cpu_updates,
                                                 f"{2 * chi}x{2 * chi} complex64 theta")
    fid = []
    if a.wide_fidelity > 0:
        # the fast path as timed (a whole sub-layer, grouped); plain gate by gate
        sub = runs[4]
        fm = make_mps(torch, fast.tensors, chi, a.method, a.theta_precision)
        fm.apply_two_site_gates(sub, workers=a.workers)
        out["fidelity_sublayer_factor_counts"] = dict(fm.factor_stats.counts)
        for b, U in sub[: a.wide_fidelity]:
            pm = make_mps(torch, fast.tensors, chi, "plain", "native")
            pm.apply_two_site_gate(b, U)
            outs = {"fast": (fm.tensors[b], fm.tensors[b + 1]),
                    "plain": (pm.tensors[b], pm.tensors[b + 1])}
            fid.append(fidelity_vs_fp64(torch, fast.tensors[b], fast.tensors[b + 1], _gate4(U),
                                        chi, outs))
    out["fidelity"] = fid
    out["max_abs_dF_fast_vs_plain"] = max((abs(f["fast"]["F"] - f["plain"]["F"]) for f in fid),
                                         default=None)
    return out


def measured_peaks(torch, dev):
    """Reference rates measured on this GPU (context for the roofline
    fractions, which use the published peaks): a 2 GiB device-to-device copy
    (read + write bytes), and rocBLAS FP64 / FP32 GEMMs at 8192^3, each the
    best of 3 after a warm-up, timed with HIP events."""
    def best(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            t.append(e0.elapsed_time(e1) * 1e-3)
        return min(t)
    src = torch.empty(1 << 28, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    t_copy = best(lambda: dst.copy_(src))
    out = {"hbm_copy_GBps": 2 * src.numel() * 8 / t_copy / 1e9}
    del src, dst
    n = 8192
    for name, dt in (("fp64_gemm_TFLOPs", torch.float64), ("fp32_gemm_TFLOPs", torch.float32)):
        A = torch.randn(n, n, dtype=dt, device=dev)
        B = torch.randn(n, n, dtype=dt, device=dev)
        t = best(lambda: torch.matmul(A, B))
        out[name] = 2.0 * n ** 3 / t / 1e12
        del A, B
    out["published"] = {"hbm_GBps": 8000.0, "fp64_mfma_TFLOPs": 78.6, "fp32_mfma_TFLOPs": 157.3}
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_slice_baseline(torch, base, sched, chi, updates, what):
    """The oracle (torch-CPU restatement of the reference update, LAPACK SVD)
    on the host cores: one warm-up update, then ``updates`` timed one by one on
    the slice ``base`` with the gates ``sched``."""
    import numpy as np
    from oracle import ref_mps as R
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    touched = 2 + max(b for b, _ in sched[: 1 + updates])
    om = R.OracleMPS(touched, bond_dim=1, eps_bond=1e-8, chi_max_per_bond=chi)
    om.tensors = [t.cpu() for t in base[:touched]]
    om.bond_dims = [chi] * (touched - 1)
    b, U = sched[0]
    om.apply_two_site_gate(b, torch.from_numpy(U).to(torch.complex64))      # warm-up
    times = []
    for b, U in sched[1: 1 + updates]:
        t0 = time.perf_counter()
        om.apply_two_site_gate(b, torch.from_numpy(U).to(torch.complex64))
        times.append(time.perf_counter() - t0)
    return {"value": updates / sum(times), "unit": "updates/s", "cores": cores,
            "cpu_model": cpu_model(), "kind": "port",
            "median_update_s": float(np.median(times)),
            "sample": f"{updates} two-site updates ({what}, LAPACK SVD, rank rule, rebuild) on "
                      f"the same chi={chi} slice, torch CPU with {cores} threads, after 1 "
                      f"warm-up update"}


def mps_fidelity(torch, ta, tb) -> float:
    """|<a|b>|^2 / (<a|a> <b|b>) of two MPS on the device (complex128
    transfer matrices, one site at a time)."""
    def overlap(x, y):
        E = torch.ones((1, 1), dtype=torch.complex128, device=x[0].device)
        for A, B in zip(x, y):
            A, B = A.to(torch.complex128), B.to(torch.complex128)
            E = sum(A[:, s, :].conj().T @ E @ B[:, s, :] for s in range(2))
        return complex(E.reshape(-1)[0])
    ab = overlap(ta, tb)
    return abs(ab) ** 2 / (overlap(ta, ta).real * overlap(tb, tb).real)


def planned_graph_mode(torch, a, dev, specs, circs, n2q, eager_bm, schmidt_bound):
    """One bound rule of the device-planned circuit: eager planned run, then a
    captured graph replayed three times from |0...0> (median), each replay
    compared bit for bit with the eager planned run, the last one with the
    host-planned run."""
    from atlas_q.batched import BatchedMPS
    from atlas_q.batched_graph import PlannedCircuit
    tag = "schmidt" if schmidt_bound else "exact"
    t0 = time.perf_counter()
    e = BatchedMPS(specs, eps_bond=1e-8, device=dev, eig_solver=a.batched_eig)
    pe = PlannedCircuit(e, e.compile_circuits(circs), schmidt_bound=schmidt_bound)
    pe.run()
    pe.finish()
    log(f"batched graph ({tag} bounds): eager planned run done in {time.perf_counter() - t0:.1f}s "
        f"(arenas {pe.arena_bytes / 2 ** 30:.2f} GiB, {len(pe.steps)} steps)")
    bm = BatchedMPS(specs, eps_bond=1e-8, device=dev, eig_solver=a.batched_eig)
    arena0 = bm.arena.clone()
    plan = PlannedCircuit(bm, bm.compile_circuits(circs), schmidt_bound=schmidt_bound)
    dims0 = plan.dims.clone()
    g = plan.capture()
    torch.cuda.synchronize()
    log(f"batched graph ({tag} bounds): captured in {time.perf_counter() - t0:.1f}s")
    times, equal = [], []
    for rep in range(3):
        bm.arena.copy_(arena0)
        plan.dims.copy_(dims0)
        plan.fails.zero_()
        plan.overflow.zero_()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
        equal.append(bool(torch.equal(bm.arena, e.arena)) and bool(torch.equal(plan.dims, pe.dims)))
    fails = int(plan.fails.item())
    plan.finish()
    def bound(m, j):
        return math.sqrt(sum(x * x for mj, x in zip(m.logs["mps"], m.logs["eps_local"])
                             if mj == j))
    light, heavy = [], []
    for j in range(len(specs)):
        eg, eh = bound(bm, j), bound(eager_bm, j)
        if eh < 1e-3:
            light.append(mps_fidelity(torch, bm.tensors(j), eager_bm.tensors(j)))
        else:
            heavy.append(abs(eg - eh) / eh)
    t = sorted(times)[1]
    log(f"batched graph ({tag} bounds): replay {t:.3f}s")
    out = {"replay_s": t, "replay_s_runs": times, "updates_per_s": n2q / t,
           "replays_bit_identical_to_eager_planned_run": equal, "eig_failures": fails,
           "ranks_lowered_to_bound": plan.clamped, "arena_gib": plan.arena_bytes / 2 ** 30,
           "vs_host_planned_run": {
               "min_fidelity_where_truncation_negligible": min(light) if light else None,
               "median_rel_diff_error_bound_where_caps_bind":
                   float(sorted(heavy)[len(heavy) // 2]) if heavy else None,
               "max_rel_diff_error_bound_where_caps_bind": max(heavy) if heavy else None}}
    del g, plan, pe, e, bm, arena0
    torch.cuda.empty_cache()
    return out


def planned_graph_bench(torch, a, dev, specs, circs, n2q, eager_bm):
    """The same population as a device-planned circuit (bond dimensions on the
    device, launch shapes from bounds) captured into one HIP graph, under
    both bound rules: the rank rule's own limits min(cap, 2 chi_l, 2 chi_r)
    (exact: the host-planned semantics) and the tighter one with the gates'
    operator Schmidt rank (round-off ranks past it lowered)."""
    return {"what": "the population's whole 8-layer circuit as one HIP graph: bond dimensions on "
                    "the device, every launch shape planned from bond-dimension bounds",
            "exact_bounds": planned_graph_mode(torch, a, dev, specs, circs, n2q, eager_bm, False),
            "schmidt_bounds": planned_graph_mode(torch, a, dev, specs, circs, n2q, eager_bm, True),
            "note": "where caps bind the two plans take the Gram matrix of different sides on some "
                    "gates; the float32 rank rule's tail weights are quantised near the storage "
                    "round-off, so single error bounds can differ by a factor of two while the "
                    "population median stays within a few per cent",
            "planned_launches_note": "the host-planned run learns the bond dimensions after every "
                                     "sub-layer (one device-to-host copy each); the planned run "
                                     "never synchronises"}


def batched_bench(torch, a, dev):
    """The batched form: a population of independent MPS (qubit counts
    log-uniform in [12, 96], caps log-uniform in [16, 256]) running the RY/RZ +
    CZ-ladder ansatz, advanced together by grouped persistent launches
    (``BatchedMPS``), then sampled with the grouped sweep captured in a HIP
    graph.  Reference: the same circuits one MPS at a time with the plain
    AdaptiveMPS method, on a subset, as updates/s."""
    import numpy as np
    from atlas_q.adaptive_mps import AdaptiveMPS
    from atlas_q.batched import BatchedMPS
    from atlas_q.circuits import batched_hea_population, hea
    specs = batched_hea_population(a.batched_mps, a.seed + 21)
    circs = [hea(n, a.batched_layers, a.seed + 100 + j) for j, (n, _) in enumerate(specs)]
    n2q = sum(sum(1 for op in c if op[0] == "2q") for c in circs)
    n1q = sum(sum(1 for op in c if op[0] == "1q") for c in circs)
    t0 = time.perf_counter()
    BatchedMPS(specs, eps_bond=1e-8, device=dev,
               eig_solver=a.batched_eig).run_hea(circs)      # warm-up / autotune
    torch.cuda.synchronize()
    log(f"batched: warm-up run {time.perf_counter() - t0:.2f}s")
    runs = []
    for _ in range(3):                 # the median of three runs from |0...0>
        bm = BatchedMPS(specs, eps_bond=1e-8, device=dev, eig_solver=a.batched_eig)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bm.run_hea(circs)
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - t0)
        log(f"batched: run {runs[-1]:.3f}s")
    t_circ = sorted(runs)[1]
    graph = None
    if not a.no_batched_graph:
        graph = planned_graph_bench(torch, a, dev, specs, circs, n2q, bm)
    S = a.batched_shots
    bm.prepare_sampler(S)
    gen = torch.Generator(device=dev).manual_seed(a.seed + 22)
    bm.uniforms.copy_(torch.rand(bm.uniforms.shape, dtype=torch.float64, device=dev,
                                 generator=gen))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        bm.sweep()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        bm.sweep()
    g.replay()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    t_samp = (time.perf_counter() - t0) / reps
    log(f"batched: sampling {t_samp * 1e3:.1f} ms per sweep")
    ref = None
    fids = []
    if a.batched_ref > 0:
        r_up, r_t = 0, 0.0
        for j in range(min(a.batched_ref, len(specs))):
            n, chi = specs[j]
            m = AdaptiveMPS(n, bond_dim=1, eps_bond=1e-8, chi_max_per_bond=chi, device=dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for op in circs[j]:
                if op[0] == "1q":
                    m.apply_single_qubit_gate(op[1], torch.from_numpy(op[2]))
                else:
                    m.apply_two_site_gate(op[1], torch.from_numpy(op[2]))
            torch.cuda.synchronize()
            r_t += time.perf_counter() - t0
            r_up += sum(1 for op in circs[j] if op[0] == "2q")
            eb = math.sqrt(sum(e * e for mj, e in zip(bm.logs["mps"], bm.logs["eps_local"])
                               if mj == j))
            ep = math.sqrt(sum(e * e for e in m.statistics.logs["eps_local"]))
            fids.append({"qubits": n, "cap": chi, "fidelity": mps_fidelity(torch, bm.tensors(j),
                                                                          m.tensors),
                         "error_bound_batched": eb, "error_bound_plain": ep})
        log(f"batched: unbatched reference on {len(fids)} MPS done")
        light = [f["fidelity"] for f in fids if f["error_bound_plain"] < 1e-3]
        heavy = [abs(f["error_bound_batched"] - f["error_bound_plain"]) / f["error_bound_plain"]
                 for f in fids if f["error_bound_plain"] >= 1e-3]
        ref = {"mps": min(a.batched_ref, len(specs)), "updates_per_s": r_up / r_t,
               "method": "plain AdaptiveMPS, one MPS and one gate at a time",
               "per_mps": fids,
               "min_fidelity_where_truncation_negligible": min(light) if light else None,
               "max_rel_diff_error_bound_where_caps_bind": max(heavy) if heavy else None,
               "note": "where caps cut through the exactly degenerate Schmidt spectra of these "
                       "circuits, any two SVD implementations keep different (equally valid) "
                       "subspaces; the reference run in complex64 and in complex128 differs "
                       "there too, so the truncated states are compared by their global error "
                       "bound sqrt(sum eps_local^2)"}
    cpu = None
    if not a.no_cpu and a.batched_cpu_mps > 0:
        from oracle import ref_mps as R
        cores = len(os.sched_getaffinity(0))
        torch.set_num_threads(cores)
        c_up, c_t = 0, 0.0
        for j in range(min(a.batched_cpu_mps, len(specs))):
            n, chi = specs[j]
            om = R.OracleMPS(n, bond_dim=1, eps_bond=1e-8, chi_max_per_bond=chi)
            om.tensors = [torch.tensor([1.0, 0.0], dtype=torch.complex64).reshape(1, 2, 1)
                          for _ in range(n)]
            om.bond_dims = [1] * (n - 1)
            t0 = time.perf_counter()
            for op in circs[j]:
                U = torch.from_numpy(op[2]).to(torch.complex64)
                if op[0] == "1q":
                    om.apply_single_qubit_gate(op[1], U)
                else:
                    om.apply_two_site_gate(op[1], U)
            c_t += time.perf_counter() - t0
            c_up += sum(1 for op in circs[j] if op[0] == "2q")
        cpu = {"value": c_up / c_t, "unit": "updates/s", "cores": cores, "cpu_model": cpu_model(),
               "kind": "port",
               "sample": f"the first {min(a.batched_cpu_mps, len(specs))} MPS of the population "
                         f"({c_up} two-site updates), one gate at a time, torch CPU with {cores} "
                         f"threads; a per-MPS rate, not extrapolated to the population"}
    chis = [int(bm.dims[j].max()) for j in range(len(specs))]
    from atlas_q.autotune import Autotuner
    shapes = {k: v for k, v in sorted(Autotuner.default().table.items()) if ":shape:" in k}
    out = {"workload": f"{len(specs)} MPS, qubits log-uniform 12-96, caps log-uniform 16-256, "
                       f"RY/RZ + CZ-ladder ansatz x {a.batched_layers} layers, complex64, "
                       f"cutoff 1e-8, {S} shots each",
           "two_site_updates": n2q, "single_qubit_gates": n1q,
           "circuit_s": t_circ, "circuit_s_runs": runs, "updates_per_s": n2q / t_circ,
           "max_bond_reached": max(chis),
           "eig_fallbacks": int(bm.eig_fallbacks),
           "eig_solver_above_order_64": a.batched_eig,
           "grouped_gemm_shapes": {
               "what": "autotuned (tile rows, work-groups) per grouped GEMM family and "
                       "log2 work-size bucket",
               "chosen": shapes,
               "rows32_share": (sum(1 for v in shapes.values() if v[0] == 32) / len(shapes)
                                if shapes else None)},
           "graph_replay": graph,
           "shots": S * len(specs), "sample_s": t_samp,
           "shots_per_s": S * len(specs) / t_samp, "sampler": "grouped sweep, HIP-graph replay",
           "unbatched_reference": ref,
           "speedup_vs_unbatched": (n2q / t_circ) / ref["updates_per_s"] if ref else None,
           "cpu_baseline": cpu}
    return out


def marker(torch):
    """One launch of the library's zero-fill kernel on one word: a dispatch
    that brackets a region in a kernel trace."""
    import ctypes
    from atlas_q._ffi import lib
    buf = torch.empty(1, dtype=torch.int64, device="cuda")
    lib().aq_zero_u64(ctypes.c_void_p(buf.data_ptr()), 1,
                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()


def traffic_from_profile(kernel: str, work_per_launch: float):
    """HBM bytes per launch of `kernel` from the committed counter summaries
    (FETCH_SIZE and WRITE_SIZE runs, calibrated on a copy of known size),
    scaled from the profiled launch to this one by the ratio of their
    algorithmic work (the profiled launch's work is recorded beside its
    bytes); (None, None) when no summary covers the kernel."""
    for path in sorted((ROOT / "profiles").glob("*traffic*.md")):
        bytes_, work = None, None
        for line in path.read_text().splitlines():
            if line.startswith(f"traffic_bytes_per_launch[{kernel}]"):
                bytes_ = float(line.split("=")[1])
            elif line.startswith(f"work_per_launch[{kernel}]"):
                work = float(line.split("=")[1])
        if bytes_ is not None and work:
            return bytes_ * (work_per_launch / work), str(path.relative_to(ROOT))
    return None, None


def eig_probe(torch, a):
    """rocSOLVER's strided-batched Hermitian eigensolvers (the rocSOLVER the
    library links) on copies of one Gram matrix G = X X^H, X complex Gaussian:
    (1) zheevd vs zheevdj vs zheevj (batch 64; orders of the batched form's
    buckets) and at order 2 chi the tridiagonalisation alone (zhetrd), the
    partial-spectrum zheevdx (top half) and complex64 cheevd; (2) zheevd vs
    batch size at order 2 chi; (3) 64 matrices as 1, 2, 4 or 8 concurrent
    calls, one host thread, rocBLAS handle and stream each.  HIP events on the
    calls' stream, best of the timed repetitions after a warm-up call."""
    import ctypes
    import threading
    from atlas_q._ffi import lib
    lib()                                               # loads the library's rocSOLVER
    rb = ctypes.CDLL("librocblas.so.5")
    rs = ctypes.CDLL("librocsolver.so.0")
    vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    FILL_UPPER, EVECT, ERANGE_INDEX, ESORT_ASC = 121, 211, 233, 252

    def handle(stream):
        h = vp()
        assert rb.rocblas_create_handle(ctypes.byref(h)) == 0
        assert rb.rocblas_set_stream(h, vp(stream.cuda_stream)) == 0
        return h

    dev = torch.device("cuda", 0)
    gen = torch.Generator(device=dev).manual_seed(a.seed)
    p = lambda t: vp(t.data_ptr())

    def gram(n):
        X = torch.randn((n, n), dtype=torch.complex128, device=dev, generator=gen)
        return X @ X.conj().T

    def call(h, meth, n, A, b, W, E, info, aux):
        nn = n * n
        if meth == "zheevd":
            r = rs.rocsolver_zheevd_strided_batched(h, i32(EVECT), i32(FILL_UPPER), i32(n), p(A),
                                                    i32(n), i64(nn), p(W), i64(n), p(E), i64(n),
                                                    p(info), i32(b))
        elif meth == "zheevdj":
            r = rs.rocsolver_zheevdj_strided_batched(h, i32(EVECT), i32(FILL_UPPER), i32(n), p(A),
                                                     i32(n), i64(nn), p(W), i64(n), p(info), i32(b))
        elif meth == "zheevj":
            r = rs.rocsolver_zheevj_strided_batched(h, i32(ESORT_ASC), i32(EVECT), i32(FILL_UPPER),
                                                    i32(n), p(A), i32(n), i64(nn), f64(0.0),
                                                    p(aux["res"]), i32(100), p(aux["sweeps"]),
                                                    p(W), i64(n), p(info), i32(b))
        elif meth == "zhetrd":
            r = rs.rocsolver_zhetrd_strided_batched(h, i32(FILL_UPPER), i32(n), p(A), i32(n),
                                                    i64(nn), p(W), i64(n), p(E), i64(n),
                                                    p(aux["tau"]), i64(n), i32(b))
        elif meth == "zheevdx":
            r = rs.rocsolver_zheevdx_strided_batched(h, i32(EVECT), i32(ERANGE_INDEX),
                                                     i32(FILL_UPPER), i32(n), p(A), i32(n),
                                                     i64(nn), f64(0.0), f64(0.0), i32(n // 2 + 1),
                                                     i32(n), p(aux["nev"]), p(W), i64(n),
                                                     p(aux["Z"]), i32(n), i64(nn), p(info), i32(b))
        elif meth == "cheevd":
            r = rs.rocsolver_cheevd_strided_batched(h, i32(EVECT), i32(FILL_UPPER), i32(n),
                                                    p(aux["A32"]), i32(n), i64(nn), p(aux["W32"]),
                                                    i64(n), p(aux["E32"]), i64(n), p(info), i32(b))
        assert r == 0, (meth, r)

    def timed(meth, n, G, b, reps):
        st = torch.cuda.current_stream(dev)
        h = handle(st)
        A = torch.empty((b, n, n), dtype=torch.complex128, device=dev)
        W = torch.empty((b, n), dtype=torch.float64, device=dev)
        E = torch.empty((b, n), dtype=torch.float64, device=dev)
        info = torch.zeros(b, dtype=torch.int32, device=dev)
        aux = {"res": torch.empty(b, dtype=torch.float64, device=dev),
               "sweeps": torch.zeros(b, dtype=torch.int32, device=dev),
               "tau": torch.empty((b, n), dtype=torch.complex128, device=dev),
               "nev": torch.zeros(b, dtype=torch.int32, device=dev)}
        if meth == "zheevdx":
            aux["Z"] = torch.empty((b, n, n), dtype=torch.complex128, device=dev)
        if meth == "cheevd":
            aux["A32"] = torch.empty((b, n, n), dtype=torch.complex64, device=dev)
            aux["W32"] = torch.empty((b, n), dtype=torch.float32, device=dev)
            aux["E32"] = torch.empty((b, n), dtype=torch.float32, device=dev)
        best = float("inf")
        for rep in range(reps + 1):
            A.copy_(G.expand(b, n, n))
            if meth == "cheevd":
                aux["A32"].copy_(A)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            call(h, meth, n, A, b, W, E, info, aux)
            e1.record()
            torch.cuda.synchronize()
            if rep:
                best = min(best, e0.elapsed_time(e1))
        rb.rocblas_destroy_handle(h)
        w = (aux["W32"].double() if meth == "cheevd" else W)[-1]
        return best, int((info != 0).sum()), w

    out = {"variants": [], "batch_scan": [], "concurrent": []}
    n2 = 2 * a.chi
    for n in (128, 256, 512, n2):
        G = gram(n)
        ref = None
        meths = (("zheevd", 64), ("zheevdj", 64), ("zheevj", 64)) if n <= 512 else \
            (("zheevd", 64), ("zheevdj", 64), ("zhetrd", 64), ("cheevd", 64), ("zheevdx", 8))
        for meth, b in meths:
            ms, bad, w = timed(meth, n, G, b, 1 if ms_slow(meth, n) else 2)
            if meth == "zheevd":
                ref = w
            dw = None if meth == "zhetrd" else float((w.sort().values - ref.sort().values).abs().max()
                                                     / ref.abs().max()) if meth != "zheevdx" else None
            row = {"method": meth, "n": n, "batch": b, "ms": ms, "ms_per_matrix": ms / b,
                   "info_nonzero": bad, "max_rel_dW_vs_zheevd": dw}
            out["variants"].append(row)
            print(f"{meth:8s} n={n:5d} batch={b:3d} {ms:10.2f} ms {ms / b:9.3f} ms/matrix"
                  f"  info!=0: {bad}" + ("" if dw is None else f"  max|dW|/max|W| vs zheevd {dw:.1e}"),
                  flush=True)
    G = gram(n2)
    for b in (1, 4, 8, 16, 32, 64):
        ms, bad, _ = timed("zheevd", n2, G, b, 2)
        out["batch_scan"].append({"n": n2, "batch": b, "ms": ms, "ms_per_matrix": ms / b})
        print(f"zheevd   n={n2:5d} batch={b:3d} {ms:10.2f} ms {ms / b:9.3f} ms/matrix", flush=True)
    total = 64
    A = torch.empty((total, n2, n2), dtype=torch.complex128, device=dev)
    W = torch.empty((total, n2), dtype=torch.float64, device=dev)
    E = torch.empty((total, n2), dtype=torch.float64, device=dev)
    info = torch.zeros(total, dtype=torch.int32, device=dev)
    for c in (1, 2, 4, 8):
        per = total // c
        streams = [torch.cuda.Stream(dev) for _ in range(c)]
        hs = [handle(s) for s in streams]
        best = float("inf")
        for rep in range(3):
            A.copy_(G.expand(total, n2, n2))
            torch.cuda.synchronize()
            t0 = time.perf_counter()

            def run(j):
                sl = slice(j * per, (j + 1) * per)
                call(hs[j], "zheevd", n2, A[sl], per, W[sl], E[sl], info[sl], None)

            th = [threading.Thread(target=run, args=(j,)) for j in range(c)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            torch.cuda.synchronize()
            if rep:
                best = min(best, 1e3 * (time.perf_counter() - t0))
        for h in hs:
            rb.rocblas_destroy_handle(h)
        out["concurrent"].append({"n": n2, "matrices": total, "calls": c, "ms": best,
                                  "ms_per_matrix": best / total})
        print(f"zheevd   n={n2:5d} {total} matrices as {c} concurrent calls of {per:2d}"
              f" {best:10.2f} ms {best / total:9.3f} ms/matrix", flush=True)
    print(json.dumps({"eig_probe": out}), flush=True)


def ms_slow(meth, n):
    """One timed repetition for the calls that take seconds."""
    return meth in ("zheevdj", "zheevdx") and n > 512 or meth == "zheevj" and n >= 512


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the two-site update and the "
                          "sampler run on a single MI355X; multi-GPU sharding is design-only",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    from atlas_q import kernels as K
    if a.gemm_probe:
        # the update path's grouped FP64 launches at the bench's shape: the
        # lower-triangle Gram X X^H of 64 thetas 2048 x 2048, then the 64
        # projections U_k^H X (k = 1024) reading eigenvector rows in place
        import numpy as np
        from atlas_q._ffi import AQ_GROUPED_HERK
        B, m, k = 64, 2 * a.chi, a.chi
        gen = torch.Generator(device="cuda").manual_seed(a.seed)
        X = torch.randn((B, m, m), dtype=torch.complex128, device="cuda", generator=gen)
        G = torch.zeros((B, m, m), dtype=torch.complex128, device="cuda")
        V = torch.empty((B, k, m), dtype=torch.complex128, device="cuda")
        gd = np.stack([K.gemm_desc(X[b], X[b].T, G[b], conj_b=True, kind=AQ_GROUPED_HERK)
                       for b in range(B)])
        pd = np.stack([K.gemm_desc(G[b], X[b], V[b], M=k, K=m, N=m,
                                   a_ptr=G[b, m - 1].data_ptr(), a_strides=(-m, 1))
                       for b in range(B)])
        for _ in range(3):
            K.grouped_gemm_descs(gd, torch.complex128, "cuda")
        for _ in range(3):
            K.grouped_gemm_descs(pd, torch.complex128, "cuda")
        torch.cuda.synchronize()
        print(json.dumps({"gemm_probe": [B, m, k], "launches": [3, 3]}), flush=True)
        return
    if a.eig_probe:
        eig_probe(torch, a)
        return
    if a.hetrd_probe:
        from atlas_q.eig import GramTridiag
        n = 2 * a.chi
        g = torch.Generator(device="cuda").manual_seed(a.seed)
        X = torch.randn((n, n), dtype=torch.complex128, device="cuda", generator=g)
        Gb = (X @ X.conj().T).expand(a.hetrd_probe, n, n).contiguous()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        GramTridiag(Gb)
        torch.cuda.synchronize()
        print(json.dumps({"hetrd_probe_batch": a.hetrd_probe, "order": n,
                          "ms_per_call": 1e3 * (time.perf_counter() - t0)}), flush=True)
        return
    if a.theta_probe:
        g = torch.Generator(device="cuda").manual_seed(a.seed)
        A = torch.randn((a.chi, 2, a.chi), dtype=torch.complex64, device="cuda", generator=g)
        B = torch.randn((a.chi, 2, a.chi), dtype=torch.complex64, device="cuda", generator=g)
        gate = np.random.default_rng(a.seed).standard_normal(32)
        X = K.theta(A, B, gate, prec="bf16x3")
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(a.theta_probe):
            K.theta(A, B, gate, prec="bf16x3", out=X)
        e.record()
        e.synchronize()
        print(json.dumps({"theta_probe_calls": a.theta_probe, "chi": a.chi,
                          "us_per_call": s.elapsed_time(e) / a.theta_probe * 1e3}), flush=True)
        return
    if a.copy_probe:
        src = torch.ones(a.copy_probe << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4, "copies": 5}), flush=True)
        return

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    from atlas_q import fastsvd
    if a.gram_chunks:
        fastsvd.GRAM_CHUNKS = tuple(float(f) for f in a.gram_chunks.split(","))
    if a.tile_rows:
        from atlas_q import autotune
        autotune.TILE_ROWS = tuple(int(r) for r in a.tile_rows.split(","))
    if a.sampler_only:
        print(json.dumps({"sampler": sampler_bench(torch, a, dev)}), flush=True)
        return
    if a.batched_only:
        print(json.dumps({"batched": batched_bench(torch, a, dev)}), flush=True)
        return
    if a.wide_only:
        print(json.dumps({"learned_rank_form": wide_bench(torch, a, dev)}), flush=True)
        return
    if a.circuit_only:
        print(json.dumps({"circuit": circuit_bench(torch, a, dev)}), flush=True)
        return
    chi, sites = a.chi, a.sites
    base = slice_state(torch, sites, chi, a.seed, dev)
    total = a.warmup + a.steps
    sched = gate_schedule(sites, total + max(a.plain_steps, a.seq_steps, sites), a.seed + 1)

    log(f"slice ready: {sites} sites at bond {chi}")

    # ---------------------------------------------------------------- fast path
    from atlas_q import roofline as RL
    fast = make_mps(torch, base, chi, a.method, a.theta_precision)
    run_gates(torch, fast, sched[: a.warmup], a.workers)
    RL.reset()
    RL.enable(True)
    if a.mark:
        marker(torch)
    t_fast = run_gates(torch, fast, sched[a.warmup: total], a.workers)
    if a.mark:
        marker(torch)
    RL.enable(False)
    ups = a.steps / t_fast
    log(f"fast path: {ups:.2f} updates/s")
    roof = RL.summary(window_s=t_fast, units=a.steps)
    if roof:
        d = roof["all"][roof["kernel"]]
        roof["traffic"], roof["traffic_source"] = traffic_from_profile(
            roof["kernel"], d["work_per_launch"])
    if roof and a.workers > 1:
        # per-update launches go out from several worker streams at once: their
        # event pairs span the other streams' kernels as well
        for name, e in roof["all"].items():
            if e["launches"] >= a.steps and not name.startswith("eig_"):   # once per update
                e["timing"] = (f"per-launch HIP events with {a.workers} worker streams running "
                               "concurrently: the span includes other streams' kernels, so "
                               "'achieved' is a lower bound on this kernel's own rate")
    eig = None
    if roof:
        eig = {"calls": {}}
        for key, e in roof["all"].items():
            if key.startswith("eig_"):
                eig["calls"][key] = {"calls": e["launches"], "mean_ms_per_call": e["mean_ms"],
                                     "total_ms": e["total_ms"],
                                     "share_of_timed_region": e["total_ms"] / (1e3 * t_fast),
                                     "achieved": e["achieved"], "unit": e["unit"],
                                     "frac": e["frac"]}
        eig["ms_per_update"] = sum(v["total_ms"] for v in eig["calls"].values()) / a.steps
        eig["solver"] = fastsvd.EIG_SOLVER
        eig["gram_chunks"] = list(fastsvd.GRAM_CHUNKS)
        orth = getattr(getattr(fast, "factor_stats", None), "orth", [])
        eig["max_orthogonality_error"] = max(orth) if orth else None
    ks = fast.statistics.logs["k_star"][a.warmup:]
    fast_seq = None
    if a.workers > 1 and a.seq_steps > 0:
        fs = make_mps(torch, fast.tensors, chi, a.method, a.theta_precision)
        run_gates(torch, fs, sched[total - 1: total], 1)             # main-thread warm-up
        fast_seq = a.seq_steps / run_gates(torch, fs, sched[total: total + a.seq_steps], 1)

    # ---------------------------------------------------------------- plain path
    plain = plain_seq = None
    if not a.no_plain and a.plain_steps > 0:
        pm = make_mps(torch, fast.tensors, chi, "plain", "native")
        run_gates(torch, pm, sched[total - a.workers: total], a.workers)  # library warm-up
        plain = a.plain_steps / run_gates(torch, pm, sched[total: total + a.plain_steps],
                                          a.workers)
        if a.workers > 1 and a.seq_steps > 0:
            ps = make_mps(torch, fast.tensors, chi, "plain", "native")
            run_gates(torch, ps, sched[total - 1: total], 1)
            plain_seq = a.seq_steps / run_gates(torch, ps, sched[total: total + a.seq_steps], 1)

    log(f"plain path: {plain} updates/s")
    # ---------------------------------------------------------------- fidelity
    fid = []
    fid_counts = None
    if a.fidelity_samples > 0:
        from atlas_q.adaptive_mps import _gate4
        state = fast.tensors
        # the fast path as timed: the whole next sub-layer through the grouped
        # factorization; the plain path gate by gate (no grouping there)
        sub = sublayers(sched[total: total + sites])[0]
        fm = make_mps(torch, state, chi, a.method, a.theta_precision)
        fm.apply_two_site_gates(sub, workers=a.workers)
        fid_counts = dict(fm.factor_stats.counts)
        step = max(1, len(sub) // a.fidelity_samples)
        for b, U in sub[::step][: a.fidelity_samples]:
            pm = make_mps(torch, state, chi, "plain", "native")
            pm.apply_two_site_gate(b, U)
            outs = {"fast": (fm.tensors[b], fm.tensors[b + 1]),
                    "plain": (pm.tensors[b], pm.tensors[b + 1])}
            fid.append(fidelity_vs_fp64(torch, state[b], state[b + 1], _gate4(U), chi, outs))
    dF = max((abs(f["fast"]["F"] - f["plain"]["F"]) for f in fid), default=None)

    log("fidelity done")
    low = None if a.no_low else low_entanglement(torch, a, dev)
    log("low-entanglement done")
    samp = None if a.no_sampler else sampler_bench(torch, a, dev)
    log("sampler done")
    bat = None if a.no_batched else batched_bench(torch, a, dev)
    log("batched done")
    wide = None if a.no_wide else wide_bench(torch, a, dev)
    log("learned-rank form (chi 2048) done")
    circ = None if a.no_circuit else circuit_bench(torch, a, dev)
    log("whole circuit done")

    peaks = None if a.no_peaks else measured_peaks(torch, dev)
    log("measured peaks done")
    # ---------------------------------------------------------------- CPU baseline
    cpu = None
    if not a.no_cpu and a.cpu_updates > 0:
        cpu = cpu_slice_baseline(torch, base, sched, chi, a.cpu_updates,
                                 f"{2 * chi}x{2 * chi} complex64 theta")
    log("CPU baseline done")

    line = {
        "metric": METRIC, "value": ups, "unit": "updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * t_fast / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64 (bf16x3 theta) + f64 factorization",
        "data": "synthetic",
        "config": {"workload": f"steady-state slice of the 128-qubit depth-60 Haar brickwork at "
                               f"chi_max={chi}: {sites} sites, all bonds {chi}, cutoff 1e-8, "
                               f"complex64; theta {2*chi}x{2*chi}, kept rank {chi}",
                   "method": a.method, "theta_precision": a.theta_precision,
                   "parallelism": f"single GPU; independent gates of a sub-layer on "
                                  f"{a.workers} HIP streams (both paths)"},
        "plain_updates_per_s": plain,
        "speedup_vs_plain": (ups / plain) if plain else None,
        "sequential": {"fast_updates_per_s": fast_seq, "plain_updates_per_s": plain_seq,
                       "speedup": (fast_seq / plain_seq) if fast_seq and plain_seq else None},
        "kept_ranks": sorted(set(ks)),
        "fidelity": {"max_abs_dF_fast_vs_plain": dF, "samples": fid,
                     "fast_sublayer_factor_counts": fid_counts},
        "factor_counts": fast.factor_stats.counts if hasattr(fast, "factor_stats") else None,
        "factor_rejects": fast.factor_stats.rejects if hasattr(fast, "factor_stats") else None,
        "low_entanglement": low,
        "circuit": circ,
        "roofline": roof,
        "eigensolver": eig,
        "cpu_baseline": cpu,
        "sampler": samp,
        "batched": bat,
        "learned_rank_form": wide,
        "measured_peaks": peaks,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
