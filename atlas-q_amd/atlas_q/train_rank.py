"""Harvest (features, kept rank) pairs from the plain method and train the
rank predictor on the GPU.

    python -m atlas_q.train_rank --pairs 20000 --out atlas_q/rank_mlp_weights.npz

Harvest: seeded random circuits run with ``svd_method="plain"`` (the
reference-compatible full SVD); before each factorization the predictor's 32
features of the theta matrix are recorded, after it the rank the reference's
rule kept.  Circuit families: Haar brickwork, weakly entangling brickwork
(exp(-i a H), H Hermitian Gaussian, a in [0.03, 0.6]) and the RY/RZ + CZ
hardware-efficient ansatz, over chain lengths 8-40, caps 8-256 and cutoffs
1e-2 .. 1e-8; and steady-state slices with flat or decaying Schmidt
profiles (``atlas_q.workloads``) at bond 32-2048 under Haar or weak gates
(bond >= 512 harvested with the exact Gram method, whose ranks are the plain
method's up to float32 rounding), so kept ranks up to 2048 are in the data.  Split by circuit (90/10), as the
reference's dataset builder
splits by file (scripts/ai_tools/build_ft_dataset.py:103-108).
"""
from __future__ import annotations

import argparse
import time

import numpy as np
import torch

from .adaptive_mps import AdaptiveMPS, DTypePolicy
from .circuits import brickwork, haar_su4, hea
from .rank_mlp import DEFAULT_WEIGHTS, RankPredictor


def weak_brickwork(n, depth, seed, angle):
    rng = np.random.default_rng(seed)
    ops = []
    for layer in range(depth):
        for b in range(layer % 2, n - 1, 2):
            Hm = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
            Hm = (Hm + Hm.conj().T) / 2
            w, v = np.linalg.eigh(Hm)
            ops.append(("2q", b, (v * np.exp(-1j * angle * w)) @ v.conj().T))
    return ops


def slice_job(rng):
    """Steady-state slices (bonds chi at both ends) with flat or decaying
    Schmidt profiles, Haar or weakly entangling gates."""
    from .workloads import gate_schedule, slice_tensors
    chi = int(2 ** rng.integers(5, 12))
    kind = "flat" if rng.random() < 0.3 else "decay"
    tau = float(rng.uniform(3.0, 40.0))
    angle = None if rng.random() < 0.3 else float(rng.uniform(0.02, 0.6))
    eps = float(10.0 ** -rng.integers(2, 9))
    big = chi >= 512
    sites = 4 if big else int(rng.integers(4, 7))
    ts = slice_tensors(sites, chi, kind, seed=int(rng.integers(1 << 30)), tau=tau)
    gates = gate_schedule(sites, int(rng.integers(3, 7) if big else rng.integers(4, 10)),
                          int(rng.integers(1 << 30)), angle)
    return chi, eps, ts, gates


def circuit(rng, cid):
    fam = cid % 3
    n = int(rng.integers(8, 41))
    cap = int(2 ** rng.integers(3, 9))
    eps = float(10.0 ** -rng.integers(2, 9))
    depth = int(rng.integers(4, 21))
    if fam == 0:
        ops = [("2q", b, U) for b, U in brickwork(n, depth, int(rng.integers(1 << 30)))]
    elif fam == 1:
        ops = weak_brickwork(n, depth, int(rng.integers(1 << 30)), float(rng.uniform(0.03, 0.6)))
    else:
        ops = hea(n, int(rng.integers(2, 9)), int(rng.integers(1 << 30)))
    return n, cap, eps, ops


def harvest(pairs: int, seed: int = 0, max_seconds: float = 600.0, device="cuda"):
    rng = np.random.default_rng(seed)
    pred = RankPredictor(device)
    F, K, C = [], [], []
    cid = 0
    t0 = last = time.time()
    while len(K) < pairs and time.time() - t0 < max_seconds:
        if cid % 4 == 3:
            cap, eps, ts, gates = slice_job(rng)
            # bond >= 512: the exact FP64 Gram factorization stands in for the
            # float32 SVD (same rule, same ranks up to float32 rounding, 10x faster)
            m = AdaptiveMPS(len(ts), bond_dim=1, eps_bond=eps, chi_max_per_bond=cap,
                            device=device, svd_method="gram" if cap >= 512 else "plain",
                            rank_predictor=pred)
            m.tensors = ts
            m.bond_dims = [cap] * (len(ts) - 1)
            ops = [("2q", b, U) for b, U in gates]
        else:
            n, cap, eps, ops = circuit(rng, cid)
            m = AdaptiveMPS(n, bond_dim=1, eps_bond=eps, chi_max_per_bond=cap, device=device,
                            svd_method="plain", rank_predictor=pred)
        m.harvest = []
        for op in ops:
            U = torch.from_numpy(op[2])
            if op[0] == "1q":
                m.apply_single_qubit_gate(op[1], U)
            else:
                m.apply_two_site_gate(op[1], U)
        for f, k in m.harvest:
            F.append(f)
            K.append(k)
            C.append(cid)
        cid += 1
        if time.time() - last > 30.0:
            last = time.time()
            print(f"  harvest: {len(K)} pairs, {cid} circuits, {last - t0:.0f}s", flush=True)
    return np.array(F, np.float32), np.array(K, np.int64), np.array(C, np.int64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--max-seconds", type=float, default=600.0)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    ap.add_argument("--data", default=None, help="save the harvested pairs (.npz)")
    a = ap.parse_args()
    t0 = time.time()
    F, K, C = harvest(a.pairs, max_seconds=a.max_seconds)
    print(f"harvested {len(K)} pairs from {C.max() + 1} circuits in {time.time() - t0:.1f}s",
          flush=True)
    if a.data:
        np.savez_compressed(a.data, F=F, k=K, circuit=C)
    test = (C % 10) == 9
    rp = RankPredictor("cuda")
    t1 = time.time()
    log = rp.fit(F[~test], K[~test], steps=a.steps)
    print(f"trained {a.steps} steps in {time.time() - t1:.1f}s; loss log {log}", flush=True)
    print("train", rp.evaluate(F[~test], K[~test]), flush=True)
    print("held-out", rp.evaluate(F[test], K[test]), flush=True)
    rp.save(a.out)
    print(f"saved {a.out}", flush=True)


if __name__ == "__main__":
    main()
