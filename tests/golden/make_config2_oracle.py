"""Write the oracle's fixture for the 64-qubit plain-GPU circuit.

    python tests/golden/make_config2_oracle.py [--threads N]

Workload: 64-qubit brickwork, Haar SU(4) gates (``circuits.brickwork``, seed
4002), chi_max = 512, cutoff 1e-8, AdaptiveMPS started at bond_dim = 1.  The
oracle (``oracle/ref_mps.OracleMPS``, the torch-CPU restatement of the
reference's AdaptiveMPS, pinned to the reference's own golden runs) runs it
with the complex128 policy (the FP64 oracle) and with the complex64 policy
(the reference's own complex64 path).

* Full depth 40 (1,260 updates): both runs' kept ranks and eps_local per
  update, and both runs' amplitudes at 1,000 uniformly drawn bitstrings
  (seed 4003).  Beyond the depth where the cap starts to bind, the two runs
  diverge: each cut through a flat spectrum turns rounding-level differences
  into a different kept subspace, and the overlap of the two final states is
  a few per cent.  The full-depth run is therefore compared by kept ranks
  and by per-update truncation fidelity, not by end-to-end amplitudes.
* Depth 8 (252 updates; bonds up to 256, below the cap): both runs are
  brought to mixed-canonical form at site 0 and the complex64 run -- the
  reference's own path in the policy the GPU run uses -- draws 10,000 shots
  from seeded uniforms with the reference's sweep (inverse CDF,
  ``batch_sweep_sample``, in chunks of 500 shots; the uniforms are
  ``default_rng(4004).random((10000, 64))`` and are regenerated by the
  test).  Stored: both runs' kept ranks, the sampled bitstrings, the
  smallest |u - p0| margin of each shot, and both runs' amplitudes at the
  sampled bitstrings and at the 1,000 uniform ones.  Even here the two
  policies differ by ~1e-3 per amplitude: under the float32-saturated rank
  rule (cutoff 1e-8 < 1e-4) the complex64 policy drops every tail below the
  float32 resolution of the total, the complex128 policy keeps it.

Bitstrings are stored as the two 32-bit halves of the 64-bit word (qubit 0 =
most significant bit).
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np
import torch

ROOT = Path(__file__).resolve().parents[2]
for p in (str(ROOT), str(ROOT / "atlas-q_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from atlas_q.circuits import brickwork  # noqa: E402
from oracle import ref_mps as R  # noqa: E402

N, DEPTH, EARLY_DEPTH, CHI, EPS = 64, 40, 8, 512, 1e-8
SEED_CIRCUIT, SEED_BITS, SEED_UNIFORMS = 4002, 4003, 4004
SHOTS, UNIFORM_BITSTRINGS, CHUNK = 10_000, 1_000, 500
OUT = Path(__file__).with_name("config2_n64_chi512.npz")


def bit_matrix(words: np.ndarray, n: int) -> np.ndarray:
    """[S, n] 0/1 matrix of 64-bit words (qubit 0 = most significant bit)."""
    w = words.astype(np.uint64)
    return ((w[:, None] >> np.arange(n - 1, -1, -1, dtype=np.uint64)) & np.uint64(1)).astype(
        np.int64)


def amplitudes(tensors, bits: np.ndarray) -> np.ndarray:
    """<b|psi> for every row of the [S, n] bit matrix: left-to-right products
    with both outcome halves formed by one matrix product per site."""
    S, n = bits.shape
    dt = tensors[0].dtype
    L = torch.ones((S, 1), dtype=dt)
    for i in range(n):
        T = tensors[i].to(dt)
        cl, _, cr = T.shape
        both = L @ T.reshape(cl, 2 * cr)                   # [S, 2 cr]
        sel = torch.from_numpy(bits[:, i]).to(torch.bool)[:, None]
        L = torch.where(sel, both[:, cr:], both[:, :cr])
    return L[:, 0].numpy().astype(np.complex128)


def run(dtype, depth):
    t0 = time.perf_counter()
    o = R.OracleMPS(N, bond_dim=1, eps_bond=EPS, chi_max_per_bond=CHI,
                    dtype_policy=R.DTypePolicy(default=dtype))
    ops = brickwork(N, depth, SEED_CIRCUIT)
    for j, (b, U) in enumerate(ops):
        o.apply_two_site_gate(b, torch.from_numpy(U))
        if j % 200 == 0:
            print(f"  {dtype}: update {j}/{len(ops)} ({time.perf_counter() - t0:.0f} s)",
                  flush=True)
    ks = np.array(o.statistics.logs["k_star"], np.int64)
    eps = np.array(o.statistics.logs["eps_local"], np.float64)
    o.to_mixed_canonical(0)
    print(f"  {dtype}: {len(ops)} updates in {time.perf_counter() - t0:.0f} s, "
          f"max bond {max(o.bond_dims)}", flush=True)
    return o, ks, eps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    if a.threads:
        torch.set_num_threads(a.threads)
    rng = np.random.default_rng(SEED_BITS)
    uni_words = rng.integers(0, np.iinfo(np.uint64).max, size=UNIFORM_BITSTRINGS,
                             dtype=np.uint64, endpoint=True)
    ub = bit_matrix(uni_words, N)
    out = dict(n=N, depth=DEPTH, early_depth=EARLY_DEPTH, chi=CHI, eps=EPS, seed=SEED_CIRCUIT,
               uniform_hi=(uni_words >> np.uint64(32)).astype(np.uint32),
               uniform_lo=(uni_words & np.uint64(0xFFFFFFFF)).astype(np.uint32),
               torch_version=torch.__version__, threads=torch.get_num_threads())
    # full depth: ranks and the end-to-end divergence of the two policies
    for tag, dt in (("c128", torch.complex128), ("c64", torch.complex64)):
        o, k, e = run(dt, DEPTH)
        out[f"k_{tag}"], out[f"eps_local_{tag}"] = k, e
        out[f"amp_uniform_{tag}"] = amplitudes(o.tensors, ub)
        del o
    # depth 12: amplitudes and shots
    hi, k128, _ = run(torch.complex128, EARLY_DEPTH)
    lo, k64, _ = run(torch.complex64, EARLY_DEPTH)
    out["early_k_c128"], out["early_k_c64"] = k128, k64
    u = np.random.default_rng(SEED_UNIFORMS).random((SHOTS, N))
    samples, margins = [], []
    t0 = time.perf_counter()
    for c in range(0, SHOTS, CHUNK):
        s, p0 = R.batch_sweep_sample(lo.tensors, u[c:c + CHUNK], return_p0=True)
        samples.append(np.array(s, np.int64))
        margins.append(np.abs(u[c:c + CHUNK] - p0).min(axis=1))
    samples = np.concatenate(samples)
    print(f"  {SHOTS} shots in {time.perf_counter() - t0:.0f} s", flush=True)
    sb = bit_matrix(samples.view(np.uint64), N)
    out.update(
        early_sample_hi=(samples.view(np.uint64) >> np.uint64(32)).astype(np.uint32),
        early_sample_lo=(samples.view(np.uint64) & np.uint64(0xFFFFFFFF)).astype(np.uint32),
        early_margin=np.concatenate(margins),
        early_amp_sample_c128=amplitudes(hi.tensors, sb),
        early_amp_sample_c64=amplitudes(lo.tensors, sb),
        early_amp_uniform_c128=amplitudes(hi.tensors, ub),
        early_amp_uniform_c64=amplitudes(lo.tensors, ub),
    )
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e6:.1f} MB)", flush=True)


if __name__ == "__main__":
    main()
