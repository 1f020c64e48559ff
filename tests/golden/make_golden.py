"""Write golden vectors for the adaptive-MPS hot path from the reference itself.

Run in the development container only (the reference checkout is not present
where the GPU tests run):

    PYTHONPATH=/root/reference/src PYTHONDONTWRITEBYTECODE=1 \
        python tests/golden/make_golden.py

It imports ATLAS-Q's ``atlas_q.adaptive_mps.AdaptiveMPS`` on CPU and records
only what closed forms and invariants cannot pin:

* the per-gate (k, eps_local, entropy, condS, dtype) sequence of
  ``apply_two_site_gate`` on seeded circuits, and the resulting amplitudes;
* the random initial state for ``bond_dim > 1`` (torch RNG + QR);
* ``_batch_sweep_sample`` / single-shot outputs with the reference's draws
  replaced by the inverse-CDF rule ``outcome = (u >= p0)`` over recorded
  uniforms (``torch.multinomial`` / ``random.random`` are swapped in-process);
* the ``to_mixed_canonical(0)`` tensors the sampler consumed.

Outputs are small ``.npz`` files next to this script (data only).
"""
from __future__ import annotations

import importlib.util
import json
import os
import random
import sys
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
# this build's circuit generators, loaded by file path: the reference package
# is also called ``atlas_q`` and must own that name in this process
_spec = importlib.util.spec_from_file_location(
    "_own_circuits", HERE.parents[1] / "atlas-q_amd" / "atlas_q" / "circuits.py")
_own = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_own)
brickwork, hea = _own.brickwork, _own.hea

from atlas_q import adaptive_mps as ref_am  # noqa: E402  (the reference)
from atlas_q.adaptive_mps import AdaptiveMPS, DTypePolicy  # noqa: E402

torch.set_num_threads(1)
META = {
    "torch": torch.__version__,
    "TRITON_AVAILABLE": bool(getattr(ref_am, "TRITON_AVAILABLE", False)),
    "threads": torch.get_num_threads(),
    "reference_file": ref_am.__file__.split("/src/")[-1],
}


def stats_arrays(mps):
    lg = mps.statistics.logs
    return dict(
        bond=np.array(lg["bond"], np.int64),
        k=np.array(lg["k_star"], np.int64),
        chi_before=np.array(lg["chi_before"], np.int64),
        eps_local=np.array(lg["eps_local"], np.float64),
        entropy=np.array(lg["entropy"], np.float64),
        condS=np.array(lg["condS"], np.float64),
        promoted=np.array([d == "torch.complex128" for d in lg["dtype"]], bool),
        driver=np.array(lg["svd_driver"]),
    )


def run_brickwork(name, n, depth, seed, dtype, chi, eps, budget=None, n_amp=None, promote=None):
    torch.manual_seed(1234)
    policy = DTypePolicy(default=dtype) if promote is None else \
        DTypePolicy(default=dtype, promote_if_cond_gt=promote)
    mps = AdaptiveMPS(n, bond_dim=1, eps_bond=eps, chi_max_per_bond=chi,
                      budget_global_mb=budget, dtype_policy=policy, device="cpu")
    ops = brickwork(n, depth, seed)
    for b, U in ops:
        mps.apply_two_site_gate(b, torch.from_numpy(U).to(dtype))
    out = stats_arrays(mps)
    psi = mps.to_statevector().numpy().astype(np.complex128)
    if n_amp is not None and psi.size > n_amp:
        idx = np.random.default_rng(seed + 7).choice(psi.size, n_amp, replace=False)
        idx.sort()
        out.update(amp_index=idx, amp=psi[idx], norm2=np.sum(np.abs(psi) ** 2))
    else:
        out.update(psi=psi)
    out["bond_dims"] = np.array(mps.bond_dims, np.int64)
    out.update(n=n, depth=depth, seed=seed, chi=chi, eps=eps,
               dtype=str(dtype), budget=-1.0 if budget is None else budget,
               promote=-1.0 if promote is None else promote)
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(name, "updates", len(out["k"]), "max k", out["k"].max(), flush=True)
    return mps


def run_hea(name, n, layers, seed, dtype, chi, eps):
    torch.manual_seed(99)
    mps = AdaptiveMPS(n, bond_dim=1, eps_bond=eps, chi_max_per_bond=chi,
                      dtype_policy=DTypePolicy(default=dtype), device="cpu")
    for op in hea(n, layers, seed):
        if op[0] == "1q":
            mps.apply_single_qubit_gate(op[1], torch.from_numpy(op[2]).to(dtype))
        else:
            mps.apply_two_site_gate(op[1], torch.from_numpy(op[2]).to(dtype))
    out = stats_arrays(mps)
    out["psi"] = mps.to_statevector().numpy().astype(np.complex128)
    out.update(n=n, layers=layers, seed=seed, chi=chi, eps=eps, dtype=str(dtype))
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(name, "updates", len(out["k"]), flush=True)


def run_init(name, n, bond_dim, dtype, seed):
    torch.manual_seed(seed)
    mps = AdaptiveMPS(n, bond_dim=bond_dim, dtype_policy=DTypePolicy(default=dtype),
                      device="cpu")
    out = {f"t{i}": t.numpy().copy() for i, t in enumerate(mps.tensors)}
    out.update(bond_dims=np.array(mps.bond_dims, np.int64), n=n, bond_dim=bond_dim,
               seed=seed, dtype=str(dtype))
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(name, "bond_dims", mps.bond_dims, flush=True)


class _UniformDraws:
    """Stand-in for torch.multinomial / random.random over recorded uniforms."""

    def __init__(self, u):
        self.u = torch.from_numpy(u)
        self.site = 0

    def multinomial(self, probs, num_samples=1, replacement=False):
        col = self.u[:, self.site]
        self.site += 1
        return (col >= probs[:, 0].double()).to(torch.int64).unsqueeze(-1)


def run_sampler(name, n, depth, seed, dtype, chi, eps, shots):
    torch.manual_seed(4321)
    mps = AdaptiveMPS(n, bond_dim=1, eps_bond=eps, chi_max_per_bond=chi,
                      dtype_policy=DTypePolicy(default=dtype), device="cpu")
    for b, U in brickwork(n, depth, seed):
        mps.apply_two_site_gate(b, torch.from_numpy(U).to(dtype))
    mps.to_mixed_canonical(0)
    tensors = {f"t{i}": t.numpy().copy() for i, t in enumerate(mps.tensors)}
    u = np.random.default_rng(seed + 11).random((shots, n))
    draws = _UniformDraws(u)
    real_multinomial = torch.multinomial
    torch.multinomial = draws.multinomial
    try:
        samples = np.array(mps.sample(shots), np.int64)
    finally:
        torch.multinomial = real_multinomial
    # single-shot path (num_shots == 1): random.random() replaced
    u1 = np.random.default_rng(seed + 13).random(n)
    it = iter(u1.tolist())
    real_random = random.random
    random.random = lambda: next(it)
    try:
        single = np.array(mps.sample(1), np.int64)
    finally:
        random.random = real_random
    np.savez_compressed(HERE / f"{name}.npz", uniforms=u, samples=samples,
                        u_single=u1, single=single, n=n, dtype=str(dtype),
                        is_canonical=bool(mps.is_canonical), **tensors)
    print(name, "shots", shots, flush=True)


RUNS = {
    # precision promotion (adaptive_mps.py:268-274): with the default 1e6 a
    # complex64 run never promotes (the float32 rule keeps cond(S) < ~4e3)
    "brick_n8_promote": lambda: run_brickwork("brick_n8_promote", 8, 8, 16, torch.complex64,
                                              16, 1e-6, promote=10.0),
}

if __name__ == "__main__" and len(sys.argv) > 1:
    # only the named runs (the other fixtures are left as they are)
    for name in sys.argv[1:]:
        RUNS[name]()
elif __name__ == "__main__":
    run_brickwork("brick_n8_c128", 8, 8, 11, torch.complex128, 4, 1e-10)
    run_brickwork("brick_n10_c64", 10, 10, 12, torch.complex64, 8, 1e-6)
    run_brickwork("brick_n10_c64_eps", 10, 6, 13, torch.complex64, 32, 3e-2)
    run_brickwork("brick_n10_c128_eps", 10, 6, 14, torch.complex128, 32, 1e-3)
    run_brickwork("brick_n8_budget", 8, 8, 15, torch.complex64, 16, 1e-6, budget=0.0025)
    run_brickwork("config1_n16_c128", 16, 20, 2024, torch.complex128, 64, 1e-10, n_amp=2048)
    run_hea("hea_n8_c64", 8, 3, 21, torch.complex64, 8, 1e-6)
    run_init("init_n6_bd4_c64", 6, 4, torch.complex64, 5)
    run_init("init_n5_bd3_c128", 5, 3, torch.complex128, 6)
    run_sampler("sampler_n12_c64", 12, 8, 31, torch.complex64, 16, 1e-6, 512)
    run_sampler("sampler_n12_c128", 12, 8, 32, torch.complex128, 16, 1e-10, 256)
    for run in RUNS.values():
        run()
    (HERE / "meta.json").write_text(json.dumps(META, indent=1) + "\n")
    print(META)
