"""Write golden vectors for the reference's tools_qih two-site update and rank
predictors from the reference itself.

Run in the development container only (the reference checkout is not present
where the GPU tests run):

    PYTHONPATH=/root/reference/src PYTHONDONTWRITEBYTECODE=1 \\
        python tests/golden/make_golden_tn_core.py

It imports ATLAS-Q's ``atlas_q.tools_qih`` modules on CPU and records what the
closed forms leave open:

* ``tn_core.mps_apply_2q`` on seeded circuits -- bond dimension and truncation
  error after every gate and the final amplitudes -- with the fixed cap, the
  adaptive rule, each AI policy (a deterministic stub predictor) and the
  reference's two predictor networks at seeded random initialisation.  On a
  CPU tensor torch rejects the ``driver`` argument and the reference's SVD
  cascade would end in its jittered CPU fallback; the module's ``svd`` is
  swapped in-process for one that drops ``driver``, so the first attempt
  succeeds as it does on the reference's GPU target;
* ``build_entangler`` gates;
* both predictors' input features and outputs on a set of spectra.

Outputs: ``tn_core_golden.npz`` next to this script (data only).
"""
from __future__ import annotations

import importlib.util
import json
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
_spec = importlib.util.spec_from_file_location(
    "_own_circuits", HERE.parents[1] / "atlas-q_amd" / "atlas_q" / "circuits.py")
_own = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_own)

from atlas_q.tools_qih import tn_core  # noqa: E402  (the reference)
from atlas_q.tools_qih.ai_rank_predictor import RankPredictor  # noqa: E402
from atlas_q.tools_qih.finetuned_predictor import FinetunedMLP, FinetunedPredictorWrapper  # noqa: E402

torch.set_num_threads(1)
_svd = torch.linalg.svd
tn_core.svd = lambda X, full_matrices=False, driver=None: _svd(X, full_matrices=full_matrices)


class FracStub:
    """Deterministic stand-in predictor: a fixed fraction of the spectrum."""

    def __init__(self, frac):
        self.frac = frac

    def predict(self, S):
        return max(1, int(self.frac * S.numel()))


def contract(cores):
    psi = cores[0].to(torch.complex128).reshape(2, -1)
    for c in cores[1:]:
        c = c.to(torch.complex128)
        psi = (psi @ c.reshape(c.shape[0], -1)).reshape(-1, c.shape[2])
    return psi.reshape(-1).numpy()


def brick_gates(n, depth, seed, kind, dtype):
    rng = np.random.default_rng(seed)
    out = []
    for layer in range(depth):
        for b in range(layer % 2, n - 1, 2):
            if kind == "su4":
                U = torch.from_numpy(_own.haar_su4(rng)).to(dtype)
            else:
                U = tn_core.build_entangler(kind, float(rng.uniform(0.1, 1.2)), "cpu", dtype)
            out.append((b, U))
    return out


def run(name, n, depth, seed, kind, dtype, chi_max, adaptive, tol, eps=1e-10, pred=None,
        policy="guardrail", alpha=0.7, out=None):
    import os
    os.environ["QIH_AI_POLICY"] = policy
    os.environ["QIH_AI_ALPHA"] = str(alpha)
    cores = tn_core.mps_init_plus(n, device="cpu", dtype=dtype)
    gates = brick_gates(n, depth, seed, kind, dtype)
    chis, errs, bonds = [], [], []
    for b, U in gates:
        cores, chi, err = tn_core.mps_apply_2q(cores, b, b + 1, U, chi_max=chi_max, eps=eps,
                                               adaptive=adaptive, tol=tol, ai_predictor=pred)
        chis.append(chi)
        errs.append(err)
        bonds.append(b)
    g = np.stack([U.numpy().astype(np.complex128) for _, U in gates])
    out[f"{name}__gates"] = g
    out[f"{name}__bonds"] = np.array(bonds, np.int64)
    out[f"{name}__chi"] = np.array(chis, np.int64)
    out[f"{name}__err"] = np.array(errs, np.float64)
    out[f"{name}__psi"] = contract(cores)
    out[f"{name}__meta"] = np.array(json.dumps(dict(
        n=n, depth=depth, seed=seed, kind=kind, dtype=str(dtype), chi_max=chi_max,
        adaptive=adaptive, tol=tol, eps=eps, policy=policy, alpha=alpha,
        predictor=None if pred is None else type(pred).__name__)))
    return cores


def main():
    out = {}
    c64, c128 = torch.complex64, torch.complex128
    run("cz_adapt_c128", 8, 6, 11, "cz", c128, None, True, 1e-3, out=out)
    run("su4_fixed_c64", 10, 6, 12, "su4", c64, 8, False, 1e-4, out=out)
    run("su4_adapt_c64", 10, 6, 13, "su4", c64, 16, True, 1e-4, out=out)
    run("rzz_eps_c64", 8, 5, 14, "rzz", c64, None, False, 1e-4, eps=1e-3, out=out)
    for pol, alpha in (("baseline", 0.7), ("guardrail", 0.7), ("blend", 0.5), ("clamp", 0.7),
                       ("clamp+blend", 0.3), ("unknown", 0.7)):
        run(f"policy_{pol.replace('+', '_')}", 10, 5, 15, "su4", c64, 12, True, 1e-3,
            pred=FracStub(0.6), policy=pol, alpha=alpha, out=out)
    torch.manual_seed(5)
    rp = RankPredictor().eval()
    run("rankpred_guardrail", 10, 5, 16, "su4", c64, 16, True, 1e-3, pred=rp, out=out)
    torch.manual_seed(6)
    ft = FinetunedMLP().eval()
    ftw = object.__new__(FinetunedPredictorWrapper)
    ftw.device, ftw.model = "cpu", ft
    run("ftpred_blend", 10, 5, 17, "su4", c64, 16, True, 1e-3, pred=ftw, policy="blend",
        alpha=0.6, out=out)

    for th, kind in ((0.3, "cz"), (0.7, "cnot"), (0.45, "rzz")):
        out[f"entangler_{kind}"] = tn_core.build_entangler(kind, th, "cpu", c128).numpy()
        out[f"entangler_{kind}_theta"] = np.float64(th)

    # predictor features and outputs on spectra of several lengths and decays
    rng = np.random.default_rng(21)
    specs = []
    for m, tau in ((5, 2.0), (20, 5.0), (64, 30.0), (100, 8.0), (128, 60.0), (300, 40.0),
                   (40, 1e9)):
        s = np.sort(np.exp(-np.arange(m) / tau) * rng.uniform(0.8, 1.2, m))[::-1].copy()
        if m == 40:
            s[30:] = 0.0                      # zero tail: non-zero count below the length
        specs.append(s)
    feats_ft, feats_lp, frac_rp, rank_rp, frac_ft, rank_ft = [], [], [], [], [], []
    cap = {}
    rp.net[0].register_forward_pre_hook(lambda mod, inp: cap.__setitem__("x", inp[0].clone()))
    for dt in (torch.float32, torch.float64):
        for s in specs:
            S = torch.tensor(s, dtype=dt)
            feats_ft.append(FinetunedPredictorWrapper._features(ftw, S).numpy()[0])
            rank_rp.append(rp.predict(S))
            feats_lp.append(cap["x"].numpy()[0])
            actual = int((S > 1e-10).sum().item())
            frac_rp.append(float(rp.forward(S, actual).item()) if actual else np.nan)
            frac_ft.append(float(ft(FinetunedPredictorWrapper._features(ftw, S)).item()))
            rank_ft.append(ftw.predict(S))
    L = max(len(s) for s in specs)
    spec_arr = np.zeros((len(specs), L))
    for k, s in enumerate(specs):
        spec_arr[k, :len(s)] = s
    out["spec_values"] = spec_arr
    out["spec_len"] = np.array([len(s) for s in specs], np.int64)
    out["feat_ft135"] = np.stack(feats_ft).astype(np.float32)
    out["feat_log1p64"] = np.stack(feats_lp).astype(np.float32)
    out["frac_rankpred"] = np.array(frac_rp)
    out["rank_rankpred"] = np.array(rank_rp, np.int64)
    out["frac_ftpred"] = np.array(frac_ft)
    out["rank_ftpred"] = np.array(rank_ft, np.int64)
    out["predictor_seeds"] = np.array([5, 6], np.int64)
    out["torch_version"] = np.array(torch.__version__)
    np.savez_compressed(HERE / "tn_core_golden.npz", **out)
    print("wrote", HERE / "tn_core_golden.npz", len(out), "arrays")


if __name__ == "__main__":
    main()
