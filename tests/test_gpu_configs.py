"""GPU runs of the three whole-circuit workloads the path is specified on.

* 16 qubits, depth 20, Haar SU(4) brickwork, complex128, chi_max 64,
  cutoff 1e-10, then mixed-canonical at site 0 and 1,000 shots: the shots
  against the oracle's sweep on the same uniforms (bit-exact where every draw
  clears p0 by 1e-10), the per-draw p0 against the Born-rule conditionals of
  the MPS's own state vector (1e-9), and each qubit's frequency against its
  exact marginal in the untruncated circuit state (5 sigma).
* 64 qubits, depth 40, Haar SU(4), complex64 on the plain GPU path, chi_max
  512, cutoff 1e-8, 10,000 shots, against the oracle's complex128 and
  complex64 runs of the same circuit (fixture ``config2_n64_chi512.npz``,
  written by ``tests/golden/make_config2_oracle.py``): kept ranks and the
  truncation fidelity of sampled full-size updates at depth 40; amplitudes
  (per-amplitude relative error with an RMS magnitude floor) and the shots
  at depth 8 against the oracle's complex64 run.
* bond 2048 (theta 4096 x 4096, the learned-rank form's cap): a brickwork
  sub-layer through the grouped fast path as benchmarked -- ranks at the
  cap, truncation fidelity against the FP64 truncation within 1e-6, kept
  weight and eps_local against the FP64 spectrum, right-orthonormal right
  cores, |F_fast - F_plain| <= 1e-6.
"""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from conftest import GOLDEN, load_golden
from test_gpu_mps import _SpectrumLog
from oracle import ref_mps as R
from oracle import statevector as SV
from atlas_q import kernels as K
from atlas_q.adaptive_mps import AdaptiveMPS, DTypePolicy, _gate4
from atlas_q.circuits import brickwork

sys.path.insert(0, str(GOLDEN))
import make_config2_oracle as C2  # noqa: E402  (fixture layout and amplitude checker)

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _layers(ops):
    """Brickwork ops grouped into sub-layers of disjoint bonds, in order."""
    out, cur, used = [], [], set()
    for b, U in ops:
        if {b, b + 1} & used:
            out.append(cur)
            cur, used = [], set()
        cur.append((b, torch.from_numpy(U)))
        used |= {b, b + 1}
    if cur:
        out.append(cur)
    return out


def _rel_amp_error(got, ref):
    ref = np.asarray(ref, np.complex128)
    floor = float(np.sqrt(np.mean(np.abs(ref) ** 2)))
    return float((np.abs(np.asarray(got) - ref) / np.maximum(np.abs(ref), floor)).max()), floor


# ---------------------------------------------------------------- config 1
def test_config1_shots_against_oracle_and_born_rule():
    g = load_golden("config1_n16_c128")
    n = int(g["n"])
    ops = brickwork(n, int(g["depth"]), int(g["seed"]))
    m = AdaptiveMPS(n, bond_dim=1, eps_bond=float(g["eps"]), chi_max_per_bond=int(g["chi"]),
                    dtype_policy=DTypePolicy(default=torch.complex128), device=DEV)
    for b, U in ops:
        m.apply_two_site_gate(b, torch.from_numpy(U))
    assert m.statistics.logs["k_star"] == g["k"].tolist()
    m.to_mixed_canonical(0)
    u = np.random.default_rng(1001).random((1000, n))
    bits, p0 = K.sample_bits(m.tensors, torch.from_numpy(u).to(DEV), want_p0=True)
    got = np.array(K.bits_to_ints(bits))
    ref, p0_ref = R.batch_sweep_sample([t.cpu() for t in m.tensors], u, return_p0=True)
    clear = (np.abs(u - p0_ref) > 1e-10).all(axis=1)
    print(f"  config 1: {int(clear.sum())}/1000 shots clear the 1e-10 margin", flush=True)
    assert clear.mean() > 0.99
    assert np.array_equal(got[clear], np.array(ref)[clear])
    p0 = p0.cpu().numpy()
    np.testing.assert_allclose(p0, p0_ref, rtol=0, atol=1e-12)
    # Born rule of the truncated state itself, per draw
    psi = m.to_statevector().cpu().numpy().astype(np.complex128)
    psi = psi / np.linalg.norm(psi)
    worst = 0.0
    for s in range(32):
        bs = [(int(got[s]) >> (n - 1 - i)) & 1 for i in range(n)]
        for i in range(n):
            worst = max(worst, abs(p0[s, i] - SV.marginal_p0(psi, n, bs[:i])))
    print(f"  config 1: max |p0 - Born conditional| {worst:.1e} (32 shots x 16 draws)", flush=True)
    assert worst < 1e-9
    # frequencies against the exact circuit state (the truncation keeps 1 - F ~ 1e-3)
    exact = SV.run_circuit(n, [("2q", b, U) for b, U in ops])
    probs = np.abs(exact) ** 2
    idx = np.arange(2 ** n)
    for q in range(n):
        p1 = float(probs[((idx >> (n - 1 - q)) & 1) == 1].sum())
        f1 = float(((got.astype(np.int64) >> (n - 1 - q)) & 1).mean())
        assert abs(f1 - p1) < 5 * np.sqrt(p1 * (1 - p1) / 1000) + 5e-3, (q, f1, p1)


# ---------------------------------------------------------------- config 2
def _words(hi, lo):
    return (hi.astype(np.uint64) << np.uint64(32)) | lo.astype(np.uint64)


def _amplitudes_gpu(tensors, words, n):
    """<b|psi> at the given 64-bit words from the device tensors, in complex128."""
    bm = torch.from_numpy(C2.bit_matrix(words, n)).to(DEV)
    L = torch.ones((len(words), 1), dtype=torch.complex128, device=DEV)
    for i, T in enumerate(tensors):
        T = T.to(torch.complex128)
        cl, _, cr = T.shape
        both = L @ T.reshape(cl, 2 * cr)
        L = torch.where(bm[:, i:i + 1].bool(), both[:, cr:], both[:, :cr])
    return L[:, 0].cpu().numpy()


def _overlap(a, b):
    return abs(np.vdot(a, b)) / np.sqrt(np.vdot(a, a).real * np.vdot(b, b).real)


def _run_config2(f, depth, capture=(), workers=4):
    n, chi, eps = int(f["n"]), int(f["chi"]), float(f["eps"])
    ops = brickwork(n, depth, int(f["seed"]))
    m = AdaptiveMPS(n, bond_dim=1, eps_bond=eps, chi_max_per_bond=chi,
                    dtype_policy=DTypePolicy(default=torch.complex64), device=DEV)
    captured = {}
    j = 0
    for layer in _layers(ops):
        ins = {}
        for q, (b, U) in enumerate(layer):
            if j + q in capture:
                ins[j + q] = (b, U, m.tensors[b].clone(), m.tensors[b + 1].clone())
        m.apply_two_site_gates(layer, workers=workers)
        for idx, (b, U, A, B) in ins.items():
            captured[idx] = (b, U, A, B, m.tensors[b].clone(), m.tensors[b + 1].clone())
        j += len(layer)
    torch.cuda.synchronize()
    return m, captured


@pytest.mark.timeout(900)
def test_config2_plain_gpu_circuit_against_fp64_oracle(monkeypatch):
    """Full depth (1,260 updates): kept ranks against the oracle's runs and
    the truncation fidelity of sampled updates at the cap (theta 1024 x
    1024) against the FP64 truncation of the same input (1 - F <= 1e-6, the
    complex64 path's budget); end-to-end amplitudes are reported: under the
    cap the truncated dynamics of any two arithmetics diverge (the oracle's
    own complex64 and complex128 runs end with an overlap of a few per cent).
    Depth 8 (bonds up to 256): per-amplitude relative error against the
    oracle's complex64 run (the reference's path in the same policy) at its
    sampled bitstrings and 1,000 uniform ones within 1e-5, widened only by
    the weight of explicitly excluded float-saturated cuts, and the 10,000
    shots bit-exact where every draw clears the oracle's p0 by more than the
    state's probability error; the distance to the complex128 run is
    reported (the complex64 policy drops tails the complex128 one keeps)."""
    import time
    f = load_golden("config2_n64_chi512")
    n = int(f["n"])
    picks = (500, 800, 1000, 1250)
    t0 = time.perf_counter()
    m, cap = _run_config2(f, int(f["depth"]), picks)
    print(f"  config 2: {len(f['k_c64'])} plain updates in {time.perf_counter() - t0:.1f} s, "
          f"max bond {max(m.bond_dims)}", flush=True)
    ks = np.array(m.statistics.logs["k_star"])
    d64 = np.nonzero(ks != f["k_c64"])[0]
    print(f"  config 2: kept ranks differ from the oracle's complex64 run at {len(d64)} and from "
          f"its complex128 run at {int((ks != f['k_c128']).sum())} of {len(ks)} updates "
          f"(oracle complex64 vs complex128: {int((f['k_c64'] != f['k_c128']).sum())}); "
          f"first at {d64[:5].tolist()}", flush=True)
    # per-update truncation against the FP64 truncation of the same input
    for idx in picks:
        b, U, A, B, A2, B2 = cap[idx]
        k = int(ks[idx])
        X = R.theta(A.cpu().to(torch.complex128), B.cpu().to(torch.complex128),
                    U.to(torch.complex128))
        Uu, S, Vh = torch.linalg.svd(X, full_matrices=False)
        k_ref = R.choose_rank_from_sigma(S.to(torch.float32), float(f["eps"]), int(f["chi"]),
                                         lambda _k: True)[0]
        P64 = (Uu[:, :k] * S[:k]) @ Vh[:k]
        P = (A2.reshape(-1, k).cpu().to(torch.complex128)
             @ B2.reshape(k, -1).cpu().to(torch.complex128))
        F = float((torch.vdot(P64.reshape(-1), P.reshape(-1)).abs() ** 2
                   / (P64.norm() ** 2 * P.norm() ** 2)).item())
        print(f"  config 2 update {idx} (bond {b}, theta {tuple(X.shape)}): k {k} (FP64 "
              f"spectrum under the complex64 rule: {k_ref}), 1 - F = {1 - F:.2e}", flush=True)
        assert 1 - F < 1e-6, (idx, F)
        if k != k_ref:
            w = (S.double() ** 2).numpy()
            lo, hi = min(k, k_ref), max(k, k_ref)
            assert w[lo:hi].sum() <= 16 * 2.0 ** -24 * w.sum(), (idx, k, k_ref)
    uni = _words(f["uniform_hi"], f["uniform_lo"])
    amp = _amplitudes_gpu(m.tensors, uni, n)
    print(f"  config 2 (depth 40): overlap estimate over 1,000 uniform bitstrings: GPU vs FP64 "
          f"oracle {_overlap(amp, f['amp_uniform_c128']):.3f}, GPU vs oracle complex64 "
          f"{_overlap(amp, f['amp_uniform_c64']):.3f}, oracle complex64 vs FP64 "
          f"{_overlap(f['amp_uniform_c64'], f['amp_uniform_c128']):.3f}", flush=True)
    del m, cap
    # ---------------- depth 8: amplitudes and shots against the oracle
    depth = int(f["early_depth"])
    log = _SpectrumLog(monkeypatch)
    m, _ = _run_config2(f, depth, workers=1)     # spectra logged in update order
    ks = np.array(m.statistics.logs["k_star"])
    excluded = []
    for j in np.nonzero(ks != f["early_k_c64"])[0]:
        w = log.spectra[j] ** 2
        lo, hi = sorted((int(ks[j]), int(f["early_k_c64"][j])))
        between = float(w[lo:hi].sum() / w.sum())
        # a float-saturated cut: the weight between the two cuts is below the
        # float32 resolution of the total (the rule's threshold equals it)
        assert between <= 16 * 2.0 ** -24, (int(j), int(ks[j]), int(f["early_k_c64"][j]))
        excluded.append(between)
    print(f"  config 2 (depth {depth}): max bond {max(m.bond_dims)}, kept ranks differ from the "
          f"oracle's complex64 run at {len(excluded)} of {len(ks)} updates, all at "
          f"float-saturated cuts (excluded weight {sum(excluded):.1e})", flush=True)
    m.to_mixed_canonical(0)
    want = _words(f["early_sample_hi"], f["early_sample_lo"])
    amp = _amplitudes_gpu(m.tensors, np.concatenate([want, uni]), n)
    ref64 = np.concatenate([f["early_amp_sample_c64"], f["early_amp_uniform_c64"]])
    ref128 = np.concatenate([f["early_amp_sample_c128"], f["early_amp_uniform_c128"]])
    err64, floor = _rel_amp_error(amp, ref64)
    tol = 1e-5 + 4.0 * float(np.sqrt(sum(excluded))) * np.abs(ref64).max() / floor
    print(f"  config 2 (depth {depth}): per-amplitude relative error vs the oracle's complex64 "
          f"run {err64:.2e} (tol {tol:.1e}, RMS floor {floor:.2e}); vs its complex128 run "
          f"{_rel_amp_error(amp, ref128)[0]:.2e} (oracle complex64 vs complex128 "
          f"{_rel_amp_error(ref64, ref128)[0]:.2e})", flush=True)
    assert err64 <= tol, (err64, tol)
    u = np.random.default_rng(C2.SEED_UNIFORMS).random((C2.SHOTS, n))
    bits, _ = K.sample_bits(m.tensors, torch.from_numpy(u).to(DEV))
    got = bits[:, 0].cpu().numpy().view(np.uint64)
    # against the oracle's shots of its own state: bit-exact where every draw
    # clears the oracle's p0 by more than the two states' amplitude distance
    mtol = max(1e-5, 2 * err64)
    clear = f["early_margin"] > mtol
    same = got[clear] == want[clear]
    print(f"  config 2 (depth {depth}): {int(clear.sum())}/{C2.SHOTS} of the oracle's shots "
          f"clear a {mtol:.1e} margin, {int(same.sum())} of them bit-exact", flush=True)
    assert clear.sum() >= 1000
    assert same.all()
    # against the oracle's sweep of the device state itself (the sampler's
    # parity): bit-exact where every draw clears p0 by 1e-5
    cs = 256
    ref, p0 = R.batch_sweep_sample([t.cpu() for t in m.tensors], u[:cs], return_p0=True)
    ok = (np.abs(u[:cs] - p0) > 1e-5).all(axis=1)
    print(f"  config 2 (depth {depth}): sampler vs the oracle's sweep of the same state: "
          f"{int(ok.sum())}/{cs} shots clear 1e-5", flush=True)
    assert ok.mean() > 0.9
    assert np.array_equal(np.array(K.bits_to_ints(bits[:cs]))[ok], np.array(ref)[ok])


# ---------------------------------------------------------------- config 5
@pytest.mark.timeout(600)
def test_chi2048_sublayer_properties():
    from atlas_q.workloads import gate_schedule, slice_tensors
    CHI, sites = 2048, 6
    ts = slice_tensors(sites, CHI, "flat", seed=2051, device=DEV, draw_on_device=True)
    gates = [(b, torch.from_numpy(U)) for b, U in gate_schedule(sites, 3, seed=2052)]
    assert [b for b, _ in gates] == [0, 2, 4]

    def mk(method, prec):
        m = AdaptiveMPS(sites, bond_dim=1, eps_bond=1e-8, chi_max_per_bond=CHI,
                        dtype_policy=DTypePolicy(default=torch.complex64), device=DEV,
                        svd_method=method, theta_precision=prec)
        m.tensors = [t.clone() for t in ts]
        m.bond_dims = [CHI] * (sites - 1)
        return m

    fast = mk("learned", "bf16x3")
    fast.apply_two_site_gates(gates, workers=4)
    torch.cuda.synchronize()
    assert fast.statistics.logs["k_star"] == [CHI] * 3
    for q, (b, U) in enumerate(gates[:2]):
        X = K.theta(ts[b].to(torch.complex128), ts[b + 1].to(torch.complex128), _gate4(U))
        Uu, S, Vh = torch.linalg.svd(X, full_matrices=False)
        P64 = (Uu[:, :CHI] * S[:CHI]) @ Vh[:CHI]
        A, B = fast.tensors[b], fast.tensors[b + 1]
        P = A.reshape(-1, CHI).to(torch.complex128) @ B.reshape(CHI, -1).to(torch.complex128)
        F = float((torch.vdot(P64.reshape(-1), P.reshape(-1)).abs() ** 2
                   / (P64.norm() ** 2 * P.norm() ** 2)).item())
        print(f"  chi 2048 gate {q}: 1 - F = {1 - F:.2e}", flush=True)
        assert 1 - F < 1e-6, (b, F)
        total = float((S ** 2).sum())
        kept = float((S[:CHI] ** 2).sum())
        assert abs(float(P.norm() ** 2) / float(X.norm() ** 2) - kept / total) < 1e-6
        np.testing.assert_allclose(fast.statistics.logs["eps_local"][q],
                                   np.sqrt(total - kept), rtol=1e-3)
        Bm = B.reshape(CHI, -1).to(torch.complex128)
        gram = Bm @ Bm.conj().T
        assert float((gram - torch.eye(CHI, dtype=gram.dtype, device=DEV)).abs().max()) < 1e-4
        if q == 0:
            plain = mk("plain", "native")
            plain.apply_two_site_gate(b, U)
            Pp = (plain.tensors[b].reshape(-1, CHI).to(torch.complex128)
                  @ plain.tensors[b + 1].reshape(CHI, -1).to(torch.complex128))
            Fp = float((torch.vdot(P64.reshape(-1), Pp.reshape(-1)).abs() ** 2
                        / (P64.norm() ** 2 * Pp.norm() ** 2)).item())
            print(f"  chi 2048 plain: 1 - F = {1 - Fp:.2e}, |dF| = {abs(F - Fp):.2e}", flush=True)
            assert abs(F - Fp) <= 1e-6, (F, Fp)
