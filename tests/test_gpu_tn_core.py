"""GPU parity of the tools_qih path (tn_core.mps_apply_2q, truncation kernel,
sqrt(S) split, predictor features and dense-stack forward) against the
oracle (oracle/tn_core_ref.py) and the reference's golden vectors."""
import json
import math
import os

import numpy as np
import pytest
import torch
import torch.nn as nn

from oracle import tn_core_ref as T

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
GOLD = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "tn_core_golden.npz")


def _runs():
    return sorted(k[:-len("__meta")] for k in GOLD.files if k.endswith("__meta"))


def _rank_net():
    h = 512
    return nn.Sequential(nn.Linear(64, h), nn.LayerNorm(h), nn.ReLU(), nn.Dropout(0.2),
                         nn.Linear(h, h), nn.LayerNorm(h), nn.ReLU(), nn.Dropout(0.2),
                         nn.Linear(h, h // 2), nn.LayerNorm(h // 2), nn.ReLU(), nn.Dropout(0.1),
                         nn.Linear(h // 2, h // 4), nn.ReLU(), nn.Linear(h // 4, 1),
                         nn.Sigmoid()).eval()


def _ft_net():
    return nn.Sequential(nn.Linear(135, 256), nn.ReLU(), nn.Linear(256, 256), nn.ReLU(),
                         nn.Linear(256, 1), nn.Sigmoid()).eval()


def _product_predictor(meta):
    from atlas_q.tools_qih.ai_rank_predictor import RankPredictorWrapper
    from atlas_q.tools_qih.finetuned_predictor import FinetunedPredictorWrapper
    if meta["predictor"] == "FracStub":
        class Stub:
            def predict(self, S):
                return max(1, int(0.6 * S.numel()))
        return Stub()
    if meta["predictor"] == "RankPredictor":
        torch.manual_seed(5)
        w = RankPredictorWrapper(device=DEV)       # seeded init = the reference's weights
        return w
    if meta["predictor"] == "FinetunedPredictorWrapper":
        torch.manual_seed(6)
        net = _ft_net()
        w = object.__new__(FinetunedPredictorWrapper)
        from atlas_q.tools_qih.finetuned_predictor import FinetunedMLP
        w.device = DEV
        w.model = FinetunedMLP(device=DEV)
        w.model.load_state_dict({f"net.{k}": v for k, v in net.state_dict().items()})
        return w
    return None


@pytest.mark.parametrize("run", _runs())
def test_mps_apply_2q_matches_reference_golden(run, monkeypatch):
    from atlas_q.tools_qih import tn_core
    meta = json.loads(str(GOLD[f"{run}__meta"]))
    print(f"[tn_core golden] {run}", flush=True)
    monkeypatch.setenv("QIH_AI_POLICY", meta["policy"])
    monkeypatch.setenv("QIH_AI_ALPHA", str(meta["alpha"]))
    dt = torch.complex128 if meta["dtype"].endswith("complex128") else torch.complex64
    cores = tn_core.mps_init_plus(meta["n"], device=DEV, dtype=dt)
    pred = _product_predictor(meta)
    chis, errs = [], []
    for b, U in zip(GOLD[f"{run}__bonds"], GOLD[f"{run}__gates"]):
        cores, chi, err = tn_core.mps_apply_2q(cores, int(b), int(b) + 1,
                                               torch.from_numpy(U).to(dt).to(DEV),
                                               chi_max=meta["chi_max"], eps=meta["eps"],
                                               adaptive=meta["adaptive"], tol=meta["tol"],
                                               ai_predictor=pred)
        chis.append(chi)
        errs.append(err)
    assert chis == GOLD[f"{run}__chi"].tolist()
    tol = 1e-10 if dt == torch.complex128 else 2e-5
    assert np.abs(np.array(errs) - GOLD[f"{run}__err"]).max() < tol
    psi = T.contract([c.cpu() for c in cores])
    ref = GOLD[f"{run}__psi"]
    atol = 1e-10 if dt == torch.complex128 else 1e-5
    assert np.abs(psi - ref).max() < atol * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_truncate_kernel_matches_oracle(dtype):
    from atlas_q import kernels as K
    rng = np.random.default_rng(7)
    cases = 0
    for trial in range(300):
        m = int(rng.integers(1, 700))
        tau = float(10 ** rng.uniform(-0.5, 2.5))
        s = np.sort(np.exp(-np.arange(m) / tau) * rng.uniform(0.5, 1.5, m))[::-1].copy()
        if trial % 7 == 0:
            s[m // 2:] = 0.0
        tol = float(10 ** rng.uniform(-5, -0.5))
        adaptive = bool(trial % 3)
        chi_max = None if trial % 4 == 0 else int(rng.integers(1, m + 2))
        eps = float(10 ** rng.uniform(-12, -1))
        S = torch.from_numpy(s).to(dtype)
        # skip spectra with a prefix sum within a few ulps of the threshold
        # (torch's CPU sum and the kernel's double-accumulated total may
        # differ there in the last bit)
        c = np.cumsum((S.numpy().astype(np.float64)) ** 2)
        thr = (1 - tol) * c[-1]
        if np.min(np.abs(c - thr)) < 1e-5 * c[-1]:
            continue
        want = T.adaptive_keep(S, tol, chi_max) if adaptive else (
            chi_max if chi_max is not None and m > chi_max else m)
        for ai in (-1, int(rng.integers(0, m + 3))):
            for pol, alpha in (("guardrail", 0.7), ("blend", float(rng.random())),
                               ("baseline", 0.7)):
                r = K.tn_truncate(S.to(DEV), tol, adaptive, chi_max, ai, pol, alpha, eps)
                if ai < 0:
                    keep = want
                else:
                    ad = T.adaptive_keep(S, tol, chi_max) if adaptive else m
                    keep = T.policy_keep(ai, ad, pol, alpha)
                keep = min(keep, m)
                assert r.keep_count == keep, (trial, ai, pol)
                St = S[:keep]
                assert r.chi == int((St > eps).sum())
                total = float((S.double() ** 2).sum())
                kept = float((St.double() ** 2).sum())
                if total > 0:
                    assert abs(r.truncation_error - (1 - kept / total)) < (
                        2e-6 if dtype == torch.float32 else 1e-12)
                cases += 1
    assert cases > 500


def test_rebuild_sqrt_matches_torch():
    from atlas_q import kernels as K
    torch.manual_seed(1)
    for dt in (torch.complex64, torch.complex128):
        X = torch.randn(24, 18, dtype=dt, device=DEV)
        U, S, Vh = torch.linalg.svd(X, full_matrices=False)
        k = 7
        A, B = K.rebuild_sqrt(U, S, Vh, k, 12, 9)
        r = torch.sqrt(S[:k])
        tol = 1e-6 if dt == torch.complex64 else 1e-13
        assert torch.allclose(A, (U[:, :k] * r).reshape(12, 2, k), atol=tol)
        assert torch.allclose(B, (r[:, None] * Vh[:k].resolve_conj()).reshape(k, 2, 9), atol=tol)


def _spectra():
    vals, lens = GOLD["spec_values"], GOLD["spec_len"]
    return [v[:L].astype(dt) for dt in (np.float32, np.float64) for v, L in zip(vals, lens)]


def test_spectrum_features_match_reference_golden():
    from atlas_q import kernels as K
    for k, s in enumerate(_spectra()):
        S = torch.from_numpy(s).to(DEV)
        f, nsv = K.spectrum_features(S, "ft135")
        assert int(nsv.item()) == T.n_sv(s)
        assert np.allclose(f.cpu().numpy()[0], GOLD["feat_ft135"][k], rtol=1e-5, atol=2e-6), k
        f, _ = K.spectrum_features(S, "log1p64")
        assert np.allclose(f.cpu().numpy()[0], GOLD["feat_log1p64"][k], rtol=1e-6,
                           atol=1e-7), k


def test_dense_forward_matches_oracle():
    from atlas_q.tools_qih.ai_rank_predictor import DenseStack
    torch.manual_seed(9)
    for net, d in ((_rank_net(), 64), (_ft_net(), 135),
                   (nn.Sequential(nn.Linear(32, 128), nn.GELU(), nn.Linear(128, 3)), 32)):
        with torch.no_grad():
            for mod in net:
                if isinstance(mod, nn.LayerNorm):
                    mod.weight.normal_(1.0, 0.2)
                    mod.bias.normal_(0.0, 0.2)
        x = torch.randn(5, d)
        got = DenseStack(net, DEV)(x.to(DEV)).cpu().numpy()
        want = T.dense_forward(T.layers_from_sequential(net), x.numpy())
        assert np.abs(got - want).max() < 2e-5


def test_reference_predictors_match_golden_and_roundtrip(tmp_path):
    from atlas_q.tools_qih.ai_rank_predictor import RankPredictorWrapper
    from atlas_q.tools_qih.finetuned_predictor import FinetunedPredictorWrapper
    torch.manual_seed(5)
    rp = RankPredictorWrapper(device=DEV)
    torch.manual_seed(6)
    net = _ft_net()
    path = tmp_path / "ft.pt"
    torch.save({f"net.{k}": v for k, v in net.state_dict().items()}, path)
    ft = FinetunedPredictorWrapper(str(path), device=DEV)
    for k, s in enumerate(_spectra()):
        S = torch.from_numpy(s).to(DEV)
        assert abs(rp.predict_fraction(S) - GOLD["frac_rankpred"][k]) < 2e-5
        assert abs(ft.predict_fraction(S) - GOLD["frac_ftpred"][k]) < 2e-5
        assert rp.predict(S) == GOLD["rank_rankpred"][k]
        assert ft.predict(S) == GOLD["rank_ftpred"][k]
    # save / load keeps the reference's state-dict keys
    p2 = tmp_path / "rp.pt"
    rp.save(p2)
    keys = set(torch.load(p2, weights_only=True))
    assert "net.0.weight" in keys and "net.14.bias" in keys
    rp2 = RankPredictorWrapper(str(p2), device=DEV)
    S = torch.from_numpy(_spectra()[3]).to(DEV)
    assert rp2.predict_fraction(S) == rp.predict_fraction(S)
    with pytest.raises(ValueError):
        FinetunedPredictorWrapper(str(tmp_path / "none_compiled.pt"), device=DEV)


def test_svd_events_match_oracle(tmp_path, monkeypatch):
    """SVD events from the device path carry the oracle's spectrum and cap."""
    import importlib
    monkeypatch.setenv("QIH_SVD_LOG_DIR", str(tmp_path))
    from atlas_q.tools_qih import ft_dataset, tn_core
    importlib.reload(tn_core)
    try:
        run = "su4_adapt_c64"
        meta = json.loads(str(GOLD[f"{run}__meta"]))
        dt = torch.complex64
        cores = tn_core.mps_init_plus(meta["n"], device=DEV, dtype=dt)
        ocores = T.mps_init_plus(meta["n"], dtype=dt)
        events = []
        for b, U in zip(GOLD[f"{run}__bonds"], GOLD[f"{run}__gates"]):
            Ut = torch.from_numpy(U).to(dt)
            cores, _, _ = tn_core.mps_apply_2q(cores, int(b), int(b) + 1, Ut.to(DEV),
                                               chi_max=16, adaptive=True, tol=1e-4)
            ocores, _, _ = T.mps_apply_2q(ocores, int(b), int(b) + 1, Ut, chi_max=16,
                                          adaptive=True, tol=1e-4, events=events)
        tn_core.SVD_EVENT_LOGGER.close()
        table = ft_dataset.read_events([tn_core.SVD_EVENT_LOGGER.path])
        assert table.unreadable == 0 and len(table) == len(events)
        for ev, oe in zip(table.events, events):
            assert ev["adaptive_keep"] == oe["adaptive_keep"]
            assert ev["pre_rank"] == oe["pre_rank"]
            assert np.allclose(ev["S_top128"], oe["S_top128"], rtol=1e-5, atol=1e-7)
            # float32 state after 27 gates: relative drift of a few 1e-5
            assert abs(ev["total_weight"] - oe["total_weight"]) < 5e-5 * oe["total_weight"]
    finally:
        monkeypatch.delenv("QIH_SVD_LOG_DIR")
        importlib.reload(tn_core)


def test_spec32_device_features_match_oracle():
    from atlas_q import kernels as K
    rng = np.random.default_rng(10)
    for dt in (torch.float32, torch.float64):
        for m in (1, 3, 24, 100, 128, 500):
            s = np.sort(rng.random(m) ** 4)[::-1].copy()
            S = torch.from_numpy(s).to(dt)
            f, _ = K.spectrum_features(S.to(DEV), "spec32")
            tot = float((S.double() ** 2).sum())
            want = T.features_spec32(S.double().numpy()[:128], tot)
            assert np.allclose(f.cpu().numpy()[0], want, rtol=1e-5, atol=2e-5), (dt, m)


class _ListLogger:
    active = True

    def __init__(self):
        self.events = []

    def log(self, **kw):
        self.events.append(kw)


def test_spectral_predictor_trains_on_events_and_plugs_into_policies(monkeypatch):
    """Harvest SVD events from the device path at one truncation setting
    (tol 1e-3, chi_max 64: predict(S) sees neither, so a predictor serves one
    setting, as the reference's do), train the library's spectral predictor
    on them with the device kernels, and compare held-out rank errors with
    the method without the network (the training set's median kept fraction
    times the non-zero count)."""
    from atlas_q.circuits import haar_su4
    from atlas_q.tools_qih import tn_core
    from atlas_q.tools_qih.spectral_predictor import SpectralRankPredictor
    from atlas_q.workloads import weak_gate
    log = _ListLogger()
    monkeypatch.setattr(tn_core, "SVD_EVENT_LOGGER", log)
    rng = np.random.default_rng(11)
    circuit_of = []
    tol, chi_max = 1e-3, 64
    ncirc = 120
    for c in range(ncirc):
        n = int(rng.integers(10, 16))
        angle = None if c % 4 == 0 else float(rng.uniform(0.05, 0.9))
        cores = tn_core.mps_init_plus(n, device=DEV, dtype=torch.complex64)
        for layer in range(int(rng.integers(6, 16))):
            for b in range(layer % 2, n - 1, 2):
                U = haar_su4(rng) if angle is None else weak_gate(rng, angle)
                before = len(log.events)
                cores, _, _ = tn_core.mps_apply_2q(cores, b, b + 1,
                                                   torch.from_numpy(U).to(torch.complex64).to(DEV),
                                                   chi_max=chi_max, adaptive=True, tol=tol)
                circuit_of += [c] * (len(log.events) - before)
    # the seam consults a predictor only for spectra of more than 10 values
    keep = [int((np.array(e["S_top128"]) > 0).sum()) > 10 for e in log.events]
    ev = [e for e, k in zip(log.events, keep) if k]
    circ = np.array(circuit_of)[np.array(keep)]
    test_c = set(range(0, ncirc, 5))
    tr = [e for e, c in zip(ev, circ) if c not in test_c]
    te = [e for e, c in zip(ev, circ) if c in test_c]
    assert len(tr) > 1000 and len(te) > 200
    sp = SpectralRankPredictor(DEV, seed=1)
    sp.fit_events(tr, steps=3000)
    F, k = sp.event_arrays(te)
    y = sp.mlp.predict_log2(torch.from_numpy(F).to(DEV)).cpu().numpy()
    kh = np.clip(np.round(2.0 ** np.clip(y, 0, 30)), 1, None)
    err_net = np.mean(np.abs(np.log2(kh) - np.log2(k)))
    nz_tr = np.array([max(1, int((np.array(e["S_top128"]) > 1e-10).sum())) for e in tr])
    ktr = np.array([e["adaptive_keep"] for e in tr])
    frac = float(np.median(ktr / nz_tr))
    nz_te = np.array([max(1, int((np.array(e["S_top128"]) > 1e-10).sum())) for e in te])
    kb = np.clip(np.round(frac * nz_te), 1, None)
    err_base = np.mean(np.abs(np.log2(kb) - np.log2(k)))
    print(f"[spectral predictor] held-out mean |log2 k error|: network {err_net:.3f}, "
          f"median-fraction baseline {err_base:.3f}, exact {np.mean(kh == k):.2f}", flush=True)
    assert err_net < 0.5 * err_base
    assert err_net < 0.15
    # the trained predictor in the reference's seam, guardrail policy
    monkeypatch.setenv("QIH_AI_POLICY", "guardrail")
    cores = tn_core.mps_init_plus(10, device=DEV, dtype=torch.complex64)
    for b in range(0, 9, 2):
        U = torch.from_numpy(haar_su4(rng)).to(torch.complex64).to(DEV)
        cores, chi, err = tn_core.mps_apply_2q(cores, b, b + 1, U, chi_max=32, adaptive=True,
                                               tol=1e-3, ai_predictor=sp)
        assert 1 <= chi <= 4 and 0.0 <= err < 1.0
