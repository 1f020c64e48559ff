"""CPU tests: the tools_qih oracle (oracle/tn_core_ref.py) pinned to closed
forms, hand-computed cases and the reference's own golden vectors; host
logic of the product's tools_qih modules (gate builders, SVD-event log,
dataset builder, policy adapter)."""
import gzip
import json
import math

import numpy as np
import pytest
import torch
import torch.nn as nn

from oracle import tn_core_ref as T

GOLD = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "tn_core_golden.npz")


def _runs():
    return sorted(k[:-len("__meta")] for k in GOLD.files if k.endswith("__meta"))


# ---------------------------------------------------------------- closed forms
def test_theta_transposed_convention_closed_form():
    rng = np.random.default_rng(0)
    li, m, rj = 3, 4, 5
    A = rng.normal(size=(li, 2, m)) + 1j * rng.normal(size=(li, 2, m))
    B = rng.normal(size=(m, 2, rj)) + 1j * rng.normal(size=(m, 2, rj))
    U = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    X = T.theta_t(torch.from_numpy(A), torch.from_numpy(B), torch.from_numpy(U)).numpy()
    ref = np.zeros((2 * li, 2 * rj), complex)
    for l in range(li):
        for p in range(2):
            for q in range(2):
                for r in range(rj):
                    acc = 0
                    for p1 in range(2):
                        for p2 in range(2):
                            acc += U[2 * p1 + p2, 2 * p + q] * sum(
                                A[l, p1, c] * B[c, p2, r] for c in range(m))
                    ref[2 * l + p, q * rj + r] = acc
    assert np.abs(X - ref).max() < 1e-12


def test_adaptive_keep_hand_computed():
    S = torch.tensor([1.0, 0.5, 0.1, 0.01], dtype=torch.float64)
    # S^2 = 1, .25, .01, .0001; total 1.2601; (1-1e-2) total = 1.247499
    # cumsum 1, 1.25, 1.26, 1.2601 -> one entry <= thresh -> keep 2
    assert T.adaptive_keep(S, 1e-2, None) == 2
    assert T.adaptive_keep(S, 1e-2, 1) == 1
    # tol 0.3: thresh 0.88207 -> no entry passes -> keep 1
    assert T.adaptive_keep(S, 0.3, None) == 1
    # tol 1e-6: thresh 1.2600987 -> 3 entries pass -> keep 4 (= len)
    assert T.adaptive_keep(S, 1e-6, None) == 4
    # tol 1e-3: thresh 1.2588399 -> 2 entries pass -> keep 3
    assert T.adaptive_keep(S, 1e-3, None) == 3


@pytest.mark.parametrize("ai,ad,policy,alpha,want", [
    (5, 3, "baseline", 0.7, 5), (5, 3, "clamp", 0.7, 5), (0, 3, "baseline", 0.7, 1),
    (5, 3, "guardrail", 0.7, 3), (2, 3, "guardrail", 0.7, 2), (2, 3, "mystery", 0.7, 2),
    (5, 3, "blend", 0.7, 3),          # round(4.4) = 4 -> min(4, 3)
    (2, 5, "blend", 0.5, 4),          # 3.5 -> 4 (half to even)
    (2, 3, "blend", 0.5, 2),          # 2.5 -> 2 (half to even)
    (1, 9, "clamp+blend", 0.25, 7),   # 0.25 + 6.75 = 7.0
])
def test_policy_arithmetic(ai, ad, policy, alpha, want):
    assert T.policy_keep(ai, ad, policy, alpha) == want


def test_svd_robust_reconstructs():
    rng = np.random.default_rng(1)
    X = torch.from_numpy(rng.normal(size=(8, 6)) + 1j * rng.normal(size=(8, 6)))
    U, S, Vh = T.svd_robust(X)
    assert torch.allclose((U * S) @ Vh, X, atol=1e-12)
    Z = torch.zeros((4, 4), dtype=torch.complex128)
    U, S, Vh = T.svd_robust(Z)
    assert float(S.abs().max()) == 0.0


def test_split_identity_and_truncation_error():
    """Cores after the split multiply back to the truncated theta, and the
    returned truncation error is 1 - kept/total."""
    rng = np.random.default_rng(2)
    A = torch.from_numpy(rng.normal(size=(4, 2, 6)) + 1j * rng.normal(size=(4, 2, 6)))
    B = torch.from_numpy(rng.normal(size=(6, 2, 3)) + 1j * rng.normal(size=(6, 2, 3)))
    U = torch.from_numpy(np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0])
    X = T.theta_t(A, B, U)
    cores, chi, err = T.mps_apply_2q([A.clone(), B.clone()], 0, 1, U, chi_max=3)
    assert chi == 3
    P = cores[0].reshape(8, chi) @ cores[1].reshape(chi, 6)
    u, s, vh = torch.linalg.svd(X, full_matrices=False)
    assert torch.allclose(P, (u[:, :3] * s[:3]) @ vh[:3], atol=1e-10)
    assert abs(err - (1 - float((s[:3] ** 2).sum() / (s ** 2).sum()))) < 1e-12


# ---------------------------------------------------------------- golden runs
def _oracle_predictor(meta):
    if meta["predictor"] == "FracStub":
        frac = 0.6

        class Stub:
            def predict(self, S):
                return max(1, int(frac * S.numel()))
        return Stub()
    if meta["predictor"] == "RankPredictor":
        torch.manual_seed(5)
        return T.SeqPredictor(reference_rank_net(), "log1p64")
    if meta["predictor"] == "FinetunedPredictorWrapper":
        torch.manual_seed(6)
        return T.SeqPredictor(reference_ft_net(), "ft135")
    return None


def reference_rank_net():
    """ai_rank_predictor.py:24-42 layer order (seeded init reproduces the
    reference's weights)."""
    h = 512
    return nn.Sequential(nn.Linear(64, h), nn.LayerNorm(h), nn.ReLU(), nn.Dropout(0.2),
                         nn.Linear(h, h), nn.LayerNorm(h), nn.ReLU(), nn.Dropout(0.2),
                         nn.Linear(h, h // 2), nn.LayerNorm(h // 2), nn.ReLU(), nn.Dropout(0.1),
                         nn.Linear(h // 2, h // 4), nn.ReLU(), nn.Linear(h // 4, 1),
                         nn.Sigmoid()).eval()


def reference_ft_net():
    """finetuned_predictor.py:11-20 layer order."""
    return nn.Sequential(nn.Linear(135, 256), nn.ReLU(), nn.Linear(256, 256), nn.ReLU(),
                         nn.Linear(256, 1), nn.Sigmoid()).eval()


@pytest.mark.parametrize("run", _runs())
def test_oracle_matches_reference_golden_runs(run):
    meta = json.loads(str(GOLD[f"{run}__meta"]))
    dt = torch.complex128 if meta["dtype"].endswith("complex128") else torch.complex64
    cores = T.mps_init_plus(meta["n"], dtype=dt)
    pred = _oracle_predictor(meta)
    pol = meta["policy"] if meta["policy"] in T.POLICIES else "guardrail"
    chis, errs = [], []
    for b, U in zip(GOLD[f"{run}__bonds"], GOLD[f"{run}__gates"]):
        cores, chi, err = T.mps_apply_2q(cores, int(b), int(b) + 1, torch.from_numpy(U).to(dt),
                                         chi_max=meta["chi_max"], eps=meta["eps"],
                                         adaptive=meta["adaptive"], tol=meta["tol"],
                                         ai_predictor=pred, policy=pol, alpha=meta["alpha"])
        chis.append(chi)
        errs.append(err)
    assert chis == GOLD[f"{run}__chi"].tolist()
    tol = 1e-12 if dt == torch.complex128 else 2e-6
    assert np.abs(np.array(errs) - GOLD[f"{run}__err"]).max() < tol
    psi = T.contract(cores)
    ref = GOLD[f"{run}__psi"]
    atol = 1e-12 if dt == torch.complex128 else 1e-5
    assert np.abs(psi - ref).max() < atol * max(1.0, np.abs(ref).max())


# ---------------------------------------------------------------- predictors
def _spectra():
    vals, lens = GOLD["spec_values"], GOLD["spec_len"]
    out = []
    for dt in (np.float32, np.float64):
        for v, L in zip(vals, lens):
            out.append(v[:L].astype(dt))
    return out


def test_feature_oracles_match_reference_golden():
    for k, s in enumerate(_spectra()):
        assert np.array_equal(T.features_ft135(s), GOLD["feat_ft135"][k]), k
        lp = T.features_log1p64(s, T.n_sv(s))
        assert np.allclose(lp, GOLD["feat_log1p64"][k], rtol=2e-7, atol=1e-7), k


def test_dense_oracle_matches_torch_modules():
    torch.manual_seed(3)
    for net, d in ((reference_rank_net(), 64), (reference_ft_net(), 135)):
        with torch.no_grad():
            for mod in net:
                if isinstance(mod, nn.LayerNorm):   # non-trivial affine
                    mod.weight.normal_(1.0, 0.2)
                    mod.bias.normal_(0.0, 0.2)
            x = torch.randn(7, d)
            want = net(x).double().numpy()
        got = T.dense_forward(T.layers_from_sequential(net), x.numpy())
        assert np.abs(got - want).max() < 1e-5


def test_predictor_oracles_match_reference_outputs():
    torch.manual_seed(5)
    rp = T.SeqPredictor(reference_rank_net(), "log1p64")
    torch.manual_seed(6)
    ft = T.SeqPredictor(reference_ft_net(), "ft135")
    for k, s in enumerate(_spectra()):
        S = torch.from_numpy(s)
        assert abs(rp.fraction(S) - GOLD["frac_rankpred"][k]) < 1e-6
        assert abs(ft.fraction(S) - GOLD["frac_ftpred"][k]) < 1e-6
        assert rp.predict(S) == GOLD["rank_rankpred"][k]
        assert ft.predict(S) == GOLD["rank_ftpred"][k]


def test_spec32_features_closed_form():
    s = np.array([2.0, 1.0, 1.0], np.float64)          # weights 4, 1, 1 of total 6
    f = T.features_spec32(s, 6.0)
    assert np.allclose(f[:3], np.log10([4 / 6, 1 / 6, 1 / 6]), atol=1e-6)
    assert np.all(f[3:24] == -12.0)
    assert np.allclose(f[24:30], -12.0)                  # nothing beyond the leading values
    p = np.array([4, 1, 1]) / 6
    assert abs(f[30] - (-(p * np.log(p)).sum())) < 1e-6
    assert abs(f[31] - math.log2(3)) < 1e-6
    g = T.features_spec32(s, 12.0)                       # half of the weight elsewhere
    assert np.allclose(g[24:30], math.log10(0.5), atol=1e-6)


# ---------------------------------------------------------------- host logic
def test_entangler_builders_match_reference():
    from atlas_q.tools_qih import tn_core
    for kind in ("cz", "cnot", "rzz"):
        th = float(GOLD[f"entangler_{kind}_theta"])
        got = tn_core.build_entangler(kind, th, "cpu", torch.complex128).numpy()
        assert np.abs(got - GOLD[f"entangler_{kind}"]).max() < 1e-6   # float32 angles
    assert torch.equal(tn_core.build_entangler("none", 0.0, "cpu", torch.complex64),
                       torch.eye(4, dtype=torch.complex64))
    with pytest.raises(ValueError):
        tn_core.build_entangler("swap", 0.1, "cpu", torch.complex64)


def test_svd_logger_format_and_dataset_builder(tmp_path, monkeypatch):
    from atlas_q.tools_qih import ft_dataset, svd_logger
    monkeypatch.setenv("QIH_SVD_LOG_DIR", str(tmp_path))
    log = svd_logger.get_logger()
    assert log.active
    evs = []
    rng = np.random.default_rng(4)
    for f in range(10):
        s = np.sort(rng.random(50))[::-1]
        cores = [torch.ones(1, 2, 1, dtype=torch.complex128)] * 2
        ev = dict(kind="svd_spectrum", i=f, j=f + 1, pre_rank=4, chi_max=None, tol=1e-3,
                  adaptive=True, entangler="unknown", total_weight=float((s ** 2).sum()),
                  S_top128=list(s) + [0.0] * 78, adaptive_keep=int(10 + f))
        log.log(**ev)
        evs.append(ev)
    log.close()
    table = ft_dataset.read_events([log.path])
    assert table.unreadable == 0 and len(table) == 10
    assert {"host", "pid", "run_ts"} <= set(table.events[0])
    assert table.events[3]["S_top128"] == evs[3]["S_top128"]
    X3 = ft_dataset.ft135_features(table.spectra[3])[0]
    want = T.features_ft135(np.array(evs[3]["S_top128"], np.float32))
    assert np.allclose(X3, want, rtol=1e-6, atol=1e-6)
    y = ft_dataset.kept_fraction(table.spectra, table.keep)
    assert abs(float(y[3]) - 13 / 50) < 1e-7
    # all rows at once equal the per-event oracle
    allx = ft_dataset.ft135_features(table.spectra)
    for r, ev in enumerate(evs):
        assert np.allclose(allx[r], T.features_ft135(np.array(ev["S_top128"], np.float32)),
                           rtol=1e-6, atol=1e-6)
    X, Y, split, files, _ = ft_dataset.build_dataset([log.path])
    assert X.shape == (10, 135) and set(split.tolist()) <= {0, 1}
    monkeypatch.delenv("QIH_SVD_LOG_DIR")
    assert not svd_logger.get_logger().active


def test_calibrated_policy_adapter():
    from atlas_q.tools_qih.ai_rank_policy import CalibratedPolicyAdapter, PolicyConfig

    class Frac:
        def predict_fraction(self, s):
            return 0.5

    class Rank:
        def predict(self, s):
            return 3

    S = torch.tensor([1.0] * 10 + [0.0] * 2)            # 10 non-zero values
    a = CalibratedPolicyAdapter(Frac(), PolicyConfig(), chi_max=100)
    assert a.predict(S) == 7                            # clamp to f_min 0.70 -> floor(7.000001)
    a = CalibratedPolicyAdapter(Frac(), PolicyConfig(f_min=0.1, chi_budget=4), chi_max=100)
    assert a.predict(S) == 4                            # 5 capped by the budget
    a = CalibratedPolicyAdapter(Rank(), PolicyConfig(f_min=0.0, f_max=1.0), chi_max=2)
    assert a.predict(S) == 2                            # 3/10 -> 3, capped by chi_max
    cal = PolicyConfig(f_min=0.0, f_max=1.0, cal_x=np.array([0.0, 1.0]),
                       cal_y=np.array([0.0, 0.5]))
    assert CalibratedPolicyAdapter(Frac(), cal, chi_max=100).predict(S) == 2   # 0.25 -> 2


def test_spectral_host_features_match_oracle():
    from atlas_q.tools_qih.spectral_predictor import spectral_features_host
    rng = np.random.default_rng(8)
    for m in (1, 5, 24, 40, 128, 300):
        s = np.sort(rng.random(m) ** 3)[::-1]
        tot = float((s ** 2).sum()) * rng.uniform(1.0, 3.0)
        assert np.array_equal(spectral_features_host(s[:128], tot),
                              T.features_spec32(s[:128], tot))
    assert np.array_equal(spectral_features_host(np.zeros(5), 0.0), T.features_spec32(np.zeros(5), 0.0))
