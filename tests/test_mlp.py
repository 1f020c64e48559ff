"""Rank-predictor MLP: the NumPy oracle pinned by finite differences and a
hand-worked AdamW step (CPU), the HIP forward/backward/AdamW kernels checked
against the oracle (GPU), and GPU training that reduces the loss."""
import numpy as np
import pytest

from oracle import mlp_ref as M


def test_oracle_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1)
    X = rng.normal(size=(5, M.IN))
    dy = rng.normal(size=5)
    g = M.backward(p, X, dy, round_weights=False)
    idx = np.concatenate([rng.choice(np.arange(M.O_W1, M.O_B1), 6, replace=False),
                          rng.choice(np.arange(M.O_W2, M.O_B2), 6, replace=False),
                          np.arange(M.O_W3, M.O_W3 + 3), [M.O_B1, M.O_B2 + 1, M.O_B3]])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = h
        fp = dy @ M.forward(p + e, X, round_weights=False)[0]
        fm = dy @ M.forward(p - e, X, round_weights=False)[0]
        num = (fp - fm) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), (i, num, g[i])


def test_oracle_adamw_first_step_is_signed_lr():
    # step 1: m_hat = g, v_hat = g^2 -> update = lr * g/(|g|+eps); weights also decay
    p = np.ones(M.NPARAM)
    g = np.full(M.NPARAM, -0.5)
    p1, m, v = M.adamw(p, g, np.zeros(M.NPARAM), np.zeros(M.NPARAM), 1, lr=0.1, wd=0.01)
    w = M.O_W1
    b = M.O_B1
    assert abs(p1[w] - (1 - 0.1 * 0.01 + 0.1 * 0.5 / (0.5 + 1e-8))) < 1e-12
    assert abs(p1[b] - (1 + 0.1 * 0.5 / (0.5 + 1e-8))) < 1e-12


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -9], np.float32)
    r = M.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert r[4] == 1.0


@pytest.mark.gpu
def test_hip_mlp_forward_backward_match_oracle():
    import torch
    from atlas_q.rank_mlp import RankPredictor
    rp = RankPredictor("cuda", seed=3)
    p = rp.master.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(77, M.IN)).astype(np.float32)
    dy = rng.normal(size=77).astype(np.float32)
    y_ref, _ = M.forward(p, X.astype(np.float64))
    y = rp.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    print(f"forward: max |y - y_oracle| = {np.abs(y - y_ref).max():.2e} "
          f"(max |y_oracle| {np.abs(y_ref).max():.2e})", flush=True)
    np.testing.assert_allclose(y, y_ref, rtol=2e-4, atol=2e-4)
    from atlas_q._ffi import lib
    import ctypes
    g = torch.zeros(rp.nparam, dtype=torch.float32, device="cuda")
    wsn = int(lib().aq_mlp_workspace_floats(77))
    ws = torch.empty(wsn, dtype=torch.float32, device="cuda")
    Xd = torch.from_numpy(X).cuda()
    dyd = torch.from_numpy(dy).cuda()
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().aq_mlp_backward(P(Xd), P(dyd), 77, P(rp.wb), P(rp.master), P(g), P(ws), wsn, s) == 0
    g_ref = M.backward(p, X.astype(np.float64), dy.astype(np.float64))
    g = g.cpu().numpy()
    print(f"backward: max |g - g_oracle| / max |g_oracle| = "
          f"{np.abs(g - g_ref).max() / np.abs(g_ref).max():.2e}", flush=True)
    np.testing.assert_allclose(g, g_ref, rtol=1e-3, atol=1e-3 * np.abs(g_ref).max())


@pytest.mark.gpu
def test_hip_adamw_matches_oracle_and_training_reduces_loss():
    import torch
    from atlas_q.rank_mlp import RankPredictor
    rng = np.random.default_rng(2)
    F = rng.normal(size=(4096, M.IN)).astype(np.float32)
    w = rng.normal(size=M.IN)
    k = np.clip(np.round(2 ** (3 + np.tanh(F @ w / 4) * 2)), 1, None)
    rp = RankPredictor("cuda", seed=0)
    log = rp.fit(F, k, steps=600, batch=1024, lr=3e-3, log_every=100)
    assert log[-1][1] < 0.2 * log[0][1], log
    ev = rp.evaluate(F, k)
    assert ev["rmse_log2"] < 0.4, ev
