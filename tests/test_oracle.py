"""The CPU oracle pinned against the reference's golden vectors, hand-computed
known answers and what the mathematics fixes (exact state vector, Born-rule
marginals).  CPU only."""
import math

import numpy as np
import pytest
import torch

from conftest import load_golden
from oracle import ref_mps as R
from oracle import statevector as SV
from atlas_q.circuits import brickwork, haar_su4, hea

torch.set_num_threads(1)


# ---------------------------------------------------------------- rank rule
def test_rank_rule_known_answer_reference_test_case():
    # truncation.py:21-83 on the reference's own test spectrum
    # (tests/unit/test_adaptive_mps.py:387-398), computed by hand:
    # E = [1, 1.25, 1.26, 1.2601, 1.260101], thresh = 0.9999*1.260101
    S = torch.tensor([1.0, 0.5, 0.1, 0.01, 0.001])
    k, eps, ent, cond = R.choose_rank_from_sigma(S, 1e-2, 5)
    assert k == 3
    assert abs(eps - math.sqrt(1.260101 - 1.26)) < 2e-5
    p = np.array([1.0, 0.25, 0.01]) / 1.26
    assert abs(ent - float(-(p * np.log(p)).sum())) < 1e-6
    assert abs(cond - 10.0) < 1e-5


def test_rank_rule_left_sided_tie():
    # thresh == E[2] exactly: searchsorted(left) -> index 2 -> k = 3
    S = torch.tensor([1.0, 1.0, 1.0, 1.0])
    assert R.choose_rank_from_sigma(S, 0.5, 10)[0] == 3


def test_rank_rule_float32_saturation_differs_from_float64():
    # weights [1, 1e-8], eps^2 = 1e-10: float64 keeps both, float32 rounds
    # 1 + 1e-8 to 1 so thresh == total == E[0] and keeps one
    S64 = torch.tensor([1.0, 1e-4], dtype=torch.float64)
    assert R.choose_rank_from_sigma(S64, 1e-5, 10)[0] == 2
    assert R.choose_rank_from_sigma(S64.float(), 1e-5, 10)[0] == 1


def test_rank_rule_special_cases():
    assert R.choose_rank_from_sigma(torch.zeros(0), 1e-3, 4) == (0, 0.0, 0.0, float("inf"))
    k, e, s, c = R.choose_rank_from_sigma(torch.tensor([1e-16, 0.0], dtype=torch.float64), 1e-3, 4)
    assert (k, e, s) == (1, 0.0, 0.0) and c == float("inf")
    assert R.choose_rank_from_sigma(torch.ones(10), 1e-10, 5)[0] == 5
    assert R.choose_rank_from_sigma(torch.ones(10), 1e-10, 10, lambda k: k <= 3)[0] == 3
    # k = 1 -> condS inf
    assert R.choose_rank_from_sigma(torch.tensor([1.0, 1e-6]), 1e-2, 4)[3] == float("inf")


def test_mps_statistics_aggregation():
    st = R.MPSStatistics()
    st.record(step=0, eps_local=1e-6)
    st.record(step=1, eps_local=2e-6)
    assert abs(st.global_error_bound() - math.sqrt(5e-12)) < 1e-18


# ---------------------------------------------------------------- theta
@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_theta_closed_form(dtype):
    g = torch.Generator().manual_seed(3)
    cl, cm, cr = 3, 4, 5
    A = torch.randn(cl, 2, cm, dtype=dtype, generator=g)
    B = torch.randn(cm, 2, cr, dtype=dtype, generator=g)
    U = torch.from_numpy(haar_su4(np.random.default_rng(1))).to(dtype)
    X = R.theta(A, B, U).to(torch.complex128)
    A64, B64, U64 = (t.to(torch.complex128) for t in (A, B, U))
    ref = torch.zeros(2 * cl, 2 * cr, dtype=torch.complex128)
    for l in range(cl):
        for p in range(2):
            for q in range(2):
                for j in range(cr):
                    s = 0
                    for p1 in range(2):
                        for p2 in range(2):
                            s += U64[2 * p + q, 2 * p1 + p2] * (A64[l, p1, :] @ B64[:, p2, j])
                    ref[2 * l + p, q * cr + j] = s
    tol = 1e-5 if dtype == torch.complex64 else 1e-12
    assert torch.allclose(X, ref, atol=tol, rtol=tol)
    # the reference's no-triton fallback applies U^T
    Xt = R.theta_transposed(A64, B64, U64.T.contiguous())
    assert torch.allclose(Xt, ref, atol=1e-12)


# ---------------------------------------------------------------- golden runs
def _replay_brickwork(g, dtype, budget=None):
    torch.manual_seed(1234)
    n = int(g["n"])
    promote = float(g["promote"]) if "promote" in g and float(g["promote"]) > 0 else None
    policy = R.DTypePolicy(default=dtype) if promote is None else \
        R.DTypePolicy(default=dtype, promote_if_cond_gt=promote)
    m = R.OracleMPS(n, bond_dim=1, eps_bond=float(g["eps"]), chi_max_per_bond=int(g["chi"]),
                    budget_global_mb=budget, dtype_policy=policy)
    for b, U in brickwork(n, int(g["depth"]), int(g["seed"])):
        m.apply_two_site_gate(b, torch.from_numpy(U).to(dtype))
    return m


@pytest.mark.parametrize("name,dtype,budget", [
    ("brick_n8_c128", torch.complex128, None),
    ("brick_n10_c64", torch.complex64, None),
    ("brick_n10_c64_eps", torch.complex64, None),
    ("brick_n10_c128_eps", torch.complex128, None),
    ("brick_n8_budget", torch.complex64, 0.0025),
    ("config1_n16_c128", torch.complex128, None),
    ("brick_n8_promote", torch.complex64, None),
])
def test_oracle_matches_reference_golden(name, dtype, budget):
    g = load_golden(name)
    m = _replay_brickwork(g, dtype, budget)
    lg = m.statistics.logs
    assert lg["bond"] == g["bond"].tolist()
    assert lg["k_star"] == g["k"].tolist()
    assert lg["chi_before"] == g["chi_before"].tolist()
    np.testing.assert_allclose(lg["eps_local"], g["eps_local"], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(lg["entropy"], g["entropy"], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(lg["condS"], g["condS"], rtol=1e-6)
    assert [d == "torch.complex128" for d in lg["dtype"]] == g["promoted"].tolist()
    assert m.bond_dims == g["bond_dims"].tolist()
    psi = m.to_statevector().numpy().astype(np.complex128)
    tol = 1e-5 if dtype == torch.complex64 else 1e-11
    if "psi" in g:
        np.testing.assert_allclose(psi, g["psi"], atol=tol)
    else:
        np.testing.assert_allclose(psi[g["amp_index"]], g["amp"], atol=tol)


def test_budget_changes_ranks():
    # the budget fixture really exercises budget_ok (adaptive_mps.py:255-264)
    g = load_golden("brick_n8_budget")
    m = _replay_brickwork(g, torch.complex64, None)
    assert m.statistics.logs["k_star"] != g["k"].tolist()


def test_oracle_hea_golden():
    g = load_golden("hea_n8_c64")
    torch.manual_seed(99)
    m = R.OracleMPS(8, bond_dim=1, eps_bond=float(g["eps"]), chi_max_per_bond=int(g["chi"]),
                    dtype_policy=R.DTypePolicy(default=torch.complex64))
    for op in hea(8, int(g["layers"]), int(g["seed"])):
        U = torch.from_numpy(op[2]).to(torch.complex64)
        if op[0] == "1q":
            m.apply_single_qubit_gate(op[1], U)
        else:
            m.apply_two_site_gate(op[1], U)
    assert m.statistics.logs["k_star"] == g["k"].tolist()
    np.testing.assert_allclose(m.to_statevector().numpy(), g["psi"], atol=1e-5)


@pytest.mark.parametrize("name,dtype", [("init_n6_bd4_c64", torch.complex64),
                                        ("init_n5_bd3_c128", torch.complex128)])
def test_oracle_init_golden(name, dtype):
    g = load_golden(name)
    torch.manual_seed(int(g["seed"]))
    m = R.OracleMPS(int(g["n"]), bond_dim=int(g["bond_dim"]),
                    dtype_policy=R.DTypePolicy(default=dtype))
    assert m.bond_dims == g["bond_dims"].tolist()
    for i, t in enumerate(m.tensors):
        np.testing.assert_allclose(t.numpy(), g[f"t{i}"], atol=1e-6)


# ---------------------------------------------------------------- exact state vector
def test_oracle_exact_without_truncation():
    n, depth = 8, 8
    ops = brickwork(n, depth, 5)
    m = R.OracleMPS(n, bond_dim=1, eps_bond=1e-14, chi_max_per_bond=64,
                    dtype_policy=R.DTypePolicy(default=torch.complex128))
    for b, U in ops:
        m.apply_two_site_gate(b, torch.from_numpy(U))
    psi_mps = m.to_statevector().numpy()
    psi = SV.run_circuit(n, [("2q", b, U) for b, U in ops])
    assert abs(abs(np.vdot(psi, psi_mps)) ** 2 - 1) < 1e-10
    np.testing.assert_allclose(psi_mps, psi, atol=1e-10)


def test_statevector_closed_forms():
    # GHZ via H (U^T convention: H is symmetric) and CNOT chain
    n = 4
    H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    CNOT = np.eye(4)[[0, 1, 3, 2]]
    ops = [("1q", 0, H)] + [("2q", i, CNOT) for i in range(n - 1)]
    psi = SV.run_circuit(n, ops)
    ref = np.zeros(16)
    ref[0] = ref[15] = 1 / np.sqrt(2)
    np.testing.assert_allclose(psi, ref, atol=1e-14)
    # the oracle MPS gives the same GHZ state with chi = 2
    m = R.OracleMPS(n, bond_dim=1, eps_bond=1e-10, chi_max_per_bond=4,
                    dtype_policy=R.DTypePolicy(default=torch.complex128))
    m.apply_single_qubit_gate(0, torch.from_numpy(H).to(torch.complex128))
    for i in range(n - 1):
        m.apply_two_site_gate(i, torch.from_numpy(CNOT).to(torch.complex128))
    assert max(m.bond_dims) == 2
    np.testing.assert_allclose(m.to_statevector().numpy(), ref, atol=1e-12)


# ---------------------------------------------------------------- sampler
@pytest.mark.parametrize("name,dtype", [("sampler_n12_c64", torch.complex64),
                                        ("sampler_n12_c128", torch.complex128)])
def test_sampler_matches_reference_golden(name, dtype):
    g = load_golden(name)
    n = int(g["n"])
    tensors = [torch.from_numpy(g[f"t{i}"]) for i in range(n)]
    out = R.batch_sweep_sample(tensors, g["uniforms"])
    assert out == g["samples"].tolist()
    assert R.single_shot_sample(tensors, g["u_single"]) == g["single"].tolist()


def test_sampler_marginals_are_born_rule():
    # with sites 1..n-1 right-canonical (to_mixed_canonical(0)) the sweep's
    # normalised p0 equals the exact conditional probability
    n = 10
    ops = brickwork(n, 6, 8)
    m = R.OracleMPS(n, bond_dim=1, eps_bond=1e-14, chi_max_per_bond=64,
                    dtype_policy=R.DTypePolicy(default=torch.complex128))
    for b, U in ops:
        m.apply_two_site_gate(b, torch.from_numpy(U))
    m.to_mixed_canonical(0)
    psi = SV.run_circuit(n, [("2q", b, U) for b, U in ops])
    u = np.random.default_rng(0).random((64, n))
    samples, p0 = R.batch_sweep_sample(m.tensors, u, return_p0=True)
    for s in range(8):
        bits = [(samples[s] >> (n - 1 - i)) & 1 for i in range(n)]
        for i in range(n):
            assert abs(p0[s, i] - SV.marginal_p0(psi, n, bits[:i])) < 1e-9
