"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point the header declares, the Python mirror exposes the
reference's names, host-side logic behaves, and nothing falls back to CPU."""
import ctypes
import inspect
import math

import numpy as np
import pytest
import torch

import atlas_q
from atlas_q import _ffi
from atlas_q import circuits as C
from atlas_q.diagnostics import MPSStatistics, bond_entropy_from_S, effective_rank
from atlas_q.truncation import check_entropy_sanity, compute_global_error_bound


def test_library_exports_every_header_symbol():
    syms = _ffi.header_symbols()
    assert {"aq_theta", "aq_rank_select", "aq_rebuild", "aq_sample", "aq_cgemm"} <= set(syms)
    L = _ffi.lib()
    for s in syms:
        assert hasattr(L, s), s
        assert s in _ffi.SIGNATURES, f"{s} has no ctypes signature"
    assert L.aq_version() >= 10000


def test_error_path_reports_message():
    L = _ffi.lib()
    rc = L.aq_theta(0, 0, None, None, None, None, 1, 1, 1, None)
    assert rc == -1
    assert b"null pointer" in L.aq_last_error()
    rc = L.aq_sample(0, 0, 0, 0, None, None, 0, None, None, None, None, 0, None)
    assert rc == -1


def test_rank_result_struct_layout():
    assert ctypes.sizeof(_ffi.RankResult) == 48


def test_export_map_matches_reference_names():
    m = atlas_q.get_adaptive_mps()
    assert set(m) == {"AdaptiveMPS", "DTypePolicy", "robust_svd", "robust_qr",
                      "condition_number", "choose_rank_from_sigma",
                      "compute_global_error_bound", "check_entropy_sanity", "MPSStatistics",
                      "bond_entropy_from_S", "effective_rank", "spectral_gap"}
    sig = inspect.signature(m["AdaptiveMPS"].__init__)
    for name, default in [("bond_dim", 8), ("eps_bond", 1e-6), ("chi_max_per_bond", 256),
                          ("budget_global_mb", None), ("device", "cuda"), ("dtype", None)]:
        assert sig.parameters[name].default == default
    assert m["DTypePolicy"]().default == torch.complex64
    assert m["DTypePolicy"]().promote_if_cond_gt == 1e6


def test_no_cpu_path():
    from atlas_q.adaptive_mps import AdaptiveMPS
    from atlas_q import kernels as K
    with pytest.raises(RuntimeError, match="HIP device"):
        AdaptiveMPS(4, device="cpu")
    with pytest.raises(RuntimeError, match="HIP device"):
        K.theta(torch.zeros(1, 2, 1, dtype=torch.complex64),
                torch.zeros(1, 2, 1, dtype=torch.complex64), np.zeros(32))


def test_statistics_and_bounds_host_logic():
    st = MPSStatistics()
    st.record(step=0, bond=0, k_star=4, chi_before=8, chi_after=4, eps_local=1e-6, entropy=1.5,
              svd_driver="torch_cuda", dtype="complex64", ms_elapsed=10.0, condS=100.0)
    st.record(step=1, bond=1, k_star=6, chi_before=8, chi_after=6, eps_local=2e-6, entropy=2.0,
              svd_driver="torch_cuda", dtype="complex64", ms_elapsed=12.0, condS=150.0)
    s = st.summary()
    assert s["total_operations"] == 2 and s["max_chi"] == 6 and s["total_time_ms"] == 22.0
    assert s["cuda_svd_pct"] == 100.0
    assert abs(st.global_error_bound() - math.sqrt(5e-12)) < 1e-18
    assert compute_global_error_bound([3.0, 4.0]) == 5.0
    assert check_entropy_sanity(1.0, 1, 1) and not check_entropy_sanity(1.1, 1, 1)
    assert abs(bond_entropy_from_S(torch.ones(4) / 2) - 2.0) < 1e-6
    assert effective_rank(torch.tensor([1.0, 0.1, 0.01, 0.001]), 0.99) <= 2


def test_circuits_are_seeded_and_haar_su4():
    a = C.brickwork(8, 4, 3)
    b = C.brickwork(8, 4, 3)
    assert len(a) == C.brickwork_count(8, 4) == 14
    for (i, U), (j, V) in zip(a, b):
        assert i == j and np.array_equal(U, V)
        assert np.allclose(U.conj().T @ U, np.eye(4), atol=1e-12)
        assert abs(np.linalg.det(U) - 1) < 1e-12
    assert [op[0] for op in a[:4]] == [0, 2, 4, 6]
    ops = C.hea(5, 2, 0)
    assert sum(op[0] == "2q" for op in ops) == 2 * 4
    pop = C.batched_hea_population(256, 0)
    assert len(pop) == 256 and min(p[0] for p in pop) >= 12 and max(p[0] for p in pop) <= 96


def test_bucket_padding_keeps_gram_eigenpairs():
    """The padding diagonal of a size bucket leaves the Gram block's
    eigenpairs unchanged and sorts every padding eigenvalue below them,
    with no two padding eigenvalues equal (host-side torch logic)."""
    from atlas_q.batched import _condition_bucket
    rng = np.random.default_rng(4)
    sz, orders = 8, np.array([5, 8, 3])
    T = torch.zeros((3, sz, sz), dtype=torch.complex128)
    grams = []
    for b, r in enumerate(orders):
        X = rng.standard_normal((r, 2)) + 1j * rng.standard_normal((r, 2))   # rank 2
        G = X @ X.conj().T
        grams.append(G)
        T[b, :r, :r] = torch.from_numpy(G)
    _condition_bucket(T, orders, None)
    for b, r in enumerate(orders):
        A = T[b].numpy()
        assert np.allclose(A[:r, :r], grams[b]) and not np.any(A[:r, r:]) and not np.any(A[r:, :r])
        lam = np.linalg.eigvalsh(A)[::-1]
        ref = np.linalg.eigvalsh(grams[b])[::-1]
        assert np.allclose(lam[:r], ref, atol=1e-12)
        pad = lam[r:]
        assert np.all(pad < 0) and len(np.unique(pad)) == len(pad)
        assert np.all(pad <= -np.abs(np.diag(grams[b])).max() + 1e-12)


def test_bucket_jitter_is_small_hermitian_and_fixed():
    from atlas_q.batched import _condition_bucket, _jitter_matrix
    E = _jitter_matrix(16, "cpu")
    assert torch.equal(E, _jitter_matrix(16, "cpu"))
    assert torch.allclose(E, E.conj().T)
    assert 0.5 < float(torch.linalg.matrix_norm(E, ord=2)) < 2.0
    rng = np.random.default_rng(5)
    X = rng.standard_normal((16, 3)) + 1j * rng.standard_normal((16, 3))
    G = torch.from_numpy(X @ X.conj().T).unsqueeze(0)
    T = G.clone()
    _condition_bucket(T, np.array([16]), E)
    d = float(G[0].diagonal().real.max())
    assert torch.allclose(T, T.conj().transpose(-2, -1))
    lam0 = np.linalg.eigvalsh(G[0].numpy())
    lam1 = np.linalg.eigvalsh(T[0].numpy())
    assert np.abs(lam1 - lam0).max() <= 2e-13 * d


def test_finite_rows_flags_each_matrix():
    from atlas_q.eig import finite_rows
    W = torch.zeros((3, 4), dtype=torch.float64)
    V = torch.zeros((3, 4, 4), dtype=torch.complex128)
    W[1, 2] = float("inf")
    V[2, 0, 3] = complex(0.0, float("nan"))
    assert finite_rows(W, V).tolist() == [True, False, False]
    assert finite_rows(W[:1], V[:1].conj()).tolist() == [True]


def test_rank_width_censored_at_training_edge():
    """A predicted log2(rank) within 0.25 of the largest training target is
    read as "at least the cap" (the network never saw larger ranks); below
    the edge the width is ceil(margin 2^y), clipped to the cap."""
    from atlas_q import rank_mlp
    p = object.__new__(rank_mlp.RankPredictor)
    p.margin, p.log2_max = 1.15, 10.0
    for y, cap, want in ((9.9, 2048, 2048), (9.75, 2048, 2048), (9.7, 2048, 957),
                         (5.0, 2048, 37), (12.0, 512, 512)):
        p.predict_log2 = lambda f, y=y: torch.tensor([y])
        assert p.predict_width(torch.zeros(32), cap, 4096) == want, (y, cap)
    d = np.load(rank_mlp.DEFAULT_WEIGHTS, allow_pickle=False)
    assert "log2_max" not in d.files or float(d["log2_max"]) >= 10.0


def test_batched_circuit_compile_and_site_merge():
    """The batched engine's host steps: circuits split into (1q block, 2q
    block on disjoint bonds) steps as a plain op-by-op walk does, and a
    block's 1q gates merged per site in op order (U1 then U2 -> U1 @ U2)."""
    from atlas_q.batched import BatchedMPS, _merge_per_site
    circs = [C.hea(n, 3, 40 + n) for n in (4, 7, 12)]
    rng = np.random.default_rng(11)
    for _ in range(6):                  # random op lists: any mix, conflicts, ragged ends
        ops = []
        for _ in range(int(rng.integers(0, 40))):
            if rng.random() < 0.5:
                ops.append(("1q", int(rng.integers(0, 8)),
                            rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))))
            else:
                ops.append(("2q", int(rng.integers(0, 7)),
                            rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))))
        circs.append(ops)
    comp = BatchedMPS.compile_circuits(circs)
    for ops, steps in zip(circs, comp):
        pos, t = 0, 0
        while pos < len(ops):
            q1 = []
            while pos < len(ops) and ops[pos][0] == "1q":
                q1.append((ops[pos][1], ops[pos][2]))
                pos += 1
            b2, used = [], set()
            while pos < len(ops) and ops[pos][0] == "2q":
                b = ops[pos][1]
                if used & {b - 1, b, b + 1}:
                    break
                used.add(b)
                b2.append((b, ops[pos][2]))
                pos += 1
            sq, su, sb, su4 = steps[t]
            assert sq.tolist() == [q for q, _ in q1] and sb.tolist() == [b for b, _ in b2]
            assert np.allclose(su, np.array([u for _, u in q1]).reshape(-1, 2, 2))
            assert np.allclose(su4, np.array([u for _, u in b2]).reshape(-1, 4, 4))
            t += 1
        assert t == len(steps)
    rng = np.random.default_rng(3)
    jj = np.array([0, 1, 0, 0, 1, 2])
    qq = np.array([3, 0, 3, 1, 0, 5])
    U = rng.standard_normal((6, 2, 2)) + 1j * rng.standard_normal((6, 2, 2))
    mj, mq, M = _merge_per_site(jj, qq, U, nmax=8)
    got = {(int(a), int(b)): m for a, b, m in zip(mj, mq, M)}
    assert set(got) == {(0, 3), (1, 0), (0, 1), (2, 5)}
    assert np.allclose(got[(0, 3)], U[0] @ U[2]) and np.allclose(got[(1, 0)], U[1] @ U[4])
    assert np.allclose(got[(0, 1)], U[3]) and np.allclose(got[(2, 5)], U[5])


def test_eigh_retry_ladder(monkeypatch):
    """eigh_desc retries on G + E with E ~ 1e-13 then 1e-11 of max|G_ii|
    (rank-deficient cores: a large zero cluster); the third attempt's
    eigenvalues stay within 1e-10 of the unperturbed ones."""
    from atlas_q import eig
    rng = np.random.default_rng(8)
    X = rng.standard_normal((80, 30)) + 1j * rng.standard_normal((80, 30))
    G = torch.from_numpy(X @ X.conj().T)                  # rank 30 of 80
    real = eig._eigh_checked
    calls = {"n": 0}

    def flaky(A):
        calls["n"] += 1
        if calls["n"] < 3:
            raise RuntimeError("linalg.eigh: non-finite eigenpairs")
        return real(A)
    monkeypatch.setattr(eig, "_eigh_checked", flaky)
    lam, V = eig.eigh_desc(G)
    assert calls["n"] == 3
    ref = np.linalg.eigvalsh(G.numpy())[::-1]
    assert np.abs(lam.numpy() - ref).max() <= 1e-10 * np.abs(ref).max()
    calls["n"] = -10                                       # every attempt fails
    monkeypatch.setattr(eig, "_eigh_checked", lambda A: (_ for _ in ()).throw(RuntimeError("x")))
    with pytest.raises(RuntimeError):
        eig.eigh_desc(G)


def test_gc_paused_restores_the_collector_state():
    """The batched engine pauses the cyclic collector for its circuit loops
    and restores the caller's setting, also when the loop raises."""
    import gc
    from atlas_q.batched import _gc_paused
    assert gc.isenabled()
    with _gc_paused():
        assert not gc.isenabled()
    assert gc.isenabled()
    try:
        with _gc_paused():
            raise ValueError
    except ValueError:
        pass
    assert gc.isenabled()
    gc.disable()
    try:
        with _gc_paused():
            pass
        assert not gc.isenabled()
    finally:
        gc.enable()
