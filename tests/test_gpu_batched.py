"""GPU checks of the batched engine (many uneven MPS per grouped launch).

* Without truncation the batched HEA state equals the exact state vector
  (complex64: fidelity within 1e-6 of 1).
* With truncation (Haar brickwork, caps binding) every MPS's rank sequence
  and amplitudes equal those of the single-MPS AdaptiveMPS Gram method run
  gate by gate (same kernels and arithmetic order per gate): ranks exact,
  amplitudes within 1e-5.
* Under truncation against the reference restatement (the oracle, CPU,
  gate by gate): every kept rank equals the oracle's complex64 run (same
  storage precision and rank rule), and each final state's fidelity with
  the oracle's complex128 run is within 1e-6 of 1 (the stated |dF| bound;
  measured ~2e-12).
* The grouped sampler returns, per MPS, exactly the bits of the single-MPS
  sampler for the same uniforms, and a HIP-graph replay of the whole sweep
  returns the eager result on every replay.
"""
import numpy as np
import pytest
import torch

from oracle import statevector as SV
from atlas_q.adaptive_mps import AdaptiveMPS
from atlas_q.autotune import Autotuner
from atlas_q.batched import BatchedMPS
from atlas_q.circuits import brickwork, hea
from atlas_q import kernels as K

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _tuner():
    return Autotuner(path="/tmp/atlasq_test_autotune.json", tune=True)


def test_batched_hea_exact_against_statevector():
    specs = [(6, 64), (9, 64), (12, 64), (7, 64), (10, 64)]
    circs = [hea(n, 3, 100 + j) for j, (n, _) in enumerate(specs)]
    bm = BatchedMPS(specs, eps_bond=1e-10, device=DEV, autotuner=_tuner())
    bm.run_hea(circs)
    for j, (n, _) in enumerate(specs):
        psi = SV.run_circuit(n, circs[j])
        t = bm.tensors(j)
        res = t[0].reshape(2, -1).to(torch.complex128)
        for q in range(1, n):
            T = t[q].to(torch.complex128)
            res = (res @ T.reshape(T.shape[0], -1)).reshape(-1, T.shape[2])
        out = res.reshape(-1).cpu().numpy()
        fid = abs(np.vdot(psi, out)) ** 2 / np.vdot(out, out).real
        assert abs(1 - fid) < 1e-6, (j, fid)


def test_batched_brickwork_matches_single_mps_gram():
    specs = [(8, 4), (10, 8), (12, 6), (9, 16)]
    depth = 8
    circs = [brickwork(n, depth, 7 + j) for j, (n, _) in enumerate(specs)]
    bm = BatchedMPS(specs, eps_bond=1e-8, device=DEV, autotuner=_tuner())
    # layer by layer: every MPS's gates of one brickwork layer form one sub-layer
    for layer in range(depth):
        ops = []
        for j, (n, _) in enumerate(specs):
            per = len(range(layer % 2, n - 1, 2))
            start = sum(len(range(l % 2, n - 1, 2)) for l in range(layer))
            ops.append([(b, U) for b, U in circs[j][start:start + per]])
        bm.apply_2q_layer(ops)
    for j, (n, chi) in enumerate(specs):
        ref = AdaptiveMPS(n, bond_dim=1, eps_bond=1e-8, chi_max_per_bond=chi, device=DEV,
                          svd_method="gram")
        for b, U in circs[j]:
            ref.apply_two_site_gate(b, torch.from_numpy(U))
        ks = [k for mj, k in zip(bm.logs["mps"], bm.logs["k"]) if mj == j]
        es = [e for mj, e in zip(bm.logs["mps"], bm.logs["eps_local"]) if mj == j]
        kr, er = ref.statistics.logs["k_star"], ref.statistics.logs["eps_local"]
        assert len(ks) == len(kr)
        for a, b, e1, e2 in zip(ks, kr, es, er):
            # a float-saturated threshold may resolve differently on the
            # rounding of tiny singular values (padded vs unpadded eigensolve)
            assert a == b or (e1 < 1e-3 and e2 < 1e-3)
        np.testing.assert_allclose(es, er, atol=1e-4)
        for bstr in (0, 1, 2 ** n - 1, 12345 % 2 ** n):
            bits = [(bstr >> (n - 1 - i)) & 1 for i in range(n)]
            t = bm.tensors(j)
            r = t[0][:, bits[0], :].to(torch.complex128)
            for q in range(1, n):
                r = r @ t[q][:, bits[q], :].to(torch.complex128)
            a = complex(r[0, 0])
            assert abs(a - ref.get_amplitude(bstr)) < 1e-5


def _batched_brickwork(specs, depth, seed0):
    circs = [brickwork(n, depth, seed0 + j) for j, (n, _) in enumerate(specs)]
    bm = BatchedMPS(specs, eps_bond=1e-8, device=DEV, autotuner=_tuner())
    for layer in range(depth):
        ops = []
        for j, (n, _) in enumerate(specs):
            per = len(range(layer % 2, n - 1, 2))
            start = sum(len(range(l % 2, n - 1, 2)) for l in range(layer))
            ops.append([(b, U) for b, U in circs[j][start:start + per]])
        bm.apply_2q_layer(ops)
    return bm, circs


def _dense(tensors):
    res = tensors[0].reshape(2, -1).to(torch.complex128)
    for T in tensors[1:]:
        T = T.to(torch.complex128)
        res = (res @ T.reshape(T.shape[0], -1)).reshape(-1, T.shape[2])
    return res.reshape(-1).cpu().numpy()


def test_batched_haar_brickwork_against_oracle():
    from oracle import ref_mps as R
    specs = [(8, 4), (10, 6), (12, 8), (9, 5), (11, 16)]
    bm, circs = _batched_brickwork(specs, 8, 31)
    worst = 0.0
    for j, (n, chi) in enumerate(specs):
        runs = {}
        for dt in (torch.complex64, torch.complex128):
            o = R.OracleMPS(n, bond_dim=1, eps_bond=1e-8, chi_max_per_bond=chi,
                            dtype_policy=R.DTypePolicy(default=dt))
            for b, U in circs[j]:
                o.apply_two_site_gate(b, torch.from_numpy(U))
            runs[dt] = o
        ks = [k for mj, k in zip(bm.logs["mps"], bm.logs["k"]) if mj == j]
        assert ks == runs[torch.complex64].statistics.logs["k_star"], j
        psi_b = _dense(bm.tensors(j))
        psi_o = runs[torch.complex128].to_statevector().numpy()
        F = abs(np.vdot(psi_o, psi_b)) ** 2 / (np.vdot(psi_o, psi_o).real
                                               * np.vdot(psi_b, psi_b).real)
        worst = max(worst, 1 - F)
        assert 1 - F < 1e-6, (j, F)
    print(f"batched vs oracle: ranks equal at every update, max 1 - F {worst:.2e}", flush=True)


def test_grouped_sampler_matches_single_and_graph_replays():
    specs = [(12, 16), (20, 8), (15, 32)]
    circs = [hea(n, 2, 50 + j) for j, (n, _) in enumerate(specs)]
    bm = BatchedMPS(specs, eps_bond=1e-8, device=DEV, autotuner=_tuner())
    bm.run_hea(circs)
    shots = 300
    bm.prepare_sampler(shots)
    rng = np.random.default_rng(3)
    eager = []
    for rep in range(3):
        for j in range(len(specs)):
            bm.uniforms_of(j).copy_(torch.from_numpy(rng.random((shots, specs[j][0]))))
        bm.sweep()
        torch.cuda.synchronize()
        eager.append([bm.sample_bits(j).clone() for j in range(len(specs))])
        for j in range(len(specs)):
            ref, _ = K.sample_bits(bm.tensors(j), bm.uniforms_of(j).clone())
            assert torch.equal(ref, bm.sample_bits(j)), (rep, j)
    # capture the sweep once, replay with the same uniform sets
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        bm.sweep()  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        bm.sweep()
    rng = np.random.default_rng(3)
    for rep in range(3):
        for j in range(len(specs)):
            bm.uniforms_of(j).copy_(torch.from_numpy(rng.random((shots, specs[j][0]))))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for j in range(len(specs)):
            assert torch.equal(eager[rep][j], bm.sample_bits(j)), (rep, j)
        # eager work between replays
        _ = (bm.arena.abs() ** 2).sum().item()


def test_svd_fill_matches_gram_eigenpairs():
    """Bucket rows whose eigensolve failed are refilled from the SVD of their
    own theta: the eigen-equation of the Gram matrix holds on the non-zero
    spectrum and the rest of the row is zero."""
    import atlas_q.batched as BM
    rng = np.random.default_rng(12)
    sizes = [(6, 10), (8, 5)]                      # (m, n): left (m <= n) and right
    xs = [rng.standard_normal(s) + 1j * rng.standard_normal(s) for s in sizes]
    m = np.array([s[0] for s in sizes])
    n = np.array([s[1] for s in sizes])
    left = m <= n
    xoff = np.array([0, m[0] * n[0]])
    X64 = torch.from_numpy(np.concatenate([x.reshape(-1) for x in xs])).to(DEV)
    sz = 8
    grams = [xs[0] @ xs[0].conj().T, xs[1].conj().T @ xs[1]]
    W = torch.full((2, sz), float("nan"), dtype=torch.float64, device=DEV)
    V = torch.full((2, sz, sz), float("nan"), dtype=torch.complex128, device=DEV)
    BM.BatchedMPS._svd_fill(W, V, [0, 1], np.array([0, 1]), X64, xoff, m, n, left)
    assert bool(torch.isfinite(W).all()) and bool(torch.isfinite(torch.view_as_real(V)).all())
    for b, G in enumerate(grams):
        r = G.shape[0]
        w = W[b, :r].cpu().numpy()
        v = V[b, :r, :r].cpu().numpy()
        assert np.allclose(np.sort(w)[::-1], np.sort(np.linalg.eigvalsh(G))[::-1], atol=1e-10)
        k = min(xs[b].shape)                      # non-zero part of the spectrum
        assert np.allclose(G @ v[:, :k], v[:, :k] * w[:k], atol=1e-9)
        assert float(W[b, r:].abs().max() if r < sz else 0.0) == 0.0


def _exact_fidelities(bm, specs, circs):
    out = []
    for j, (n, _) in enumerate(specs):
        psi = SV.run_circuit(n, circs[j])
        t = bm.tensors(j)
        res = t[0].reshape(2, -1).to(torch.complex128)
        for q in range(1, n):
            T = t[q].to(torch.complex128)
            res = (res @ T.reshape(T.shape[0], -1)).reshape(-1, T.shape[2])
        v = res.reshape(-1).cpu().numpy()
        out.append(abs(np.vdot(psi, v)) ** 2 / np.vdot(v, v).real)
    return out


def test_batched_wide_buckets_exact_against_statevector(monkeypatch):
    """Orders above the Jacobi limit (tridiagonal-solver buckets of order 128
    and 256) without truncation equal the exact state; one bucket row whose
    vectors are reported as failing the orthogonality re-check goes through
    the theta-SVD refill and the state is still exact."""
    from atlas_q.eig import GramTridiag
    specs = [(16, 256), (14, 128)]
    circs = [hea(n, 9, 300 + j) for j, (n, _) in enumerate(specs)]
    real = GramTridiag.vectors
    seen = {"calls": 0}

    def one_failure(self, ks):
        Et, err = real(self, ks)
        seen["calls"] += 1
        if seen["calls"] == 1:
            err = err.clone()
            err[0] = 1.0
        return Et, err
    monkeypatch.setattr(GramTridiag, "vectors", one_failure)
    bm = BatchedMPS(specs, eps_bond=1e-10, device=DEV, autotuner=_tuner())
    bm.run_hea(circs)
    assert seen["calls"] > 0 and bm.eig_fallbacks >= 1
    assert max(int(bm.dims[j].max()) for j in range(len(specs))) > 64
    fids = _exact_fidelities(bm, specs, circs)
    print("fidelities vs state vector", fids)
    for j, fid in enumerate(fids):
        assert abs(1 - fid) < 1e-5, (j, fid)


def test_batched_hea_tridiag_buckets_no_fallbacks():
    """The CZ-ladder states' Gram matrices repeat eigenvalues exactly; with
    buckets of order 128-512 on the tridiagonal solver every bucket row
    passes the orthogonality re-check (no theta-SVD refill) and the states
    equal the exact state vector (cutoff 1e-10, caps not binding)."""
    specs = [(16, 256), (17, 256), (15, 128), (18, 512)]
    circs = [hea(n, 8, 900 + j) for j, (n, _) in enumerate(specs)]
    bm = BatchedMPS(specs, eps_bond=1e-10, device=DEV, autotuner=_tuner())
    bm.run_hea(circs)
    fids = _exact_fidelities(bm, specs, circs)
    print("bond max", [int(bm.dims[j].max()) for j in range(len(specs))],
          "fallbacks", bm.eig_fallbacks, "fidelities", fids)
    assert max(int(bm.dims[j].max()) for j in range(len(specs))) > 64
    assert bm.eig_fallbacks == 0
    for j, fid in enumerate(fids):
        assert abs(1 - fid) < 1e-5, (j, fid)


def test_batched_wide_buckets_rocsolver_comparison_path(monkeypatch):
    """The rocSOLVER divide-and-conquer buckets (eig_solver="heevd", the
    comparison point) give the exact state too, a failed row included."""
    import atlas_q.batched as BM
    specs = [(16, 256), (14, 128)]
    circs = [hea(n, 9, 300 + j) for j, (n, _) in enumerate(specs)]
    real = BM.heevd_batched
    seen = {"calls": 0}

    def one_failure(T, inplace=False, return_ok=False):
        W, V, ok = real(T, inplace=inplace, return_ok="device")
        seen["calls"] += 1
        if seen["calls"] == 1:
            ok = ok.clone()
            ok[0] = False
        return W, V, ok
    monkeypatch.setattr(BM, "heevd_batched", one_failure)
    bm = BatchedMPS(specs, eps_bond=1e-10, device=DEV, autotuner=_tuner(), eig_solver="heevd")
    bm.run_hea(circs)
    assert seen["calls"] > 0 and bm.eig_fallbacks >= 1
    for j, fid in enumerate(_exact_fidelities(bm, specs, circs)):
        assert abs(1 - fid) < 1e-5, (j, fid)


def test_batched_tridiag_and_rocsolver_buckets_agree():
    """Truncating brickwork (caps 40-72, buckets of order 128 and 256): the
    library's tridiagonal buckets and rocSOLVER's divide-and-conquer keep
    the same ranks and states (fidelity 1 within 1e-6 per MPS), with no
    fallback in either."""
    specs = [(12, 40), (13, 50), (14, 72)]
    depth = 12
    circs = [brickwork(n, depth, 170 + j) for j, (n, _) in enumerate(specs)]
    runs = []
    for solver in ("tridiag", "heevd"):
        bm = BatchedMPS(specs, eps_bond=1e-8, device=DEV, autotuner=_tuner(), eig_solver=solver)
        for layer in range(depth):
            ops = []
            for j, (n, _) in enumerate(specs):
                per = len(range(layer % 2, n - 1, 2))
                start = sum(len(range(l % 2, n - 1, 2)) for l in range(layer))
                ops.append([(b, U) for b, U in circs[j][start:start + per]])
            bm.apply_2q_layer(ops)
        assert bm.eig_fallbacks == 0, solver
        runs.append(bm)
    a, b = runs
    assert a.logs["k"] == b.logs["k"]
    for j, (n, _) in enumerate(specs):
        va = _dense(a.tensors(j))
        vb = _dense(b.tensors(j))
        fid = abs(np.vdot(va, vb)) ** 2 / (np.vdot(va, va).real * np.vdot(vb, vb).real)
        print(f"mps {j}: fidelity tridiag vs rocSOLVER buckets {fid:.15f}")
        assert abs(1 - fid) < 1e-6


def _dense(ts):
    res = ts[0].reshape(2, -1).to(torch.complex128)
    for T in ts[1:]:
        T = T.to(torch.complex128)
        res = (res @ T.reshape(T.shape[0], -1)).reshape(-1, T.shape[2])
    return res.reshape(-1).cpu().numpy()


def test_batched_padded_buckets_match_single_mps_gram():
    """Caps that are not powers of two put Gram matrices of order 2 cap into
    padded buckets above the Jacobi limit (orders 80, 100, 144 in buckets
    128 and 256); ranks and amplitudes equal the single-MPS Gram method."""
    specs = [(12, 40), (13, 50), (14, 72)]
    depth = 14
    circs = [brickwork(n, depth, 70 + j) for j, (n, _) in enumerate(specs)]
    bm = BatchedMPS(specs, eps_bond=1e-8, device=DEV, autotuner=_tuner())
    for layer in range(depth):
        ops = []
        for j, (n, _) in enumerate(specs):
            per = len(range(layer % 2, n - 1, 2))
            start = sum(len(range(l % 2, n - 1, 2)) for l in range(layer))
            ops.append([(b, U) for b, U in circs[j][start:start + per]])
        bm.apply_2q_layer(ops)
    assert bm.eig_fallbacks == 0
    for j, (n, chi) in enumerate(specs):
        assert int(bm.dims[j].max()) == chi
        ref = AdaptiveMPS(n, bond_dim=1, eps_bond=1e-8, chi_max_per_bond=chi, device=DEV,
                          svd_method="gram")
        for b, U in circs[j]:
            ref.apply_two_site_gate(b, torch.from_numpy(U))
        ks = [k for mj, k in zip(bm.logs["mps"], bm.logs["k"]) if mj == j]
        es = [e for mj, e in zip(bm.logs["mps"], bm.logs["eps_local"]) if mj == j]
        kr, er = ref.statistics.logs["k_star"], ref.statistics.logs["eps_local"]
        assert len(ks) == len(kr)
        for a, b, e1, e2 in zip(ks, kr, es, er):
            assert a == b or (e1 < 1e-3 and e2 < 1e-3)
        for bstr in (0, 1, 2 ** n - 1, 12345 % 2 ** n):
            bits = [(bstr >> (n - 1 - i)) & 1 for i in range(n)]
            t = bm.tensors(j)
            r = t[0][:, bits[0], :].to(torch.complex128)
            for q in range(1, n):
                r = r @ t[q][:, bits[q], :].to(torch.complex128)
            assert abs(complex(r[0, 0]) - ref.get_amplitude(bstr)) < 1e-5


def test_concurrent_bucket_eigensolves_match_one_stream():
    """The tridiagonal-solver size buckets of a sub-layer solved on three worker
    streams (each waiting for the Gram launch, the caller's stream waiting
    for each of them) give bit-identical states and logs to one stream."""
    specs = [(16, 256), (14, 128), (13, 50)]
    circs = [hea(n, 9, 500 + j) for j, (n, _) in enumerate(specs)]
    runs = []
    for workers in (1, 3):
        bm = BatchedMPS(specs, eps_bond=1e-8, device=DEV, autotuner=_tuner(),
                        eig_workers=workers)
        bm.run_hea(circs)
        torch.cuda.synchronize()
        runs.append((bm.arena.clone(), {k: list(v) for k, v in bm.logs.items()},
                     bm.eig_fallbacks, max(int(bm.dims[j].max()) for j in range(len(specs)))))
    (a1, l1, f1, d1), (a3, l3, f3, d3) = runs
    assert d1 > 64 and d1 == d3          # buckets above the Jacobi limit were used
    assert f1 == f3
    assert l1 == l3
    assert torch.equal(a1, a3)
