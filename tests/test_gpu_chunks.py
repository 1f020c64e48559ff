"""The exact Gram path with a sub-layer's Gram cores split into chunks on their
own streams (each chunk's vectors, projection and cores overlapping the other
chunks' reductions) against the same sub-layer factored as one batch.

Every matrix goes through the same kernels with the same per-matrix arithmetic
either way (the reduction, bisection, inverse iteration and back-transformation
work matrix by matrix inside a batch), so kept ranks, statistics and cores
must agree bit for bit.  The cores are also checked against the FP64
truncation of the same theta (fidelity within 1e-6 of 1, the stated
tolerance).
"""
import pytest
import torch

from atlas_q import fastsvd
from atlas_q import kernels as K
from atlas_q.adaptive_mps import AdaptiveMPS, DTypePolicy, _gate4

pytestmark = pytest.mark.gpu
DEV = "cuda"
CHI = 192     # theta 384 x 384: above the single-work-group reduction's limit (256)


def _mps(ts):
    m = AdaptiveMPS(len(ts), bond_dim=1, eps_bond=1e-8, chi_max_per_bond=CHI,
                    dtype_policy=DTypePolicy(default=torch.complex64), device=DEV,
                    svd_method="gram", theta_precision="native")
    m.tensors = [t.clone() for t in ts]
    m.bond_dims = [CHI] * (len(ts) - 1)
    return m


def test_chunked_sublayer_matches_one_batch():
    from atlas_q.workloads import gate_schedule, slice_tensors
    sites = 52
    ts = slice_tensors(sites, CHI, "flat", seed=5, device=DEV, draw_on_device=True)
    gates = [(b, torch.from_numpy(U)) for b, U in gate_schedule(sites, 26, seed=6)]
    assert [b for b, _ in gates] == list(range(0, sites - 1, 2))
    split = (0.25, 0.3125, 0.4375)
    runs = {}
    for label, chunks in (("one batch", ()), ("chunked", split)):
        old = fastsvd.GRAM_CHUNKS
        fastsvd.GRAM_CHUNKS = chunks
        try:
            m = _mps(ts)
            m.apply_two_site_gates(gates, workers=4)
            torch.cuda.synchronize()
            if chunks:
                sizes = fastsvd._chunk_sizes(len(gates))
        finally:
            fastsvd.GRAM_CHUNKS = old
        runs[label] = m
    assert len(sizes) == 3
    one, ch = runs["one batch"], runs["chunked"]
    print("chunk sizes", sizes, "factor counts",
          ch.factor_stats.counts, flush=True)
    assert ch.factor_stats.counts["gram"] == len(gates)
    for key in ("k_star", "eps_local", "entropy"):
        assert one.statistics.logs[key] == ch.statistics.logs[key], key
    for q in range(sites):
        assert torch.equal(one.tensors[q], ch.tensors[q]), q
    worst = 0.0
    for b, U in gates[:4]:
        X = K.theta(ts[b].to(torch.complex128), ts[b + 1].to(torch.complex128), _gate4(U))
        Uu, S, Vh = torch.linalg.svd(X, full_matrices=False)
        k = ch.tensors[b].shape[2]
        P64 = (Uu[:, :k] * S[:k]) @ Vh[:k]
        P = (ch.tensors[b].reshape(-1, k).to(torch.complex128)
             @ ch.tensors[b + 1].reshape(k, -1).to(torch.complex128))
        F = float((torch.vdot(P64.reshape(-1), P.reshape(-1)).abs() ** 2
                   / (P64.norm() ** 2 * P.norm() ** 2)).item())
        worst = max(worst, 1 - F)
    print(f"chunked cores: max 1 - F vs the FP64 truncation {worst:.2e}", flush=True)
    assert worst < 1e-6
