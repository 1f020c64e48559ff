"""GPU checks at the headline shapes, through properties that hold at any size.

* A brickwork sub-layer at bond 1024 (theta 2048 x 2048, the chi_max = 1024
  steady state), factored as the bench times it (learned width, BF16x3 theta,
  grouped Gram path): every kept rank is the cap; per gate the truncation
  fidelity against the FP64 truncation (complex128 SVD of the complex128
  theta) is within 1e-6 of 1 (the stated tolerance), the kept weight and
  eps_local match the FP64 spectrum, the right core is right-orthonormal,
  and |F_fast - F_plain| <= 1e-6 against the plain GPU SVD path.
* The sampler at the 64-qubit, chi_max = 512 shape with 10,000 shots: the
  oracle's bits on a subset of the shots (bit-exact where every CDF margin
  exceeds 1e-5), and the first qubit's frequency against its exact marginal.
"""
import numpy as np
import pytest
import torch

from oracle import ref_mps as R
from atlas_q import kernels as K
from atlas_q.adaptive_mps import AdaptiveMPS, DTypePolicy, _gate4

pytestmark = pytest.mark.gpu
DEV = "cuda"
CHI = 1024


def _mps(ts, method, prec):
    m = AdaptiveMPS(len(ts), bond_dim=1, eps_bond=1e-8, chi_max_per_bond=CHI,
                    dtype_policy=DTypePolicy(default=torch.complex64), device=DEV,
                    svd_method=method, theta_precision=prec)
    m.tensors = [t.clone() for t in ts]
    m.bond_dims = [CHI] * (len(ts) - 1)
    return m


def test_chi1024_sublayer_properties():
    from atlas_q.workloads import gate_schedule, slice_tensors
    sites = 8
    ts = slice_tensors(sites, CHI, "flat", seed=2031, device=DEV, draw_on_device=True)
    gates = [(b, torch.from_numpy(U)) for b, U in gate_schedule(sites, 4, seed=2032)]
    assert [b for b, _ in gates] == [0, 2, 4, 6]
    fast = _mps(ts, "learned", "bf16x3")
    fast.apply_two_site_gates(gates, workers=4)
    torch.cuda.synchronize()
    assert fast.statistics.logs["k_star"] == [CHI] * 4
    assert fast.factor_stats.counts["gram"] == 4          # the grouped exact path
    for q, (b, U) in enumerate(gates[:2]):
        X = K.theta(ts[b].to(torch.complex128), ts[b + 1].to(torch.complex128), _gate4(U))
        Uu, S, Vh = torch.linalg.svd(X, full_matrices=False)
        P64 = (Uu[:, :CHI] * S[:CHI]) @ Vh[:CHI]
        A, B = fast.tensors[b], fast.tensors[b + 1]
        P = A.reshape(-1, CHI).to(torch.complex128) @ B.reshape(CHI, -1).to(torch.complex128)
        F = float((torch.vdot(P64.reshape(-1), P.reshape(-1)).abs() ** 2
                   / (P64.norm() ** 2 * P.norm() ** 2)).item())
        assert 1 - F < 1e-6, (b, F)
        total = float((S ** 2).sum())
        kept = float((S[:CHI] ** 2).sum())
        assert abs(float(P.norm() ** 2) / float(X.norm() ** 2) - kept / total) < 1e-6
        np.testing.assert_allclose(fast.statistics.logs["eps_local"][q],
                                   np.sqrt(total - kept), rtol=1e-3)
        Bm = B.reshape(CHI, -1).to(torch.complex128)
        gram = Bm @ Bm.conj().T
        assert float((gram - torch.eye(CHI, dtype=gram.dtype, device=gram.device)).abs().max()) < 1e-4
        if q == 0:
            plain = _mps(ts, "plain", "native")
            plain.apply_two_site_gate(b, U)
            Pp = (plain.tensors[b].reshape(-1, CHI).to(torch.complex128)
                  @ plain.tensors[b + 1].reshape(CHI, -1).to(torch.complex128))
            Fp = float((torch.vdot(P64.reshape(-1), Pp.reshape(-1)).abs() ** 2
                        / (P64.norm() ** 2 * Pp.norm() ** 2)).item())
            assert abs(F - Fp) <= 1e-6, (F, Fp)


def test_sampler_64q_chi512_shape():
    from atlas_q.workloads import random_right_isometry
    n, cmax, S = 64, 512, 10000
    chi = [1] + [min(cmax, 2 ** min(b + 1, n - b - 1)) for b in range(n - 1)] + [1]
    gen = torch.Generator(device=DEV).manual_seed(77)
    ts = [random_right_isometry(chi[i], chi[i + 1], gen).to(torch.complex64).contiguous()
          for i in range(n)]
    u = torch.rand((S, n), dtype=torch.float64, device=DEV, generator=gen)
    bits, _ = K.sample_bits(ts, u)
    torch.cuda.synchronize()
    # first qubit: right-canonical chain, so P(q0 = s) = ||A_0[:, s, :]||^2
    p1 = float(ts[0][:, 1, :].abs().square().sum() / ts[0].abs().square().sum())
    ones = np.array(K.bits_to_ints(bits)) < 0          # qubit 0 is the sign bit of the int64
    sigma = np.sqrt(p1 * (1 - p1) / S)
    assert abs(ones.mean() - p1) < 5 * sigma + 1e-12
    cs = 64
    tc = [t.cpu() for t in ts]
    uc = u[:cs].cpu().numpy()
    exp, p0 = R.batch_sweep_sample(tc, uc, return_p0=True)
    got = np.array(K.bits_to_ints(bits[:cs]))
    clear = (np.abs(uc - p0) > 1e-5).all(axis=1)
    assert clear.sum() >= cs - 2
    assert np.array_equal(got[clear], np.array(exp)[clear])
