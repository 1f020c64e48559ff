"""GPU parity of the callers next to the two-site path (MPO application and
expectation value, environment scans, TDVP split) against the oracle
(oracle/neighbors_ref.py)."""
import numpy as np
import pytest
import torch

from oracle import neighbors_ref as N
from oracle import tn_core_ref as T

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _rand_mps(n, chi, rng, dtype):
    dims = [1] + [min(chi, 2 ** min(i + 1, n - i - 1)) for i in range(n - 1)] + [1]
    return [torch.from_numpy(rng.normal(size=(dims[i], 2, dims[i + 1]))
                             + 1j * rng.normal(size=(dims[i], 2, dims[i + 1]))).to(dtype)
            for i in range(n)]


def _rand_mpo(n, D, rng):
    dims = [1] + [D] * (n - 1) + [1]
    return [torch.from_numpy(rng.normal(size=(dims[i], 2, 2, dims[i + 1]))
                             + 1j * rng.normal(size=(dims[i], 2, 2, dims[i + 1])))
            for i in range(n)]


def _on_device(mps_tensors, n, dtype):
    from atlas_q.adaptive_mps import AdaptiveMPS, DTypePolicy
    m = AdaptiveMPS(n, bond_dim=1, device=DEV, dtype_policy=DTypePolicy(default=dtype))
    m.tensors = [t.to(DEV) for t in mps_tensors]
    m.bond_dims = [t.shape[2] for t in mps_tensors[:-1]]
    return m


@pytest.mark.parametrize("dtype,tol", [(torch.complex64, 2e-5), (torch.complex128, 1e-11)])
def test_apply_mpo_to_mps_matches_oracle(dtype, tol):
    from atlas_q import kernels as K
    from atlas_q.mpo_ops import MPO, apply_mpo_to_mps
    rng = np.random.default_rng(3)
    n = 8
    A = _rand_mps(n, 16, rng, dtype)
    W = _rand_mpo(n, 3, rng)
    for w, a in zip(W, A):
        got = K.mpo_apply_site(w.to(DEV), a.to(DEV)).cpu()
        want = N.mpo_apply_sites([w], [a])[0]
        assert torch.allclose(got, want, atol=tol * float(want.abs().max()))
    res = apply_mpo_to_mps(MPO(W), _on_device(A, n, dtype))
    phi = T.contract([t.cpu() for t in res.tensors])
    want = T.contract(N.mpo_apply_sites(W, A))
    assert np.abs(phi - want).max() < 10 * tol * np.abs(want).max()


@pytest.mark.parametrize("dtype,tol", [(torch.complex64, 1e-4), (torch.complex128, 1e-10)])
def test_expectation_value_matches_oracle(dtype, tol):
    from atlas_q.mpo_ops import MPO, expectation_value
    rng = np.random.default_rng(4)
    for n, chi, D in ((6, 8, 2), (10, 64, 3)):
        A = _rand_mps(n, chi, rng, dtype)
        W = _rand_mpo(n, D, rng)
        got = expectation_value(MPO(W), _on_device(A, n, dtype))
        want = N.expectation_value(W, A)
        assert abs(got - want) < tol * max(1.0, abs(want)), (n, chi, got, want)
    # dense shortcut (mpo_ops.py:1107-1125)
    n = 5
    A = _rand_mps(n, 4, rng, dtype)
    H = torch.from_numpy(rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32)))
    H = (H + H.conj().T) / 2
    mpo = MPO(_rand_mpo(n, 2, rng))
    mpo.full_matrix = H
    got = expectation_value(mpo, _on_device(A, n, dtype))
    psi = torch.from_numpy(T.contract(A))
    psi = psi / torch.linalg.norm(psi)
    want = complex(psi.conj() @ (H @ psi))
    assert abs(got - want) < tol * max(1.0, abs(want))


@pytest.mark.parametrize("dtype,tol", [(torch.complex64, 2e-5), (torch.complex128, 1e-12)])
def test_env_scans_match_oracle(dtype, tol):
    from atlas_q.sim_env_scan import build_left_envs, build_right_envs
    rng = np.random.default_rng(5)
    for B, L, chi, tile in ((3, 9, 5, 4), (2, 6, 96, 3)):
        E = torch.from_numpy(rng.normal(size=(B, L, chi, chi))
                             + 1j * rng.normal(size=(B, L, chi, chi))).to(dtype) / chi ** 0.5
        eye = torch.eye(chi, dtype=dtype)
        for fn, ref in ((build_left_envs, N.build_left_envs),
                        (build_right_envs, N.build_right_envs)):
            got = fn(E.to(DEV), eye.to(DEV), tile_size=tile).cpu()
            want = ref(E, eye, tile_size=tile)
            scale = want.abs().amax(dim=(-2, -1), keepdim=True)
            assert ((got - want).abs() / scale).max() < tol * chi, (fn.__name__, B, L, chi)


def test_tdvp_split_matches_oracle():
    from atlas_q.tdvp import evolve_split_into, two_site_split
    rng = np.random.default_rng(6)
    for dtype, tol in ((torch.complex64, 2e-5), (torch.complex128, 1e-11)):
        for chiL, chiR, chi_max, eps in ((8, 8, 6, 1e-3), (32, 16, 64, 1e-2), (16, 16, 40, 1e-8)):
            s = np.exp(-np.arange(min(2 * chiL, 2 * chiR)) / 3.0)
            U, _ = np.linalg.qr(rng.normal(size=(2 * chiL, len(s))) + 1j * rng.normal(size=(2 * chiL, len(s))))
            V, _ = np.linalg.qr(rng.normal(size=(2 * chiR, len(s))) + 1j * rng.normal(size=(2 * chiR, len(s))))
            th = torch.from_numpy((U * s) @ V.conj().T).to(dtype).reshape(chiL, 2, 2, chiR)
            A, B, k = two_site_split(th.to(DEV), chi_max, eps)
            Ar, Br, kr = N.tdvp_two_site_split(th, chi_max, eps)
            assert k == kr
            P = (A.reshape(-1, k) @ B.reshape(k, -1)).cpu()
            Pr = Ar.reshape(-1, kr) @ Br.reshape(kr, -1)
            assert torch.allclose(P, Pr, atol=tol * float(Pr.abs().max()) * 10)
    m = _on_device(_rand_mps(4, 4, rng, torch.complex64), 4, torch.complex64)
    th = torch.einsum("abc,cde->abde", m.tensors[1], m.tensors[2])
    k = evolve_split_into(m, 1, th, 8, 1e-6)
    assert m.bond_dims[1] == k == m.tensors[1].shape[2] == m.tensors[2].shape[0]
