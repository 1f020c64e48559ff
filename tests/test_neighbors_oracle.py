"""CPU tests: the oracle of the callers next to the two-site path
(oracle/neighbors_ref.py) pinned to dense closed forms."""
import numpy as np
import torch

from oracle import neighbors_ref as N
from oracle import tn_core_ref as T


def _rand_mps(n, chi, rng, dtype=torch.complex128):
    dims = [1] + [min(chi, 2 ** min(i + 1, n - i - 1)) for i in range(n - 1)] + [1]
    return [torch.from_numpy(rng.normal(size=(dims[i], 2, dims[i + 1]))
                             + 1j * rng.normal(size=(dims[i], 2, dims[i + 1]))).to(dtype)
            for i in range(n)]


def _rand_mpo(n, D, rng):
    dims = [1] + [D] * (n - 1) + [1]
    return [torch.from_numpy(rng.normal(size=(dims[i], 2, 2, dims[i + 1]))
                             + 1j * rng.normal(size=(dims[i], 2, 2, dims[i + 1])))
            for i in range(n)]


def _dense_mpo(W):
    """O[s, t] (rows: output bits s, columns: input bits t; qubit 0 most
    significant) of an MPO with tensors [l, s, t, r]."""
    O = W[0][0]                                    # [s, t, r]
    for w in W[1:]:
        O = torch.einsum("STl,lstr->SsTtr", O, w)
        S2, s2, T2, t2, r = O.shape
        O = O.reshape(S2 * s2, T2 * t2, r)
    return O[..., 0]


def test_apply_mpo_matches_dense_operator():
    rng = np.random.default_rng(0)
    n = 5
    A = _rand_mps(n, 4, rng)
    W = _rand_mpo(n, 3, rng)
    M = N.mpo_apply_sites(W, A)
    psi = torch.from_numpy(T.contract(A))
    phi = torch.from_numpy(T.contract(M))
    O = _dense_mpo(W)
    assert torch.allclose(phi, O @ psi, atol=1e-10)


def test_expectation_value_matches_dense_closed_form():
    rng = np.random.default_rng(1)
    n = 5
    A = _rand_mps(n, 4, rng)
    W = _rand_mpo(n, 2, rng)
    psi = torch.from_numpy(T.contract(A))
    O = _dense_mpo(W)
    # the reference's index order pairs conj(psi) with the operator's second
    # index: psi^H O^T psi / psi^H psi
    want = (psi.conj() @ (O.T @ psi)) / (psi.conj() @ psi)
    got = N.expectation_value(W, A)
    assert abs(got - complex(want)) < 1e-10 * max(1.0, abs(complex(want)))


def test_env_scans_match_explicit_products():
    rng = np.random.default_rng(2)
    B, L, chi = 2, 7, 3
    E = torch.from_numpy(rng.normal(size=(B, L, chi, chi)) + 1j * rng.normal(size=(B, L, chi, chi)))
    eye = torch.eye(chi, dtype=E.dtype)
    Le = N.build_left_envs(E, eye, tile_size=3)
    Re = N.build_right_envs(E, eye, tile_size=3)
    for b in range(B):
        P = eye.clone()
        for k in range(L):
            P = P @ E[b, k]
            if (k + 1) % 3 == 0:
                P = P / torch.linalg.norm(P)
            assert torch.allclose(Le[b, k + 1], P, atol=1e-10)
        P = eye.clone()
        for k in range(L - 1, -1, -1):
            P = E[b, k] @ P
            if k % 3 == 0:
                P = P / torch.linalg.norm(P)
            assert torch.allclose(Re[b, k], P, atol=1e-10)


def test_tdvp_split_threshold_is_one_minus_eps():
    """A diagonal theta with singular values 1, .5, .1, .01: the TDVP rule's
    threshold is (1 - eps) total, so eps = 0.01 keeps 2 values (the gate
    update's (1 - eps^2) would keep 4)."""
    s = torch.tensor([1.0, 0.5, 0.1, 0.01], dtype=torch.float64)
    X = torch.diag(s).to(torch.complex128)             # [4, 4] = theta [2, 2, 2, 2]
    A, B, k = N.tdvp_two_site_split(X.reshape(2, 2, 2, 2), 10, 1e-2)
    # S^2 = 1, .25, .01, 1e-4: total 1.2601, 0.99 total = 1.247499;
    # cumsum 1, 1.25, ... -> first index >= threshold is 1 -> k = 2
    assert k == 2
    assert torch.allclose(A.reshape(4, k) @ B.reshape(k, 4),
                          torch.diag(torch.tensor([1.0, 0.5, 0, 0], dtype=torch.complex128)),
                          atol=1e-12)
    assert N.tdvp_two_site_split(X.reshape(2, 2, 2, 2), 1, 1e-2)[2] == 1
