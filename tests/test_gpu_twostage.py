"""GPU checks of the two-stage tridiagonalisation's stages against LAPACK
(numpy float64 on the host).

Stage 1 (aq_he2hb_own, dense -> bandwidth 32): the band's eigenvalues equal
the input's (1e-13 relative to the largest), entries beyond the band are
absent from the band array, and Q1 B Q1^H rebuilt on the host from the
stored reflectors and T factors reproduces the input (1e-13 relative to
max |G|).
"""
import numpy as np
import pytest
import torch

from atlas_q import kernels as K
from atlas_q._ffi import check, lib

pytestmark = pytest.mark.gpu
DEV = "cuda"
PB = 32


def _gram(rng, n, cols):
    X = rng.standard_normal((n, cols)) + 1j * rng.standard_normal((n, cols))
    return X @ X.conj().T


def _stage1(Gs):
    B, n, _ = Gs.shape
    L = lib()
    A = torch.from_numpy(Gs).to(DEV).contiguous()
    # only the lower triangle is read: poison the strict upper triangle
    iu = torch.triu_indices(n, n, 1)
    A[:, iu[0], iu[1]] = complex(float("nan"), 0.0)
    npan = int(L.aq_he2hb_panels(n))
    AB = torch.empty((B, n, 2 * PB), dtype=torch.complex128, device=DEV)
    Tall = torch.zeros((B, max(npan, 1), PB, PB), dtype=torch.complex128, device=DEV)
    nws = int(L.aq_he2hb_workspace_elems(n, B))
    ws = torch.empty(nws, dtype=torch.complex128, device=DEV)
    check(L.aq_he2hb_own(n, K._ptr(A), n * n, K._ptr(AB), K._ptr(Tall), K._ptr(ws), nws, B,
                         K._stream(A)), "aq_he2hb_own")
    torch.cuda.synchronize()
    return A.cpu().numpy(), AB.cpu().numpy(), Tall.cpu().numpy(), npan


def _band_dense(ab):
    n = ab.shape[0]
    Bm = np.zeros((n, n), complex)
    for c in range(n):
        for d in range(min(PB + 1, n - c)):
            Bm[c + d, c] = ab[c, d]
    return np.tril(Bm) + np.tril(Bm, -1).conj().T


def _q1(A, Tall, npan, n):
    Q = np.eye(n, dtype=complex)
    for p in range(npan):
        j = p * PB
        m = n - j - PB
        P = A[j + PB:, j:j + PB]
        V = np.tril(P, -1)[:, :PB].copy()
        for c in range(min(PB, m)):
            V[c, c] = 1.0
        Qp = np.eye(n, dtype=complex)
        Qp[j + PB:, j + PB:] -= V @ Tall[p] @ V.conj().T
        Q = Q @ Qp
    return Q


@pytest.mark.parametrize("n", [34, 40, 65, 130, 512])
def test_band_reduction_matches_lapack(n):
    rng = np.random.default_rng(300 + n)
    Gs = np.stack([_gram(rng, n, n), _gram(rng, n, max(1, n // 3))])
    A, AB, Tall, npan = _stage1(Gs)
    for b in range(2):
        Bm = _band_dense(AB[b])
        assert np.all(AB[b][:, PB + 1:] == 0)
        ref = np.linalg.eigvalsh(Gs[b])
        scale = np.abs(ref).max()
        ew = np.abs(np.linalg.eigvalsh(Bm) - ref).max() / scale
        msg = f"  n={n}[{b}]: eig(B) vs eig(G) {ew:.1e}"
        if n <= 130:
            Q = _q1(A[b], Tall[b], npan, n)
            rec = np.abs(Q @ Bm @ Q.conj().T - Gs[b]).max() / np.abs(Gs[b]).max()
            orth = np.abs(Q.conj().T @ Q - np.eye(n)).max()
            msg += f", |Q B Q^H - G| {rec:.1e}, |Q^H Q - I| {orth:.1e}"
            assert rec < 1e-13 and orth < 1e-13, (rec, orth)
        print(msg, flush=True)
        assert ew < 1e-13


def _stage2(AB, n):
    B = AB.shape[0]
    L = lib()
    ab = torch.from_numpy(AB).to(DEV).contiguous()
    nlog = int(L.aq_hb2st_log_tasks(n))
    Vlog = torch.zeros((B, nlog, PB), dtype=torch.complex128, device=DEV)
    Tlog = torch.zeros((B, nlog), dtype=torch.complex128, device=DEV)
    D = torch.empty((B, n), dtype=torch.float64, device=DEV)
    E = torch.empty((B, max(n - 1, 1)), dtype=torch.float64, device=DEV)
    check(L.aq_hb2st_own(n, K._ptr(ab), K._ptr(Vlog), K._ptr(Tlog), K._ptr(D), K._ptr(E), B,
                         K._stream(ab)), "aq_hb2st_own")
    torch.cuda.synchronize()
    return D.cpu().numpy(), E.cpu().numpy(), Vlog.cpu().numpy(), Tlog.cpu().numpy()


def _q2(Vlog, Tlog, n):
    smax = -(-(n - 1) // PB)
    Q = np.eye(n, dtype=complex)
    for i in range(n - 1):
        r0, s = i + 1, 0
        while r0 < n:
            L = min(PB, n - r0)
            v = Vlog[i * smax + s][:L]
            H = np.eye(n, dtype=complex)
            H[r0:r0 + L, r0:r0 + L] -= Tlog[i * smax + s] * np.outer(v, v.conj())
            Q = Q @ H
            r0 += L
            s += 1
    return Q


@pytest.mark.parametrize("n", [2, 3, 33, 34, 65, 97, 200, 600, 2048])
def test_bulge_chasing_matches_lapack(n):
    """Stage 2 on the band of stage 1: eigenvalues of tridiag(D, E) equal the
    input's (1e-13 relative); for n <= 200 the logged reflectors rebuild the
    band, Q2 T Q2^H = B (1e-13 relative)."""
    rng = np.random.default_rng(400 + n)
    Gs = np.stack([_gram(rng, n, n), _gram(rng, n, max(1, n // 4))])
    if n < PB + 2:
        AB = np.zeros((2, n, 2 * PB), complex)
        for b in range(2):
            for c in range(n):
                for d in range(n - c):
                    AB[b, c, d] = Gs[b][c + d, c]
    else:
        _, AB, _, _ = _stage1(Gs)
    D, E, Vlog, Tlog = _stage2(AB, n)
    for b in range(2):
        T = np.diag(D[b]) + np.diag(E[b][:n - 1], 1) + np.diag(E[b][:n - 1], -1)
        ref = np.linalg.eigvalsh(Gs[b])
        ew = np.abs(np.linalg.eigvalsh(T) - ref).max() / np.abs(ref).max()
        msg = f"  n={n}[{b}]: eig(T) vs eig(G) {ew:.1e}"
        if n <= 200:
            Bm = _band_dense(AB[b])
            Q = _q2(Vlog[b], Tlog[b], n)
            rec = np.abs(Q @ T @ Q.conj().T - Bm).max() / np.abs(Bm).max()
            msg += f", |Q2 T Q2^H - B| {rec:.1e}"
            assert rec < 1e-13, rec
        print(msg, flush=True)
        assert ew < 1e-13


def test_stage_times_order_2048_batch_64():
    """Times of both stages and of the one-stage reduction on the update's
    shape (64 Gram matrices of order 2048; printed), with the tridiagonal
    eigenvalues of both routes against each other (1e-13 relative)."""
    from atlas_q.eig import GramTridiag
    n, B = 2048, 64
    L = lib()
    g = torch.Generator(device=DEV).manual_seed(9)
    X = torch.randn((n, n), dtype=torch.complex128, device=DEV, generator=g)
    G = (X @ X.conj().T).expand(B, n, n).contiguous()
    npan = int(L.aq_he2hb_panels(n))
    AB = torch.empty((B, n, 2 * PB), dtype=torch.complex128, device=DEV)
    Tall = torch.empty((B, npan, PB, PB), dtype=torch.complex128, device=DEV)
    nws = int(L.aq_he2hb_workspace_elems(n, B))
    ws = torch.empty(nws, dtype=torch.complex128, device=DEV)
    nlog = int(L.aq_hb2st_log_tasks(n))
    Vlog = torch.empty((B, nlog, PB), dtype=torch.complex128, device=DEV)
    Tlog = torch.empty((B, nlog), dtype=torch.complex128, device=DEV)
    D = torch.empty((B, n), dtype=torch.float64, device=DEV)
    E = torch.empty((B, n - 1), dtype=torch.float64, device=DEV)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    A = G.clone()
    torch.cuda.synchronize()
    ev[0].record()
    check(L.aq_he2hb_own(n, K._ptr(A), n * n, K._ptr(AB), K._ptr(Tall), K._ptr(ws), nws, B,
                         K._stream(A)), "aq_he2hb_own")
    ev[1].record()
    check(L.aq_hb2st_own(n, K._ptr(AB), K._ptr(Vlog), K._ptr(Tlog), K._ptr(D), K._ptr(E), B,
                         K._stream(A)), "aq_hb2st_own")
    ev[2].record()
    A2 = G.clone()
    ev[3].record()
    tri = GramTridiag(A2)
    ev[4].record()
    torch.cuda.synchronize()
    t1, t2, t0 = (ev[0].elapsed_time(ev[1]) / B, ev[1].elapsed_time(ev[2]) / B,
                  ev[3].elapsed_time(ev[4]) / B)
    print(f"  per matrix: band reduction {t1:.2f} ms, bulge chasing {t2:.2f} ms, "
          f"one-stage reduction {t0:.2f} ms", flush=True)
    T2 = np.diag(D[0].cpu().numpy()) + np.diag(E[0].cpu().numpy(), 1) + \
        np.diag(E[0].cpu().numpy(), -1)
    d1, e1 = tri.D[0].cpu().numpy(), tri.E[0].cpu().numpy()
    T1 = np.diag(d1) + np.diag(e1, 1) + np.diag(e1, -1)
    w1, w2 = np.linalg.eigvalsh(T1), np.linalg.eigvalsh(T2)
    assert np.abs(w1 - w2).max() < 1e-13 * np.abs(w1).max()
