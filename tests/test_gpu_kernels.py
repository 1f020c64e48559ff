"""GPU parity of the individual HIP kernels against the CPU oracle.

Tolerances: complex64 results within 1e-5 of the largest magnitude (FP32
MFMA), complex128 within 1e-12; BF16x3 within 5e-5 (hi/lo bf16 split, about
2^-17 relative per product); rank rule: k exact, diagnostics to 1e-6.
"""
import numpy as np
import pytest
import torch

from oracle import ref_mps as R
from atlas_q import kernels as K
from atlas_q.circuits import haar_su4

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _rand(shape, dtype, gen):
    real = torch.float32 if dtype == torch.complex64 else torch.float64
    return torch.complex(torch.randn(*shape, generator=gen, dtype=real),
                         torch.randn(*shape, generator=gen, dtype=real))


def _relerr(a, b):
    a = a.detach().cpu().to(torch.complex128)
    b = b.detach().cpu().to(torch.complex128)
    return float((a - b).abs().max() / max(b.abs().max(), 1e-300))


TOL = {torch.complex64: 1e-5, torch.complex128: 1e-12}


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 4, 5), (16, 32, 24), (33, 65, 70),
                                   (64, 64, 64), (130, 100, 7)])
def test_theta_matches_oracle(dtype, shape):
    cl, cm, cr = shape
    print(f"theta {dtype} {shape}", flush=True)
    g = torch.Generator().manual_seed(cl * 1000 + cm * 10 + cr)
    A, B = _rand((cl, 2, cm), dtype, g), _rand((cm, 2, cr), dtype, g)
    U = haar_su4(np.random.default_rng(cm))
    ref = R.theta(A.to(torch.complex128), B.to(torch.complex128), torch.from_numpy(U))
    X = K.theta(A.to(DEV), B.to(DEV), K.gate_to_host(U, (4, 4)))
    assert X.shape == (2 * cl, 2 * cr)
    assert _relerr(X, ref) < TOL[dtype]


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 4, 5), (33, 65, 70), (130, 100, 7),
                                   (64, 96, 64), (256, 200, 300)])
def test_theta_presplit_planes_equal_inline_split(shape):
    """aq_theta_planes (operands split once into BF16 planes, zero-padded to
    whole tiles) returns exactly the bits of aq_theta's BF16x3 path (split
    while staging each tile): the same splits, products and accumulation
    order, the padding adding exact zeros."""
    import ctypes
    from atlas_q._ffi import AQ_C64, AQ_PREC_BF16X3, check, lib
    cl, cm, cr = shape
    g = torch.Generator().manual_seed(cl + 7 * cm + 13 * cr)
    A = _rand((cl, 2, cm), torch.complex64, g).to(DEV)
    B = _rand((cm, 2, cr), torch.complex64, g).to(DEV)
    U = haar_su4(np.random.default_rng(cr))
    gh = np.ascontiguousarray(K.gate_to_host(U, (4, 4)), np.float64)
    X1 = torch.empty((2 * cl, 2 * cr), dtype=torch.complex64, device=DEV)
    check(lib().aq_theta(AQ_C64, AQ_PREC_BF16X3, K._ptr(A), K._ptr(B),
                         gh.ctypes.data_as(ctypes.c_void_p), K._ptr(X1), cl, cm, cr,
                         K._stream(A)), "aq_theta")
    X2 = K.theta(A, B, gh, prec="bf16x3")
    assert K.THETA_PRESPLIT
    assert torch.equal(X1, X2), float((X1 - X2).abs().max())


@pytest.mark.parametrize("shape", [(16, 32, 24), (64, 96, 64), (128, 128, 128)])
def test_theta_bf16x3_error_bound(shape):
    cl, cm, cr = shape
    g = torch.Generator().manual_seed(7)
    A, B = _rand((cl, 2, cm), torch.complex64, g), _rand((cm, 2, cr), torch.complex64, g)
    U = haar_su4(np.random.default_rng(1))
    ref = R.theta(A.to(torch.complex128), B.to(torch.complex128), torch.from_numpy(U))
    X = K.theta(A.to(DEV), B.to(DEV), K.gate_to_host(U, (4, 4)), prec="bf16x3")
    err = _relerr(X, ref)
    print(f"bf16x3 theta {shape}: max rel err {err:.2e}")
    assert err < 5e-5


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (64, 128, 16), (67, 130, 33), (300, 257, 129)])
@pytest.mark.parametrize("layout", ["nn", "tn", "nt", "conj"])
def test_cgemm_matches_matmul(dtype, mnk, layout):
    M, N, K_ = mnk
    g = torch.Generator().manual_seed(M + N + K_)
    A = _rand((M, K_), dtype, g)
    B = _rand((K_, N), dtype, g)
    Ad, Bd = A.to(DEV), B.to(DEV)
    ca = cb = False
    if layout == "tn":
        Ad = A.T.contiguous().to(DEV).T  # column-major A
    elif layout == "nt":
        Bd = B.T.contiguous().to(DEV).T
    elif layout == "conj":
        ca = cb = True
    C = K.cgemm(Ad, Bd, conj_a=ca, conj_b=cb)
    A64, B64 = A.to(torch.complex128), B.to(torch.complex128)
    if ca:
        A64, B64 = A64.conj(), B64.conj()
    assert _relerr(C, A64 @ B64) < 10 * TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("mnk", [(16, 16, 2048), (2048, 16, 2048), (250, 250, 2048),
                                 (250, 2048, 1999), (5, 3, 700)])
@pytest.mark.parametrize("layout", ["nn", "tn", "conj"])
def test_cgemm_splitk_matches_matmul(dtype, mnk, layout):
    """Skinny outputs over long reductions take the split-K path; the slices
    are added in a fixed order, so two calls agree bit for bit."""
    M, N, K_ = mnk
    assert K.splitk_for(M, N, K_) > 1
    g = torch.Generator().manual_seed(M * 7 + N + K_)
    A = _rand((M, K_), dtype, g)
    B = _rand((K_, N), dtype, g)
    Ad, Bd = A.to(DEV), B.to(DEV)
    ca = cb = layout == "conj"
    if layout == "tn":
        Ad = A.T.contiguous().to(DEV).T
    C1 = K.cgemm(Ad, Bd, conj_a=ca, conj_b=cb)
    C2 = K.cgemm(Ad, Bd, conj_a=ca, conj_b=cb)
    A64, B64 = A.to(torch.complex128), B.to(torch.complex128)
    if ca:
        A64, B64 = A64.conj(), B64.conj()
    assert _relerr(C1, A64 @ B64) < 10 * TOL[dtype]
    assert torch.equal(C1, C2)


def test_cgemm_splitk_bf16x3_error_bound():
    g = torch.Generator().manual_seed(5)
    A = _rand((2048, 2048), torch.complex64, g)
    B = _rand((2048, 24), torch.complex64, g)
    C = K.cgemm(A.to(DEV), B.to(DEV), prec="bf16x3")
    ref = A.to(torch.complex128) @ B.to(torch.complex128)
    assert float((C.cpu().to(torch.complex128) - ref).norm() / ref.norm()) < 2e-5


def test_cgemm_bf16x3_error_bound():
    g = torch.Generator().manual_seed(3)
    A = _rand((256, 512), torch.complex64, g)
    B = _rand((512, 200), torch.complex64, g)
    C = K.cgemm(A.to(DEV), B.to(DEV), prec="bf16x3")
    ref = A.to(torch.complex128) @ B.to(torch.complex128)
    fro = float((C.cpu().to(torch.complex128) - ref).norm() / ref.norm())
    assert fro < 2e-5


def _spectra():
    rng = np.random.default_rng(11)
    out = [np.array([1.0, 0.5, 0.1, 0.01, 0.001]), np.ones(4), np.array([1.0, 1e-4]),
           np.array([1.0]), np.array([1.0, 1e-6])]
    for m in (7, 100, 513, 2048, 3000):
        s = np.sort(rng.exponential(1.0, m) ** 3)[::-1]
        out.append(s / s[0])
        out.append(np.sort(np.exp(-rng.uniform(0, 30, m)))[::-1])
    return out


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_rank_rule_matches_oracle(dt):
    n = 0
    for S in _spectra():
        St = torch.tensor(np.ascontiguousarray(S), dtype=dt)
        for eps in (1e-1, 1e-2, 1e-5, 1e-8, 1e-10):
            for cap in (1, 3, len(S) // 2 + 1, 10 ** 6):
                ref = R.choose_rank_from_sigma(St, eps, cap)
                r = K.rank_select(St.to(DEV), eps, cap)
                assert r.k == ref[0], (len(S), eps, cap, r.k, ref[0])
                np.testing.assert_allclose(r.eps_local, ref[1], rtol=1e-6, atol=1e-30)
                np.testing.assert_allclose(r.entropy, ref[2], rtol=1e-5, atol=1e-6)
                if np.isfinite(ref[3]):
                    np.testing.assert_allclose(r.condS, ref[3], rtol=1e-6)
                else:
                    assert not np.isfinite(r.condS)
                n += 1
    assert n > 100


def test_rank_rule_degenerate_total():
    r = K.rank_select(torch.tensor([1e-16, 0.0], dtype=torch.float64, device=DEV), 1e-3, 4)
    assert r.k == 1 and r.flags == 1 and r.eps_local == 0.0


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_rebuild_matches_torch(dtype):
    g = torch.Generator().manual_seed(5)
    X = _rand((40, 56), dtype, g).to(DEV)
    U, S, Vh = torch.linalg.svd(X, full_matrices=False)
    k = 17
    A, B = K.rebuild(U, S, Vh, k, 20, 28)
    refA = (U[:, :k] * S[:k]).reshape(20, 2, k)
    refB = Vh[:k].reshape(k, 2, 28)
    assert torch.equal(A, refA.contiguous()) or _relerr(A, refA) < TOL[dtype]
    assert torch.equal(B, refB.contiguous())


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_apply_1q_transposed_convention(dtype):
    g = torch.Generator().manual_seed(9)
    T = _rand((5, 2, 7), dtype, g)
    U = torch.from_numpy(haar_su4(np.random.default_rng(2))[:2, :2]).to(dtype)
    ref = torch.einsum("st,asb->atb", U, T)
    out = K.apply_1q(T.to(DEV), K.gate_to_host(U, (2, 2)))
    assert _relerr(out, ref) < TOL[dtype]


@pytest.mark.parametrize("rows32", [False, True])
@pytest.mark.parametrize("xcd", [False, True])
@pytest.mark.parametrize("M,N", [(128, 256), (100, 73), (200, 200), (130, 520)])
def test_grouped_herk_lower_triangle(M, N, xcd, rows32):
    """AQ_GROUPED_HERK computes every element with row >= col of X X^H (the
    triangle the eigensolver reads); problems of different orders in one
    launch; with 32-row tiles as a full product."""
    from atlas_q._ffi import AQ_GROUPED_HERK
    torch.manual_seed(M + N)
    Xs = [torch.randn(M, N, dtype=torch.complex128, device=DEV),
          torch.randn(M // 2 + 1, N, dtype=torch.complex128, device=DEV)]
    Gs = [torch.zeros(x.shape[0], x.shape[0], dtype=torch.complex128, device=DEV) for x in Xs]
    descs = np.stack([K.gemm_desc(x, x.T, g, conj_b=True, kind=AQ_GROUPED_HERK)
                      for x, g in zip(Xs, Gs)])
    K.grouped_gemm_descs(descs, torch.complex128, DEV, xcd_ranges=xcd, rows32=rows32)
    for x, g in zip(Xs, Gs):
        ref = x @ x.conj().T
        low = torch.tril(torch.ones_like(ref.real, dtype=torch.bool))
        assert torch.allclose(g[low], ref[low], atol=1e-10 * N)


@pytest.mark.parametrize("rows32", [False, True])
@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("b_layout", ["row", "col"])
def test_grouped_gemm_mixed_problems(dtype, b_layout, rows32):
    """One grouped launch over problems of unrelated shapes (ragged M, N, K,
    K below one k tile, K = 1), per-problem conjugation flags, and A read with
    a negative row stride (rows reversed, as the projection reads the
    eigensolver's output): every element matches the float64 product, with
    either tile height."""
    shapes = [(1, 1, 1), (67, 3, 130), (130, 33, 257), (64, 16, 128), (5, 200, 9)]
    flags = [(False, False), (True, False), (False, True), (True, True), (True, False)]
    g = torch.Generator().manual_seed(11 + (b_layout == "col"))
    descs, refs, outs, keep = [], [], [], []
    for (M, K_, N), (ca, cb) in zip(shapes, flags):
        A = _rand((M, K_), dtype, g).to(DEV)
        Bm = _rand((K_, N), dtype, g)
        if b_layout == "col":                  # column-major, column stride K + 1
            B = torch.zeros(N * (K_ + 1), dtype=dtype, device=DEV).as_strided((K_, N), (1, K_ + 1))
            B.copy_(Bm.to(DEV))
        else:
            B = Bm.to(DEV)
        C = torch.full((M, N), float("nan"), dtype=dtype, device=DEV)
        Af = A.flip(0).contiguous()            # A = rows of Af read bottom-up
        descs.append(K.gemm_desc(Af, B, C, conj_a=ca, conj_b=cb, M=M, K=K_, N=N,
                                 a_ptr=Af[M - 1].data_ptr(), a_strides=(-K_, 1)))
        A64, B64 = A.cpu().to(torch.complex128), Bm.to(torch.complex128)
        refs.append((A64.conj() if ca else A64) @ (B64.conj() if cb else B64))
        outs.append(C)
        keep += [A, Af, B]
    K.grouped_gemm_descs(np.stack(descs), dtype, DEV, rows32=rows32)
    for C, ref in zip(outs, refs):
        assert _relerr(C, ref) < 10 * TOL[dtype]


def test_heevd_reads_only_the_lower_triangle():
    """The row-major lower triangle is all aq_heevd_batched reads (column-major
    upper): the other triangle may hold anything."""
    from atlas_q.eig import heevd_batched
    rng = np.random.default_rng(3)
    Z = rng.standard_normal((2, 96, 96)) + 1j * rng.standard_normal((2, 96, 96))
    G = Z @ np.conj(np.transpose(Z, (0, 2, 1)))
    Gl = np.tril(G) + np.triu(np.full_like(G, 7.0 + 3.0j), 1)
    lam, V = heevd_batched(torch.from_numpy(Gl).to(DEV))
    ref = np.linalg.eigvalsh(G)[:, ::-1]
    assert np.abs(lam.cpu().numpy() - ref).max() < 1e-10 * np.abs(ref).max()
