"""Typed wrappers over the HIP entry points, taking torch tensors on a HIP
device.  torch provides device memory and the current stream; every
computation runs in the hand-written kernels of ``libatlasq.so``.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _ffi
from ._ffi import RankResult, check, lib

PREC = {"native": _ffi.AQ_PREC_NATIVE, "fp32": _ffi.AQ_PREC_NATIVE,
        "bf16x3": _ffi.AQ_PREC_BF16X3}


def _code(dtype: torch.dtype) -> int:
    if dtype == torch.complex64:
        return _ffi.AQ_C64
    if dtype == torch.complex128:
        return _ffi.AQ_C128
    raise TypeError(f"atlas_q kernels take complex64/complex128, got {dtype}")


def _real_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return _ffi.AQ_C64
    if dtype == torch.float64:
        return _ffi.AQ_C128
    raise TypeError(f"singular values must be float32/float64, got {dtype}")


def _dev(t: torch.Tensor, name: str) -> None:
    if t.device.type != "cuda":
        raise RuntimeError(f"atlas_q: {name} must live on a HIP device (got {t.device}); "
                           "this build has no CPU path")


def _stream(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _plain(t: torch.Tensor) -> torch.Tensor:
    """Materialise torch's lazy conjugate / negative views (e.g. the Vh that
    torch.linalg.svd returns): the kernels read raw memory."""
    if t.is_conj():
        t = t.resolve_conj()
    if t.is_neg():
        t = t.resolve_neg()
    return t


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    assert not t.is_conj() and not t.is_neg(), "pass resolved tensors to the kernels"
    return ctypes.c_void_p(t.data_ptr())


def gate_to_host(U, shape: Tuple[int, ...]) -> np.ndarray:
    """Gate values as interleaved float64 (re, im), row-major, on the host."""
    if isinstance(U, torch.Tensor):
        arr = U.detach().resolve_conj().resolve_neg().to("cpu").numpy()
    else:
        arr = np.asarray(U)
    arr = np.ascontiguousarray(arr.astype(np.complex128).reshape(shape))
    return arr.reshape(-1).view(np.float64).copy()


def _dptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- theta
def theta(A: torch.Tensor, B: torch.Tensor, gate64: np.ndarray, prec: str = "native",
          out: torch.Tensor = None) -> torch.Tensor:
    """X[2l+p, q*chiR+j] = sum U[2p+q, 2p1+p2] sum_m A[l,p1,m] B[m,p2,j]."""
    _dev(A, "A")
    _dev(B, "B")
    if A.dtype != B.dtype:
        raise TypeError("theta: A and B must share a dtype")
    if A.dim() != 3 or B.dim() != 3 or A.shape[1] != 2 or B.shape[1] != 2 or A.shape[2] != B.shape[0]:
        raise ValueError(f"theta: bad shapes {tuple(A.shape)} x {tuple(B.shape)}")
    A = _plain(A).contiguous()
    B = _plain(B).contiguous()
    cl, cm, cr = A.shape[0], A.shape[2], B.shape[2]
    if out is None:
        out = torch.empty((2 * cl, 2 * cr), dtype=A.dtype, device=A.device)
    assert out.is_contiguous() and tuple(out.shape) == (2 * cl, 2 * cr) and out.dtype == A.dtype
    g = np.ascontiguousarray(gate64, dtype=np.float64)
    assert g.size == 32
    if prec == "bf16x3" and A.dtype == torch.complex64 and THETA_PRESPLIT:
        # operands split once into BF16 planes (workspace from the caching
        # allocator, released in stream order after the launch)
        nbytes = int(lib().aq_theta_planes_bytes(cl, cm, cr))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=A.device)
        check(lib().aq_theta_planes(_ptr(A), _ptr(B), _dptr(g), _ptr(out), cl, cm, cr, _ptr(ws),
                                    nbytes, _stream(A)), "aq_theta_planes")
        return out
    check(lib().aq_theta(_code(A.dtype), PREC[prec], _ptr(A), _ptr(B), _dptr(g), _ptr(out),
                         cl, cm, cr, _stream(A)), "aq_theta")
    return out


THETA_PRESPLIT = True   # BF16x3 theta on operands split once (aq_theta_planes)


# ---------------------------------------------------------------- rank rule
_RANK_BYTES = ctypes.sizeof(RankResult)


def rank_select_async(S: torch.Tensor, eps: float, cap: int, mode: int = _ffi.AQ_RANK_REF
                      ) -> torch.Tensor:
    """Launch the rank rule; returns the device result buffer (uint8[48])."""
    _dev(S, "S")
    S = _plain(S).contiguous()
    buf = torch.empty(_RANK_BYTES, dtype=torch.uint8, device=S.device)
    check(lib().aq_rank_select(_real_code(S.dtype), _ptr(S), S.numel(), float(eps), int(cap),
                               int(mode), _ptr(buf), _stream(S)), "aq_rank_select")
    return buf


def rank_result(buf: torch.Tensor) -> RankResult:
    host = buf.cpu().numpy()  # the one device->host sync of a gate update
    return RankResult.from_buffer_copy(host.tobytes())


def rank_select(S: torch.Tensor, eps: float, cap: int, mode: int = _ffi.AQ_RANK_REF
                ) -> RankResult:
    return rank_result(rank_select_async(S, eps, cap, mode))


# ---------------------------------------------------------------- rebuild
def rebuild(U: torch.Tensor, S: torch.Tensor, Vh: torch.Tensor, k: int, chiL: int, chiR: int):
    """A' = (U[:, :k] * S[:k]).reshape(chiL, 2, k), B' = Vh[:k].reshape(k, 2, chiR)."""
    _dev(U, "U")
    U, Vh = _plain(U), _plain(Vh)
    dt = U.dtype
    rows, cols = 2 * chiL, 2 * chiR
    A = torch.empty((chiL, 2, k), dtype=dt, device=U.device)
    B = torch.empty((k, 2, chiR), dtype=dt, device=U.device)
    S = S.contiguous()
    check(lib().aq_rebuild(_code(dt), _ptr(U), U.stride(0), U.stride(1), _ptr(S), _ptr(Vh),
                           Vh.stride(0), Vh.stride(1), rows, cols, int(k), _ptr(A), _ptr(B),
                           _stream(U)), "aq_rebuild")
    return A, B


# ---------------------------------------------------------------- 1q gate
def apply_1q(T: torch.Tensor, gate64: np.ndarray) -> torch.Tensor:
    """T'[a,t,b] = sum_s U[s,t] T[a,s,b] (AdaptiveMPS convention)."""
    _dev(T, "T")
    T = _plain(T).contiguous()
    out = torch.empty_like(T)
    g = np.ascontiguousarray(gate64, dtype=np.float64)
    check(lib().aq_apply_1q(_code(T.dtype), _ptr(T), _ptr(out), T.shape[0], T.shape[2],
                            _dptr(g), _stream(T)), "aq_apply_1q")
    return out


# ---------------------------------------------------------------- GEMM
def cgemm(A: torch.Tensor, B: torch.Tensor, conj_a: bool = False, conj_b: bool = False,
          prec: str = "native", out: torch.Tensor = None) -> torch.Tensor:
    """C = op(A) @ op(B) for 2-D complex tensors (any strides)."""
    _dev(A, "A")
    if A.dtype != B.dtype or A.dim() != 2 or B.dim() != 2 or A.shape[1] != B.shape[0]:
        raise ValueError(f"cgemm: bad operands {tuple(A.shape)} {A.dtype} x {tuple(B.shape)} {B.dtype}")
    A, B = _plain(A), _plain(B)
    M, K = A.shape
    N = B.shape[1]
    if out is None:
        out = torch.empty((M, N), dtype=A.dtype, device=A.device)
    splits = splitk_for(M, N, K)
    if splits > 1:
        ws = torch.empty(splits * M * N, dtype=A.dtype, device=A.device)
        check(lib().aq_cgemm_splitk(_code(A.dtype), PREC[prec], M, N, K, _ptr(A), A.stride(0),
                                    A.stride(1), int(conj_a), _ptr(B), B.stride(0), B.stride(1),
                                    int(conj_b), _ptr(out), out.stride(0), out.stride(1), splits,
                                    _ptr(ws), ws.numel() * ws.element_size(), _stream(A)),
              "aq_cgemm_splitk")
        return out
    check(lib().aq_cgemm(_code(A.dtype), PREC[prec], M, N, K, _ptr(A), A.stride(0), A.stride(1),
                         int(conj_a), _ptr(B), B.stride(0), B.stride(1), int(conj_b), _ptr(out),
                         out.stride(0), out.stride(1), _stream(A)), "aq_cgemm")
    return out


SPLITK_TILES = 128     # output tiles (64 x 128) below which the reduction is split
SPLITK_TARGET = 512    # work-groups aimed at (two per CU)
SPLITK_MIN_CHUNK = 64  # reduction length per slice


def splitk_for(M: int, N: int, K: int) -> int:
    """Reduction slices for an M x N x K product: 1 unless the output tiles
    cannot fill the GPU and the reduction is long."""
    tiles = -(-M // 64) * -(-N // 128)
    if tiles >= SPLITK_TILES or K < 2 * SPLITK_MIN_CHUNK:
        return 1
    return int(max(1, min(K // SPLITK_MIN_CHUNK, max(2, SPLITK_TARGET // tiles))))


# ---------------------------------------------------------------- sampler
def sample_bits(tensors: Sequence[torch.Tensor], uniforms: torch.Tensor, prec: str = "native",
                want_p0: bool = False, single_shot_rules: bool = False):
    """Run the fused sweep sampler; returns (bits uint64 [S, ceil(n/64)] as int64
    tensor, p0 float64 [S, n] or None)."""
    n = len(tensors)
    dt = tensors[0].dtype
    dev = tensors[0].device
    _dev(tensors[0], "site tensors")
    ts = [_plain(t).to(dt).contiguous() for t in tensors]
    chi = [1] + [int(t.shape[2]) for t in ts]
    for i, t in enumerate(ts):
        if int(t.shape[0]) != chi[i] or int(t.shape[1]) != 2:
            raise ValueError(f"sampler: site {i} has shape {tuple(t.shape)}, expected left bond {chi[i]}")
    if chi[-1] != 1:
        raise ValueError("sampler: last site must have right bond 1")
    S = int(uniforms.shape[0])
    assert uniforms.dtype == torch.float64 and tuple(uniforms.shape) == (S, n)
    uniforms = _plain(uniforms).contiguous()
    chi_arr = (ctypes.c_int64 * (n + 1))(*chi)
    ptrs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in ts])
    ws_bytes = lib().aq_sample_workspace_bytes(_code(dt), n, chi_arr, S)
    if ws_bytes < 0:
        raise RuntimeError("aq_sample_workspace_bytes failed")
    ws = torch.empty(int(ws_bytes), dtype=torch.uint8, device=dev)
    words = (n + 63) // 64
    bits = torch.empty((S, words), dtype=torch.int64, device=dev)
    p0 = torch.empty((S, n), dtype=torch.float64, device=dev) if want_p0 else None
    check(lib().aq_sample(_code(dt), PREC[prec], int(bool(single_shot_rules)), n, ptrs, chi_arr, S, _ptr(uniforms), _ptr(bits),
                          _ptr(p0) if p0 is not None else None, _ptr(ws), int(ws_bytes),
                          _stream(tensors[0])), "aq_sample")
    return bits, p0


def bits_to_ints(bits: torch.Tensor) -> List[int]:
    """The reference's sample ints: word 0 (low 64 bits) as signed int64."""
    return bits[:, 0].cpu().tolist()


# ---------------------------------------------------------------- tn_core
def tn_truncate(S: torch.Tensor, tol: float, adaptive: bool, chi_max, ai_keep: int,
                policy: str, alpha: float, eps: float) -> "_ffi.TnTrunc":
    """Truncation of tools_qih/tn_core.py:233-372 on the device; returns the
    host copy of the result (one device->host transfer)."""
    _dev(S, "S")
    S = _plain(S).contiguous()
    buf = torch.empty(ctypes.sizeof(_ffi.TnTrunc), dtype=torch.uint8, device=S.device)
    pol = _ffi.AQ_TN_POLICY.get(policy, _ffi.AQ_TN_POLICY["guardrail"])
    check(lib().aq_tn_truncate(_real_code(S.dtype), _ptr(S), S.numel(), float(tol),
                               int(bool(adaptive)), -1 if chi_max is None else int(chi_max),
                               int(ai_keep), pol, float(alpha), float(eps), _ptr(buf),
                               _stream(S)), "aq_tn_truncate")
    return _ffi.TnTrunc.from_buffer_copy(buf.cpu().numpy().tobytes())


def rebuild_sqrt(U: torch.Tensor, S: torch.Tensor, Vh: torch.Tensor, k: int, chiL: int,
                 chiR: int):
    """A' = (U[:, :k] * sqrt(S[:k])).reshape(chiL, 2, k),
    B' = (sqrt(S[:k])[:, None] * Vh[:k]).reshape(k, 2, chiR)."""
    _dev(U, "U")
    U, Vh = _plain(U), _plain(Vh)
    dt = U.dtype
    A = torch.empty((chiL, 2, k), dtype=dt, device=U.device)
    B = torch.empty((k, 2, chiR), dtype=dt, device=U.device)
    S = S.contiguous()
    check(lib().aq_rebuild_sqrt(_code(dt), _ptr(U), U.stride(0), U.stride(1), _ptr(S), _ptr(Vh),
                                Vh.stride(0), Vh.stride(1), 2 * chiL, 2 * chiR, int(k), _ptr(A),
                                _ptr(B), _stream(U)), "aq_rebuild_sqrt")
    return A, B


def spectrum_features(S: torch.Tensor, kind: str, limit: int = -1):
    """Predictor inputs from a spectrum on the device: kind "log1p64"
    (RankPredictor, 64 floats over the first min(limit, 64) values; limit < 0
    means #(S > 1e-10)) or "ft135" (FinetunedMLP, 135 floats).
    Returns (features float32 [1, d], n_sv int64 [1]) on the device."""
    _dev(S, "S")
    S = _plain(S).contiguous()
    code, d = {"log1p64": (_ffi.AQ_FEAT_LOG1P64, 64), "ft135": (_ffi.AQ_FEAT_FT135, 135),
               "spec32": (_ffi.AQ_FEAT_SPEC32, 32)}[kind]
    feat = torch.empty((1, d), dtype=torch.float32, device=S.device)
    nsv = torch.empty(1, dtype=torch.int64, device=S.device)
    check(lib().aq_spectrum_features(code, _real_code(S.dtype), _ptr(S), S.numel(), int(limit),
                                     _ptr(feat), _ptr(nsv), _stream(S)), "aq_spectrum_features")
    return feat, nsv


def dense_forward(layers, x: torch.Tensor) -> torch.Tensor:
    """FP32 dense stack; layers: list of dicts with device tensors W [d_out,
    d_in], b, ln_g, ln_b (optional), act name and ln_eps."""
    _dev(x, "x")
    x = x.to(torch.float32).contiguous()
    arr = (_ffi.DenseLayer * len(layers))()
    for l, L in enumerate(layers):
        W = L["W"]
        assert W.dtype == torch.float32 and W.is_contiguous() and W.device == x.device
        arr[l].W = W.data_ptr()
        arr[l].b = L["b"].data_ptr() if L.get("b") is not None else None
        arr[l].ln_g = L["ln_g"].data_ptr() if L.get("ln_g") is not None else None
        arr[l].ln_b = L["ln_b"].data_ptr() if L.get("ln_b") is not None else None
        arr[l].d_out, arr[l].d_in = int(W.shape[0]), int(W.shape[1])
        arr[l].act = _ffi.AQ_ACT[L.get("act", "none")]
        arr[l].ln_eps = float(L.get("ln_eps", 1e-5))
    if x.shape[-1] != arr[0].d_in:
        raise ValueError(f"dense_forward: input width {x.shape[-1]} != {arr[0].d_in}")
    batch = x.numel() // arr[0].d_in
    y = torch.empty((batch, arr[len(layers) - 1].d_out), dtype=torch.float32, device=x.device)
    check(lib().aq_dense_forward(len(layers), ctypes.cast(arr, ctypes.c_void_p), _ptr(x), batch,
                                 _ptr(y), _stream(x)), "aq_dense_forward")
    return y


# ---------------------------------------------------------------- MPO / envs
def mpo_apply_site(W: torch.Tensor, A: torch.Tensor) -> torch.Tensor:
    """M[(l a), s, (r b)] = sum_t W[l, s, t, r] A[a, t, b]."""
    _dev(A, "A")
    W = _plain(W).to(A.dtype).to(A.device).contiguous()
    A = _plain(A).contiguous()
    Dl, Dr = int(W.shape[0]), int(W.shape[3])
    if tuple(W.shape[1:3]) != (2, 2) or A.shape[1] != 2:
        raise ValueError(f"mpo_apply_site: W {tuple(W.shape)}, A {tuple(A.shape)}")
    cl, cr = int(A.shape[0]), int(A.shape[2])
    M = torch.empty((Dl * cl, 2, Dr * cr), dtype=A.dtype, device=A.device)
    check(lib().aq_mpo_apply_site(_code(A.dtype), _ptr(W), Dl, Dr, _ptr(A), cl, cr, _ptr(M),
                                  _stream(A)), "aq_mpo_apply_site")
    return M


def mpo_env_combine(Z: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """E'[R, r, B] = sum_{L,s,t} W[L,s,t,R] Z[L, t, r, s*chiB + B]."""
    _dev(Z, "Z")
    Z = _plain(Z).contiguous()
    W = _plain(W).to(Z.dtype).to(Z.device).contiguous()
    DL, _, chir, two_chiB = Z.shape
    DR = int(W.shape[3])
    chiB = two_chiB // 2
    E = torch.empty((DR, chir, chiB), dtype=Z.dtype, device=Z.device)
    check(lib().aq_mpo_env_combine(_code(Z.dtype), _ptr(Z), _ptr(W), int(DL), DR, int(chir),
                                   int(chiB), _ptr(E), _stream(Z)), "aq_mpo_env_combine")
    return E


def fro_normalize_(X: torch.Tensor, floor: float = 1e-12) -> torch.Tensor:
    """In place: each X[b] (leading index) divided by max(||X[b]||_F, floor)."""
    _dev(X, "X")
    assert X.is_contiguous() and not X.is_conj()
    batch = int(X.shape[0])
    count = X.numel() // max(batch, 1)
    check(lib().aq_fro_normalize_batched(_code(X.dtype), _ptr(X), batch, count, count,
                                         float(floor), _stream(X)), "aq_fro_normalize_batched")
    return X


def herk_tiles(M):
    """Tiles of an AQ_GROUPED_HERK problem of order M (numpy-vectorised)."""
    tm = -(-np.asarray(M, np.int64) // 64)
    q = tm // 2
    return q * (q + 1) + (tm % 2) * (q + 1)


def gemm_desc(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, conj_a=False, conj_b=False,
              kind: int = _ffi.AQ_GROUPED_GEMM, M=None, N=None, K=None, a_ptr=None,
              a_strides=None):
    """One aq_gemm_desc_t (numpy record) for C = op(A) op(B); A may be given
    by pointer and strides (e.g. a reversed row order) with explicit sizes."""
    d = np.zeros(1, _ffi.GEMM_DESC)[0]
    d["A"] = A.data_ptr() if a_ptr is None else a_ptr
    d["B"], d["C"] = B.data_ptr(), C.data_ptr()
    d["M"] = A.shape[0] if M is None else M
    d["K"] = A.shape[1] if K is None else K
    d["N"] = B.shape[1] if N is None else N
    d["a_s0"], d["a_s1"] = A.stride() if a_strides is None else a_strides
    d["b_s0"], d["b_s1"] = B.stride()
    d["c_s0"], d["c_s1"] = C.stride()
    d["conj_a"], d["conj_b"], d["kind"] = int(conj_a), int(conj_b), int(kind)
    return d


def grouped_gemm_descs(descs: np.ndarray, dtype: torch.dtype, device, keep=None,
                       xcd_ranges: bool = False, rows32: bool = False) -> None:
    """Launch prepared descriptors (one dtype, one operand layout) in one
    grouped launch; the uploads are appended to ``keep`` (the caller holds
    them until the stream has run) or synchronised here.  ``rows32``: 32-row
    tiles of 128 threads (an AQ_GROUPED_HERK problem is then a full
    product)."""
    launch = grouped_gemm_prepare(descs, dtype, device, keep, xcd_ranges, rows32)
    if launch is not None:
        launch()


def grouped_gemm_prepare(descs: np.ndarray, dtype: torch.dtype, device, keep=None,
                         xcd_ranges: bool = False, rows32: bool = False):
    """``grouped_gemm_descs`` in two steps: the uploads now (they synchronise
    the current stream), the launch when the returned function is called (on
    the stream current then); None for an empty list."""
    if len(descs) == 0:
        return None
    ka = bool(np.all(descs["a_s1"] == 1))
    nb = bool(np.all(descs["b_s1"] == 1))
    if not ka and np.any(descs["a_s1"] == 1) or not nb and np.any(descs["b_s1"] == 1):
        raise ValueError("grouped_gemm: operands must share their layouts")
    herk = descs["kind"] == _ffi.AQ_GROUPED_HERK
    if rows32:
        paired = descs["kind"] == _ffi.AQ_GROUPED_THETA
        tn = np.where(paired, -(-(descs["N"] // 2) // 64), -(-descs["N"] // 128))
        tiles = (-(-descs["M"] // 32)) * tn
    else:
        tiles = np.where(herk, herk_tiles(descs["M"]),
                         (-(-descs["M"] // 64)) * (-(-descs["N"] // 128)))
    pre = np.zeros(len(tiles), np.int64)
    pre[1:] = np.cumsum(tiles[:-1])
    total = int(tiles.sum())
    dd = torch.from_numpy(np.ascontiguousarray(descs).view(np.uint8).copy()).to(device)
    pd = torch.from_numpy(pre).to(device)
    grid = int(min(total, 65536))
    if xcd_ranges and grid >= 8:
        grid -= grid % 8          # XCD-contiguous tile ranges need a multiple of 8
    no_theta = not bool(np.any(descs["kind"] == _ffi.AQ_GROUPED_THETA))
    layout = (int(ka) | (int(nb) << 1) | (int(no_theta) << 2) | (int(bool(xcd_ranges)) << 3)
              | (_ffi.AQ_GROUPED_LAYOUT_ROWS32 if rows32 else 0))
    if keep is not None:
        keep.extend([dd, pd])

    def launch():
        st = torch.cuda.current_stream(device)
        check(lib().aq_grouped_gemm(_code(dtype), ctypes.c_void_p(dd.data_ptr()),
                                    ctypes.c_void_p(pd.data_ptr()), len(descs), total, layout,
                                    grid, ctypes.c_void_p(st.cuda_stream)), "aq_grouped_gemm")
        if keep is None:
            st.synchronize()
    return launch


def grouped_gemm(problems, dtype: torch.dtype, device) -> None:
    """C_p = op(A_p) op(B_p) for a list of (A, B, C, conj_a, conj_b) 2-D
    tensor views in one grouped launch (all of one dtype; every A shares its
    fast axis, as every B does)."""
    if not problems:
        return
    descs = np.stack([gemm_desc(A, B, C, ca, cb) for A, B, C, ca, cb in problems])
    for A, B, C, _, _ in problems:
        assert not A.is_conj() and not B.is_conj() and C.dtype == dtype
    grouped_gemm_descs(descs, dtype, device)




QR_MAX_ROWS = 4096   # rows of a factored matrix (the panel kernel keeps one column in LDS)


def qr(X: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """Reduced QR of a complex matrix [m, n] or batch [b, m, n] on the device
    by the library's blocked Householder factorisation (aq_geqrf + aq_ungqr,
    LAPACK's zgeqrf / zungqr): Q [.., m, k], R [.., k, n], k = min(m, n), in
    X's dtype.  The factorisation runs in complex128 (complex64 input is
    promoted and the factors rounded back); as LAPACK's, R has a real
    diagonal of either sign."""
    _dev(X, "X")
    if X.dtype not in (torch.complex64, torch.complex128) or X.dim() not in (2, 3):
        raise ValueError(f"qr: need a complex [m, n] or [b, m, n] tensor, got {tuple(X.shape)} "
                         f"{X.dtype}")
    batched = X.dim() == 3
    A = (X if batched else X.unsqueeze(0)).to(torch.complex128).contiguous().clone()
    b, m, n = A.shape
    if m > QR_MAX_ROWS:
        raise ValueError(f"qr: {m} rows exceed the factorisation's limit of {QR_MAX_ROWS}")
    k = min(m, n)
    L = lib()
    npan = int(L.aq_geqrf_panels(m, n))
    T = torch.empty((b, npan, 32, 32), dtype=torch.complex128, device=A.device)
    nws = int(L.aq_geqrf_workspace_elems(m, n, b))
    ws = torch.empty(nws, dtype=torch.complex128, device=A.device)
    st = _stream(A)
    check(L.aq_geqrf(m, n, _ptr(A), m * n, _ptr(T), _ptr(ws), nws, b, st), "aq_geqrf")
    Q = torch.empty((b, m, k), dtype=torch.complex128, device=A.device)
    check(L.aq_ungqr(m, n, _ptr(A), m * n, _ptr(T), _ptr(Q), m * k, _ptr(ws), nws, b, st),
          "aq_ungqr")
    R = torch.triu(A[:, :k, :])
    Q, R = Q.to(X.dtype), R.to(X.dtype)
    return (Q, R) if batched else (Q[0], R[0])
