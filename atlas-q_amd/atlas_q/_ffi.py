"""ctypes binding of the C ABI in ``include/atlasq.h``.

The shared library ``libatlasq.so`` sits next to this file (built in-tree by
``atlas-q_amd/build.py``).  Loading it needs only the ROCm runtime, not a GPU,
so symbol checks run anywhere; every compute call needs a HIP device and fails
loudly otherwise -- there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

LIB_PATH = Path(__file__).with_name("libatlasq.so")
HEADER = Path(__file__).resolve().parents[2] / "include" / "atlasq.h"

AQ_OK = 0
AQ_C64, AQ_C128 = 0, 1
AQ_PREC_NATIVE, AQ_PREC_BF16X3 = 0, 1
AQ_RANK_REF = 0
AQ_RANK_TDVP = 1
AQ_SAMPLE_SCRATCH_BYTES = 64 << 20

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_vp = ctypes.c_void_p
c_dp = ctypes.POINTER(ctypes.c_double)


import numpy as _np

GEMM_DESC = _np.dtype([("A", "u8"), ("B", "u8"), ("C", "u8"), ("M", "i8"), ("N", "i8"),
                       ("K", "i8"), ("a_s0", "i8"), ("a_s1", "i8"), ("b_s0", "i8"),
                       ("b_s1", "i8"), ("c_s0", "i8"), ("c_s1", "i8"), ("conj_a", "i4"),
                       ("conj_b", "i4"), ("kind", "i4"), ("pad", "i4"), ("gate", "f8", (32,))])
SITE1Q_DESC = _np.dtype([("T_in", "u8"), ("T_out", "u8"), ("chiL", "i8"), ("chiR", "i8"),
                         ("gate", "f8", (8,))])
SAMPLE_DESC = _np.dtype([("L", "u8"), ("off", "u8"), ("site", "u8"), ("Tout", "u8"),
                         ("part", "u8"), ("uniforms", "u8"), ("bits", "u8"),
                         ("off_next", "u8"), ("ldL", "i8"), ("S", "i8"), ("chiL", "i8"),
                         ("chiR", "i8"), ("n", "i8"), ("site_index", "i8"), ("words", "i8")])
COPY_DESC = _np.dtype([("src", "u8"), ("dst", "u8"), ("row_scale", "u8"), ("col_scale", "u8"),
                       ("rows", "i8"), ("cols", "i8"), ("s0", "i8"), ("s1", "i8"),
                       ("conj", "i4"), ("pad", "i4")])
AQ_GROUPED_GEMM, AQ_GROUPED_THETA, AQ_GROUPED_HERK = 0, 1, 2
AQ_GROUPED_LAYOUT_ROWS32 = 16   # aq_grouped_gemm layout bit: 32-row tiles of 128 threads
LAYER_GATE = _np.dtype([("siteA", "u8"), ("siteB", "u8"), ("X", "u8"), ("X64", "u8"), ("G", "u8"),
                        ("V", "u8"), ("S64", "u8"), ("Sinv", "u8"), ("P", "u8"), ("dims", "u8"),
                        ("bond", "i8"), ("sz", "i8"), ("left", "i8"), ("vfmt", "i8"),
                        ("kslot", "i8"), ("kbound", "i8"), ("gate", "f8", (32,))])


class RankResult(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("eps_local", ctypes.c_double), ("entropy", ctypes.c_double),
                ("condS", ctypes.c_double), ("total_weight", ctypes.c_double),
                ("kept_weight", ctypes.c_double)]


class TnTrunc(ctypes.Structure):
    _fields_ = [("keep_count", ctypes.c_int64), ("chi", ctypes.c_int64),
                ("adaptive_keep", ctypes.c_int64), ("prefix", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("total_weight", ctypes.c_double),
                ("kept_weight", ctypes.c_double), ("truncation_error", ctypes.c_double)]


class DenseLayer(ctypes.Structure):
    _fields_ = [("W", ctypes.c_void_p), ("b", ctypes.c_void_p), ("ln_g", ctypes.c_void_p),
                ("ln_b", ctypes.c_void_p), ("d_in", ctypes.c_int32), ("d_out", ctypes.c_int32),
                ("act", ctypes.c_int32), ("pad", ctypes.c_int32), ("ln_eps", ctypes.c_double)]


AQ_TN_POLICY = {"baseline": 0, "guardrail": 1, "blend": 2, "clamp": 3, "clamp+blend": 4}
AQ_FEAT_LOG1P64, AQ_FEAT_FT135, AQ_FEAT_SPEC32 = 0, 1, 2
AQ_ACT = {"none": 0, "relu": 1, "gelu": 2, "sigmoid": 3}

# name -> (restype, argtypes)
SIGNATURES = {
    "aq_version": (c_int, []),
    "aq_last_error": (ctypes.c_char_p, []),
    "aq_theta": (c_int, [c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp]),
    "aq_theta_planes_bytes": (c_i64, [c_i64, c_i64, c_i64]),
    "aq_gram_svd_small": (c_int, [c_i64, c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64,
                                  c_int, ctypes.c_double, c_vp, c_vp]),
    "aq_hetrd_max_groups": (c_int, [c_int]),
    "aq_layer_gate_size": (c_i64, []),
    "aq_layer_pre": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp, c_vp]),
    "aq_layer_pad": (c_int, [c_vp, c_i64, c_i64, c_vp, c_vp, c_vp]),
    "aq_layer_post": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                              c_vp]),
    "aq_layer_1q": (c_int, [c_vp, c_vp, c_vp, c_int, c_vp]),
    "aq_cholqr_workspace_elems": (c_i64, [c_i64, c_i64]),
    "aq_cholqr": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, ctypes.c_double, c_vp, c_i64,
                          c_vp, c_vp]),
    "aq_theta_planes": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp]),
    "aq_rank_select": (c_int, [c_int, c_vp, c_i64, ctypes.c_double, c_i64, c_int, c_vp, c_vp]),
    "aq_rebuild": (c_int, [c_int, c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64,
                           c_i64, c_vp, c_vp, c_vp]),
    "aq_apply_1q": (c_int, [c_int, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "aq_cgemm": (c_int, [c_int, c_int, c_i64, c_i64, c_i64, c_vp, c_i64, c_i64, c_int,
                         c_vp, c_i64, c_i64, c_int, c_vp, c_i64, c_i64, c_vp]),
    "aq_cgemm_splitk": (c_int, [c_int, c_int, c_i64, c_i64, c_i64, c_vp, c_i64, c_i64, c_int,
                                c_vp, c_i64, c_i64, c_int, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64,
                                c_vp]),
    "aq_sample_workspace_bytes": (c_i64, [c_int, c_i64, c_vp, c_i64]),
    "aq_sample": (c_int, [c_int, c_int, c_int, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp,
                          c_i64, c_vp]),
    "aq_sample_noalloc": (c_int, [c_int, c_int, c_int, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp,
                                  c_vp, c_vp]),
    "aq_heev_small": (c_int, [c_i64, c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_int,
                              ctypes.c_double, c_vp, c_vp]),
    "aq_heev_small_orders": (c_int, [c_i64, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64,
                                     c_i64, c_int, ctypes.c_double, c_vp, c_vp]),
    "aq_heevd_batched": (c_int, [c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp]),
    "aq_bgemm": (c_int, [c_int, c_i64, c_i64, c_i64, ctypes.c_double, c_vp, c_i64, c_i64, c_i64,
                         c_int, c_vp, c_i64, c_i64, c_i64, c_int, ctypes.c_double, c_vp, c_i64,
                         c_i64, c_i64, c_i64, c_vp]),
    "aq_bgemm_lower": (c_int, [c_i64, c_i64, ctypes.c_double, c_vp, c_i64, c_i64, c_i64, c_int,
                               c_vp, c_i64, c_i64, c_i64, c_int, ctypes.c_double, c_vp, c_i64,
                               c_i64, c_i64, c_i64, c_vp]),
    "aq_hetrd_batched": (c_int, [c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "aq_hetrd_workspace_elems": (c_i64, [c_i64, c_i64, c_i64]),
    "aq_hetrd_own": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64,
                             c_vp]),
    "aq_hetrd_own_groups": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64,
                                    c_i64, c_int, c_vp]),
    "aq_he2hb_panels": (c_i64, [c_i64]),
    "aq_he2hb_workspace_elems": (c_i64, [c_i64, c_i64]),
    "aq_he2hb_own": (c_int, [c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_i64, c_vp]),
    "aq_hb2st_log_tasks": (c_i64, [c_i64]),
    "aq_hb2st_own": (c_int, [c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "aq_geqrf_panels": (c_i64, [c_i64, c_i64]),
    "aq_geqrf_workspace_elems": (c_i64, [c_i64, c_i64, c_i64]),
    "aq_geqrf": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_i64, c_vp]),
    "aq_ungqr": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp]),
    "aq_tridiag_eigvals": (c_int, [c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_i64,
                                   c_vp]),
    "aq_tridiag_eigvecs_scratch_doubles": (c_i64, [c_i64, c_i64, c_i64]),
    "aq_tridiag_eigvecs": (c_int, [c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64,
                                   c_vp, c_i64, c_vp, c_i64, c_i64, c_vp]),
    "aq_hetrd_vblocks": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "aq_hetrd_tfactors": (c_int, [c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_vp]),
    "aq_orth_band_reduce": (c_int, [c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp]),
    "aq_mlp_num_params": (c_int, []),
    "aq_mlp_sync_bf16": (c_int, [c_vp, c_vp, c_vp]),
    "aq_mlp_forward": (c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "aq_mlp_workspace_floats": (c_i64, [c_i64]),
    "aq_mlp_backward": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "aq_mlp_adamw": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             c_int, c_vp]),
    "aq_desc_sizes": (c_int, [c_vp]),
    "aq_grouped_gemm": (c_int, [c_int, c_vp, c_vp, c_int, c_i64, c_int, c_int, c_vp]),
    "aq_grouped_1q": (c_int, [c_vp, c_vp, c_int, c_i64, c_int, c_vp]),
    "aq_rank_select_batched": (c_int, [c_int, c_vp, c_vp, c_vp, ctypes.c_double, c_int, c_vp,
                                       c_vp]),
    "aq_sample_grouped_site": (c_int, [c_vp, c_vp, c_vp, c_int, c_i64, c_i64, c_int, c_vp]),
    "aq_zero_u64": (c_int, [c_vp, c_i64, c_vp]),
    "aq_grouped_copy": (c_int, [c_vp, c_vp, c_int, c_i64, c_int, c_vp]),
    "aq_tn_truncate": (c_int, [c_int, c_vp, c_i64, ctypes.c_double, c_int, c_i64, c_i64, c_int,
                               ctypes.c_double, ctypes.c_double, c_vp, c_vp]),
    "aq_rebuild_sqrt": (c_int, [c_int, c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64,
                                c_i64, c_i64, c_vp, c_vp, c_vp]),
    "aq_spectrum_features": (c_int, [c_int, c_int, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp]),
    "aq_dense_forward": (c_int, [c_int, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "aq_mpo_apply_site": (c_int, [c_int, c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "aq_mpo_env_combine": (c_int, [c_int, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp]),
    "aq_fro_normalize_batched": (c_int, [c_int, c_vp, c_i64, c_i64, c_i64, ctypes.c_double,
                                         c_vp]),
}

_lib = None


def header_symbols() -> list:
    """Entry points declared in include/atlasq.h."""
    txt = HEADER.read_text()
    decl = r"^(?:const\s+)?(?:int|int64_t|char|void)\s*\**\s*(aq_\w+)\s*\("
    return sorted(set(re.findall(decl, txt, re.M)))


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"atlas_q: HIP library {LIB_PATH} is missing; build it with "
                "`python atlas-q_amd/build.py` (no CPU fallback exists)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check_layouts() -> None:
    """Assert that the NumPy descriptor layouts match the library's structs."""
    out = (ctypes.c_int64 * 7)()
    check(lib().aq_desc_sizes(out), "aq_desc_sizes")
    want = [GEMM_DESC.itemsize, SITE1Q_DESC.itemsize, SAMPLE_DESC.itemsize,
            COPY_DESC.itemsize, ctypes.sizeof(RankResult), ctypes.sizeof(TnTrunc),
            ctypes.sizeof(DenseLayer)]
    if list(out) != want:
        raise RuntimeError(f"descriptor layout mismatch: library {list(out)} vs binding {want}")
    if int(lib().aq_layer_gate_size()) != LAYER_GATE.itemsize:
        raise RuntimeError("aq_layer_gate_t layout mismatch")


def check(status: int, what: str) -> None:
    if status != AQ_OK:
        msg = lib().aq_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {status}): {msg}")
