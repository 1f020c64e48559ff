"""aq_sample_noalloc (fixed device scratch, no allocation) against aq_sample
(caller workspace), both through the C ABI: bits and p0 identical on two
calls in a row, the second smaller than the first, and a sample whose
workspace exceeds the fixed scratch rejected with AQ_ERR_ARG."""
import ctypes

import numpy as np
import pytest
import torch

from atlas_q import _ffi
from atlas_q.workloads import random_right_isometry

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _chain(n, cmax, seed, dtype):
    chi = [1] + [min(cmax, 2 ** min(b + 1, n - b - 1)) for b in range(n - 1)] + [1]
    gen = torch.Generator(device="cpu").manual_seed(seed)
    ts = [random_right_isometry(chi[i], chi[i + 1], gen).to(dtype).contiguous().to(DEV)
          for i in range(n)]
    return ts, chi


def _both(ts, chi, shots, seed, dtype):
    L = _ffi.lib()
    n = len(ts)
    code = _ffi.AQ_C64 if dtype == torch.complex64 else _ffi.AQ_C128
    chi_arr = (ctypes.c_int64 * (n + 1))(*chi)
    ptrs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in ts])
    gen = torch.Generator(device="cpu").manual_seed(seed)
    u = torch.rand((shots, n), dtype=torch.float64, generator=gen).to(DEV)
    words = (n + 63) // 64
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    need = L.aq_sample_workspace_bytes(code, n, chi_arr, shots)
    assert 0 < need <= _ffi.AQ_SAMPLE_SCRATCH_BYTES
    ws = torch.empty(need, dtype=torch.uint8, device=DEV)
    out = {}
    for name in ("ref", "noalloc"):
        bits = torch.full((shots, words), -1, dtype=torch.int64, device=DEV)
        p0 = torch.full((shots, n), -1.0, dtype=torch.float64, device=DEV)
        if name == "ref":
            st = L.aq_sample(code, 0, 0, n, ptrs, chi_arr, shots, ctypes.c_void_p(u.data_ptr()),
                             ctypes.c_void_p(bits.data_ptr()), ctypes.c_void_p(p0.data_ptr()),
                             ctypes.c_void_p(ws.data_ptr()), need, stream)
        else:
            st = L.aq_sample_noalloc(code, 0, 0, n, ptrs, chi_arr, shots,
                                     ctypes.c_void_p(u.data_ptr()),
                                     ctypes.c_void_p(bits.data_ptr()),
                                     ctypes.c_void_p(p0.data_ptr()), stream)
        _ffi.check(st, name)
        torch.cuda.synchronize()
        out[name] = (bits.cpu().numpy(), p0.cpu().numpy())
    return out


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_noalloc_sampler_matches_workspace_sampler(dtype):
    for n, cmax, shots, seed in ((24, 64, 3000, 1), (10, 16, 500, 2)):
        print(f"noalloc sampler {dtype} n={n} chi<={cmax} shots={shots}", flush=True)
        ts, chi = _chain(n, cmax, seed, dtype)
        out = _both(ts, chi, shots, seed + 10, dtype)
        assert np.array_equal(out["ref"][0], out["noalloc"][0])
        assert np.array_equal(out["ref"][1], out["noalloc"][1])


def test_noalloc_sampler_rejects_oversize():
    L = _ffi.lib()
    n, shots = 4, 1 << 20
    chi = [1, 2, 4, 2, 1]
    chi_arr = (ctypes.c_int64 * (n + 1))(*chi)
    need = L.aq_sample_workspace_bytes(_ffi.AQ_C128, n, chi_arr, shots)
    assert need > _ffi.AQ_SAMPLE_SCRATCH_BYTES
    ptrs = (ctypes.c_void_p * n)(*([1] * n))
    st = L.aq_sample_noalloc(_ffi.AQ_C128, 0, 0, n, ptrs, chi_arr, shots, ctypes.c_void_p(8),
                             ctypes.c_void_p(8), None,
                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == -1
    assert b"scratch" in L.aq_last_error()
