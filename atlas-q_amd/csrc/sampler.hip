// Batched sweep sampler over all shots (mps_pytorch.py:247-355 semantics).
//
// Per site i, two launches:
//   1. sample_gemm: T = L_i [S, chiL] x A_i [chiL, 2 chiR] on MFMA with the
//      outcome halves (s = 0, 1) paired per lane; the epilogue writes T and
//      the per-row partial sums of |T|^2 of each half for its 32 columns.
//      L_i is not materialised: row s of L_i is the half of the previous
//      site's T selected by that shot's outcome (a per-row gather offset),
//      so the reference's [chiL, S, chiR] gather never exists.
//   2. sample_select: per shot, sum the partials in a fixed order, form the
//      reference's probabilities (|.|^2 sums, +1e-15 floor, normalise, in the
//      tensor's real type), draw outcome = (u >= p0), set the bit and the
//      gather offset for the next site.
#include <mutex>
#include <type_traits>

#include "bf16x3.h"
#include "gemm_core.h"
#include "sampler_core.h"

namespace aq {

template <typename T>
__global__ __launch_bounds__(kThreads, 2) void sample_gemm_kernel(
    const c2_t<T>* __restrict__ Lbase, const int32_t* __restrict__ off, int64_t ldL,
    const c2_t<T>* __restrict__ site, c2_t<T>* __restrict__ Tout, T* __restrict__ part,
    int64_t S, int64_t chiL, int64_t chiR, int ntiles) {
  __shared__ Smem<T> sm;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t j0 = (int64_t)blockIdx.y * kBJ;
  AGather<T> aop{Lbase, off, ldL, S, chiL};
  BStrided<T> bop{site, 2 * chiR, 1, chiL, 2 * chiR, 0};
  ColMap cm{1, chiR};
  Frag<T> fr;
  cgemm_mainloop<T, true, true, std::is_same<T, float>::value>(aop, bop, cm, chiL, m0, j0, sm,
                                                               fr);
  sampler_epilogue<T>(fr, Tout, part, S, chiR, m0, j0, ntiles);
}

__global__ __launch_bounds__(kThreads) void sample_gemm_bf16x3_kernel(
    const float2* __restrict__ Lbase, const int32_t* __restrict__ off, int64_t ldL,
    const float2* __restrict__ site, float2* __restrict__ Tout, float* __restrict__ part,
    int64_t S, int64_t chiL, int64_t chiR, int ntiles) {
  __shared__ SmemBf16 sm;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t j0 = (int64_t)blockIdx.y * kBJ;
  AGather<float> aop{Lbase, off, ldL, S, chiL};
  BStrided<float> bop{site, 2 * chiR, 1, chiL, 2 * chiR, 0};
  ColMap cm{1, chiR};
  Frag<float> fr;
  cgemm_mainloop_bf16x3<true, true>(aop, bop, cm, chiL, m0, j0, sm, fr);
  sampler_epilogue<float>(fr, Tout, part, S, chiR, m0, j0, ntiles);
}

template <typename T>
__global__ void sample_select_kernel(const T* __restrict__ part, int ntiles,
                                     const double* __restrict__ uniforms, int64_t n, int64_t site,
                                     int64_t S, int64_t chiR, int32_t* __restrict__ off,
                                     uint64_t* __restrict__ bits, int64_t words,
                                     double* __restrict__ p0_out, int single) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  select_one<T>(part, ntiles, uniforms, n, site, s, chiR, off, bits, words, p0_out, single);
}

// Single-shot first site (mps_pytorch.py:198-201): p_s = |sum_b T0[0,s,b]|^2,
// written as tile 0 of the partials (other tiles zeroed) for every shot.
template <typename T>
__global__ __launch_bounds__(256) void first_site_amp_kernel(const c2_t<T>* __restrict__ T0,
                                                             int64_t chi1, T* __restrict__ part,
                                                             int ntiles, int64_t S) {
  __shared__ T red[4][256];
  const int tid = threadIdx.x;
  T sr0 = 0, si0 = 0, sr1 = 0, si1 = 0;
  for (int64_t b = tid; b < chi1; b += 256) {
    sr0 += T0[b].x;
    si0 += T0[b].y;
    sr1 += T0[chi1 + b].x;
    si1 += T0[chi1 + b].y;
  }
  red[0][tid] = sr0;
  red[1][tid] = si0;
  red[2][tid] = sr1;
  red[3][tid] = si1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int q = 0; q < 4; ++q) red[q][tid] += red[q][tid + w];
    __syncthreads();
  }
  const T p0 = red[0][0] * red[0][0] + red[1][0] * red[1][0];
  const T p1 = red[2][0] * red[2][0] + red[3][0] * red[3][0];
  for (int64_t i = tid; i < S * ntiles; i += 256) {
    const bool first = (i % ntiles) == 0;
    part[2 * i + 0] = first ? p0 : (T)0;
    part[2 * i + 1] = first ? p1 : (T)0;
  }
}

__global__ void zero_u64_kernel(uint64_t* p, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct SampleWs {
  size_t t0, t1, part, off, total;
};

template <typename T>
static SampleWs sample_ws_layout(int64_t n, const int64_t* chi, int64_t S) {
  int64_t cmax = 1;
  for (int64_t i = 0; i <= n; ++i) cmax = chi[i] > cmax ? chi[i] : cmax;
  SampleWs w;
  const size_t tbytes = align256((size_t)S * 2 * cmax * sizeof(c2_t<T>));
  const int64_t nt = cdiv(cmax, 32);
  w.t0 = 0;
  w.t1 = tbytes;
  w.part = 2 * tbytes;
  w.off = w.part + align256((size_t)S * nt * 2 * sizeof(T));
  w.total = w.off + align256((size_t)S * sizeof(int32_t));
  return w;
}

template <typename T>
static int sample_impl(int prec, int flags, int64_t n, const void* const* sites,
                       const int64_t* chi, int64_t S, const double* uniforms, uint64_t* bits,
                       double* p0_out, void* ws, int64_t ws_bytes, hipStream_t stream) {
  const int single = (flags & AQ_SAMPLE_SINGLE_SHOT_RULES) ? 1 : 0;
  const SampleWs L = sample_ws_layout<T>(n, chi, S);
  AQ_REQUIRE((int64_t)L.total <= ws_bytes, AQ_ERR_ARG,
             "aq_sample: workspace too small (%lld < %lld bytes)", (long long)ws_bytes,
             (long long)L.total);
  char* base = (char*)ws;
  c2_t<T>* tbuf[2] = {(c2_t<T>*)(base + L.t0), (c2_t<T>*)(base + L.t1)};
  T* part = (T*)(base + L.part);
  int32_t* off = (int32_t*)(base + L.off);
  const int64_t words = cdiv(n, 64);
  {
    const int64_t cnt = S * words;
    zero_u64_kernel<<<(unsigned)cdiv(cnt, 256), 256, 0, stream>>>(bits, cnt);
    AQ_LAUNCH_CHECK();
  }
  for (int64_t i = 0; i < n; ++i) {
    const int64_t cl = chi[i], cr = chi[i + 1];
    const int ntiles = (int)cdiv(cr, 32);
    const c2_t<T>* Lb = (i == 0) ? nullptr : tbuf[(i + 1) & 1];
    const int64_t ldL = 2 * cl;  // row stride of the previous site's T
    c2_t<T>* Tout = tbuf[i & 1];
    dim3 grid((unsigned)cdiv(S, kBM), (unsigned)cdiv(cr, kBJ));
    if constexpr (std::is_same<T, float>::value) {
      if (prec == AQ_PREC_BF16X3) {
        sample_gemm_bf16x3_kernel<<<grid, kThreads, 0, stream>>>(
            Lb, i == 0 ? nullptr : off, ldL, (const float2*)sites[i], Tout, part, S, cl, cr,
            ntiles);
      } else {
        sample_gemm_kernel<T><<<grid, kThreads, 0, stream>>>(
            Lb, i == 0 ? nullptr : off, ldL, (const c2_t<T>*)sites[i], Tout, part, S, cl, cr,
            ntiles);
      }
    } else {
      sample_gemm_kernel<T><<<grid, kThreads, 0, stream>>>(
          Lb, i == 0 ? nullptr : off, ldL, (const c2_t<T>*)sites[i], Tout, part, S, cl, cr,
          ntiles);
    }
    AQ_LAUNCH_CHECK();
    if (single && i == 0) {
      first_site_amp_kernel<T><<<1, 256, 0, stream>>>((const c2_t<T>*)sites[0], cr, part,
                                                      ntiles, S);
      AQ_LAUNCH_CHECK();
    }
    sample_select_kernel<T><<<(unsigned)cdiv(S, 256), 256, 0, stream>>>(
        part, ntiles, uniforms, n, i, S, cr, off, bits, words, p0_out, single);
    AQ_LAUNCH_CHECK();
  }
  return AQ_OK;
}

}  // namespace aq

using namespace aq;

extern "C" int64_t aq_sample_workspace_bytes(int dtype, int64_t n, const int64_t* chi,
                                             int64_t shots) {
  if (!chi || n <= 0 || shots <= 0) return -1;
  return dtype == AQ_C128 ? (int64_t)sample_ws_layout<double>(n, chi, shots).total
                          : (int64_t)sample_ws_layout<float>(n, chi, shots).total;
}

extern "C" int aq_sample(int dtype, int prec, int flags, int64_t n, const void* const* sites,
                         const int64_t* chi, int64_t shots, const double* uniforms,
                         uint64_t* bits, double* p0_out, void* workspace,
                         int64_t workspace_bytes, hipStream_t stream) {
  AQ_REQUIRE(sites && chi && uniforms && bits && workspace, AQ_ERR_ARG,
             "aq_sample: null pointer");
  AQ_REQUIRE(n > 0 && shots > 0, AQ_ERR_ARG, "aq_sample: n and shots must be positive");
  AQ_REQUIRE(chi[0] == 1 && chi[n] == 1, AQ_ERR_ARG,
             "aq_sample: boundary bonds must be 1 (got %lld, %lld)", (long long)chi[0],
             (long long)chi[n]);
  for (int64_t i = 0; i < n; ++i) {
    AQ_REQUIRE(sites[i] != nullptr, AQ_ERR_ARG, "aq_sample: site %lld is null", (long long)i);
    AQ_REQUIRE(chi[i + 1] > 0 && chi[i + 1] < (1 << 24), AQ_ERR_ARG,
               "aq_sample: bad bond size at %lld", (long long)(i + 1));
  }
  AQ_REQUIRE(cdiv(shots, kBM) < (1LL << 31), AQ_ERR_ARG, "aq_sample: too many shots");
  if (dtype == AQ_C64)
    return sample_impl<float>(prec, flags, n, sites, chi, shots, uniforms, bits, p0_out, workspace,
                              workspace_bytes, stream);
  AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_sample: unknown dtype %d", dtype);
  AQ_REQUIRE(prec == AQ_PREC_NATIVE, AQ_ERR_ARG, "aq_sample: complex128 is FP64 only");
  return sample_impl<double>(prec, flags, n, sites, chi, shots, uniforms, bits, p0_out, workspace,
                             workspace_bytes, stream);
}

// ------------------------------------------------------------ no-allocation
// aq_sample_noalloc: the sampler with its workspace in a file-scope device
// array instead of caller memory.  Each GPU holds its own copy of `scratch`
// for the life of the process; in host code the name is only a stand-in, so
// the array is reached through the device address hipGetSymbolAddress gives
// for the current device (looked up once per device, kept in scratch_ptr).
// Nothing in the workspace has to start at zero: every site's T rows and
// |T|^2 partials are written by the GEMM before the select reads them, and
// the gather offsets are written by a site's select before the next site's
// GEMM reads them (the first site reads none).
namespace aq {
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char scratch[AQ_SAMPLE_SCRATCH_BYTES];
static void* scratch_ptr[kMaxDevices];
static std::mutex scratch_lock;

static int scratch_address(void** out) {
  int dev = 0;
  AQ_HIP(hipGetDevice(&dev));
  AQ_REQUIRE(dev >= 0 && dev < kMaxDevices, AQ_ERR_UNSUPPORTED,
             "aq_sample_noalloc: device index %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> g(scratch_lock);
  if (scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    AQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    AQ_REQUIRE(p != nullptr, AQ_ERR_HIP, "aq_sample_noalloc: no device address for scratch");
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return AQ_OK;
}
}  // namespace aq

extern "C" int aq_sample_noalloc(int dtype, int prec, int flags, int64_t n,
                                 const void* const* sites, const int64_t* chi, int64_t shots,
                                 const double* uniforms, uint64_t* bits, double* p0_out,
                                 hipStream_t stream) {
  AQ_REQUIRE(chi && n > 0 && shots > 0, AQ_ERR_ARG,
             "aq_sample_noalloc: n, shots and chi are required");
  AQ_REQUIRE(dtype == AQ_C64 || dtype == AQ_C128, AQ_ERR_ARG,
             "aq_sample_noalloc: unknown dtype %d", dtype);
  const int64_t need = aq_sample_workspace_bytes(dtype, n, chi, shots);
  AQ_REQUIRE(need > 0 && need <= (int64_t)AQ_SAMPLE_SCRATCH_BYTES, AQ_ERR_ARG,
             "aq_sample_noalloc: this sample needs %lld workspace bytes, above the fixed "
             "%lld-byte scratch (fewer shots per call, or aq_sample with a caller workspace)",
             (long long)need, (long long)AQ_SAMPLE_SCRATCH_BYTES);
  void* ws = nullptr;
  const int st = scratch_address(&ws);
  if (st != AQ_OK) return st;
  return aq_sample(dtype, prec, flags, n, sites, chi, shots, uniforms, bits, p0_out, ws,
                   (int64_t)AQ_SAMPLE_SCRATCH_BYTES, stream);
}
