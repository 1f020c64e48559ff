// Rank rule, core rebuild, single-qubit gate, and the library's error state.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include "aq_common.h"

namespace aq {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------- rank rule
// choose_rank_from_sigma (truncation.py:21-83) with the reference's CPU
// rounding.  One work-group: the singular values are staged in LDS chunks;
// the cumulative sum is a single sequential double chain (the CPU cumsum's
// accumulation order) run by lane 0; the entropy is a parallel reduction.
template <typename T>
__device__ void rank_block(const T* __restrict__ S, int64_t m, T c, int64_t cap,
                           aq_rank_result_t* out) {
  constexpr int CH = 2048;  // multiple of 8
  __shared__ T sw[CH];
  __shared__ double red[256];
  __shared__ int64_t s_k;
  __shared__ double s_total_d;
  __shared__ T s_total, s_ekm1, s_ecap;
  const int tid = threadIdx.x;

  // pass 1: total = E[m-1].  One lane runs the sequential double chain (the
  // CPU cumsum's accumulation order); it reads the staged weights eight at a
  // time so the LDS loads are issued ahead of the dependent adds.
  double acc = 0.0;
  for (int64_t base = 0; base < m; base += CH) {
    const int64_t len = (m - base) < CH ? (m - base) : CH;
    __syncthreads();
    for (int64_t i = tid; i < CH; i += 256) {
      T w = 0;
      if (i < len) {
        const T s = S[base + i];
        w = s * s;
      }
      sw[i] = w;  // zero-padded to a multiple of 8
    }
    __syncthreads();
    if (tid == 0) {
      for (int64_t i = 0; i < len; i += 8) {
        T w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = sw[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += (double)w[u];
      }
    }
  }
  if (tid == 0) {
    s_total = (T)acc;
    s_total_d = acc;
  }
  __syncthreads();
  const T total = s_total;
  if (total < (T)1e-30) {
    if (tid == 0) {
      out->k = 1;
      out->flags = 1;
      out->eps_local = 0.0;
      out->entropy = 0.0;
      out->condS = INFINITY;
      out->total_weight = (double)total;
      out->kept_weight = (double)total;
    }
    return;
  }
  const T thresh = c * total;
  // pass 2: first index with E[i] >= thresh (left searchsorted); also keep
  // E[cap-1] for the capped case.  Chunks of 8: the eight prefix values are
  // formed first, then tested in order.
  if (tid == 0) {
    s_k = -1;
    s_ekm1 = total;
    s_ecap = total;
  }
  acc = 0.0;
  for (int64_t base = 0; base < m; base += CH) {
    const int64_t len = (m - base) < CH ? (m - base) : CH;
    __syncthreads();
    if (s_k >= 0) break;
    for (int64_t i = tid; i < CH; i += 256) {
      T w = 0;
      if (i < len) {
        const T s = S[base + i];
        w = s * s;
      }
      sw[i] = w;
    }
    __syncthreads();
    if (tid == 0) {
      for (int64_t i = 0; i < len; i += 8) {
        T w[8], e[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = sw[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          acc += (double)w[u];
          e[u] = (T)acc;
        }
        int hit = -1;
#pragma unroll
        for (int u = 7; u >= 0; --u)
          if (i + u < len && e[u] >= thresh) hit = u;
        const int64_t cidx = cap - 1 - (base + i);
        if (cidx >= 0 && cidx < 8 && (hit < 0 || cidx <= hit)) s_ecap = e[cidx];
        if (hit >= 0) {
          s_k = base + i + hit;
          s_ekm1 = e[hit];
          break;
        }
      }
    }
  }
  __syncthreads();
  int64_t ktol = (s_k < 0 ? m - 1 : s_k) + 1;
  if (ktol > m) ktol = m;
  int64_t k = ktol < cap ? ktol : cap;
  if (k < 1) k = 1;
  const T e_km1 = (k == ktol) ? s_ekm1 : s_ecap;
  // entropy over kept values: p = w / sum(w)
  double part = 0.0;
  for (int64_t i = tid; i < k; i += 256) {
    const T s = S[i];
    part += (double)(T)(s * s);
  }
  red[tid] = part;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double wsum = red[0];
  __syncthreads();
  double ent = 0.0;
  for (int64_t i = tid; i < k; i += 256) {
    const T s = S[i];
    double p = (double)(T)(s * s) / wsum;
    if (p < 1e-30) p = 1e-30;
    ent -= ((double)(T)(s * s) / wsum) * log(p);
  }
  red[tid] = ent;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    T d = total - e_km1;
    if (d < (T)0) d = 0;
    out->k = (int32_t)k;
    out->flags = 0;
    out->eps_local = (double)sqrt(d);
    out->entropy = red[0];
    const T sk = S[k - 1];
    out->condS = (k > 1 && sk > (T)0) ? (double)(S[0] / sk) : INFINITY;
    out->total_weight = s_total_d;
    out->kept_weight = wsum;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void rank_kernel(const T* __restrict__ S, int64_t m, T c,
                                                   int64_t cap, aq_rank_result_t* out) {
  rank_block<T>(S, m, c, cap, out);
}

template <typename T>
__global__ __launch_bounds__(256) void rank_batched_kernel(const void* const* __restrict__ Sp,
                                                           const int64_t* __restrict__ m,
                                                           const int64_t* __restrict__ cap, T c,
                                                           aq_rank_result_t* out) {
  const int b = blockIdx.x;
  rank_block<T>((const T*)Sp[b], m[b], c, cap[b] < 1 ? 1 : cap[b], out + b);
}

// ---------------------------------------------------------------- rebuild
// A'[r, c] = U[r, c] * S[c] (r < rows, c < k), B'[r, c] = Vh[r, c] (r < k,
// c < cols); 32x32 tiles through LDS so both the read (either storage order
// of the factor) and the row-major write are coalesced.
template <typename T>
__global__ __launch_bounds__(256) void scale_copy_kernel(const c2_t<T>* __restrict__ src,
                                                         int64_t s0, int64_t s1,
                                                         const T* __restrict__ scale,
                                                         c2_t<T>* __restrict__ dst, int64_t R,
                                                         int64_t Cn) {
  __shared__ c2_t<T> tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const bool col_major = (s0 == 1);
  for (int i = ty; i < 32; i += 8) {
    // read: consecutive lanes follow the unit-stride index
    const int64_t r = col_major ? r0 + tx : r0 + i;
    const int64_t c = col_major ? c0 + i : c0 + tx;
    if (r < R && c < Cn) {
      c2_t<T> v = src[r * s0 + c * s1];
      if (scale) {
        const T sc = scale[c];
        v.x *= sc;
        v.y *= sc;
      }
      if (col_major) tile[tx][i] = v;
      else tile[i][tx] = v;
    }
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int64_t r = r0 + i, c = c0 + tx;
    if (r < R && c < Cn) dst[r * Cn + c] = tile[i][tx];
  }
}

template <typename T>
static int rebuild_impl(const void* U, int64_t us0, int64_t us1, const void* S, const void* Vh,
                        int64_t vs0, int64_t vs1, int64_t rows, int64_t cols, int64_t k,
                        void* Ao, void* Bo, hipStream_t s) {
  dim3 ga((unsigned)cdiv(k, 32), (unsigned)cdiv(rows, 32));
  scale_copy_kernel<T><<<ga, 256, 0, s>>>((const c2_t<T>*)U, us0, us1, (const T*)S,
                                          (c2_t<T>*)Ao, rows, k);
  AQ_LAUNCH_CHECK();
  dim3 gb((unsigned)cdiv(cols, 32), (unsigned)cdiv(k, 32));
  scale_copy_kernel<T><<<gb, 256, 0, s>>>((const c2_t<T>*)Vh, vs0, vs1, nullptr,
                                          (c2_t<T>*)Bo, k, cols);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

// ---------------------------------------------------------------- 1q gate
template <typename T>
__global__ void apply_1q_kernel(const c2_t<T>* __restrict__ in, c2_t<T>* __restrict__ out,
                                int64_t chiL, int64_t chiR, c2_t<T> u00, c2_t<T> u01,
                                c2_t<T> u10, c2_t<T> u11) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= chiL * chiR) return;
  const int64_t a = idx / chiR, b = idx % chiR;
  const c2_t<T> x0 = in[(a * 2 + 0) * chiR + b];
  const c2_t<T> x1 = in[(a * 2 + 1) * chiR + b];
  // T'[a,t,b] = sum_s U[s,t] T[a,s,b]
  out[(a * 2 + 0) * chiR + b] = cadd(cmul(u00, x0), cmul(u10, x1));
  out[(a * 2 + 1) * chiR + b] = cadd(cmul(u01, x0), cmul(u11, x1));
}

template <typename T>
static c2_t<T> mk(const double* g, int i) {
  c2_t<T> v;
  v.x = (T)g[2 * i];
  v.y = (T)g[2 * i + 1];
  return v;
}

}  // namespace aq

using namespace aq;

extern "C" int aq_version(void) { return 10000; }

extern "C" const char* aq_last_error(void) { return g_err; }

extern "C" int aq_desc_sizes(int64_t* out) {
  AQ_REQUIRE(out, AQ_ERR_ARG, "aq_desc_sizes: null pointer");
  out[0] = (int64_t)sizeof(aq_gemm_desc_t);
  out[1] = (int64_t)sizeof(aq_site1q_desc_t);
  out[2] = (int64_t)sizeof(aq_sample_desc_t);
  out[3] = (int64_t)sizeof(aq_copy_desc_t);
  out[4] = (int64_t)sizeof(aq_rank_result_t);
  out[5] = (int64_t)sizeof(aq_tn_trunc_t);
  out[6] = (int64_t)sizeof(aq_dense_layer_t);
  return AQ_OK;
}

extern "C" int aq_rank_select(int s_dtype, const void* S, int64_t m, double eps, int64_t cap,
                              int mode, aq_rank_result_t* out, hipStream_t stream) {
  AQ_REQUIRE(S && out, AQ_ERR_ARG, "aq_rank_select: null pointer");
  AQ_REQUIRE(m > 0, AQ_ERR_ARG, "aq_rank_select: empty spectrum");
  AQ_REQUIRE(cap >= 1, AQ_ERR_ARG, "aq_rank_select: cap must be >= 1");
  AQ_REQUIRE(mode == AQ_RANK_REF || mode == AQ_RANK_TDVP, AQ_ERR_ARG,
             "aq_rank_select: unknown mode %d", mode);
  // Python-float (1 - eps**2), truncation.py:58; (1 - eps), tdvp.py:553
  const double c = mode == AQ_RANK_TDVP ? 1.0 - eps : 1.0 - eps * eps;
  if (s_dtype == AQ_C64) {
    rank_kernel<float><<<1, 256, 0, stream>>>((const float*)S, m, (float)c, cap, out);
  } else {
    AQ_REQUIRE(s_dtype == AQ_C128, AQ_ERR_ARG, "aq_rank_select: unknown dtype %d", s_dtype);
    rank_kernel<double><<<1, 256, 0, stream>>>((const double*)S, m, c, cap, out);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_rebuild(int dtype, const void* U, int64_t u_s0, int64_t u_s1, const void* S,
                          const void* Vh, int64_t v_s0, int64_t v_s1, int64_t rows, int64_t cols,
                          int64_t k, void* A_out, void* B_out, hipStream_t stream) {
  AQ_REQUIRE(U && S && Vh && A_out && B_out, AQ_ERR_ARG, "aq_rebuild: null pointer");
  AQ_REQUIRE(rows > 0 && cols > 0 && k > 0, AQ_ERR_ARG, "aq_rebuild: sizes must be positive");
  AQ_REQUIRE(rows % 2 == 0 && cols % 2 == 0, AQ_ERR_ARG, "aq_rebuild: rows/cols must be even");
  if (dtype == AQ_C64)
    return rebuild_impl<float>(U, u_s0, u_s1, S, Vh, v_s0, v_s1, rows, cols, k, A_out, B_out,
                               stream);
  AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_rebuild: unknown dtype %d", dtype);
  return rebuild_impl<double>(U, u_s0, u_s1, S, Vh, v_s0, v_s1, rows, cols, k, A_out, B_out,
                              stream);
}

extern "C" int aq_apply_1q(int dtype, const void* T_in, void* T_out, int64_t chiL, int64_t chiR,
                           const double* g, hipStream_t stream) {
  AQ_REQUIRE(T_in && T_out && g, AQ_ERR_ARG, "aq_apply_1q: null pointer");
  AQ_REQUIRE(chiL > 0 && chiR > 0, AQ_ERR_ARG, "aq_apply_1q: sizes must be positive");
  const int64_t cnt = chiL * chiR;
  const unsigned blocks = (unsigned)cdiv(cnt, 256);
  if (dtype == AQ_C64) {
    apply_1q_kernel<float><<<blocks, 256, 0, stream>>>(
        (const float2*)T_in, (float2*)T_out, chiL, chiR, mk<float>(g, 0), mk<float>(g, 1),
        mk<float>(g, 2), mk<float>(g, 3));
  } else {
    AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_apply_1q: unknown dtype %d", dtype);
    apply_1q_kernel<double><<<blocks, 256, 0, stream>>>(
        (const double2*)T_in, (double2*)T_out, chiL, chiR, mk<double>(g, 0), mk<double>(g, 1),
        mk<double>(g, 2), mk<double>(g, 3));
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_rank_select_batched(int s_dtype, const void* const* S_ptrs, const int64_t* m,
                                      const int64_t* cap, double eps, int count,
                                      aq_rank_result_t* out, hipStream_t stream) {
  AQ_REQUIRE(S_ptrs && m && cap && out, AQ_ERR_ARG, "aq_rank_select_batched: null pointer");
  AQ_REQUIRE(count >= 1, AQ_ERR_ARG, "aq_rank_select_batched: count must be >= 1");
  const double c = 1.0 - eps * eps;
  if (s_dtype == AQ_C64) {
    rank_batched_kernel<float><<<count, 256, 0, stream>>>(S_ptrs, m, cap, (float)c, out);
  } else {
    AQ_REQUIRE(s_dtype == AQ_C128, AQ_ERR_ARG, "aq_rank_select_batched: unknown dtype");
    rank_batched_kernel<double><<<count, 256, 0, stream>>>(S_ptrs, m, cap, c, out);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}
