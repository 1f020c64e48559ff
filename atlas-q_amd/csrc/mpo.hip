// Matrix-product-operator steps next to the two-site path: MPO x MPS site
// contraction (mpo_ops.py:1034-1082), the combine step of the MPO
// expectation-value transfer (mpo_ops.py:1083-1150) and the batched
// Frobenius renormalisation of the environment scans
// (sim_env_scan.py:53-130).  All three are HBM-bound elementwise /
// reduction kernels; the GEMM-shaped parts of those callers run on the MFMA
// GEMM (aq_cgemm / aq_grouped_gemm).
#include <math.h>

#include "aq_common.h"

namespace aq {

// M[(l*chiL + a), s, (r*chiR + b)] = sum_t W[l, s, t, r] * A[a, t, b]
// (einsum "lstr, atb -> lasrb" reshaped), one thread per output element.
template <typename T>
__global__ __launch_bounds__(256) void mpo_apply_site_kernel(const c2_t<T>* __restrict__ W,
                                                             int64_t Dl, int64_t Dr,
                                                             const c2_t<T>* __restrict__ A,
                                                             int64_t chiL, int64_t chiR,
                                                             c2_t<T>* __restrict__ M) {
  const int64_t total = Dl * chiL * 2 * Dr * chiR;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  int64_t rem = idx;
  const int64_t b = rem % chiR;
  rem /= chiR;
  const int64_t r = rem % Dr;
  rem /= Dr;
  const int64_t s = rem % 2;
  rem /= 2;
  const int64_t a = rem % chiL;
  const int64_t l = rem / chiL;
  c2_t<T> acc;
  acc.x = 0;
  acc.y = 0;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const c2_t<T> w = W[((l * 2 + s) * 2 + t) * Dr + r];
    const c2_t<T> x = A[(a * 2 + t) * chiR + b];
    acc = cadd(acc, cmul(w, x));
  }
  M[idx] = acc;
}

// E'[R, r, B] = sum_{L, s, t} W[L, s, t, R] * Z[L, t, r, s*chiB + B]
// with Z [DL, 2, chir, 2 chiB] (Z[L, t] = conj(A_t)^T E_L A, both outcome
// halves side by side), one thread per output element, fixed summation order.
template <typename T>
__global__ __launch_bounds__(256) void mpo_env_combine_kernel(const c2_t<T>* __restrict__ Z,
                                                              const c2_t<T>* __restrict__ W,
                                                              int64_t DL, int64_t DR,
                                                              int64_t chir, int64_t chiB,
                                                              c2_t<T>* __restrict__ E) {
  const int64_t total = DR * chir * chiB;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int64_t B = idx % chiB;
  const int64_t r = (idx / chiB) % chir;
  const int64_t R = idx / (chiB * chir);
  c2_t<T> acc;
  acc.x = 0;
  acc.y = 0;
  for (int64_t L = 0; L < DL; ++L)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const c2_t<T> w = W[((L * 2 + s) * 2 + t) * DR + R];
        const c2_t<T> z = Z[((L * 2 + t) * chir + r) * (2 * chiB) + s * chiB + B];
        acc = cadd(acc, cmul(w, z));
      }
  E[idx] = acc;
}

// X_b <- X_b / max(||X_b||_F, floor) for `batch` matrices of `count`
// contiguous complex elements at X + b * stride; one work-group per matrix.
template <typename T>
__global__ __launch_bounds__(256) void fro_normalize_kernel(c2_t<T>* __restrict__ X,
                                                            int64_t stride, int64_t count,
                                                            double floor_) {
  __shared__ double red[256];
  c2_t<T>* x = X + (int64_t)blockIdx.x * stride;
  double part = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += 256) {
    const double re = (double)x[i].x, im = (double)x[i].y;
    part += re * re + im * im;
  }
  red[threadIdx.x] = part;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  double nrm = sqrt(red[0]);
  if (nrm < floor_) nrm = floor_;
  const T d = (T)nrm;
  for (int64_t i = threadIdx.x; i < count; i += 256) {
    x[i].x = x[i].x / d;
    x[i].y = x[i].y / d;
  }
}

}  // namespace aq

using namespace aq;

extern "C" int aq_mpo_apply_site(int dtype, const void* W, int64_t Dl, int64_t Dr, const void* A,
                                 int64_t chiL, int64_t chiR, void* M, hipStream_t stream) {
  AQ_REQUIRE(W && A && M, AQ_ERR_ARG, "aq_mpo_apply_site: null pointer");
  AQ_REQUIRE(Dl > 0 && Dr > 0 && chiL > 0 && chiR > 0, AQ_ERR_ARG,
             "aq_mpo_apply_site: sizes must be positive");
  const int64_t total = Dl * chiL * 2 * Dr * chiR;
  AQ_REQUIRE(cdiv(total, 256) < (1LL << 31), AQ_ERR_ARG, "aq_mpo_apply_site: too large");
  const unsigned blocks = (unsigned)cdiv(total, 256);
  if (dtype == AQ_C64) {
    mpo_apply_site_kernel<float><<<blocks, 256, 0, stream>>>(
        (const float2*)W, Dl, Dr, (const float2*)A, chiL, chiR, (float2*)M);
  } else {
    AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_mpo_apply_site: unknown dtype %d", dtype);
    mpo_apply_site_kernel<double><<<blocks, 256, 0, stream>>>(
        (const double2*)W, Dl, Dr, (const double2*)A, chiL, chiR, (double2*)M);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_mpo_env_combine(int dtype, const void* Z, const void* W, int64_t DL,
                                  int64_t DR, int64_t chir, int64_t chiB, void* E,
                                  hipStream_t stream) {
  AQ_REQUIRE(Z && W && E, AQ_ERR_ARG, "aq_mpo_env_combine: null pointer");
  AQ_REQUIRE(DL > 0 && DR > 0 && chir > 0 && chiB > 0, AQ_ERR_ARG,
             "aq_mpo_env_combine: sizes must be positive");
  const int64_t total = DR * chir * chiB;
  AQ_REQUIRE(cdiv(total, 256) < (1LL << 31), AQ_ERR_ARG, "aq_mpo_env_combine: too large");
  const unsigned blocks = (unsigned)cdiv(total, 256);
  if (dtype == AQ_C64) {
    mpo_env_combine_kernel<float><<<blocks, 256, 0, stream>>>(
        (const float2*)Z, (const float2*)W, DL, DR, chir, chiB, (float2*)E);
  } else {
    AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_mpo_env_combine: unknown dtype %d", dtype);
    mpo_env_combine_kernel<double><<<blocks, 256, 0, stream>>>(
        (const double2*)Z, (const double2*)W, DL, DR, chir, chiB, (double2*)E);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_fro_normalize_batched(int dtype, void* X, int64_t batch, int64_t stride,
                                        int64_t count, double floor_, hipStream_t stream) {
  AQ_REQUIRE(X, AQ_ERR_ARG, "aq_fro_normalize_batched: null pointer");
  AQ_REQUIRE(batch > 0 && batch < (1LL << 31) && count > 0 && stride >= count, AQ_ERR_ARG,
             "aq_fro_normalize_batched: bad sizes");
  if (dtype == AQ_C64) {
    fro_normalize_kernel<float><<<(unsigned)batch, 256, 0, stream>>>((float2*)X, stride, count,
                                                                     floor_);
  } else {
    AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_fro_normalize_batched: unknown dtype %d", dtype);
    fro_normalize_kernel<double><<<(unsigned)batch, 256, 0, stream>>>((double2*)X, stride, count,
                                                                      floor_);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}
