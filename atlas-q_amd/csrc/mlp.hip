// Rank-predictor MLP: 32 -> 128 -> 128 -> 1, GELU (erf form), BF16 weights,
// FP32 activations / accumulation, FP32 master weights for AdamW.
//
// Forward (per sample s):  z1 = W1 x + b1, h1 = gelu(z1)
//                          z2 = W2 h1 + b2, h2 = gelu(z2)
//                          y  = w3 . h2 + b3
// Backward from dy (dL/dy per sample): gradients of all weights / biases,
// reduced deterministically (per-block partials, then a fixed-order sum).
//
// One work-group of 128 threads (two waves) handles a tile of 32 samples.  The
// layer products run on v_mfma_f32_32x32x16_bf16 (32 x 32 output blocks, 16 k
// per instruction, FP32 accumulation): the weights enter as their BF16
// copies, exactly; an FP32 activation or gradient operand is split into
// BF16 hi + lo (x = hi + lo to 2^-16 relative), so a product with a weight is
// two MFMAs and a weight gradient (both operands FP32) the three-product
// BF16x3 form.  Layer sizes are fixed by the model, so every index is a
// compile-time bound.
#include <math.h>

#include "aq_common.h"
#include "bf16x3.h"

namespace aq {

constexpr int kIn = 32, kH = 128, kTile = 32;

// parameter block (FP32 master) layout, also the gradient layout
//   W1 [kH][kIn] | b1 [kH] | W2 [kH][kH] | b2 [kH] | W3 [kH] | b3 [1]
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + kH * kIn;
constexpr int oW2 = oB1 + kH;
constexpr int oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH;
constexpr int oB3 = oW3 + kH;
constexpr int kNParam = oB3 + 1;

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}
__device__ __forceinline__ float bf16r(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// LDS of a tile.  Row-major planes with k contiguous and 8 elements of row
// padding (16-byte aligned rows for the MFMA operand loads).
constexpr int kLdIn = kIn + 8;
constexpr int kLdH = kH + 8;

struct MlpSmem {
  uint16_t w1[kH][kLdIn];          // W1 (BF16): [unit][input]
  uint16_t w2[kH][kLdH];           // W2 (BF16): [output unit][input unit]
  uint16_t x[2][kTile][kLdIn];     // X hi / lo: [sample][input]
  uint16_t h1[2][kTile][kLdH];     // gelu(z1) hi / lo: [sample][unit]
  float z1[kTile][kH + 1];
  float z2[kTile][kH + 1];         // the backward pass overwrites it with dL/dz2
  float w3[kH];
  float dys[kTile];
};
struct MlpBwdSmem {
  MlpSmem f;
  uint16_t d2[2][kTile][kLdH];     // dL/dz2 hi / lo: [sample][unit]
  uint16_t d1[2][kTile][kLdH];     // dL/dz1 hi / lo: [sample][unit]
};

// MFMA operand fragments: lane (r = lane & 31, q = lane >> 5) holds 8
// consecutive k (k0 + 8q .. k0 + 8q + 7) of row / column r.
__device__ __forceinline__ bf16x8 frag_rows(const uint16_t* base, int ld, int k0) {
  const int lane = threadIdx.x & 63;
  return load8(base + (lane & 31) * ld + k0 + 8 * (lane >> 5));
}
// the same from a [k][r] layout (k strided by ld)
__device__ __forceinline__ bf16x8 frag_cols(const uint16_t* base, int ld, int k0) {
  const int lane = threadIdx.x & 63;
  const uint16_t* p = base + (int64_t)(k0 + 8 * (lane >> 5)) * ld + (lane & 31);
  i16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (short)p[e * ld];
  return __builtin_bit_cast(bf16x8, v);
}
// accumulator register r of this lane -> (row, column) of the 32 x 32 block
__device__ __forceinline__ int acc_row(int r) {
  return 8 * (r >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (r & 3);
}
__device__ __forceinline__ int acc_col() { return threadIdx.x & 31; }

__device__ __forceinline__ void put_split(uint16_t (*pl)[kTile][kLdH], int s, int j, float v) {
  uint16_t h, l;
  split_bf16(v, h, l);
  pl[0][s][j] = h;
  pl[1][s][j] = l;
}

// weights (BF16 copies) and the tile's inputs into LDS
__device__ void load_tile(MlpSmem& sm, const uint16_t* __restrict__ wb,
                          const float* __restrict__ X, int64_t s0, int cnt) {
  for (int i = threadIdx.x; i < kH * kIn; i += blockDim.x) sm.w1[i / kIn][i % kIn] = wb[oW1 + i];
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) sm.w2[i / kH][i % kH] = wb[oW2 + i];
  for (int i = threadIdx.x; i < kH; i += blockDim.x) sm.w3[i] = bf16r(wb[oW3 + i]);
  for (int i = threadIdx.x; i < kTile * kIn; i += blockDim.x) {
    const int s = i / kIn, k = i % kIn;
    uint16_t h, l;
    split_bf16(s < cnt ? X[(s0 + s) * kIn + k] : 0.f, h, l);
    sm.x[0][s][k] = h;
    sm.x[1][s][k] = l;
  }
}

// forward of one tile on MFMA (wave w computes units 64 w .. 64 w + 63);
// leaves z1, h1, z2 in LDS and returns y for samples s < cnt
__device__ void tile_forward(MlpSmem& sm, const float* __restrict__ bias, int cnt,
                             float* __restrict__ y_out, int64_t s0) {
  const int wave = threadIdx.x >> 6;
  // layer 1: z1 = X W1^T + b1 (K = 32)
  for (int jb = 2 * wave; jb < 2 * wave + 2; ++jb) {
    f32x16 acc = {};
#pragma unroll
    for (int k0 = 0; k0 < kIn; k0 += 16) {
      const bf16x8 b = frag_rows(&sm.w1[32 * jb][0], kLdIn, k0);
      acc = mma_bf16(frag_rows(&sm.x[1][0][0], kLdIn, k0), b, acc);
      acc = mma_bf16(frag_rows(&sm.x[0][0][0], kLdIn, k0), b, acc);
    }
    const int j = 32 * jb + acc_col();
    const float b1 = bias[oB1 + j];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int s = acc_row(r);
      const float z = acc[r] + b1;
      sm.z1[s][j] = z;
      put_split(sm.h1, s, j, gelu_f(z));
    }
  }
  __syncthreads();
  // layer 2: z2 = H1 W2^T + b2 (K = 128)
  for (int ib = 2 * wave; ib < 2 * wave + 2; ++ib) {
    f32x16 acc = {};
#pragma unroll
    for (int k0 = 0; k0 < kH; k0 += 16) {
      const bf16x8 b = frag_rows(&sm.w2[32 * ib][0], kLdH, k0);
      acc = mma_bf16(frag_rows(&sm.h1[1][0][0], kLdH, k0), b, acc);
      acc = mma_bf16(frag_rows(&sm.h1[0][0][0], kLdH, k0), b, acc);
    }
    const int i = 32 * ib + acc_col();
    const float b2 = bias[oB2 + i];
#pragma unroll
    for (int r = 0; r < 16; ++r) sm.z2[acc_row(r)][i] = acc[r] + b2;
  }
  __syncthreads();
  // y = w3 . gelu(z2) + b3: thread s < cnt reduces sample s (fixed order)
  const int t = threadIdx.x;
  if (t < cnt && y_out) {
    float a = bias[oB3];
    for (int k = 0; k < kH; ++k) a += sm.w3[k] * gelu_f(sm.z2[t][k]);
    y_out[s0 + t] = a;
  }
}

__global__ __launch_bounds__(128) void mlp_forward_kernel(const float* __restrict__ X, int64_t B,
                                                          const uint16_t* __restrict__ wb,
                                                          const float* __restrict__ master,
                                                          float* __restrict__ y) {
  extern __shared__ __align__(16) unsigned char smraw[];
  MlpSmem& sm = *reinterpret_cast<MlpSmem*>(smraw);
  const int64_t s0 = (int64_t)blockIdx.x * kTile;
  const int cnt = (int)((B - s0) < kTile ? (B - s0) : kTile);
  load_tile(sm, wb, X, s0, cnt);
  __syncthreads();
  tile_forward(sm, master, cnt, y, s0);
}

// Per-block gradient partials: part[blockIdx.x][kNParam]
__global__ __launch_bounds__(128) void mlp_backward_kernel(const float* __restrict__ X,
                                                           const float* __restrict__ dy,
                                                           int64_t B, const uint16_t* __restrict__ wb,
                                                           const float* __restrict__ master,
                                                           float* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smraw[];
  MlpBwdSmem& bs = *reinterpret_cast<MlpBwdSmem*>(smraw);
  MlpSmem& sm = bs.f;
  const int64_t s0 = (int64_t)blockIdx.x * kTile;
  const int cnt = (int)((B - s0) < kTile ? (B - s0) : kTile);
  load_tile(sm, wb, X, s0, cnt);
  if (threadIdx.x < kTile) sm.dys[threadIdx.x] = (threadIdx.x < cnt) ? dy[s0 + threadIdx.x] : 0.f;
  __syncthreads();
  tile_forward(sm, master, cnt, nullptr, s0);
  const int j = threadIdx.x;            // unit j of the element-wise steps
  const int wave = threadIdx.x >> 6;
  float* P = part + (int64_t)blockIdx.x * kNParam;
  // layer 3 and dL/dz2 (rows past cnt zero)
  float gw3 = 0.f, gb3 = 0.f, gb2 = 0.f;
  for (int s = 0; s < kTile; ++s) {
    float d = 0.f;
    if (s < cnt) {
      const float z2 = sm.z2[s][j];
      gw3 += sm.dys[s] * gelu_f(z2);
      d = sm.dys[s] * sm.w3[j] * gelu_grad(z2);
      gb3 += sm.dys[s];
    }
    gb2 += d;
    put_split(bs.d2, s, j, d);
  }
  P[oW3 + j] = gw3;
  P[oB2 + j] = gb2;
  if (j == 0) P[oB3] = gb3;
  __syncthreads();
  // layer 2 weights: dW2[i][k] = sum_s d2[s][i] h1[s][k] (K = 32 samples, BF16x3)
  for (int blk = wave; blk < 16; blk += 2) {
    const int ib = blk >> 2, kb = blk & 3;
    f32x16 acc = {};
#pragma unroll
    for (int k0 = 0; k0 < kTile; k0 += 16) {
      const bf16x8 ah = frag_cols(&bs.d2[0][0][32 * ib], kLdH, k0);
      const bf16x8 al = frag_cols(&bs.d2[1][0][32 * ib], kLdH, k0);
      const bf16x8 bh = frag_cols(&sm.h1[0][0][32 * kb], kLdH, k0);
      const bf16x8 bl = frag_cols(&sm.h1[1][0][32 * kb], kLdH, k0);
      acc = mma_bf16(ah, bl, acc);
      acc = mma_bf16(al, bh, acc);
      acc = mma_bf16(ah, bh, acc);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
      P[oW2 + (32 * ib + acc_row(r)) * kH + 32 * kb + acc_col()] = acc[r];
  }
  // dL/dz1[s][k] = gelu'(z1[s][k]) sum_i d2[s][i] W2[i][k] (K = 128 units)
  for (int kb = 2 * wave; kb < 2 * wave + 2; ++kb) {
    f32x16 acc = {};
#pragma unroll
    for (int k0 = 0; k0 < kH; k0 += 16) {
      const bf16x8 b = frag_cols(&sm.w2[0][32 * kb], kLdH, k0);
      acc = mma_bf16(frag_rows(&bs.d2[1][0][0], kLdH, k0), b, acc);
      acc = mma_bf16(frag_rows(&bs.d2[0][0][0], kLdH, k0), b, acc);
    }
    const int k = 32 * kb + acc_col();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int s = acc_row(r);
      put_split(bs.d1, s, k, s < cnt ? acc[r] * gelu_grad(sm.z1[s][k]) : 0.f);
    }
  }
  __syncthreads();
  float gb1 = 0.f;
  for (int s = 0; s < kTile; ++s) gb1 += bf16_to_f32(bs.d1[0][s][j]) + bf16_to_f32(bs.d1[1][s][j]);
  P[oB1 + j] = gb1;
  // layer 1 weights: dW1[k][c] = sum_s d1[s][k] x[s][c] (K = 32 samples, BF16x3)
  for (int kb = 2 * wave; kb < 2 * wave + 2; ++kb) {
    f32x16 acc = {};
#pragma unroll
    for (int k0 = 0; k0 < kTile; k0 += 16) {
      const bf16x8 ah = frag_cols(&bs.d1[0][0][32 * kb], kLdH, k0);
      const bf16x8 al = frag_cols(&bs.d1[1][0][32 * kb], kLdH, k0);
      const bf16x8 bh = frag_cols(&sm.x[0][0][0], kLdIn, k0);
      const bf16x8 bl = frag_cols(&sm.x[1][0][0], kLdIn, k0);
      acc = mma_bf16(ah, bl, acc);
      acc = mma_bf16(al, bh, acc);
      acc = mma_bf16(ah, bh, acc);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
      P[oW1 + (32 * kb + acc_row(r)) * kIn + acc_col()] = acc[r];
  }
}

// grad[p] = sum over blocks of part[b][p] (fixed order)
__global__ void mlp_reduce_kernel(const float* __restrict__ part, int64_t nblocks,
                                  float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  float a = 0.f;
  for (int64_t b = 0; b < nblocks; ++b) a += part[b * kNParam + p];
  grad[p] = a;
}

__device__ __forceinline__ uint16_t to_bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// AdamW (decoupled weight decay on weights, not biases), then refresh the
// BF16 copies.
__global__ void adamw_kernel(float* __restrict__ w, const float* __restrict__ g,
                             float* __restrict__ m, float* __restrict__ v,
                             uint16_t* __restrict__ wb, float lr, float beta1, float beta2,
                             float eps, float wd, float grad_scale, int step) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  const bool is_bias = (p >= oB1 && p < oW2) || (p >= oB2 && p < oW3) || p == oB3;
  const float gi = g[p] * grad_scale;
  const float mi = beta1 * m[p] + (1.f - beta1) * gi;
  const float vi = beta2 * v[p] + (1.f - beta2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  const float bc1 = 1.f - powf(beta1, (float)step);
  const float bc2 = 1.f - powf(beta2, (float)step);
  float wi = w[p];
  if (!is_bias) wi -= lr * wd * wi;
  wi -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[p] = wi;
  wb[p] = to_bf16_rne(wi);
}

__global__ void to_bf16_kernel(const float* __restrict__ w, uint16_t* __restrict__ wb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNParam) wb[p] = to_bf16_rne(w[p]);
}

static bool g_attr_set = false;
static int set_attrs() {
  if (g_attr_set) return AQ_OK;
  AQ_HIP(hipFuncSetAttribute((const void*)mlp_forward_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize, sizeof(MlpSmem)));
  AQ_HIP(hipFuncSetAttribute((const void*)mlp_backward_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize, sizeof(MlpBwdSmem)));
  g_attr_set = true;
  return AQ_OK;
}

}  // namespace aq

using namespace aq;

extern "C" int aq_mlp_num_params(void) { return kNParam; }

extern "C" int aq_mlp_sync_bf16(const float* master, uint16_t* wb, hipStream_t stream) {
  AQ_REQUIRE(master && wb, AQ_ERR_ARG, "aq_mlp_sync_bf16: null pointer");
  to_bf16_kernel<<<(kNParam + 255) / 256, 256, 0, stream>>>(master, wb);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_mlp_forward(const float* X, int64_t batch, const uint16_t* wb,
                              const float* master, float* y, hipStream_t stream) {
  AQ_REQUIRE(X && wb && master && y, AQ_ERR_ARG, "aq_mlp_forward: null pointer");
  AQ_REQUIRE(batch > 0, AQ_ERR_ARG, "aq_mlp_forward: empty batch");
  int rc = set_attrs();
  if (rc) return rc;
  const unsigned blocks = (unsigned)cdiv(batch, kTile);
  mlp_forward_kernel<<<blocks, 128, sizeof(MlpSmem), stream>>>(X, batch, wb, master, y);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int64_t aq_mlp_workspace_floats(int64_t batch) {
  return cdiv(batch, kTile) * (int64_t)kNParam;
}

extern "C" int aq_mlp_backward(const float* X, const float* dy, int64_t batch, const uint16_t* wb,
                               const float* master, float* grad, float* workspace,
                               int64_t workspace_floats, hipStream_t stream) {
  AQ_REQUIRE(X && dy && wb && master && grad && workspace, AQ_ERR_ARG,
             "aq_mlp_backward: null pointer");
  AQ_REQUIRE(batch > 0, AQ_ERR_ARG, "aq_mlp_backward: empty batch");
  const int64_t nb = cdiv(batch, kTile);
  AQ_REQUIRE(workspace_floats >= nb * kNParam, AQ_ERR_ARG, "aq_mlp_backward: workspace too small");
  int rc = set_attrs();
  if (rc) return rc;
  mlp_backward_kernel<<<(unsigned)nb, 128, sizeof(MlpBwdSmem), stream>>>(X, dy, batch, wb,
                                                                         master, workspace);
  AQ_LAUNCH_CHECK();
  mlp_reduce_kernel<<<(kNParam + 255) / 256, 256, 0, stream>>>(workspace, nb, grad);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_mlp_adamw(float* master, const float* grad, float* m, float* v, uint16_t* wb,
                            double lr, double beta1, double beta2, double eps, double wd,
                            double grad_scale, int step, hipStream_t stream) {
  AQ_REQUIRE(master && grad && m && v && wb, AQ_ERR_ARG, "aq_mlp_adamw: null pointer");
  AQ_REQUIRE(step >= 1, AQ_ERR_ARG, "aq_mlp_adamw: step counts from 1");
  adamw_kernel<<<(kNParam + 255) / 256, 256, 0, stream>>>(
      master, grad, m, v, wb, (float)lr, (float)beta1, (float)beta2, (float)eps, (float)wd,
      (float)grad_scale, step);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}
