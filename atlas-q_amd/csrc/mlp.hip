// Rank-predictor MLP: 32 -> 128 -> 128 -> 1, GELU (erf form), BF16 weights,
// FP32 activations / accumulation, FP32 master weights for AdamW.
//
// Forward (per sample s):  z1 = W1 x + b1, h1 = gelu(z1)
//                          z2 = W2 h1 + b2, h2 = gelu(z2)
//                          y  = w3 . h2 + b3
// Backward from dy (dL/dy per sample): gradients of all weights / biases,
// reduced deterministically (per-block partials, then a fixed-order sum).
//
// One work-group of 128 threads handles a tile of 32 samples; thread j owns
// hidden unit j.  The BF16 weights are converted to FP32 into LDS once per
// work-group.  Layer sizes are fixed by the model, so every index is a
// compile-time bound.
#include <math.h>

#include "aq_common.h"

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

struct MlpSmem {
  float w1[kH][kIn + 1];
  float w2[kH][kH + 1];
  float w3[kH];
  float x[kTile][kIn];
  float z1[kTile][kH];
  float h1[kTile][kH];
  float z2[kTile][kH];
};

// weights: bf16 copies (W1|W2|W3 in the same layout as the master, biases FP32)
__device__ void load_weights(MlpSmem& sm, const uint16_t* __restrict__ wb) {
  for (int i = threadIdx.x; i < kH * kIn; i += blockDim.x) sm.w1[i / kIn][i % kIn] = bf16r(wb[oW1 + i]);
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) sm.w2[i / kH][i % kH] = bf16r(wb[oW2 + i]);
  for (int i = threadIdx.x; i < kH; i += blockDim.x) sm.w3[i] = bf16r(wb[oW3 + i]);
}

// forward of one tile; leaves z1, z2 in LDS and returns y for samples s < cnt
__device__ void tile_forward(MlpSmem& sm, const float* __restrict__ bias, int cnt,
                             float* __restrict__ y_out, int64_t s0) {
  const int j = threadIdx.x;
  const float b1 = bias[oB1 + j], b2 = bias[oB2 + j];
  for (int s = 0; s < cnt; ++s) {
    float a = b1;
#pragma unroll
    for (int k = 0; k < kIn; ++k) a += sm.w1[j][k] * sm.x[s][k];
    sm.z1[s][j] = a;
    sm.h1[s][j] = gelu_f(a);
  }
  __syncthreads();
  for (int s = 0; s < cnt; ++s) {
    float a = b2;
    for (int k = 0; k < kH; ++k) a += sm.w2[j][k] * sm.h1[s][k];
    sm.z2[s][j] = a;
  }
  __syncthreads();
  // y = w3 . gelu(z2) + b3: thread s < cnt reduces sample s (fixed order)
  if (j < cnt && y_out) {
    float a = bias[oB3];
    for (int k = 0; k < kH; ++k) a += sm.w3[k] * gelu_f(sm.z2[j][k]);
    y_out[s0 + j] = a;
  }
}

__global__ __launch_bounds__(128) void mlp_forward_kernel(const float* __restrict__ X, int64_t B,
                                                          const uint16_t* __restrict__ wb,
                                                          const float* __restrict__ master,
                                                          float* __restrict__ y) {
  extern __shared__ unsigned char smraw[];
  MlpSmem& sm = *reinterpret_cast<MlpSmem*>(smraw);
  load_weights(sm, wb);
  const int64_t s0 = (int64_t)blockIdx.x * kTile;
  const int cnt = (int)((B - s0) < kTile ? (B - s0) : kTile);
  for (int i = threadIdx.x; i < kTile * kIn; i += blockDim.x) {
    const int s = i / kIn, k = i % kIn;
    sm.x[s][k] = (s < cnt) ? X[(s0 + s) * kIn + k] : 0.f;
  }
  __syncthreads();
  tile_forward(sm, master, cnt, y, s0);
}

// Per-block gradient partials: part[blockIdx.x][kNParam]
__global__ __launch_bounds__(128) void mlp_backward_kernel(const float* __restrict__ X,
                                                           const float* __restrict__ dy,
                                                           int64_t B, const uint16_t* __restrict__ wb,
                                                           const float* __restrict__ master,
                                                           float* __restrict__ part) {
  extern __shared__ unsigned char smraw[];
  MlpSmem& sm = *reinterpret_cast<MlpSmem*>(smraw);
  __shared__ float d2[kTile][kH];   // dL/dz2
  __shared__ float dys[kTile];
  load_weights(sm, wb);
  const int64_t s0 = (int64_t)blockIdx.x * kTile;
  const int cnt = (int)((B - s0) < kTile ? (B - s0) : kTile);
  for (int i = threadIdx.x; i < kTile * kIn; i += blockDim.x) {
    const int s = i / kIn, k = i % kIn;
    sm.x[s][k] = (s < cnt) ? X[(s0 + s) * kIn + k] : 0.f;
  }
  if (threadIdx.x < kTile) dys[threadIdx.x] = (threadIdx.x < cnt) ? dy[s0 + threadIdx.x] : 0.f;
  __syncthreads();
  tile_forward(sm, master, cnt, nullptr, s0);
  const int j = threadIdx.x;
  float* P = part + (int64_t)blockIdx.x * kNParam;
  // layer 3
  float gw3 = 0.f, gb3 = 0.f;
  for (int s = 0; s < cnt; ++s) {
    const float z2 = sm.z2[s][j];
    gw3 += dys[s] * gelu_f(z2);
    d2[s][j] = dys[s] * sm.w3[j] * gelu_grad(z2);
    gb3 += dys[s];
  }
  P[oW3 + j] = gw3;
  if (j == 0) P[oB3] = gb3;
  __syncthreads();
  // layer 2: dW2[j][k] = sum_s d2[s][j] h1[s][k]; thread j owns row j
  float gb2 = 0.f;
  for (int s = 0; s < cnt; ++s) gb2 += d2[s][j];
  P[oB2 + j] = gb2;
  for (int k = 0; k < kH; ++k) {
    float a = 0.f;
    for (int s = 0; s < cnt; ++s) a += d2[s][j] * sm.h1[s][k];
    P[oW2 + j * kH + k] = a;
  }
  // dL/dz1[s][j] = gelu'(z1[s][j]) * sum_i d2[s][i] W2[i][j]  (thread j = unit j)
  float d1[kTile];
  for (int s = 0; s < kTile; ++s) {
    float a = 0.f;
    if (s < cnt)
      for (int i = 0; i < kH; ++i) a += d2[s][i] * sm.w2[i][j];
    d1[s] = (s < cnt) ? a * gelu_grad(sm.z1[s][j]) : 0.f;
  }
  float gb1 = 0.f;
  for (int s = 0; s < cnt; ++s) gb1 += d1[s];
  P[oB1 + j] = gb1;
  for (int k = 0; k < kIn; ++k) {
    float a = 0.f;
    for (int s = 0; s < cnt; ++s) a += d1[s] * sm.x[s][k];
    P[oW1 + j * kIn + k] = a;
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
                             hipFuncAttributeMaxDynamicSharedMemorySize, sizeof(MlpSmem)));
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
  mlp_backward_kernel<<<(unsigned)nb, 128, sizeof(MlpSmem), stream>>>(X, dy, batch, wb, master,
                                                                      workspace);
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
