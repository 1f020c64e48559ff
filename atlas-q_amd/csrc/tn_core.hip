// The reference's second two-site update, tools_qih/tn_core.py:mps_apply_2q
// (tn_core.py:186-384), and its learned-rank predictors
// (tools_qih/ai_rank_predictor.py:14-89, tools_qih/finetuned_predictor.py:8-81):
//
//   tn_truncate_kernel   keep count (adaptive rule, AI policy, fixed cap), the
//                        S > eps filter, kept / total weight and the
//                        truncation error, one work-group per spectrum
//   sqrt_split_kernel    cores[i] = U diag(sqrt S), cores[j] = diag(sqrt S) Vh
//   spectrum_features    the predictors' inputs computed from S on the device
//   dense_forward_kernel Linear (+LayerNorm) (+activation) stacks in FP32, one
//                        work-group per sample, activations in LDS
#include <math.h>

#include "aq_common.h"

namespace aq {

// ---------------------------------------------------------------- truncation
// Arithmetic follows the reference run by torch on the CPU: S**2 rounded to
// S's type, cumsum accumulated in double and rounded per element (torch's CPU
// cumsum), threshold (1 - tol) * total computed as float(1 - tol) * total in
// S's type, mask cumsum <= threshold, S > eps with eps rounded to S's type.
// The policy blend round(alpha * ai + (1 - alpha) * adaptive) is evaluated in
// double without contraction and rounded half to even (Python's round()).
template <typename T>
__global__ __launch_bounds__(256) void tn_truncate_kernel(const T* __restrict__ S, int64_t m,
                                                          T one_minus_tol, int use_adaptive,
                                                          int64_t chi_max, int64_t ai_keep,
                                                          int policy, double alpha, T eps,
                                                          aq_tn_trunc_t* out) {
  constexpr int CH = 2048;
  __shared__ T sw[CH];
  __shared__ int64_t s_cnt;
  __shared__ int s_prefix;
  const int tid = threadIdx.x;

  // pass 1: total = cumsum[m-1]
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
      sw[i] = w;
    }
    __syncthreads();
    if (tid == 0)
      for (int64_t i = 0; i < len; ++i) acc += (double)sw[i];
  }
  __shared__ T s_total;
  if (tid == 0) s_total = (T)acc;
  __syncthreads();
  const T total = s_total;
  const T thresh = one_minus_tol * total;

  // pass 2: number of prefix sums <= thresh (the cumsum is non-decreasing, so
  // the mask is a prefix and its count is the first index above thresh)
  if (tid == 0) s_cnt = -1;
  acc = 0.0;
  for (int64_t base = 0; base < m; base += CH) {
    const int64_t len = (m - base) < CH ? (m - base) : CH;
    __syncthreads();
    if (s_cnt >= 0) break;
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
      for (int64_t i = 0; i < len; ++i) {
        acc += (double)sw[i];
        if ((T)acc > thresh) {
          s_cnt = base + i;
          break;
        }
      }
    }
  }
  __syncthreads();
  const int64_t cnt = s_cnt < 0 ? m : s_cnt;  // entries with cumsum <= thresh

  // adaptive keep: count + 1 (at most m), 1 when no entry passes; capped
  int64_t adaptive_keep = m;
  if (use_adaptive) {
    adaptive_keep = cnt > 0 ? (cnt + 1 < m ? cnt + 1 : m) : 1;
    if (chi_max >= 0 && adaptive_keep > chi_max) adaptive_keep = chi_max;
  }
  int64_t keep;
  if (ai_keep >= 0) {
    int64_t blend = 0;
    if (policy == AQ_TN_POLICY_BLEND || policy == AQ_TN_POLICY_CLAMP_BLEND) {
      const double a = __dmul_rn(alpha, (double)ai_keep);
      const double b = __dmul_rn(__dsub_rn(1.0, alpha), (double)adaptive_keep);
      blend = (int64_t)rint(__dadd_rn(a, b));
    }
    switch (policy) {
      case AQ_TN_POLICY_BASELINE:
      case AQ_TN_POLICY_CLAMP:
        keep = ai_keep;
        break;
      case AQ_TN_POLICY_BLEND:
      case AQ_TN_POLICY_CLAMP_BLEND:
        keep = blend < adaptive_keep ? blend : adaptive_keep;
        break;
      default:  // guardrail
        keep = ai_keep < adaptive_keep ? ai_keep : adaptive_keep;
    }
    if (keep < 1) keep = 1;
  } else if (use_adaptive) {
    keep = adaptive_keep;
  } else {
    keep = (chi_max >= 0 && m > chi_max) ? chi_max : m;
  }
  if (keep > m) keep = m;

  // kept weight, S > eps count and whether the survivors form a prefix
  double kept_acc = 0.0;
  int64_t chi = 0;
  if (tid == 0) s_prefix = 1;
  __syncthreads();
  for (int64_t base = 0; base < keep; base += CH) {
    const int64_t len = (keep - base) < CH ? (keep - base) : CH;
    __syncthreads();
    for (int64_t i = tid; i < CH; i += 256) sw[i] = i < len ? S[base + i] : (T)0;
    __syncthreads();
    if (tid == 0) {
      for (int64_t i = 0; i < len; ++i) {
        const T s = sw[i];
        kept_acc += (double)(T)(s * s);
        if (s > eps) {
          if (chi != base + i) s_prefix = 0;
          ++chi;
        }
      }
    }
  }
  if (tid == 0) {
    const T kept = (T)kept_acc;
    out->keep_count = keep;
    out->chi = chi;
    out->adaptive_keep = adaptive_keep;
    out->prefix = s_prefix;
    out->pad = 0;
    out->total_weight = (double)total;
    out->kept_weight = (double)kept;
    out->truncation_error = total > (T)0 ? 1.0 - (double)(T)(kept / total) : 0.0;
  }
}

// ---------------------------------------------------------------- sqrt split
// dst[r, c] = src[r*s0 + c*s1] * sqrt(S[c]) (scale_cols) or * sqrt(S[r]),
// 32x32 tiles through LDS so the read (either storage order) and the
// row-major write are both coalesced.
template <typename T>
__global__ __launch_bounds__(256) void sqrt_split_kernel(const c2_t<T>* __restrict__ src,
                                                         int64_t s0, int64_t s1,
                                                         const T* __restrict__ S, int scale_cols,
                                                         c2_t<T>* __restrict__ dst, int64_t R,
                                                         int64_t Cn) {
  __shared__ c2_t<T> tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const bool col_major = (s0 == 1);
  for (int i = ty; i < 32; i += 8) {
    const int64_t r = col_major ? r0 + tx : r0 + i;
    const int64_t c = col_major ? c0 + i : c0 + tx;
    if (r < R && c < Cn) {
      c2_t<T> v = src[r * s0 + c * s1];
      const T sc = sqrt(S[scale_cols ? c : r]);
      v.x *= sc;
      v.y *= sc;
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

// ---------------------------------------------------------------- features
// AQ_FEAT_LOG1P64 (ai_rank_predictor.py:55-69,79-85): n_sv = #(S > 1e-10),
//   n = min(limit (n_sv when limit < 0), m, 64), x[i] = log1p(S[i] / S[0])
//   for i < n (0 elsewhere, all zero when S[0] <= 0); 64 floats.
// AQ_FEAT_FT135 (finetuned_predictor.py:51-75): sv = S[:128] zero-padded as
//   float32, mass = sv^2 / (sum sv^2 + 1e-12), entropy H of the mass, logs =
//   log(sv + 1e-12) - max, cumulative mass at 8,16,32,64,96,128; 135 floats.
// AQ_FEAT_SPEC32 (no reference counterpart; the inputs of this library's own
//   spectral predictor): with p_i = S_i^2 / sum S^2 over the leading <= 128
//   values, log10 max(p_i, 1e-12) for i < 24 (-12 past the spectrum), log10
//   max(1 - sum_{i<c} p_i, 1e-12) at c = 8,16,32,64,96,128, the entropy
//   -sum p ln p, log2 max(1, #(S_i > 1e-10)); 32 floats, all-zero tail
//   features for an all-zero spectrum.
// One work-group of 128 threads per spectrum.
template <typename T>
__global__ __launch_bounds__(128) void spectrum_features_kernel(const T* __restrict__ S, int64_t m,
                                                                int kind, int64_t limit,
                                                                float* __restrict__ feat,
                                                                int64_t* __restrict__ n_sv) {
  __shared__ float red[128];
  __shared__ int64_t cnt_s[128];
  const int tid = threadIdx.x;
  int64_t c = 0;
  for (int64_t i = tid; i < m; i += 128) c += (S[i] > (T)1e-10) ? 1 : 0;
  cnt_s[tid] = c;
  __syncthreads();
  for (int w = 64; w > 0; w >>= 1) {
    if (tid < w) cnt_s[tid] += cnt_s[tid + w];
    __syncthreads();
  }
  const int64_t nsv = cnt_s[0];
  if (tid == 0) *n_sv = nsv;
  if (kind == AQ_FEAT_LOG1P64) {
    const int64_t lim = limit >= 0 ? limit : nsv;
    const int64_t n = lim < 64 ? (lim < m ? lim : m) : (m < 64 ? m : 64);
    const T s0 = S[0];
    if (tid < 64) {
      float x = 0.0f;
      if (tid < n && s0 > (T)0) x = (float)log1p((T)(S[tid] / s0));
      feat[tid] = x;
    }
    return;
  }
  if (kind == AQ_FEAT_SPEC32) {
    // weights p_i = S_i^2 / total over the leading (<= 128) values
    __shared__ double dred[128];
    double part = 0.0;
    for (int64_t i = tid; i < m; i += 128) {
      const double v = (double)S[i];
      part += v * v;
    }
    dred[tid] = part;
    __syncthreads();
    for (int w = 64; w > 0; w >>= 1) {
      if (tid < w) dred[tid] += dred[tid + w];
      __syncthreads();
    }
    const double tot = dred[0];
    __syncthreads();
    const double v = tid < m ? (double)S[tid] : 0.0;
    const double p = tot > 0.0 ? v * v / tot : 0.0;
    dred[tid] = p;
    if (tid < 24) feat[tid] = (tot > 0.0 && tid < m) ? (float)log10(fmax(p, 1e-12)) : -12.0f;
    __syncthreads();
    if (tid == 0) {
      if (tot <= 0.0) {
        for (int q = 24; q < 32; ++q) feat[q] = 0.0f;
      } else {
        const int cuts[6] = {8, 16, 32, 64, 96, 128};
        const int lead = m < 128 ? (int)m : 128;
        double run = 0.0, ent = 0.0;
        int ci = 0, nz = 0;
        for (int i = 0; i < lead; ++i) {
          const double pi = dred[i];
          run += pi;
          if (pi > 0.0) ent -= pi * log(pi);
          if ((T)S[i] > (T)1e-10) ++nz;
          while (ci < 6 && (i + 1 == cuts[ci] || (i + 1 == lead && cuts[ci] > lead))) {
            feat[24 + ci] = (float)log10(fmax(1.0 - run, 1e-12));
            ++ci;
          }
        }
        feat[30] = (float)ent;
        feat[31] = (float)log2((double)(nz > 1 ? nz : 1));
      }
    }
    return;
  }
  // AQ_FEAT_FT135
  const float sv = tid < m ? (float)S[tid] : 0.0f;
  const float sv2 = sv * sv;
  red[tid] = sv2;
  __syncthreads();
  for (int w = 64; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const float sw = red[0] + 1e-12f;
  __syncthreads();
  const float mass = sv2 / sw;
  red[tid] = mass > 0.0f ? mass * logf(mass + 1e-12f) : 0.0f;
  __syncthreads();
  for (int w = 64; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const float H = -red[0];
  __syncthreads();
  const float lg = logf(sv + 1e-12f);
  red[tid] = lg;
  __syncthreads();
  for (int w = 64; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmaxf(red[tid], red[tid + w]);
    __syncthreads();
  }
  const float lmax = red[0];
  __syncthreads();
  feat[tid] = lg - lmax;
  // cumulative mass landmarks: prefix sums of the mass in thread order
  red[tid] = mass;
  __syncthreads();
  if (tid == 0) {
    const int cuts[6] = {8, 16, 32, 64, 96, 128};
    float run = 0.0f;
    int ci = 0;
    for (int i = 0; i < 128; ++i) {
      run += red[i];
      if (i + 1 == cuts[ci]) {
        feat[129 + ci] = run;
        ++ci;
      }
    }
    feat[128] = H;
  }
}

// ---------------------------------------------------------------- dense MLP
constexpr int kDenseMaxDim = 512;
constexpr int kDenseMaxLayers = 8;

struct DenseStack {
  aq_dense_layer_t layer[kDenseMaxLayers];
  int n;
};

__device__ __forceinline__ float act_apply(int act, float z) {
  switch (act) {
    case AQ_ACT_RELU: return z > 0.0f ? z : 0.0f;
    case AQ_ACT_GELU: return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
    case AQ_ACT_SIGMOID: return 1.0f / (1.0f + expf(-z));
    default: return z;
  }
}

__global__ __launch_bounds__(256) void dense_forward_kernel(DenseStack st,
                                                            const float* __restrict__ x,
                                                            int64_t d0, float* __restrict__ y,
                                                            int64_t d_last) {
  __shared__ float buf[2][kDenseMaxDim];
  __shared__ float red[2][4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t s = blockIdx.x;
  for (int i = tid; i < d0; i += 256) buf[0][i] = x[s * d0 + i];
  __syncthreads();
  int cur = 0;
  for (int l = 0; l < st.n; ++l) {
    const aq_dense_layer_t L = st.layer[l];
    const float* in = buf[cur];
    float* o = buf[cur ^ 1];
    // one wave per output row, lanes across the input
    for (int r = wave; r < L.d_out; r += 4) {
      const float* w = L.W + (int64_t)r * L.d_in;
      float acc = 0.0f;
      for (int k = lane; k < L.d_in; k += 64) acc += w[k] * in[k];
      for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
      if (lane == 0) o[r] = acc + (L.b ? L.b[r] : 0.0f);
    }
    __syncthreads();
    if (L.ln_g) {
      // LayerNorm over the d_out outputs: biased variance, then affine
      float sum = 0.0f, sq = 0.0f;
      for (int i = tid; i < L.d_out; i += 256) sum += o[i];
      for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
      if (lane == 0) red[0][wave] = sum;
      __syncthreads();
      const float mean = (red[0][0] + red[0][1] + red[0][2] + red[0][3]) / (float)L.d_out;
      for (int i = tid; i < L.d_out; i += 256) {
        const float d = o[i] - mean;
        sq += d * d;
      }
      for (int off = 32; off >= 1; off >>= 1) sq += __shfl_xor(sq, off, 64);
      if (lane == 0) red[1][wave] = sq;
      __syncthreads();
      const float var = (red[1][0] + red[1][1] + red[1][2] + red[1][3]) / (float)L.d_out;
      const float inv = 1.0f / sqrtf(var + (float)L.ln_eps);
      for (int i = tid; i < L.d_out; i += 256)
        o[i] = (o[i] - mean) * inv * L.ln_g[i] + (L.ln_b ? L.ln_b[i] : 0.0f);
      __syncthreads();
    }
    for (int i = tid; i < L.d_out; i += 256) o[i] = act_apply(L.act, o[i]);
    __syncthreads();
    cur ^= 1;
  }
  for (int i = tid; i < d_last; i += 256) y[s * d_last + i] = buf[cur][i];
}

}  // namespace aq

using namespace aq;

extern "C" int aq_tn_truncate(int s_dtype, const void* S, int64_t m, double tol, int adaptive,
                              int64_t chi_max, int64_t ai_keep, int policy, double alpha,
                              double eps, aq_tn_trunc_t* out, hipStream_t stream) {
  AQ_REQUIRE(S && out, AQ_ERR_ARG, "aq_tn_truncate: null pointer");
  AQ_REQUIRE(m > 0, AQ_ERR_ARG, "aq_tn_truncate: empty spectrum");
  AQ_REQUIRE(policy >= AQ_TN_POLICY_BASELINE && policy <= AQ_TN_POLICY_CLAMP_BLEND, AQ_ERR_ARG,
             "aq_tn_truncate: unknown policy %d", policy);
  const int use_adaptive = (adaptive && tol > 0.0) ? 1 : 0;
  if (s_dtype == AQ_C64) {
    tn_truncate_kernel<float><<<1, 256, 0, stream>>>((const float*)S, m, (float)(1.0 - tol),
                                                     use_adaptive, chi_max, ai_keep, policy,
                                                     alpha, (float)eps, out);
  } else {
    AQ_REQUIRE(s_dtype == AQ_C128, AQ_ERR_ARG, "aq_tn_truncate: unknown dtype %d", s_dtype);
    tn_truncate_kernel<double><<<1, 256, 0, stream>>>((const double*)S, m, 1.0 - tol,
                                                      use_adaptive, chi_max, ai_keep, policy,
                                                      alpha, eps, out);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_rebuild_sqrt(int dtype, const void* U, int64_t u_s0, int64_t u_s1,
                               const void* S, const void* Vh, int64_t v_s0, int64_t v_s1,
                               int64_t rows, int64_t cols, int64_t k, void* A_out, void* B_out,
                               hipStream_t stream) {
  AQ_REQUIRE(U && S && Vh && A_out && B_out, AQ_ERR_ARG, "aq_rebuild_sqrt: null pointer");
  AQ_REQUIRE(rows > 0 && cols > 0 && k > 0, AQ_ERR_ARG, "aq_rebuild_sqrt: sizes must be positive");
  AQ_REQUIRE(rows % 2 == 0 && cols % 2 == 0, AQ_ERR_ARG, "aq_rebuild_sqrt: rows/cols must be even");
  AQ_REQUIRE(cdiv(rows, 32) < 65536 && cdiv(k, 32) < 65536, AQ_ERR_ARG,
             "aq_rebuild_sqrt: sizes beyond the grid limit");
  dim3 ga((unsigned)cdiv(k, 32), (unsigned)cdiv(rows, 32));
  dim3 gb((unsigned)cdiv(cols, 32), (unsigned)cdiv(k, 32));
  if (dtype == AQ_C64) {
    sqrt_split_kernel<float><<<ga, 256, 0, stream>>>((const float2*)U, u_s0, u_s1,
                                                     (const float*)S, 1, (float2*)A_out, rows, k);
    AQ_LAUNCH_CHECK();
    sqrt_split_kernel<float><<<gb, 256, 0, stream>>>((const float2*)Vh, v_s0, v_s1,
                                                     (const float*)S, 0, (float2*)B_out, k, cols);
  } else {
    AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_rebuild_sqrt: unknown dtype %d", dtype);
    sqrt_split_kernel<double><<<ga, 256, 0, stream>>>((const double2*)U, u_s0, u_s1,
                                                      (const double*)S, 1, (double2*)A_out, rows,
                                                      k);
    AQ_LAUNCH_CHECK();
    sqrt_split_kernel<double><<<gb, 256, 0, stream>>>((const double2*)Vh, v_s0, v_s1,
                                                      (const double*)S, 0, (double2*)B_out, k,
                                                      cols);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_spectrum_features(int kind, int s_dtype, const void* S, int64_t m,
                                    int64_t limit, float* feat, int64_t* n_sv,
                                    hipStream_t stream) {
  AQ_REQUIRE(S && feat && n_sv, AQ_ERR_ARG, "aq_spectrum_features: null pointer");
  AQ_REQUIRE(m > 0, AQ_ERR_ARG, "aq_spectrum_features: empty spectrum");
  AQ_REQUIRE(kind == AQ_FEAT_LOG1P64 || kind == AQ_FEAT_FT135 || kind == AQ_FEAT_SPEC32,
             AQ_ERR_ARG,
             "aq_spectrum_features: unknown kind %d", kind);
  if (s_dtype == AQ_C64) {
    spectrum_features_kernel<float><<<1, 128, 0, stream>>>((const float*)S, m, kind, limit, feat,
                                                           n_sv);
  } else {
    AQ_REQUIRE(s_dtype == AQ_C128, AQ_ERR_ARG, "aq_spectrum_features: unknown dtype %d", s_dtype);
    spectrum_features_kernel<double><<<1, 128, 0, stream>>>((const double*)S, m, kind, limit,
                                                            feat, n_sv);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_dense_forward(int n_layers, const aq_dense_layer_t* layers, const float* x,
                                int64_t batch, float* y, hipStream_t stream) {
  AQ_REQUIRE(layers && x && y, AQ_ERR_ARG, "aq_dense_forward: null pointer");
  AQ_REQUIRE(n_layers >= 1 && n_layers <= kDenseMaxLayers, AQ_ERR_ARG,
             "aq_dense_forward: 1..%d layers supported (got %d)", kDenseMaxLayers, n_layers);
  AQ_REQUIRE(batch >= 1 && batch < (1LL << 31), AQ_ERR_ARG, "aq_dense_forward: bad batch");
  DenseStack st;
  st.n = n_layers;
  for (int l = 0; l < n_layers; ++l) {
    const aq_dense_layer_t& L = layers[l];
    AQ_REQUIRE(L.W != nullptr, AQ_ERR_ARG, "aq_dense_forward: layer %d has no weights", l);
    AQ_REQUIRE(L.d_in >= 1 && L.d_in <= kDenseMaxDim && L.d_out >= 1 && L.d_out <= kDenseMaxDim,
               AQ_ERR_ARG, "aq_dense_forward: layer %d dims %d x %d outside 1..%d", l, L.d_out,
               L.d_in, kDenseMaxDim);
    AQ_REQUIRE(l == 0 || L.d_in == layers[l - 1].d_out, AQ_ERR_ARG,
               "aq_dense_forward: layer %d input %d != previous output %d", l, L.d_in,
               layers[l - 1].d_out);
    AQ_REQUIRE(L.act >= AQ_ACT_NONE && L.act <= AQ_ACT_SIGMOID, AQ_ERR_ARG,
               "aq_dense_forward: layer %d unknown activation %d", l, L.act);
    st.layer[l] = L;
  }
  dense_forward_kernel<<<(unsigned)batch, 256, 0, stream>>>(st, x, layers[0].d_in, y,
                                                            layers[n_layers - 1].d_out);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}
