// Cholesky QR of tall-skinny complex128 matrices on the library's own kernels
// (FP64): the orthonormalisation of the randomized SVD's range-finder
// sketches.
//
//   W = Y^H Y (+ shift I)                 split-K MFMA GEMM (the reduction
//                                          over the m rows spread over up
//                                          to kCqSplits work-group columns)
//   W = L L^H                              blocked by 64: each diagonal block
//                                          factored and inverted in LDS by
//                                          potrf_small_kernel, the panel below
//                                          and the trailing block by GEMMs
//   Q = Y L^{-H}                           blocked forward substitution over
//                                          column blocks of 64:
//                                          Q_J = (Y_J - Q_{<J} L_{J,<J}^H) L_JJ^{-H}
//
// One call is one pass; the caller chains passes (a shifted pass, then two
// plain ones: shiftedCholeskyQR3, Fukaya, Kannan, Nakatsukasa, Yamamoto,
// Yanagisawa, SIAM J. Sci. Comput. 42 (2020)).  A pivot that is not positive
// and finite sets *info (1 + its index) and is replaced by 1, so the pass
// finishes with finite numbers and the caller discards the result.
#include <math.h>

#include <algorithm>

#include "aq_common.h"

extern "C" int aq_bgemm(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                        int64_t a_s0, int64_t a_s1, int64_t a_sb, int conj_a, const void* B,
                        int64_t b_s0, int64_t b_s1, int64_t b_sb, int conj_b, double beta,
                        void* C, int64_t c_s0, int64_t c_s1, int64_t c_sb, int64_t batch,
                        hipStream_t stream);
extern "C" int aq_cgemm_splitk(int dtype, int prec, int64_t M, int64_t N, int64_t K,
                               const void* A, int64_t a_s0, int64_t a_s1, int conj_a,
                               const void* B, int64_t b_s0, int64_t b_s1, int conj_b, void* C,
                               int64_t c_s0, int64_t c_s1, int64_t splits, void* workspace,
                               int64_t workspace_bytes, hipStream_t stream);

namespace aq {

constexpr int kCqB = 64;        // block order
constexpr int kCqThreads = 256;
constexpr int kCqSplits = 16;   // most K chunks of the Gram matrix

// K chunks of W = Y^H Y: about 128 rows each, at most kCqSplits, and the
// partial products within 2^24 complex elements
__host__ inline int64_t cq_splits(int64_t m, int64_t l) {
  int64_t sp = std::max<int64_t>(1, std::min<int64_t>(kCqSplits, m / 128));
  while (sp > 1 && sp * l * l > (1LL << 24)) --sp;
  return sp;
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zmul_conj_b(double2 a, double2 b) {  // a conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// Cholesky factor and its inverse of the leading n x n block (n <= 64) of a
// Hermitian positive definite matrix W (row-major, lower triangle read):
// L (lower, W = L L^H) and Li = L^{-1} (lower), both written with zero upper
// triangles; one work-group, everything in LDS.
__global__ __launch_bounds__(kCqThreads) void potrf_small_kernel(
    const double2* __restrict__ W, int64_t ldw, double2* __restrict__ L, int64_t ldl,
    double2* __restrict__ Li, int64_t ldi, int n, int info_base, int* __restrict__ info) {
  __shared__ double2 a[kCqB][kCqB + 1];
  __shared__ double2 x[kCqB][kCqB + 1];
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = 0;
  for (int e = tid; e < n * n; e += kCqThreads) {
    const int r = e / n, c = e % n;
    a[r][c] = r >= c ? W[(int64_t)r * ldw + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      double d = a[j][j].x;
      if (!(d > 0.0) || !isfinite(d)) {
        if (!bad) bad = info_base + j + 1;
        d = 1.0;
      }
      a[j][j] = make_double2(sqrt(d), 0.0);
    }
    __syncthreads();
    const double inv = 1.0 / a[j][j].x;
    for (int i = j + 1 + tid; i < n; i += kCqThreads) a[i][j] = make_double2(a[i][j].x * inv,
                                                                             a[i][j].y * inv);
    __syncthreads();
    const int t = n - j - 1;
    for (int e = tid; e < t * t; e += kCqThreads) {
      const int i = j + 1 + e / t, c = j + 1 + e % t;
      if (c <= i) {
        const double2 p = zmul_conj_b(a[i][j], a[c][j]);
        a[i][c] = make_double2(a[i][c].x - p.x, a[i][c].y - p.y);
      }
    }
    __syncthreads();
  }
  // L^{-1} row by row: x[i][c] = (delta_ic - sum_{c <= p < i} L[i][p] x[p][c]) / L[i][i]
  // for all c <= i at once, four threads per column (lanes 4c .. 4c + 3)
  // splitting the sum, combined by lane shuffles
  for (int e = tid; e < n * n; e += kCqThreads) x[e / n][e % n] = make_double2(0.0, 0.0);
  __syncthreads();
  {
    const int c = tid >> 2, part = tid & 3;   // 64 columns x 4 (kCqThreads = 256)
    for (int i = 0; i < n; ++i) {
      double2 s = make_double2(0.0, 0.0);
      if (c <= i && c < n)
        for (int p = c + part; p < i; p += 4) {
          const double2 q = zmul(a[i][p], x[p][c]);
          s.x += q.x;
          s.y += q.y;
        }
      s.x += __shfl_xor(s.x, 1);
      s.y += __shfl_xor(s.y, 1);
      s.x += __shfl_xor(s.x, 2);
      s.y += __shfl_xor(s.y, 2);
      if (part == 0 && c <= i && c < n) {
        const double inv = 1.0 / a[i][i].x;
        x[i][c] = make_double2(((c == i ? 1.0 : 0.0) - s.x) * inv, -s.y * inv);
      }
      __syncthreads();
    }
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += kCqThreads) {
    const int r = e / n, c = e % n;
    L[(int64_t)r * ldl + c] = a[r][c];
    Li[(int64_t)r * ldi + c] = x[r][c];
  }
  if (tid == 0 && bad && *info == 0) *info = bad;
}

// W += shift I with shift = scale * trace(W) (real part), one work-group.
__global__ __launch_bounds__(kCqThreads) void diag_shift_kernel(double2* __restrict__ W,
                                                                int64_t ldw, int n, double scale) {
  __shared__ double red[kCqThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kCqThreads) s += W[(int64_t)i * ldw + i].x;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kCqThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double sh = scale * red[0];
  for (int i = threadIdx.x; i < n; i += kCqThreads) W[(int64_t)i * ldw + i].x += sh;
}

// dst[r, c] = src[r, c] for r < rows, c < cols (row strides lds / ldd)
__global__ __launch_bounds__(256) void copy_block_kernel(const double2* __restrict__ src,
                                                         int64_t lds, double2* __restrict__ dst,
                                                         int64_t ldd, int64_t rows, int64_t cols) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / cols, c = e % cols;
    dst[r * ldd + c] = src[r * lds + c];
  }
}

}  // namespace aq

using namespace aq;

extern "C" int64_t aq_cholqr_workspace_elems(int64_t m, int64_t l) {
  if (m <= 0 || l <= 0) return 0;
  // W, L, block inverses (l x 64 each: block J at rows 64 J), T (m x 64),
  // the split-K partial products of W
  return 2 * l * l + l * kCqB + m * kCqB + cq_splits(m, l) * l * l;
}

extern "C" int aq_cholqr(int64_t m, int64_t l, const void* Yv, int64_t ldy, void* Qv, int64_t ldq,
                         double shift_scale, void* ws, int64_t ws_elems, int* info,
                         hipStream_t stream) {
  AQ_REQUIRE(Yv && Qv && ws && info, AQ_ERR_ARG, "aq_cholqr: null pointer");
  AQ_REQUIRE(m >= 1 && l >= 1 && l <= m, AQ_ERR_ARG, "aq_cholqr: need 1 <= l <= m (got m %lld, l %lld)",
             (long long)m, (long long)l);
  AQ_REQUIRE(l <= 4096 && ldy >= l && ldq >= l, AQ_ERR_ARG, "aq_cholqr: bad width or strides");
  AQ_REQUIRE(shift_scale >= 0.0 && isfinite(shift_scale), AQ_ERR_ARG,
             "aq_cholqr: shift_scale must be finite, >= 0");
  AQ_REQUIRE(ws_elems >= aq_cholqr_workspace_elems(m, l), AQ_ERR_ARG,
             "aq_cholqr: workspace too small");
  AQ_REQUIRE(Yv != Qv, AQ_ERR_ARG, "aq_cholqr: Q must not alias Y");
  const double2* Y = (const double2*)Yv;
  double2* Q = (double2*)Qv;
  double2* W = (double2*)ws;
  double2* L = W + l * l;
  double2* Li = L + l * l;      // [l][64]: inverse of diagonal block J at rows 64 J
  double2* T = Li + l * kCqB;   // [m][64]
  AQ_HIP(hipMemsetAsync(info, 0, sizeof(int), stream));
  // W = Y^H Y: A(i, k) = conj(Y[k, i]), B(k, j) = Y[k, j]
  const int64_t sp = cq_splits(m, l);
  double2* part = T + m * kCqB;
  int st = aq_cgemm_splitk(AQ_C128, AQ_PREC_NATIVE, l, l, m, Y, 1, ldy, 1, Y, ldy, 1, 0, W, l, 1,
                           sp, part, sp * l * l * (int64_t)sizeof(double2), stream);
  if (st != AQ_OK) return st;
  if (shift_scale > 0.0) {
    diag_shift_kernel<<<1, kCqThreads, 0, stream>>>(W, l, (int)l, shift_scale);
    AQ_LAUNCH_CHECK();
  }
  AQ_HIP(hipMemsetAsync(L, 0, sizeof(double2) * l * l, stream));
  for (int64_t j0 = 0; j0 < l; j0 += kCqB) {
    const int64_t bj = std::min<int64_t>(kCqB, l - j0);
    const int64_t r0 = j0 + bj, rb = l - r0;
    double2* LiJ = Li + j0 * kCqB;
    potrf_small_kernel<<<1, kCqThreads, 0, stream>>>(W + j0 * l + j0, l, L + j0 * l + j0, l, LiJ,
                                                     kCqB, (int)bj, (int)j0, info);
    AQ_LAUNCH_CHECK();
    if (rb > 0) {
      // L[r0:, J] = W[r0:, J] L_JJ^{-H}: B(k, c) = conj(Li[c, k])
      st = aq_bgemm(AQ_C128, rb, bj, bj, 1.0, W + r0 * l + j0, l, 1, 0, 0, LiJ, 1, kCqB, 0, 1, 0.0,
                    L + r0 * l + j0, l, 1, 0, 1, stream);
      if (st != AQ_OK) return st;
      // W[r0:, r0:] -= L[r0:, J] L[r0:, J]^H
      st = aq_bgemm(AQ_C128, rb, rb, bj, -1.0, L + r0 * l + j0, l, 1, 0, 0, L + r0 * l + j0, 1, l,
                    0, 1, 1.0, W + r0 * l + r0, l, 1, 0, 1, stream);
      if (st != AQ_OK) return st;
    }
  }
  // Q_J = (Y_J - Q_{<J} L_{J,<J}^H) L_JJ^{-H}
  for (int64_t j0 = 0; j0 < l; j0 += kCqB) {
    const int64_t bj = std::min<int64_t>(kCqB, l - j0);
    const unsigned grid = (unsigned)std::min<int64_t>(cdiv(m * bj, 256), 4096);
    copy_block_kernel<<<grid, 256, 0, stream>>>(Y + j0, ldy, T, kCqB, m, bj);
    AQ_LAUNCH_CHECK();
    if (j0 > 0) {
      // T -= Q[:, :j0] L[J, :j0]^H: B(k, c) = conj(L[j0 + c, k])
      st = aq_bgemm(AQ_C128, m, bj, j0, -1.0, Q, ldq, 1, 0, 0, L + j0 * l, 1, l, 0, 1, 1.0, T, kCqB,
                    1, 0, 1, stream);
      if (st != AQ_OK) return st;
    }
    // Q[:, J] = T L_JJ^{-H}
    st = aq_bgemm(AQ_C128, m, bj, bj, 1.0, T, kCqB, 1, 0, 0, Li + j0 * kCqB, 1, kCqB, 0, 1, 0.0,
                  Q + j0, ldq, 1, 0, 1, stream);
    if (st != AQ_OK) return st;
  }
  return AQ_OK;
}
