// One-sided (Hestenes) Jacobi SVD of the randomized SVD's small projected
// core, with wavefront-level reductions (FP64, orders <= 64).
//
// The core's Gram matrix G = B B^H (l x l, Hermitian positive semidefinite)
// is factored G = L L^H in LDS (Cholesky; a pivot below the round-off level
// of the largest diagonal entry ends the factor there: the remaining columns
// are zero, the semidefinite case), then the columns of L are rotated until
// they are mutually orthogonal: L J = U Sigma, so that G = U Sigma^2 U^H --
// the eigenvalues of G are the squared column norms and its eigenvectors the
// normalised columns.  Sweeps use the round-robin (tournament) ordering; a
// round's ne/2 disjoint column pairs are spread over the work-group's waves,
// lane i of a wave holding row i of both columns, so the three inner
// products of a pair (||a_p||^2, ||a_q||^2, a_p^H a_q) are wavefront
// reductions (xor shuffles, no LDS traffic, no barrier).  The 2 x 2 rotation
// is the one of the two-sided Jacobi kernel (jacobi.hip) applied to the
// pair's Gram block: new a_p = c a_p - s e a_q, new a_q = s a_p + c e a_q,
// e = conj(z) / |z|.  Iteration stops when every pair of a sweep has
// |a_p^H a_q| <= tol ||a_p|| ||a_q|| (or after max_sweeps).  The FP64
// orthogonality of the returned vectors is re-checked in the kernel:
// err[b] = max |U^H U - I| over the columns with non-zero norm.  Columns past
// the numerical rank come back as zero vectors with eigenvalue 0; a caller
// that needs them takes its fallback.
#include <float.h>
#include <math.h>

#include "aq_common.h"

namespace aq {

constexpr int kSvdN = 64;
constexpr int kSvdThreads = 512;   // 8 waves

__device__ __forceinline__ double2 zc(double2 x) { return make_double2(x.x, -x.y); }
__device__ __forceinline__ double2 zm(double2 x, double2 y) {
  return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ __launch_bounds__(kSvdThreads) void gram_svd_small_kernel(
    int n, const double2* __restrict__ G, int64_t ldg, int64_t sG, double* __restrict__ Wout,
    double2* __restrict__ U, int64_t ldu, int64_t sU, int max_sweeps, double tol,
    double* __restrict__ err) {
  __shared__ double2 a[kSvdN][kSvdN + 1];     // a[r][c]: L, then the rotated columns
  __shared__ double nrm2[kSvdN];
  __shared__ int order[kSvdN];
  __shared__ int rank_end, rotated;
  __shared__ double red[kSvdThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t b = blockIdx.x;
  const double2* Gb = G + b * sG;
  const int ne = n + (n & 1);
  for (int i = tid; i < kSvdN * kSvdN; i += kSvdThreads) {
    const int r = i / kSvdN, c = i % kSvdN;
    a[r][c] = (r < n && c < n && r >= c) ? Gb[(int64_t)r * ldg + c] : make_double2(0.0, 0.0);
  }
  if (tid == 0) rank_end = n;
  __syncthreads();
  // ---- semidefinite Cholesky (lower), columns past the first tiny pivot zero
  double dmax = 0.0;
  for (int i = 0; i < n; ++i) dmax = fmax(dmax, a[i][i].x);
  const double pmin = (double)n * DBL_EPSILON * dmax;
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      const double d = a[j][j].x;
      if (!(d > pmin) || rank_end < n) {
        if (rank_end == n) rank_end = j;
        a[j][j] = make_double2(0.0, 0.0);
      } else {
        a[j][j] = make_double2(sqrt(d), 0.0);
      }
    }
    __syncthreads();
    if (rank_end <= j) {
      for (int i = j + 1 + tid; i < n; i += kSvdThreads) a[i][j] = make_double2(0.0, 0.0);
      __syncthreads();
      continue;
    }
    const double inv = 1.0 / a[j][j].x;
    for (int i = j + 1 + tid; i < n; i += kSvdThreads)
      a[i][j] = make_double2(a[i][j].x * inv, a[i][j].y * inv);
    __syncthreads();
    const int t = n - j - 1;
    for (int e = tid; e < t * t; e += kSvdThreads) {
      const int i = j + 1 + e / t, c = j + 1 + e % t;
      if (c <= i) {
        const double2 p = zm(a[i][j], zc(a[c][j]));
        a[i][c] = make_double2(a[i][c].x - p.x, a[i][c].y - p.y);
      }
    }
    __syncthreads();
  }
  // zero the upper triangle (it held G's unused entries)
  for (int i = tid; i < kSvdN * kSvdN; i += kSvdThreads) {
    const int r = i / kSvdN, c = i % kSvdN;
    if (c > r) a[r][c] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  // ---- Hestenes sweeps over the columns of L
  const int half = ne / 2;
  const int nw = kSvdThreads / 64;
  for (int sweep = 0; sweep < max_sweeps && ne >= 2; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < ne - 1; ++round) {
      for (int k = wave; k < half; k += nw) {
        const int m1 = ne - 1;
        int pp = k - 1 - round;
        pp += pp < 0 ? m1 : 0;
        int qq = (ne - 1 - k) - 1 - round;
        qq += qq < 0 ? m1 : 0;
        int p = k == 0 ? 0 : 1 + pp;
        int q = 1 + qq;
        if (p > q) { const int tt = p; p = q; q = tt; }
        const double2 ap = a[lane][p], aq = a[lane][q];
        const double al = wave_sum(ap.x * ap.x + ap.y * ap.y);
        const double be = wave_sum(aq.x * aq.x + aq.y * aq.y);
        const double gx = wave_sum(ap.x * aq.x + ap.y * aq.y);   // a_p^H a_q
        const double gy = wave_sum(ap.x * aq.y - ap.y * aq.x);
        const double gm = fmax(fabs(gx), fabs(gy));
        double az = 0.0;
        if (gm > 0.0) {
          const double ux = gx / gm, uy = gy / gm;
          az = gm * sqrt(ux * ux + uy * uy);
        }
        if (az > tol * sqrt(al * be) && az > 0.0) {
          const double tau = (be - al) / (2.0 * az);
          const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          const double c = 1.0 / sqrt(1.0 + t * t);
          const double s = t * c;
          const double2 e = make_double2(gx / az, -gy / az);   // e^{-i phi}
          const double2 se = make_double2(s * e.x, s * e.y);
          const double2 eq = zm(e, aq), seq = zm(se, aq);
          a[lane][p] = make_double2(c * ap.x - seq.x, c * ap.y - seq.y);
          a[lane][q] = make_double2(s * ap.x + c * eq.x, s * ap.y + c * eq.y);
          if (lane == 0) rotated = 1;
        }
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();
    if (!any) break;
  }
  // ---- norms, order, vectors
  for (int c = wave; c < n; c += nw) {
    const double2 x = a[lane][c];
    const double s2 = wave_sum(x.x * x.x + x.y * x.y);
    if (lane == 0) nrm2[c] = s2;
  }
  __syncthreads();
  if (tid < n) {
    // descending rank of column tid (ties by index)
    const double v = nrm2[tid];
    int pos = 0;
    for (int c = 0; c < n; ++c) pos += (nrm2[c] > v) || (nrm2[c] == v && c < tid);
    order[pos] = tid;
  }
  __syncthreads();
  double* Wb = Wout + b * (int64_t)n;
  double2* Ub = U + b * sU;
  for (int i = tid; i < n * n; i += kSvdThreads) {
    const int r = i / n, j = i % n;
    const int c = order[j];
    const double s2 = nrm2[c];
    const double inv = s2 > 0.0 ? 1.0 / sqrt(s2) : 0.0;
    const double2 x = a[r][c];
    Ub[(int64_t)r * ldu + j] = make_double2(x.x * inv, x.y * inv);
    if (r == 0) Wb[j] = s2;
  }
  __syncthreads();
  // ---- FP64 re-check: max |U^H U - I| over the non-zero columns
  double m = 0.0;
  for (int i = tid; i < n * n; i += kSvdThreads) {
    const int p = i / n, q = i % n;
    if (q < p) continue;
    const int cp = order[p], cq = order[q];
    const double np = nrm2[cp], nq = nrm2[cq];
    if (!(np > 0.0) || !(nq > 0.0)) continue;   // beyond the numerical rank
    double gx = 0.0, gy = 0.0;
    for (int r = 0; r < n; ++r) {
      const double2 x = a[r][cp], y = a[r][cq];
      gx += x.x * y.x + x.y * y.y;
      gy += x.x * y.y - x.y * y.x;
    }
    const double sc = 1.0 / sqrt(np * nq);
    gx *= sc;
    gy *= sc;
    if (p == q) gx -= 1.0;
    m = fmax(m, sqrt(gx * gx + gy * gy));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if (lane == 0) red[wave] = m;
  __syncthreads();
  if (tid == 0) {
    double mm = 0.0;
    for (int w = 0; w < nw; ++w) mm = fmax(mm, red[w]);
    err[b] = mm;
  }
}

}  // namespace aq

using namespace aq;

extern "C" int aq_gram_svd_small(int64_t n, const void* G, int64_t ldg, int64_t strideG, double* W,
                                 void* U, int64_t ldu, int64_t strideU, int64_t batch,
                                 int max_sweeps, double tol, double* err, hipStream_t stream) {
  AQ_REQUIRE(G && W && U && err, AQ_ERR_ARG, "aq_gram_svd_small: null pointer");
  AQ_REQUIRE(n >= 1 && n <= kSvdN, AQ_ERR_ARG, "aq_gram_svd_small: order %lld not in [1, 64]",
             (long long)n);
  AQ_REQUIRE(ldg >= n && ldu >= n && batch >= 1 && batch <= (1LL << 31), AQ_ERR_ARG,
             "aq_gram_svd_small: bad strides or batch");
  AQ_REQUIRE(max_sweeps >= 1 && tol > 0.0, AQ_ERR_ARG, "aq_gram_svd_small: bad sweep limits");
  gram_svd_small_kernel<<<(unsigned)batch, kSvdThreads, 0, stream>>>(
      (int)n, (const double2*)G, ldg, strideG, W, (double2*)U, ldu, strideU, max_sweeps, tol, err);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}
