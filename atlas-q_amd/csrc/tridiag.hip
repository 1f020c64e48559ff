// Top-k eigenpairs of batched Hermitian Gram cores from their tridiagonal
// form (FP64):
//
//   aq_hetrd_batched       dense -> tridiagonal by rocSOLVER's zhetrd (strided
//                          batched; Householder vectors left in the matrix),
//                          the comparison point of the library's own blocked
//                          reduction (hetrd.hip), which produces the same
//                          storage
//   aq_tridiag_eigvals     all or the leading eigenvalues by bisection on
//                          Sturm counts, one thread per eigenvalue
//   aq_tridiag_eigvecs     eigenvectors of the leading eigenvalues by one
//                          step of twisted inverse iteration (twisted LDL^T
//                          factorisation at the computed eigenvalue, twist
//                          index at the smallest |gamma|), one thread each;
//                          the tridiagonal splits where an off-diagonal is
//                          negligible and copies of a repeated eigenvalue
//                          are taken from different blocks
//   aq_hetrd_vblocks       the Householder vectors regrouped into blocks of
//                          nb for the MFMA back-transformation, with the
//                          blocks' compact-WY factors
//   aq_orth_band_reduce    the FP64 orthogonality re-check of the computed
//                          vectors (Gram blocks along the diagonal)
//
// Layout.  The Gram matrices are row-major n x n.  rocSOLVER reads them as
// column-major, i.e. as conj(G); with fill mode "upper" it works on our lower
// triangle (the one the Gram launch computes) and leaves H_i's vector v_i
// (v_i[i] = 1, v_i[r] = 0 for r > i) in row i + 1, columns 0 .. i-1, with
// conj(G) = Q T Q^H, Q = H_{n-2} ... H_1 H_0, T real symmetric tridiagonal
// (D, E).  An eigenvector z of T gives the eigenvector conj(Q z) of G.  The
// eigenvectors are kept as rows: Zt[j, :] = z_j^T, and the back-transform
// forms Et = Zt Q^T, whose row j is (Q z_j)^T = the conjugate transpose of
// G's j-th eigenvector -- the row of U_k^H that the projection U_k^H X reads.
#include <float.h>
#include <math.h>

#include <algorithm>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "aq_common.h"

namespace aq {

constexpr int kTriThreads = 256;
constexpr int kClThreads = 1024;   // cluster pass: one thread per member, 16 waves for the projections
constexpr int kTriMaxN = 4096;  // d, e^2 and e of one matrix in LDS (96 KB at n = 4096)

__device__ __forceinline__ double guard_pivot(double q, double pivmin) {
  return fabs(q) < pivmin ? -pivmin : q;
}

// Number of eigenvalues of the tridiagonal (d, e2 = e^2) block [s, t] below x.
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int s, int t,
                                           double x, double pivmin) {
  double q = guard_pivot(d[s] - x, pivmin);
  int c = q < 0.0;
  for (int i = s + 1; i <= t; ++i) {
    q = guard_pivot(d[i] - x - e2[i - 1] / q, pivmin);
    c += q < 0.0;
  }
  return c;
}

struct TriBounds {
  double gl, gu, pivmin, tnorm;
};

// Loads d and e of one matrix into LDS (e2 = e^2 with split entries zeroed
// when `split`), and reduces the Gershgorin bounds over the work-group.
__device__ TriBounds load_tridiag(const double* __restrict__ D, const double* __restrict__ E,
                                  int n, double* d, double* e2, double* e, bool split) {
  __shared__ double red[3][kClThreads / 64];     // up to the cluster kernel's 16 waves
  double lo = INFINITY, hi = -INFINITY, emax = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double di = D[i];
    const double em = i > 0 ? fabs(E[i - 1]) : 0.0;
    const double ep = i + 1 < n ? fabs(E[i]) : 0.0;
    d[i] = di;
    if (i + 1 < n) {
      double ei = E[i];
      if (split) {
        // negligible coupling (the bisection criterion of LAPACK's dstebz):
        // e_i^2 <= (eps^2 |d_i d_{i+1}|) + tiny
        const double dn = D[i + 1];
        if (ei * ei <= DBL_EPSILON * DBL_EPSILON * fabs(di * dn) + DBL_MIN) ei = 0.0;
      }
      if (e) e[i] = ei;
      e2[i] = ei * ei;
      emax = fmax(emax, ei * ei);
    }
    lo = fmin(lo, di - em - ep);
    hi = fmax(hi, di + em + ep);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off));
    hi = fmax(hi, __shfl_xor(hi, off));
    emax = fmax(emax, __shfl_xor(emax, off));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = lo;
    red[1][w] = hi;
    red[2][w] = emax;
  }
  __syncthreads();
  TriBounds b;
  b.gl = red[0][0];
  b.gu = red[1][0];
  double em2 = red[2][0];
  for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
    b.gl = fmin(b.gl, red[0][q]);
    b.gu = fmax(b.gu, red[1][q]);
    em2 = fmax(em2, red[2][q]);
  }
  b.tnorm = fmax(fabs(b.gl), fabs(b.gu));
  b.pivmin = DBL_MIN * fmax(1.0, em2);
  return b;
}

// ---------------------------------------------------------------- bisection
// W[b, j] = j-th largest eigenvalue, j < nval.
template <int NMAX>
__global__ __launch_bounds__(kTriThreads) void tri_bisect_kernel(
    int n, const double* __restrict__ D, int64_t sD, const double* __restrict__ E, int64_t sE,
    double* __restrict__ W, int64_t sW, int nval) {
  __shared__ double d[NMAX], e2[NMAX];
  const int64_t b = blockIdx.y;
  const TriBounds tb = load_tridiag(D + b * sD, E + b * sE, n, d, e2, nullptr, false);
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nval) return;
  const int a = n - 1 - j;  // ascending index of the eigenvalue
  const double width = tb.gu - tb.gl;
  double lo = tb.gl - 2.0 * DBL_EPSILON * tb.tnorm - 4.0 * tb.pivmin;
  double hi = tb.gu + 2.0 * DBL_EPSILON * tb.tnorm + 4.0 * tb.pivmin;
  const double atol = 2.0 * DBL_EPSILON * tb.tnorm + 4.0 * tb.pivmin;
  // invariant: count(lo) <= a < count(hi)
  for (int it = 0; it < 128 && hi - lo > atol; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_count(d, e2, 0, n - 1, mid, tb.pivmin) <= a) lo = mid;
    else hi = mid;
  }
  (void)width;
  W[b * sW + j] = 0.5 * (lo + hi);
}

// ---------------------------------------------------------------- inverse iteration
// The unreduced block [s, t] (split at zero off-diagonals) of the j-th
// largest eigenvalue: copies of a repeated eigenvalue (within delta) are
// taken from different blocks, in block order.
__device__ __forceinline__ void find_block(int n, const double* d, const double* e2,
                                           const double* e, const double* Wb, int j,
                                           double delta, double pivmin, int& s, int& t) {
  const double lam = Wb[j];
  int j0 = j;
  while (j0 > 0 && Wb[j0 - 1] - lam <= delta) --j0;
  const int copy = j - j0;
  int acc = 0, bs = 0;
  s = 0;
  t = n - 1;
  for (int i = 0; i < n; ++i) {
    if (i == n - 1 || e[i] == 0.0) {  // block [bs, i]
      const int c = sturm_count(d, e2, bs, i, lam + delta, pivmin) -
                    sturm_count(d, e2, bs, i, lam - delta, pivmin);
      if (acc + c > copy) {
        s = bs;
        t = i;
        return;
      }
      acc += c;
      bs = i + 1;
    }
  }
}

// Zt[b, j, :] = unit eigenvector (real, stored as complex) of W[b, j] for
// j < k[b]; rows k[b] .. kmax-1 are zeroed.  scratch: 3 * n * batch * kmax
// doubles (L, U, D+ per thread, thread-minor so that steps coalesce).
template <int NMAX>
__global__ __launch_bounds__(kTriThreads) void tri_invit_kernel(
    int n, const double* __restrict__ D, int64_t sD, const double* __restrict__ E, int64_t sE,
    const double* __restrict__ W, int64_t sW, const int32_t* __restrict__ kvec, int kmax,
    double2* __restrict__ Z, int64_t sZ, double* __restrict__ scratch, int64_t batch) {
  __shared__ double d[NMAX], e2[NMAX], e[NMAX];
  __shared__ int nsplit;
  const int64_t b = blockIdx.y;
  const TriBounds tb = load_tridiag(D + b * sD, E + b * sE, n, d, e2, e, true);
  if (threadIdx.x == 0) nsplit = 0;
  __syncthreads();
  for (int i = threadIdx.x; i + 1 < n; i += blockDim.x)
    if (e[i] == 0.0) atomicAdd(&nsplit, 1);
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= kmax) return;
  double2* z = Z + b * sZ + (int64_t)j * n;
  const int kb = kvec ? kvec[b] : kmax;
  if (j >= kb) {
    for (int i = 0; i < n; ++i) z[i] = make_double2(0.0, 0.0);
    return;
  }
  const double* Wb = W + b * sW;
  const double lam = Wb[j];
  const double delta = 8.0 * DBL_EPSILON * tb.tnorm + 8.0 * tb.pivmin;
  // block of the tridiagonal this vector lives in
  int s = 0, t = n - 1;
  if (nsplit > 0) find_block(n, d, e2, e, Wb, j, delta, tb.pivmin, s, t);
  const int64_t T = batch * (int64_t)kmax;
  const int64_t tid = b * kmax + j;
  double* Ls = scratch;
  double* Us = scratch + (int64_t)n * T;
  double* Ps = scratch + 2 * (int64_t)n * T;
#define AQ_SC(arr, i) arr[(int64_t)(i) * T + tid]
  const double piv = tb.pivmin;
  // stationary factorisation T - lam = L+ D+ L+^T (top down)
  double dp = guard_pivot(d[s] - lam, piv);
  AQ_SC(Ps, s) = dp;
  for (int i = s; i < t; ++i) {
    const double l = e[i] / dp;
    AQ_SC(Ls, i) = l;
    dp = guard_pivot(d[i + 1] - lam - l * e[i], piv);
    AQ_SC(Ps, i + 1) = dp;
  }
  // The loops below read this thread's factors back from scratch; one wave
  // runs per SIMD here, so nothing hides a load's latency but other loads:
  // each loop fetches kPf entries ahead of the dependent recurrence (the
  // arithmetic and its order are unchanged).
  constexpr int kPf = 8;
  // progressive factorisation T - lam = U- D- U-^T (bottom up) and the twist
  double dm = guard_pivot(d[t] - lam, piv);
  int r = t;
  double best = fabs(AQ_SC(Ps, t));
  for (int i = t - 1; i >= s;) {
    double pf[kPf];
#pragma unroll
    for (int u = 0; u < kPf; ++u) pf[u] = (i - u >= s) ? AQ_SC(Ps, i - u) : 0.0;
#pragma unroll
    for (int u = 0; u < kPf; ++u, --i) {
      if (i < s) break;
      const double uu = e[i] / dm;
      AQ_SC(Us, i) = uu;
      dm = guard_pivot(d[i] - lam - uu * e[i], piv);
      const double g = fabs(pf[u] + dm - (d[i] - lam));
      if (g < best) {
        best = g;
        r = i;
      }
    }
  }
  // z_r = 1, z_i = -L_i z_{i+1} above, z_{i+1} = -U_i z_i below
  double nrm2 = 1.0, zi = 1.0;
  for (int i = r - 1; i >= s;) {
    double pf[kPf];
#pragma unroll
    for (int u = 0; u < kPf; ++u) pf[u] = (i - u >= s) ? AQ_SC(Ls, i - u) : 0.0;
#pragma unroll
    for (int u = 0; u < kPf; ++u, --i) {
      if (i < s) break;
      zi = -pf[u] * zi;
      nrm2 += zi * zi;
    }
  }
  zi = 1.0;
  for (int i = r; i < t;) {
    double pf[kPf];
#pragma unroll
    for (int u = 0; u < kPf; ++u) pf[u] = (i + u < t) ? AQ_SC(Us, i + u) : 0.0;
#pragma unroll
    for (int u = 0; u < kPf; ++u, ++i) {
      if (i >= t) break;
      zi = -pf[u] * zi;
      nrm2 += zi * zi;
    }
  }
  const double inv = 1.0 / sqrt(nrm2);
  for (int i = 0; i < s; ++i) z[i] = make_double2(0.0, 0.0);
  for (int i = t + 1; i < n; ++i) z[i] = make_double2(0.0, 0.0);
  z[r] = make_double2(inv, 0.0);
  zi = 1.0;
  for (int i = r - 1; i >= s;) {
    double pf[kPf];
#pragma unroll
    for (int u = 0; u < kPf; ++u) pf[u] = (i - u >= s) ? AQ_SC(Ls, i - u) : 0.0;
#pragma unroll
    for (int u = 0; u < kPf; ++u, --i) {
      if (i < s) break;
      zi = -pf[u] * zi;
      z[i] = make_double2(zi * inv, 0.0);
    }
  }
  zi = 1.0;
  for (int i = r; i < t;) {
    double pf[kPf];
#pragma unroll
    for (int u = 0; u < kPf; ++u) pf[u] = (i + u < t) ? AQ_SC(Us, i + u) : 0.0;
#pragma unroll
    for (int u = 0; u < kPf; ++u, ++i) {
      if (i >= t) break;
      zi = -pf[u] * zi;
      z[i + 1] = make_double2(zi * inv, 0.0);
    }
  }
#undef AQ_SC
}

// ---------------------------------------------------------------- clusters
// Vectors of clustered eigenvalues.  One twisted-factorisation step gives an
// eigenvector accurate to about eps ||T|| / gap (measured: 1e-10 .. 3e-10
// non-orthogonality at gaps of 1e-6 .. 2e-6 ||T||), and copies of one value
// in one block the same vector.  Kept eigenvalues of one unreduced block
// that chain within gtol = kClusterRel ||T|| of each other form a cluster;
// every cluster of two or more is recomputed as a block (LAPACK dstein's
// inverse iteration, its members at once): one thread per member -- the
// members of all clusters together -- factors T_block - lambda_r (guarded
// L D L^T) and runs three solves from a fixed pseudo-random start,
// normalising its own vector, which is written to Z's row.  The clusters'
// rows are then orthonormalised by multi-work-group passes over Z
// (big_project_kernel, big_block_kernel), twice: the rows of a cluster of
// more than kBigR members are cleared of the vectors of their unreduced
// block that are in no cluster, then the clustered rows are
// orthonormalised in steps of kBigR rows in
// eigenvalue order (Cholesky QR of the step's rows, modified Gram-Schmidt
// where they are too close to dependent for it, then their projection out
// of every later clustered row of the block, a work-group per kBigR target
// rows).  Whole-spectrum chains -- the near-null tail of a Gram matrix,
// hundreds of members long, where the guarded factorisations are least
// accurate -- cost O(m^2 n) there, spread over the GPU.  The factors and
// vectors of the solves live in scratch as [n][kmax] (member j in column j).
constexpr double kClusterRel = 1e-5;

__device__ __forceinline__ double start_value(int j, int i) {
  uint32_t h = (uint32_t)j * 0x9E3779B1u ^ ((uint32_t)i * 0x85EBCA77u + 0x165667B1u);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (double)(h & 0xFFFFFu) / 524288.0 - 1.0;
}

__device__ double wg_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q];
  return s;
}

// Doubles of scratch per matrix: three [n][kmax] column blocks (L factors,
// reciprocal pivots, vectors) and ints (block start / end, cluster id, next
// member, roots, member list per cluster, chain tail per block start).
__host__ __device__ inline int64_t cluster_scratch_doubles(int64_t n, int64_t kmax) {
  const int64_t ints = 7 * kmax + n + 8;
  return 3 * n * kmax + (ints + 1) / 2 + 2;
}

template <int NMAX>
__global__ __launch_bounds__(kClThreads) void tri_cluster_kernel(
    int n, const double* __restrict__ D, int64_t sD, const double* __restrict__ E, int64_t sE,
    const double* __restrict__ W, int64_t sW, const int32_t* __restrict__ kvec, int kmax,
    double2* __restrict__ Z, int64_t sZ, double* __restrict__ scratch, int64_t sScr) {
  __shared__ double d[NMAX], e2[NMAX], e[NMAX];
  __shared__ int nsplit, nroots;
  __shared__ int s_rows[NMAX], s_nrows;
  const int64_t b = blockIdx.x;
  const TriBounds tb = load_tridiag(D + b * sD, E + b * sE, n, d, e2, e, true);
  const int kb = kvec ? min((int)kvec[b], kmax) : kmax;
  if (kb < 2) return;
  if (threadIdx.x == 0) nsplit = 0;
  __syncthreads();
  for (int i = threadIdx.x; i + 1 < n; i += blockDim.x)
    if (e[i] == 0.0) atomicAdd(&nsplit, 1);
  __syncthreads();
  const double* Wb = W + b * sW;
  // three disjoint [n][kmax] blocks (restrict: the loads of the solves'
  // next rows may be issued ahead of the current row's store)
  double* __restrict__ LF = scratch + b * sScr;
  double* __restrict__ DF = LF + (int64_t)n * kmax;
  double* __restrict__ X = DF + (int64_t)n * kmax;
  int* bs = (int*)(X + (int64_t)n * kmax);
  int* bt = bs + kmax;
  int* cl = bt + kmax;
  int* nx = cl + kmax;
  int* roots = nx + kmax;
  int* mem = roots + kmax;
  int* last = mem + kmax;                            // [n]
  int* bigroot = last + n;                           // [kmax]: root of a deferred cluster, or -1
  const int64_t ld = kmax;
  const double delta = 8.0 * DBL_EPSILON * tb.tnorm + 8.0 * tb.pivmin;
  const double gabs = kClusterRel * tb.tnorm;
  const double pfloor = fmax(DBL_EPSILON * tb.tnorm, tb.pivmin);
  for (int j = threadIdx.x; j < kb; j += blockDim.x) {
    int s = 0, t = n - 1;
    if (nsplit > 0) find_block(n, d, e2, e, Wb, j, delta, tb.pivmin, s, t);
    bs[j] = s;
    bt[j] = t;
    nx[j] = -1;
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) last[i] = -1;
  for (int j = threadIdx.x; j < kmax; j += blockDim.x) bigroot[j] = -1;
  __threadfence_block();
  __syncthreads();
  if (threadIdx.x == 0) {
    // chains per block: cl[j] = root, nx[] links the members in order
    int nr = 0;
    for (int j = 0; j < kb; ++j) {
      const int p = last[bs[j]];
      if (p >= 0 && Wb[p] - Wb[j] <= gabs) {
        cl[j] = cl[p];
        if (nx[cl[j]] < 0) roots[nr++] = cl[j];   // second member: the root becomes a cluster
        nx[p] = j;
      } else {
        cl[j] = j;
      }
      last[bs[j]] = j;
    }
    nroots = nr;
    // member count per cluster, at its root (mem[root])
    for (int q = 0; q < nr; ++q) {
      int c = 0;
      for (int j = roots[q]; j >= 0; j = nx[j]) ++c;
      mem[roots[q]] = c;
    }
  }
  __threadfence_block();
  __syncthreads();
  double2* Zb = Z + b * sZ;
  // every member of every cluster at once, one thread per member, its vector
  // in column j of the scratch blocks
  for (int j = threadIdx.x; j < kb; j += blockDim.x) {
    if (cl[j] == j && nx[j] < 0) continue;           // not in a cluster
    const int s = bs[j], t = bt[j];
    const double lam = Wb[j];
    double dp = d[s] - lam;
    if (fabs(dp) < pfloor) dp = dp < 0.0 ? -pfloor : pfloor;
    DF[(int64_t)s * ld + j] = 1.0 / dp;
    for (int i = s; i < t; ++i) {
      const double l = e[i] / dp;
      LF[(int64_t)i * ld + j] = l;
      dp = d[i + 1] - lam - l * e[i];
      if (fabs(dp) < pfloor) dp = dp < 0.0 ? -pfloor : pfloor;
      DF[(int64_t)(i + 1) * ld + j] = 1.0 / dp;
    }
    for (int i = s; i <= t; ++i) X[(int64_t)i * ld + j] = start_value(j, i);
    for (int it = 0; it < 3; ++it) {
      double xp = X[(int64_t)s * ld + j];
      double mx = fabs(xp);
      for (int i = s + 1; i <= t; ++i) {
        xp = X[(int64_t)i * ld + j] - LF[(int64_t)(i - 1) * ld + j] * xp;
        X[(int64_t)i * ld + j] = xp;
        mx = fmax(mx, fabs(xp));
      }
      double sc = mx > 0.0 ? 1.0 / mx : 1.0;
      mx = 0.0;
      for (int i = s; i <= t; ++i) {
        const double v = X[(int64_t)i * ld + j] * sc * DF[(int64_t)i * ld + j];
        X[(int64_t)i * ld + j] = v;
        mx = fmax(mx, fabs(v));
      }
      sc = mx > 0.0 ? 1.0 / mx : 1.0;
      xp = X[(int64_t)t * ld + j] * sc;
      X[(int64_t)t * ld + j] = xp;
      double ss = xp * xp;
      for (int i = t - 1; i >= s; --i) {
        xp = X[(int64_t)i * ld + j] * sc - LF[(int64_t)i * ld + j] * xp;
        X[(int64_t)i * ld + j] = xp;
        ss += xp * xp;
      }
      const double inv = ss > 0.0 ? 1.0 / sqrt(ss) : 0.0;
      for (int i = s; i <= t; ++i) X[(int64_t)i * ld + j] *= inv;
    }
    bigroot[j] = cl[j];
  }
  __threadfence_block();
  __syncthreads();
  // the members' rows of Z, written by the whole work-group (coalesced): the
  // members listed in LDS first (any order: the rows are independent), so
  // the work-group walks those rows only
  if (threadIdx.x == 0) s_nrows = 0;
  __syncthreads();
  for (int j = threadIdx.x; j < kb; j += blockDim.x)
    if (bigroot[j] >= 0) s_rows[atomicAdd(&s_nrows, 1)] = j;
  __syncthreads();
  const int nrows = s_nrows;
  for (int q = 0; q < nrows; ++q) {
    const int j = s_rows[q], s = bs[j], t = bt[j];
    for (int i = threadIdx.x; i < n; i += blockDim.x)
      Zb[(int64_t)j * n + i] = make_double2(i >= s && i <= t ? X[(int64_t)i * ld + j] : 0.0, 0.0);
  }
}

// ---------------------------------------------------------------- deferred clusters
constexpr int kBigR = 32;   // rows per step and per target tile
constexpr int kBigC = 64;   // vector entries per LDS tile

// The cluster view of matrix b's cluster scratch: block starts bs[kmax],
// each clustered row's root big[kmax] (-1 for a row in no cluster) and the
// member count at each root csz[] (the layout of tri_cluster_kernel's ints:
// bs, bt, cl, nx, roots, mem [kmax], last [n], bigroot [kmax], after three
// [n][kmax] double blocks).
__device__ inline void big_view(double* scratch, int64_t sScr, int64_t b, int n, int kmax,
                                const int*& bs, const int*& big, const int*& csz) {
  const double* base = scratch + b * sScr + 3 * (int64_t)n * kmax;
  bs = (const int*)base;
  csz = bs + 5 * (int64_t)kmax;           // mem[]: member count at each root
  big = bs + 6 * (int64_t)kmax + n;
}

// Target rows t0 .. t0 + kBigR - 1 of matrix blockIdx.y (clustered rows)
// lose their components along source rows: mode 0, the rows of clusters of
// more than kBigR members lose those along every kept row of the same
// unreduced block that is in no cluster; mode 1, every clustered row loses
// those along the clustered rows of the same block in [row0 - kBigR, row0)
// (the step just orthonormalised; targets are rows >= row0).  Per chunk of kBigR sources: H = T S^T (thread (a, c),
// vectors staged through LDS), then T -= H S.  Sources are never targets of
// the same launch.
__global__ __launch_bounds__(1024) void big_project_kernel(int n, int kmax, double2* __restrict__ Z,
                                                           int64_t sZ, const int32_t* __restrict__ kvec,
                                                           double* __restrict__ scratch, int64_t sScr,
                                                           int mode, int row0) {
  __shared__ double Ta[kBigR][kBigC + 1];
  __shared__ double Sa[kBigR][kBigC + 1];
  __shared__ double H[kBigR][kBigR + 1];
  __shared__ int tact[kBigR], tblk[kBigR];
  __shared__ int any;
  const int64_t b = blockIdx.y;
  const int kb = kvec ? min((int)kvec[b], kmax) : kmax;
  const int *bs, *big, *csz;
  big_view(scratch, sScr, b, n, kmax, bs, big, csz);
  const int t0 = row0 + (int)blockIdx.x * kBigR;
  double2* Zb = Z + b * sZ;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  if (threadIdx.x < kBigR) {
    const int r = t0 + (int)threadIdx.x;
    // mode 0 clears only the rows of clusters longer than a step (a short
    // cluster's vectors are as accurate as the others', eps ||T|| / gap)
    const int act = r < kb && big[r] >= 0 && (mode != 0 || csz[big[r]] > kBigR);
    tact[threadIdx.x] = act;
    tblk[threadIdx.x] = act ? bs[r] : -1;
    if (act) any = 1;
  }
  __syncthreads();
  if (!any) return;
  const int jlo = mode == 0 ? 0 : row0 - kBigR;
  const int jhi = mode == 0 ? kb : min(row0, kb);
  const int a = threadIdx.x >> 5, c = threadIdx.x & 31;
  for (int j0 = jlo; j0 < jhi; j0 += kBigR) {
    const int j = j0 + c;
    bool ok = false;
    if (tact[a] && j < jhi)
      ok = (mode == 0 ? big[j] < 0 : big[j] >= 0) && bs[j] == tblk[a];
    __syncthreads();
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (ok) any = 1;
    __syncthreads();
    if (!any) continue;
    double acc = 0.0;
    for (int i0 = 0; i0 < n; i0 += kBigC) {
      const int nc = min(kBigC, n - i0);
      for (int e = threadIdx.x; e < kBigR * kBigC; e += blockDim.x) {
        const int rr = e / kBigC, ii = e % kBigC;
        const int tr = t0 + rr, sr = j0 + rr;
        Ta[rr][ii] = (ii < nc && tact[rr]) ? Zb[(int64_t)tr * n + i0 + ii].x : 0.0;
        Sa[rr][ii] = (ii < nc && sr < jhi) ? Zb[(int64_t)sr * n + i0 + ii].x : 0.0;
      }
      __syncthreads();
      for (int ii = 0; ii < nc; ++ii) acc += Ta[a][ii] * Sa[c][ii];
      __syncthreads();
    }
    H[a][c] = ok ? acc : 0.0;
    __syncthreads();
    for (int i0 = 0; i0 < n; i0 += kBigC) {
      const int nc = min(kBigC, n - i0);
      for (int e = threadIdx.x; e < kBigR * kBigC; e += blockDim.x) {
        const int rr = e / kBigC, ii = e % kBigC;
        const int sr = j0 + rr;
        Sa[rr][ii] = (ii < nc && sr < jhi) ? Zb[(int64_t)sr * n + i0 + ii].x : 0.0;
      }
      __syncthreads();
      for (int e = threadIdx.x; e < kBigR * kBigC; e += blockDim.x) {
        const int rr = e / kBigC, ii = e % kBigC;
        if (ii < nc && tact[rr]) {
          double v = 0.0;
#pragma unroll 8
          for (int q = 0; q < kBigR; ++q) v += H[rr][q] * Sa[q][ii];
          Zb[(int64_t)(t0 + rr) * n + i0 + ii].x -= v;
        }
      }
      __syncthreads();
    }
  }
}

// The deferred rows among row0 .. row0 + kBigR - 1 of matrix blockIdx.x,
// orthonormalised per unreduced block: M = Z_s Z_s^T restricted to pairs of
// one block (identity on the other rows), M = R^T R, Z_s <- R^-T Z_s (one
// thread per vector entry); modified Gram-Schmidt over the rows, in order,
// where a pivot falls to 1e-14 of its diagonal entry or below.
__global__ __launch_bounds__(256) void big_block_kernel(int n, int kmax, double2* __restrict__ Z,
                                                        int64_t sZ, const int32_t* __restrict__ kvec,
                                                        double* __restrict__ scratch, int64_t sScr,
                                                        int row0) {
  constexpr int kP = kBigR * kBigR / 256;          // entries of M per thread
  __shared__ double T[kBigR][kBigC + 1];
  __shared__ double M[kBigR][kBigR + 1];
  __shared__ double Ri[kBigR][kBigR + 1];
  __shared__ double dg[kBigR];
  __shared__ double red[256 / 64];
  __shared__ int act[kBigR], blk[kBigR];
  __shared__ int any;
  const int64_t b = blockIdx.x;
  const int kb = kvec ? min((int)kvec[b], kmax) : kmax;
  const int *bs, *big, *csz;
  big_view(scratch, sScr, b, n, kmax, bs, big, csz);
  double2* Zb = Z + b * sZ;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  if (threadIdx.x < kBigR) {
    const int r = row0 + (int)threadIdx.x;
    const int ac = r < kb && big[r] >= 0;
    act[threadIdx.x] = ac;
    blk[threadIdx.x] = ac ? bs[r] : -1 - (int)threadIdx.x;    // inactive rows pair with none
    if (ac) any = 1;
  }
  __syncthreads();
  if (!any) return;
  double acc[kP];
#pragma unroll
  for (int u = 0; u < kP; ++u) acc[u] = 0.0;
  for (int i0 = 0; i0 < n; i0 += kBigC) {
    const int nc = min(kBigC, n - i0);
    for (int e = threadIdx.x; e < kBigR * kBigC; e += blockDim.x) {
      const int rr = e / kBigC, ii = e % kBigC;
      T[rr][ii] = (ii < nc && act[rr]) ? Zb[(int64_t)(row0 + rr) * n + i0 + ii].x : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kP; ++u) {
      const int e = (int)threadIdx.x + u * 256, a = e >> 5, c = e & 31;
      for (int ii = 0; ii < nc; ++ii) acc[u] += T[a][ii] * T[c][ii];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < kP; ++u) {
    const int e = (int)threadIdx.x + u * 256, a = e >> 5, c = e & 31;
    const bool pair = act[a] && act[c] && blk[a] == blk[c];
    M[a][c] = pair ? acc[u] : (a == c ? 1.0 : 0.0);
    if (a == c) dg[a] = M[a][c];
  }
  __syncthreads();
  bool fail = false;
  for (int k = 0; k < kBigR; ++k) {
    const double d = M[k][k];
    if (!(d > 1e-14 * dg[k])) {        // the same decision in every thread
      fail = true;
      break;
    }
    const double rkk = sqrt(d);
    __syncthreads();
    if (threadIdx.x >= (unsigned)k && threadIdx.x < (unsigned)kBigR)
      M[k][threadIdx.x] = threadIdx.x == (unsigned)k ? rkk : M[k][threadIdx.x] / rkk;
    __syncthreads();
    for (int e = threadIdx.x; e < kBigR * kBigR; e += blockDim.x) {
      const int a = e >> 5, c = e & 31;
      if (a > k && c >= a) M[a][c] -= M[k][a] * M[k][c];
    }
    __syncthreads();
  }
  if (!fail) {
    // Ri = R^-1 (upper triangular; thread j < kBigR builds column j), then
    // row r of the result is sum_{q <= r} Ri[q][r] z_q
    if (threadIdx.x < (unsigned)kBigR) {
      const int j = threadIdx.x;
      for (int i = kBigR - 1; i >= 0; --i) {
        double v = i == j ? 1.0 : 0.0;
        for (int k = i + 1; k <= j; ++k) v -= M[i][k] * Ri[k][j];
        Ri[i][j] = i <= j ? v / M[i][i] : 0.0;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      double z[kBigR];
#pragma unroll
      for (int r = 0; r < kBigR; ++r) z[r] = act[r] ? Zb[(int64_t)(row0 + r) * n + i].x : 0.0;
#pragma unroll 1
      for (int r = 0; r < kBigR; ++r) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < kBigR; ++q) v += Ri[q][r] * z[q];
        if (act[r]) Zb[(int64_t)(row0 + r) * n + i].x = v;
      }
    }
    return;
  }
  __syncthreads();
  // modified Gram-Schmidt over the deferred rows of the step, in order
  for (int r = 0; r < kBigR; ++r) {
    if (!act[r]) continue;
    double2* zr = Zb + (int64_t)(row0 + r) * n;
    double ss = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) ss += zr[i].x * zr[i].x;
    ss = wg_sum(ss, red);
    const double inv = ss > 0.0 ? 1.0 / sqrt(ss) : 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) zr[i].x *= inv;
    __threadfence_block();
    __syncthreads();
    for (int q = r + 1; q < kBigR; ++q) {
      if (!act[q] || blk[q] != blk[r]) continue;
      double2* zq = Zb + (int64_t)(row0 + q) * n;
      double h = 0.0;
      for (int i = threadIdx.x; i < n; i += blockDim.x) h += zr[i].x * zq[i].x;
      h = wg_sum(h, red);
      for (int i = threadIdx.x; i < n; i += blockDim.x) zq[i].x -= h * zr[i].x;
      __threadfence_block();
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------- back-transform blocks
// Vblk[b, p, j, r] = v_i[r] (i = p*nb + j; zero row when i > n-2).
__global__ __launch_bounds__(256) void hetrd_vblk_kernel(int n, int nb, int nblk,
                                                         const double2* __restrict__ A,
                                                         int64_t sA, double2* __restrict__ V) {
  const int64_t b = blockIdx.z;
  const int p = blockIdx.y;
  const int64_t row0 = (int64_t)p * nb;
  double2* Vp = V + (b * nblk + p) * (int64_t)nb * n;
  const double2* Ab = A + b * sA;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)nb * n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / n), r = (int)(idx % n);
    const int64_t i = row0 + j;
    double2 v = make_double2(0.0, 0.0);
    if (i <= n - 2) {
      if (r < i) v = Ab[(i + 1) * n + r];
      else if (r == i) v.x = 1.0;
    }
    Vp[idx] = v;
  }
}

// Forward compact-WY factor of P^H = H_{i0}^H ... H_{i0+nb-1}^H from the
// reflectors' Gram matrix Gv[a, c] = v_a^H v_c (nb x nb, row-major):
// T'[j, j] = conj(tau_j), T'[0:j, j] = -conj(tau_j) T'[0:j, 0:j] Gv[0:j, j].
// One work-group of nb threads per (matrix, block, sub-block): the blocks are
// ld x ld (Gv, Tf), block p's reflectors start at p * ld, and sub-block z
// covers its reflectors z nb .. z nb + nb - 1 (the diagonal nb x nb block of
// Gv and Tf at (z nb, z nb)).
__global__ void hetrd_tfactor_kernel(int n, int nb, int nblk, const double2* __restrict__ tau,
                                     int64_t stau, const double2* __restrict__ Gv,
                                     double2* __restrict__ Tf, int ld) {
  __shared__ double2 t[64][65];
  const int64_t b = blockIdx.y;
  const int p = blockIdx.x;
  const int off = (int)blockIdx.z * nb;
  const int a = threadIdx.x;
  const int64_t base = (b * nblk + p) * (int64_t)ld * ld + (int64_t)off * ld + off;
  const double2* G = Gv + base;
  for (int c = 0; c < nb; ++c) t[a][c] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const int64_t i = (int64_t)p * ld + off + j;
    double2 tj = make_double2(0.0, 0.0);
    if (i <= n - 2) {
      tj = tau[b * stau + i];
      tj.y = -tj.y;  // conj(tau)
    }
    // row a < j of column j: -conj(tau_j) sum_{c < j} T'[a, c] Gv[c, j]
    if (a < j) {
      double2 acc = make_double2(0.0, 0.0);
      for (int c = a; c < j; ++c) acc = cadd(acc, cmul(t[a][c], G[(int64_t)c * ld + j]));
      const double2 v = cmul(tj, acc);
      t[a][j] = make_double2(-v.x, -v.y);
    } else if (a == j) {
      t[a][j] = tj;
    }
    __syncthreads();
  }
  for (int c = 0; c < nb; ++c) Tf[base + (int64_t)a * ld + c] = t[a][c];
}

// Zeroes the lower-left h x h block of every 2h x 2h factor (its scratch
// during the join below).
__global__ __launch_bounds__(256) void hetrd_tlow_zero_kernel(int h, int64_t nblocks,
                                                              double2* __restrict__ Tf) {
  const int ld = 2 * h;
  const int64_t total = nblocks * h * h;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t blk = e / (h * h);
    const int r = (int)(e % (h * h)) / h, c = (int)(e % h);
    Tf[blk * ld * ld + (int64_t)(h + r) * ld + c] = make_double2(0.0, 0.0);
  }
}

// err[b] = max over a, c < k_b of |Q[b, tile, a, c] - delta_ac| where the
// tiles are the diagonal Gram blocks Q[tile] = Zt[64 tile : 64 tile + 64]
// Zt[64 tile : 64 tile + 128]^H (64 x 128, row-major) of each matrix.
__global__ __launch_bounds__(256) void orth_band_reduce_kernel(const double2* __restrict__ Q,
                                                               int ntile,
                                                               const int32_t* __restrict__ kvec,
                                                               int kmax, double* __restrict__ err) {
  __shared__ double red[4];
  const int64_t b = blockIdx.x;
  const int kb = kvec ? kvec[b] : kmax;
  double m = 0.0;
  for (int64_t idx = threadIdx.x; idx < (int64_t)ntile * 64 * 128; idx += blockDim.x) {
    const int tile = (int)(idx / (64 * 128));
    const int rem = (int)(idx % (64 * 128));
    const int a = tile * 64 + rem / 128, c = tile * 64 + rem % 128;
    if (a >= kb || c >= kb) continue;
    double2 q = Q[(b * ntile) * 64 * 128 + idx];
    if (a == c) q.x -= 1.0;
    m = fmax(m, sqrt(q.x * q.x + q.y * q.y));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int q = 1; q < 4; ++q) r = fmax(r, red[q]);
    err[b] = r;
  }
}

namespace {
struct Handle {
  rocblas_handle h = nullptr;
  ~Handle() {
    if (h) rocblas_destroy_handle(h);
  }
};
thread_local Handle t_handle;
}  // namespace

}  // namespace aq

using namespace aq;

extern "C" int aq_hetrd_batched(int64_t n, void* A, int64_t strideA, double* D, double* E,
                                void* tau, int64_t batch, hipStream_t stream) {
  AQ_REQUIRE(A && D && E && tau, AQ_ERR_ARG, "aq_hetrd_batched: null pointer");
  AQ_REQUIRE(n >= 2 && n <= kTriMaxN, AQ_ERR_ARG, "aq_hetrd_batched: n out of range (%lld)",
             (long long)n);
  AQ_REQUIRE(batch >= 1 && batch <= 65535 && strideA >= n * n, AQ_ERR_ARG,
             "aq_hetrd_batched: bad batch or stride");
  if (!t_handle.h) {
    rocblas_status s = rocblas_create_handle(&t_handle.h);
    AQ_REQUIRE(s == rocblas_status_success, AQ_ERR_HIP,
               "aq_hetrd_batched: rocblas_create_handle failed (%d)", (int)s);
  }
  rocblas_status s = rocblas_set_stream(t_handle.h, stream);
  AQ_REQUIRE(s == rocblas_status_success, AQ_ERR_HIP, "aq_hetrd_batched: set_stream (%d)",
             (int)s);
  s = rocsolver_zhetrd_strided_batched(
      t_handle.h, rocblas_fill_upper, (rocblas_int)n,
      reinterpret_cast<rocblas_double_complex*>(A), (rocblas_int)n, (rocblas_stride)strideA, D,
      (rocblas_stride)n, E, (rocblas_stride)(n - 1),
      reinterpret_cast<rocblas_double_complex*>(tau), (rocblas_stride)(n - 1),
      (rocblas_int)batch);
  AQ_REQUIRE(s == rocblas_status_success, AQ_ERR_HIP,
             "aq_hetrd_batched: rocsolver_zhetrd_strided_batched failed (%d)", (int)s);
  return AQ_OK;
}

extern "C" int aq_tridiag_eigvals(int64_t n, const double* D, int64_t sD, const double* E,
                                  int64_t sE, double* W, int64_t sW, int64_t nval, int64_t batch,
                                  hipStream_t stream) {
  AQ_REQUIRE(D && E && W, AQ_ERR_ARG, "aq_tridiag_eigvals: null pointer");
  AQ_REQUIRE(n >= 1 && n <= kTriMaxN && nval >= 1 && nval <= n, AQ_ERR_ARG,
             "aq_tridiag_eigvals: n %lld / nval %lld out of range", (long long)n,
             (long long)nval);
  AQ_REQUIRE(batch >= 1 && batch <= 65535, AQ_ERR_ARG, "aq_tridiag_eigvals: bad batch");
  dim3 grid((unsigned)cdiv(nval, kTriThreads), (unsigned)batch);
  if (n <= 2048)
    tri_bisect_kernel<2048><<<grid, kTriThreads, 0, stream>>>((int)n, D, sD, E, sE, W, sW,
                                                              (int)nval);
  else
    tri_bisect_kernel<kTriMaxN><<<grid, kTriThreads, 0, stream>>>((int)n, D, sD, E, sE, W, sW,
                                                                  (int)nval);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int64_t aq_tridiag_eigvecs_scratch_doubles(int64_t n, int64_t kmax, int64_t batch) {
  const int64_t kc = std::min<int64_t>(kmax, n);
  return std::max<int64_t>(3 * n * kmax, cluster_scratch_doubles(n, kc)) * batch;
}

extern "C" int aq_tridiag_eigvecs(int64_t n, const double* D, int64_t sD, const double* E,
                                  int64_t sE, const double* W, int64_t sW, const int32_t* kvec,
                                  int64_t kmax, void* Z, int64_t sZ, double* scratch,
                                  int64_t scratch_doubles, int64_t batch, hipStream_t stream) {
  AQ_REQUIRE(D && E && W && Z && scratch, AQ_ERR_ARG, "aq_tridiag_eigvecs: null pointer");
  // rows k[b] .. kmax-1 are zeroed, so kmax (the row count of Z, padded by the
  // caller) may exceed n when every k[b] <= n; without kvec every row up to
  // kmax is computed and kmax <= n
  AQ_REQUIRE(n >= 1 && n <= kTriMaxN && kmax >= 1 && (kvec ? true : kmax <= n), AQ_ERR_ARG,
             "aq_tridiag_eigvecs: n %lld / kmax %lld out of range", (long long)n,
             (long long)kmax);
  AQ_REQUIRE(batch >= 1 && batch <= 65535 && sZ >= kmax * n, AQ_ERR_ARG,
             "aq_tridiag_eigvecs: bad batch or stride");
  AQ_REQUIRE(scratch_doubles >= aq_tridiag_eigvecs_scratch_doubles(n, kmax, batch), AQ_ERR_ARG,
             "aq_tridiag_eigvecs: scratch too small");
  dim3 grid((unsigned)cdiv(kmax, kTriThreads), (unsigned)batch);
  if (n <= 2048)
    tri_invit_kernel<2048><<<grid, kTriThreads, 0, stream>>>(
        (int)n, D, sD, E, sE, W, sW, kvec, (int)kmax, (double2*)Z, sZ, scratch, batch);
  else
    tri_invit_kernel<kTriMaxN><<<grid, kTriThreads, 0, stream>>>(
        (int)n, D, sD, E, sE, W, sW, kvec, (int)kmax, (double2*)Z, sZ, scratch, batch);
  AQ_LAUNCH_CHECK();
  // clustered eigenvalues: the scratch of the step above is free again
  const int64_t kc = std::min<int64_t>(kmax, n);
  const int64_t sScr = cluster_scratch_doubles(n, kc);
  if (kc >= 2) {
    if (n <= 2048)
      tri_cluster_kernel<2048><<<(unsigned)batch, kClThreads, 0, stream>>>(
          (int)n, D, sD, E, sE, W, sW, kvec, (int)kc, (double2*)Z, sZ, scratch, sScr);
    else
      tri_cluster_kernel<kTriMaxN><<<(unsigned)batch, kClThreads, 0, stream>>>(
          (int)n, D, sD, E, sE, W, sW, kvec, (int)kc, (double2*)Z, sZ, scratch, sScr);
    AQ_LAUNCH_CHECK();
    // the clusters' rows, orthonormalised by the passes below (a work-group
    // with no clustered row in its range exits at once)
    {
      const int nstep = (int)cdiv(kc, kBigR);
      for (int pass = 0; pass < 2; ++pass) {
        big_project_kernel<<<dim3((unsigned)nstep, (unsigned)batch), 1024, 0, stream>>>(
            (int)n, (int)kc, (double2*)Z, sZ, kvec, scratch, sScr, 0, 0);
        AQ_LAUNCH_CHECK();
        for (int p = 0; p < nstep; ++p) {
          big_block_kernel<<<(unsigned)batch, 256, 0, stream>>>(
              (int)n, (int)kc, (double2*)Z, sZ, kvec, scratch, sScr, p * kBigR);
          AQ_LAUNCH_CHECK();
          const int row0 = (p + 1) * kBigR;
          if (row0 < kc) {
            big_project_kernel<<<dim3((unsigned)cdiv(kc - row0, kBigR), (unsigned)batch), 1024, 0,
                                 stream>>>((int)n, (int)kc, (double2*)Z, sZ, kvec, scratch, sScr, 1,
                                           row0);
            AQ_LAUNCH_CHECK();
          }
        }
      }
    }
  }
  return AQ_OK;
}

extern "C" int aq_hetrd_vblocks(int64_t n, int64_t nb, const void* A, int64_t strideA,
                                void* V, int64_t batch, hipStream_t stream) {
  AQ_REQUIRE(A && V, AQ_ERR_ARG, "aq_hetrd_vblocks: null pointer");
  AQ_REQUIRE(n >= 2 && nb >= 1 && nb <= 128 && batch >= 1 && batch <= 65535, AQ_ERR_ARG,
             "aq_hetrd_vblocks: bad sizes");
  const int64_t nblk = cdiv(n - 1, nb);
  dim3 grid((unsigned)std::min<int64_t>(cdiv(nb * n, 256), 64), (unsigned)nblk,
            (unsigned)batch);
  hetrd_vblk_kernel<<<grid, 256, 0, stream>>>((int)n, (int)nb, (int)nblk, (const double2*)A,
                                               strideA, (double2*)V);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_bgemm(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                        int64_t a_s0, int64_t a_s1, int64_t a_sb, int conj_a, const void* B,
                        int64_t b_s0, int64_t b_s1, int64_t b_sb, int conj_b, double beta, void* C,
                        int64_t c_s0, int64_t c_s1, int64_t c_sb, int64_t batch,
                        hipStream_t stream);

extern "C" int aq_hetrd_tfactors(int64_t n, int64_t nb, const void* tau, int64_t stau,
                                 const void* Gv, void* Tf, int64_t batch, hipStream_t stream) {
  AQ_REQUIRE(tau && Gv && Tf, AQ_ERR_ARG, "aq_hetrd_tfactors: null pointer");
  AQ_REQUIRE(n >= 2 && nb >= 1 && (nb <= 64 || (nb <= 128 && nb % 2 == 0)) && batch >= 1 &&
                 batch <= 65535,
             AQ_ERR_ARG, "aq_hetrd_tfactors: bad sizes");
  const int64_t nblk = cdiv(n - 1, nb);
  if (nb <= 64) {
    dim3 grid((unsigned)nblk, (unsigned)batch, 1);
    hetrd_tfactor_kernel<<<grid, (unsigned)nb, 0, stream>>>((int)n, (int)nb, (int)nblk,
                                                            (const double2*)tau, stau,
                                                            (const double2*)Gv, (double2*)Tf,
                                                            (int)nb);
    AQ_LAUNCH_CHECK();
    return AQ_OK;
  }
  // blocks of 65..128 reflectors: the two halves' factors, then their join
  const int h = (int)(nb / 2);
  dim3 grid((unsigned)nblk, (unsigned)batch, 2);
  hetrd_tfactor_kernel<<<grid, (unsigned)h, 0, stream>>>((int)n, h, (int)nblk,
                                                         (const double2*)tau, stau,
                                                         (const double2*)Gv, (double2*)Tf,
                                                         (int)nb);
  AQ_LAUNCH_CHECK();
  // the join T'_12 = -T'_11 Gv_12 T'_22 (the column recursion split at h) on
  // the batched MFMA GEMM, with the lower-left block as scratch for
  // Gv_12 T'_22, then that block zeroed
  const int64_t nbb = nb * nb, blocks = nblk * batch;
  for (int64_t b0 = 0; b0 < blocks; b0 += 65535) {   // the GEMM's batch limit
    const int64_t cnt = std::min<int64_t>(65535, blocks - b0);
    const double2* G = (const double2*)Gv + b0 * nbb;
    double2* T = (double2*)Tf + b0 * nbb;
    int rc = aq_bgemm(AQ_C128, h, h, h, 1.0, G + h, nb, 1, nbb, 0, T + (int64_t)h * nb + h, nb,
                      1, nbb, 0, 0.0, T + (int64_t)h * nb, nb, 1, nbb, cnt, stream);
    if (rc != AQ_OK) return rc;
    rc = aq_bgemm(AQ_C128, h, h, h, -1.0, T, nb, 1, nbb, 0, T + (int64_t)h * nb, nb, 1, nbb, 0,
                  0.0, T + h, nb, 1, nbb, cnt, stream);
    if (rc != AQ_OK) return rc;
  }
  double2* T = (double2*)Tf;
  hetrd_tlow_zero_kernel<<<(unsigned)std::min<int64_t>(cdiv(blocks * h * h, 256), 65536), 256, 0,
                           stream>>>(h, blocks, T);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_orth_band_reduce(const void* Q, int64_t ntile, const int32_t* kvec,
                                   int64_t kmax, double* err, int64_t batch,
                                   hipStream_t stream) {
  AQ_REQUIRE(Q && err, AQ_ERR_ARG, "aq_orth_band_reduce: null pointer");
  AQ_REQUIRE(ntile >= 1 && batch >= 1, AQ_ERR_ARG, "aq_orth_band_reduce: bad sizes");
  orth_band_reduce_kernel<<<(unsigned)batch, 256, 0, stream>>>(
      (const double2*)Q, (int)ntile, kvec, (int)kmax, err);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}
