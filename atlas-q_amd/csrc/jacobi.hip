// Batched two-sided Jacobi eigensolver for small Hermitian FP64 matrices
// (n <= 64), one work-group per matrix, everything resident in LDS.
//
// Cyclic Jacobi with the round-robin (tournament) ordering: each round
// rotates n/2 disjoint index pairs at once; a sweep is n-1 rounds.  For a
// pair (p, q) with a = A_pp, b = A_qq, z = A_pq = |z| e^{i phi}:
//   tau = (b - a) / (2|z|),  t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
//   c = 1/sqrt(1+t^2), s = t c,  J = [[c, s], [-s e^{-i phi}, c e^{-i phi}]]
// and A <- J^H A J, V <- V J; a pair with |z| <= 1e-18 (|a| + |b|) is not
// rotated, only its off-diagonal cleared.  The off-diagonal Frobenius norm is reduced
// across the work-group after every sweep; iteration stops when it falls
// below tol * ||A||_F or after max_sweeps.  Eigenvalues are returned in
// descending order with their eigenvectors (columns of V).
//
// Used for the projected cores of the randomized SVD, the Gram matrices of
// small two-site updates, the sketch spectra of the rank-predictor features
// and the batched small updates.
#include <math.h>

#include "aq_common.h"

namespace aq {

constexpr int kJmax = 64;

// LDS image of one matrix of order <= JM (JM = 16, 32, 64: 13, 38, 139 KB, so
// that orders <= 32 run several work-groups per CU; order 64 runs one
// 512-thread work-group per CU)
template <int JM>
struct JacobiSmem {
  double2 a[JM][JM + 1];
  double2 v[JM][JM + 1];
  double cs[JM / 2];       // c per pair
  double2 sp[JM / 2];      // s e^{-i phi} per pair
  double2 ep[JM / 2];      // e^{-i phi} per pair
  int pp[JM / 2], pq[JM / 2];
  double red[512];
  double lam[JM];
  int order[JM];
};

__device__ __forceinline__ double2 cm(double2 x, double2 y) {
  return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
}
__device__ __forceinline__ double2 cconj(double2 x) { return make_double2(x.x, -x.y); }
__device__ __forceinline__ double2 cadd2(double2 x, double2 y) { return make_double2(x.x + y.x, x.y + y.y); }
__device__ __forceinline__ double2 cscale(double2 x, double s) { return make_double2(x.x * s, x.y * s); }

__device__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// n: the batch's order (output row length); orders (nullable): matrix b's own
// order nb <= n -- only its leading nb x nb block is read and diagonalised;
// W[nb:n] = 0 and the trailing rows/columns of V are those of the identity.
template <int JM, int NT>
__global__ __launch_bounds__(NT) void heev_jacobi_kernel(int n, const int64_t* __restrict__ orders,
                                                         const double2* __restrict__ A,
                                                         int64_t lda, int64_t sA,
                                                         double* __restrict__ W,
                                                         double2* __restrict__ V, int64_t ldv,
                                                         int64_t sV, int max_sweeps, double tol,
                                                         int* __restrict__ sweeps_out) {
  extern __shared__ unsigned char smraw[];
  JacobiSmem<JM>& sm = *reinterpret_cast<JacobiSmem<JM>*>(smraw);
  const int tid = threadIdx.x;
  const int64_t b = blockIdx.x;
  const int nfull = n;
  if (orders) {
    const int64_t o = orders[b];
    n = (int)(o < 1 ? 1 : (o > nfull ? nfull : o));
  }
  const int ne = n + (n & 1);  // even order for the tournament (padding index = n)
  const double2* Ab = A + b * sA;
  for (int i = tid; i < ne * ne; i += blockDim.x) {
    const int r = i / ne, c = i % ne;
    double2 x = make_double2(0.0, 0.0);
    if (r < n && c < n) x = Ab[(int64_t)r * lda + c];
    sm.a[r][c] = x;
    sm.v[r][c] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  // symmetrize (the input is Hermitian up to rounding)
  for (int i = tid; i < ne * ne; i += blockDim.x) {
    const int r = i / ne, c = i % ne;
    if (r < c) {
      const double2 u = sm.a[r][c], l = cconj(sm.a[c][r]);
      const double2 h = make_double2(0.5 * (u.x + l.x), 0.5 * (u.y + l.y));
      sm.a[r][c] = h;
      sm.a[c][r] = cconj(h);
    } else if (r == c) {
      sm.a[r][c].y = 0.0;
    }
  }
  __syncthreads();
  double fro = 0.0;
  for (int i = tid; i < ne * ne; i += blockDim.x) {
    const double2 x = sm.a[i / ne][i % ne];
    fro += x.x * x.x + x.y * x.y;
  }
  fro = block_sum(fro, sm.red);
  const double thresh2 = tol * tol * fro;
  int sweep = 0;
  const int half = ne / 2;
  // this thread's items of the A pass (blocks kr, kc) and of the V pass
  // (pair k, row r) are the same in every round: decompose them once
  // (integer division by the runtime order costs tens of instructions)
  // A pass: the blocks (kr, kc) with kr >= kc, numbered row by row
  constexpr int AIT = (JM / 2 * (JM / 2 + 1) / 2 + NT - 1) / NT;
  constexpr int VIT = (JM * (JM / 2) + NT - 1) / NT;
  int akr[AIT], akc[AIT], vk[VIT], vr[VIT];
#pragma unroll
  for (int j = 0; j < AIT; ++j) {
    const int it = tid + j * NT;
    akr[j] = -1;
    akc[j] = 0;
    if (it < half * (half + 1) / 2) {
      int r = (int)((sqrt(8.0 * it + 1.0) - 1.0) * 0.5);
      while (r * (r + 1) / 2 > it) --r;
      while ((r + 1) * (r + 2) / 2 <= it) ++r;
      akr[j] = r;
      akc[j] = it - r * (r + 1) / 2;
    }
  }
#pragma unroll
  for (int j = 0; j < VIT; ++j) {
    const int it = tid + j * NT;
    vk[j] = it < ne * half ? it / ne : -1;
    vr[j] = it < ne * half ? it % ne : 0;
  }
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0;
    for (int i = tid; i < ne * ne; i += blockDim.x) {
      const int r = i / ne, c = i % ne;
      if (r != c) {
        const double2 x = sm.a[r][c];
        off += x.x * x.x + x.y * x.y;
      }
    }
    off = block_sum(off, sm.red);
    if (off <= thresh2 || ne < 2) break;
    for (int round = 0; round < ne - 1; ++round) {
      // pairs of this round: (pos(k), pos(ne-1-k)) of the round-robin
      // tournament, position 0 fixed and the others rotated by `round`:
      // pos(0) = 0, pos(i) = 1 + ((i - 1 - round) mod (ne - 1))
      if (tid < half) {
        // tid - 1 - round and i2 - 1 - round lie in (-m1, m1): one
        // conditional add is the modulo
        const int m1 = ne - 1;
        const int i2 = ne - 1 - tid;
        int pp = tid - 1 - round;
        pp += pp < 0 ? m1 : 0;
        int qq = i2 - 1 - round;
        qq += qq < 0 ? m1 : 0;
        int p = tid == 0 ? 0 : 1 + pp;
        int q = 1 + qq;
        if (p > q) { const int t = p; p = q; q = t; }
        sm.pp[tid] = p;
        sm.pq[tid] = q;
        const double a = sm.a[p][p].x, bq = sm.a[q][q].x;
        const double2 z = sm.a[p][q];
        // |z| scaled so that no square underflows: near convergence |z|
        // reaches 1e-160 and below, where z.x^2 is subnormal, |z| loses
        // digits and e = z/|z| is no longer unimodular (J not unitary)
        const double zm = fmax(fabs(z.x), fabs(z.y));
        double az = 0.0;
        if (zm > 0.0) {
          const double zx = z.x / zm, zy = z.y / zm;
          az = zm * sqrt(zx * zx + zy * zy);
        }
        double c = 1.0, s = 0.0;
        double2 e = make_double2(1.0, 0.0);
        // a pair already diagonal to 1e-18 (|a| + |b|) is only cleaned below
        if (az > 1e-18 * (fabs(a) + fabs(bq)) && az > 0.0) {
          const double tau = (bq - a) / (2.0 * az);
          const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          s = t * c;
          e = make_double2(z.x / az, -z.y / az);  // e^{-i phi}
        }
        sm.cs[tid] = c;
        sm.ep[tid] = e;
        sm.sp[tid] = cscale(e, s);
      }
      __syncthreads();
      // A <- J^H A J by 2 x 2 blocks: the rotations of a round act on disjoint
      // pairs, so block (pair kr, pair kc) -- rows {p_r, q_r} x columns
      // {p_c, q_c} -- maps to J_r^H A[block] J_c from its own old values
      // (one pass over A, no intermediate barrier).  Only blocks with
      // kr >= kc are computed; block (kc, kr) is written as the conjugate
      // transpose, which keeps A exactly Hermitian.
#pragma unroll
      for (int j = 0; j < AIT; ++j) {
        const int kr = akr[j], kc = akc[j];
        if (kr < 0) continue;
        const int pr = sm.pp[kr], qr = sm.pq[kr], pc = sm.pp[kc], qc = sm.pq[kc];
        const double cc = sm.cs[kc];
        const double sc = sm.sp[kc].x * sm.ep[kc].x + sm.sp[kc].y * sm.ep[kc].y;
        const double2 ec = sm.ep[kc], sec = sm.sp[kc];
        const double cr = sm.cs[kr];
        const double sr = sm.sp[kr].x * sm.ep[kr].x + sm.sp[kr].y * sm.ep[kr].y;
        const double2 erc = cconj(sm.ep[kr]), serc = cconj(sm.sp[kr]);
        const double2 a00 = sm.a[pr][pc], a01 = sm.a[pr][qc];
        const double2 a10 = sm.a[qr][pc], a11 = sm.a[qr][qc];
        // columns (A J_c): new p = c ap - s e aq ; new q = s ap + c e aq
        const double2 b00 = cadd2(cscale(a00, cc), cscale(cm(sec, a01), -1.0));
        const double2 b01 = cadd2(cscale(a00, sc), cscale(cm(ec, a01), cc));
        const double2 b10 = cadd2(cscale(a10, cc), cscale(cm(sec, a11), -1.0));
        const double2 b11 = cadd2(cscale(a10, sc), cscale(cm(ec, a11), cc));
        // rows (J_r^H .): new p = c bp - s conj(e) bq ; new q = s bp + c conj(e) bq
        double2 n00 = cadd2(cscale(b00, cr), cscale(cm(serc, b10), -1.0));
        double2 n10 = cadd2(cscale(b00, sr), cscale(cm(erc, b10), cr));
        double2 n01 = cadd2(cscale(b01, cr), cscale(cm(serc, b11), -1.0));
        double2 n11 = cadd2(cscale(b01, sr), cscale(cm(erc, b11), cr));
        if (kr == kc) {
          // the rotated pair is exactly diagonalised: clear the rounding in
          // its off-diagonal entries and the imaginary parts of its diagonal
          n01 = make_double2(0.0, 0.0);
          n10 = make_double2(0.0, 0.0);
          n00.y = 0.0;
          n11.y = 0.0;
        }
        sm.a[pr][pc] = n00;
        sm.a[pr][qc] = n01;
        sm.a[qr][pc] = n10;
        sm.a[qr][qc] = n11;
        if (kr != kc) {
          sm.a[pc][pr] = cconj(n00);
          sm.a[qc][pr] = cconj(n01);
          sm.a[pc][qr] = cconj(n10);
          sm.a[qc][qr] = cconj(n11);
        }
      }
      // V <- V J   (item = row r, pair k; V is not read by the A pass)
#pragma unroll
      for (int j = 0; j < VIT; ++j) {
        const int k = vk[j], r = vr[j];
        if (k < 0) continue;
        const int p = sm.pp[k], q = sm.pq[k];
        const double c = sm.cs[k];
        const double s = sm.sp[k].x * sm.ep[k].x + sm.sp[k].y * sm.ep[k].y;  // |s|, sign kept
        const double2 e = sm.ep[k];
        const double2 se = sm.sp[k];
        const double2 vp = sm.v[r][p], vq = sm.v[r][q];
        sm.v[r][p] = cadd2(cscale(vp, c), cscale(cm(se, vq), -1.0));
        sm.v[r][q] = cadd2(cscale(vp, s), cscale(cm(e, vq), c));
      }
      __syncthreads();
    }
  }
  // eigenvalues, descending order by rank (ties broken by index)
  for (int i = tid; i < n; i += blockDim.x) sm.lam[i] = sm.a[i][i].x;
  __syncthreads();
  for (int i = tid; i < n; i += blockDim.x) {
    const double li = sm.lam[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double lj = sm.lam[j];
      rank += (lj > li) || (lj == li && j < i);
    }
    sm.order[rank] = i;
  }
  __syncthreads();
  double* Wb = W + b * nfull;
  double2* Vb = V + b * sV;
  for (int i = tid; i < nfull; i += blockDim.x) Wb[i] = i < n ? sm.lam[sm.order[i]] : 0.0;
  for (int i = tid; i < nfull * nfull; i += blockDim.x) {
    const int r = i / nfull, c = i % nfull;
    double2 x = make_double2(r == c ? 1.0 : 0.0, 0.0);
    if (r < n && c < n) x = sm.v[r][sm.order[c]];
    Vb[(int64_t)r * ldv + c] = x;
  }
  if (tid == 0 && sweeps_out) sweeps_out[b] = sweep;
}

template <int JM, int NT>
static int launch_jacobi(int n, const int64_t* orders, const void* A, int64_t lda, int64_t sA,
                         double* W, void* V, int64_t ldv, int64_t sV, int64_t batch,
                         int max_sweeps, double tol, int* sweeps, hipStream_t stream) {
  static bool attr = false;
  if (!attr) {
    AQ_HIP(hipFuncSetAttribute((const void*)heev_jacobi_kernel<JM, NT>,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               sizeof(JacobiSmem<JM>)));
    attr = true;
  }
  heev_jacobi_kernel<JM, NT><<<(unsigned)batch, NT, sizeof(JacobiSmem<JM>), stream>>>(
      n, orders, (const double2*)A, lda, sA, W, (double2*)V, ldv, sV, max_sweeps, tol, sweeps);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

static int heev_small_impl(int64_t n, const int64_t* orders, const void* A, int64_t lda,
                           int64_t strideA, double* W, void* V, int64_t ldv, int64_t strideV,
                           int64_t batch, int max_sweeps, double tol, int* sweeps,
                           hipStream_t stream) {
  AQ_REQUIRE(A && W && V, AQ_ERR_ARG, "aq_heev_small: null pointer");
  AQ_REQUIRE(n >= 1 && n <= kJmax, AQ_ERR_ARG, "aq_heev_small: n must be in [1, %d] (got %lld)",
             kJmax, (long long)n);
  AQ_REQUIRE(batch >= 1 && batch < (1LL << 31), AQ_ERR_ARG, "aq_heev_small: bad batch");
  AQ_REQUIRE(lda >= n && ldv >= n, AQ_ERR_ARG, "aq_heev_small: leading dimension < n");
  AQ_REQUIRE(max_sweeps >= 1 && max_sweeps <= 100, AQ_ERR_ARG, "aq_heev_small: max_sweeps");
  // the tournament runs on the even order n + (n & 1) <= JM
  if (n <= 16)
    return launch_jacobi<16, 64>((int)n, orders, A, lda, strideA, W, V, ldv, strideV, batch,
                                 max_sweeps, tol, sweeps, stream);
  if (n <= 32)
    return launch_jacobi<32, 256>((int)n, orders, A, lda, strideA, W, V, ldv, strideV, batch,
                                  max_sweeps, tol, sweeps, stream);
  return launch_jacobi<64, 512>((int)n, orders, A, lda, strideA, W, V, ldv, strideV, batch,
                                max_sweeps, tol, sweeps, stream);
}

}  // namespace aq

using namespace aq;

extern "C" int aq_heev_small(int64_t n, const void* A, int64_t lda, int64_t strideA, double* W,
                             void* V, int64_t ldv, int64_t strideV, int64_t batch, int max_sweeps,
                             double tol, int* sweeps, hipStream_t stream) {
  return heev_small_impl(n, nullptr, A, lda, strideA, W, V, ldv, strideV, batch, max_sweeps, tol,
                         sweeps, stream);
}

extern "C" int aq_heev_small_orders(int64_t n, const int64_t* orders, const void* A, int64_t lda,
                                    int64_t strideA, double* W, void* V, int64_t ldv,
                                    int64_t strideV, int64_t batch, int max_sweeps, double tol,
                                    int* sweeps, hipStream_t stream) {
  AQ_REQUIRE(orders, AQ_ERR_ARG, "aq_heev_small_orders: null orders");
  return heev_small_impl(n, orders, A, lda, strideA, W, V, ldv, strideV, batch, max_sweeps, tol,
                         sweeps, stream);
}
