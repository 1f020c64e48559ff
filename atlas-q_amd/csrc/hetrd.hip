// Batched reduction of Hermitian matrices to real tridiagonal form (FP64),
// the library's own replacement of LAPACK's blocked zhetrd on the device.
//
// Output format: the LAPACK / rocSOLVER one for fill mode "upper" of the
// matrix read column-major (the conventions at the top of tridiag.hip): with
// A(r, c) = mem[r + c n] (our row-major buffer read column-major, i.e.
// conj(G)), the reflector H_i (i = 0 .. n-2) has v_i[i] = 1, v_i[0:i] in
// A(0:i, i+1) = memory row i+1, elements 0 .. i-1, tau[i], and
// A = Q T Q^H with Q = H_{n-2} ... H_0, T = tridiag(D, E) real.
//
// Algorithm: panels of nb columns from the last column down (LAPACK's zlatrd
// with uplo = upper).  Within a panel, column ii is first updated with the
// panel's earlier reflectors (A(0:ii+1, ii) -= V_p W_p(ii,:)^H + W_p V_p(ii,:)^H),
// then its reflector is formed, and W(:, ii) = tau (A_0 v - V_p (W_p^H v)
// - W_p (V_p^H v)) - 1/2 tau^2 (..)^H v v, where A_0 is the leading ii x ii
// block as it stood at the start of the panel.  After the panel the leading
// block takes the rank-2nb update A_0 -= V W^H + W V^H on MFMA.
//
// Per column the work is one streaming Hermitian matrix-vector product over
// the leading block -- each stored element (the upper triangle, r <= c) read
// once from HBM, for every matrix of the batch in one launch of 64 x 64
// tiles (hetrd_hemv_kernel: partial sums per tile, the tile's share of the
// quadratic form v^H A_0 v, and the dot products with the panel's earlier
// columns) -- then one wide kernel that forms W(:, ii) and updates the next
// column (hetrd_finish_kernel) and one per-matrix kernel for that column's
// reflector (hetrd_reflect_kernel).  The reduction reads half the data a
// full-square matrix-vector product would.
#include <float.h>
#include <math.h>

#include <algorithm>
#include <mutex>

#include "aq_common.h"

namespace aq {

constexpr int kHvTile = 64;     // hemv tile edge
constexpr int kStepThreads = 1024;

__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 nt_load(const double2* p) {
  typedef double dv2 __attribute__((ext_vector_type(2)));
  const dv2 t = __builtin_nontemporal_load((const dv2*)p);
  return make_double2(t.x, t.y);
}
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  // c + a b
  return make_double2(c.x + a.x * b.x - a.y * b.y, c.y + a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cfma_conj(double2 a, double2 b, double2 c) {
  // c + conj(a) b
  return make_double2(c.x + a.x * b.x + a.y * b.y, c.y + a.x * b.y - a.y * b.x);
}

// ------------------------------------------------------------ hemv partials
// For the leading ii x ii block of A (A(r, c) = mem[c n + r], stored for
// r <= c) and v = A(0:ii, ii) (memory row ii), tile (C, R), R <= C, of
// memory rows c in [64C, 64C+64) and elements r in [64R, 64R+64) writes
//   P1[C][r] = sum_c A(r, c) v_c          (r in tile)
//   P2[R][c] = sum_{r (< c)} conj(A(r, c)) v_r   (c in tile)
// to the partial buffer part[b][slot][n]: slot C for P1, slot nblk + R for P2.
// One wave per tile; lane (i, j) = (lane / 8, lane % 8) reads the 8 x 8
// sub-blocks at rows 8 bc + i, elements 8 br + j (each load instruction is 8
// full 128-byte rows), accumulates P1 over its rows in registers (finished
// by 3 shuffle steps at the end) and finishes each memory row's P2 as soon as
// that row is read (3 shuffle steps, one store).
__global__ __launch_bounds__(256) void hetrd_hemv_kernel(const double2* __restrict__ A,
                                                         int64_t sA, int n, int ii, int nblk,
                                                         int tiles, double2* __restrict__ part,
                                                         int64_t sPart, const double2* __restrict__ W,
                                                         int64_t sW, int iw, int later,
                                                         double2* __restrict__ dots,
                                                         int64_t sDots, double* __restrict__ qpart,
                                                         int64_t sQ) {
  const int64_t b = blockIdx.y;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) {
    // chunk q of rows r in [64q, 64q + 64): partial dot products of v with
    // the panel's later W columns (lane j < later) and reflectors (lane
    // 32 + j), dots[b][q][lane]
    const int q = t - tiles;
    const int lane = threadIdx.x & 63;
    if (q >= nblk) return;
    const double2* Ab = A + b * sA;
    const double2* v = Ab + (int64_t)ii * n;
    const int j = lane & 31;
    double2 s = make_double2(0.0, 0.0);
    if (j < later) {
      const double2* x = lane < 32 ? W + b * sW + (int64_t)(iw + 1 + j) * n
                                   : Ab + (int64_t)(ii + 1 + j) * n;
      const int r0 = kHvTile * q, r1 = min(ii, r0 + kHvTile);
      int r = r0;
      for (; r + 8 <= r1; r += 8) {
        double2 a[8], c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          a[u] = x[r + u];
          c[u] = v[r + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s = cfma_conj(a[u], c[u], s);
      }
      for (; r < r1; ++r) s = cfma_conj(x[r], v[r], s);
    }
    dots[b * sDots + (int64_t)q * 64 + lane] = s;
    return;
  }
  // triangular enumeration t = C (C + 1) / 2 + R
  int C = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (C * (C + 1) / 2 > t) --C;
  while ((C + 1) * (C + 2) / 2 <= t) ++C;
  const int R = t - C * (C + 1) / 2;
  const double2* Ab = A + b * sA;
  const double2* v = Ab + (int64_t)ii * n;  // memory row ii
  const int lane = threadIdx.x & 63, li = lane >> 3, lj = lane & 7;
  const bool diag = (R == C);
  const double2 zero = make_double2(0.0, 0.0);
  double2 vr[8], acc1[8];
  double qs = 0.0;  // this tile's share of v^H A_0 v
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int r = kHvTile * R + 8 * q + lj;
    vr[q] = r < ii ? v[r] : zero;
    acc1[q] = zero;
  }
  double2* Pb = part + b * sPart;
  // two rows of the tile in flight at a time (187 registers: two waves per
  // SIMD, so one wave's loads overlap another's arithmetic; with all eight
  // rows' loads hoisted the kernel held 256 registers and ran one wave per
  // SIMD, 10 % slower)
#pragma unroll 2
  for (int bc = 0; bc < 8; ++bc) {
    const int c = kHvTile * C + 8 * bc + li;
    const bool crow = c < ii;
    const double2 vc = crow ? v[c] : zero;
    const double2* rowp = Ab + (int64_t)(crow ? c : 0) * n + kHvTile * R + lj;
    double2 a[8];
#pragma unroll
    for (int br = 0; br < 8; ++br) {
      const int r = kHvTile * R + 8 * br + lj;
      const bool ok = crow && r < ii && (!diag || r <= c);
      // the triangle is streamed once per column: non-temporal loads keep
      // it from displacing the partial sums and vectors in the caches
      a[br] = ok ? nt_load(rowp + 8 * br) : zero;
    }
    double2 p2 = zero;
#pragma unroll
    for (int br = 0; br < 8; ++br) {
      const int r = kHvTile * R + 8 * br + lj;
      double2 x = a[br];
      const bool on_diag = diag && r == c;
      if (on_diag) x.y = 0.0;  // the diagonal of a Hermitian matrix is real
      acc1[br] = cfma(x, vc, acc1[br]);
      if (!on_diag) p2 = cfma_conj(x, vr[br], p2);
    }
    // p2 is this lane's part of P2[c]: Re(conj(v_c) p2) = Re(off-diagonal part)
    qs += vc.x * p2.x + vc.y * p2.y;
    // P2[c] over lj (lanes 8 li .. 8 li + 7), written at once
    for (int off = 1; off < 8; off <<= 1) {
      p2.x += __shfl_xor(p2.x, off);
      p2.y += __shfl_xor(p2.y, off);
    }
    if (lj == 0 && crow) Pb[(int64_t)(nblk + R) * n + c] = p2;
  }
  // v^H A_0 v over the tile's stored elements (r <= c): Re(sum conj(v_r) P1_r)
  // holds the diagonal and the real part of the off-diagonal sum once, the
  // P2 side above the off-diagonal part a second time
  {
#pragma unroll
    for (int q = 0; q < 8; ++q) qs += vr[q].x * acc1[q].x + vr[q].y * acc1[q].y;
    for (int off = 32; off > 0; off >>= 1) qs += __shfl_xor(qs, off);
    if (lane == 0) qpart[b * sQ + t] = qs;
  }
  // P1 over li (lanes j, j+8, ..)
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    for (int off = 8; off < 64; off <<= 1) {
      acc1[q].x += __shfl_xor(acc1[q].x, off);
      acc1[q].y += __shfl_xor(acc1[q].y, off);
    }
  }
  if (li == 0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int r = kHvTile * R + 8 * q + lj;
      if (r < ii) Pb[(int64_t)C * n + r] = acc1[q];
    }
  }
}

// ------------------------------------------------------------ block reductions
__device__ double2 block_sum(double2 x, double2* scratch) {
  // scratch: >= kStepThreads / 64 entries; all threads receive the sum
  for (int off = 32; off > 0; off >>= 1) {
    x.x += __shfl_xor(x.x, off);
    x.y += __shfl_xor(x.y, off);
  }
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[w] = x;
  __syncthreads();
  double2 s = make_double2(0.0, 0.0);
  for (int q = 0; q < nw; ++q) s = cadd(s, scratch[q]);
  return s;
}

// Householder reflector of LAPACK's zlarfg for x = A(0:m-1, col), alpha =
// A(m-1, col) (memory row col, elements 0 .. m-1): on return memory row col
// holds v (v[m-1] = 1 in place of beta), tau and E = beta are written.
__device__ void make_reflector(double2* rowcol, int m, double2* tau_out, double* e_out,
                               double2* scratch) {
  double2 loc = make_double2(0.0, 0.0);
  for (int r = threadIdx.x; r < m - 1; r += blockDim.x) {
    const double2 x = rowcol[r];
    loc.x += x.x * x.x + x.y * x.y;
  }
  const double xn2 = block_sum(loc, scratch).x;
  const double2 alpha = rowcol[m - 1];
  double2 tau;
  double beta;
  double2 scale;
  if (xn2 == 0.0 && alpha.y == 0.0) {
    tau = make_double2(0.0, 0.0);
    beta = alpha.x;
    scale = make_double2(1.0, 0.0);
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    // scale = 1 / (alpha - beta)
    const double dr = alpha.x - beta, di = alpha.y;
    const double den = dr * dr + di * di;
    scale = make_double2(dr / den, -di / den);
  }
  __syncthreads();
  for (int r = threadIdx.x; r < m - 1; r += blockDim.x) rowcol[r] = cmul(rowcol[r], scale);
  if (threadIdx.x == 0) {
    rowcol[m - 1] = make_double2(1.0, 0.0);
    *tau_out = tau;
    *e_out = beta;
  }
  __syncthreads();
}

// ------------------------------------------------------------ small orders
// Orders up to kSmallN: the whole reduction by one work-group per matrix
// (LAPACK's unblocked zhetd2, uplo "upper" on the column-major view, the
// output format above), with no launch per column.  The leading block is
// first made fully Hermitian in memory (the elements below the stored
// triangle get the conjugates of the stored ones), so that per column the
// matrix-vector product w = tau A v reads coalesced memory rows (1024 /
// kSmallN threads per row of A, partial sums in LDS) and the rank-2 update
// A -= v w^H + w v^H writes both triangles; three passes over the leading
// block per column, from L2 at these sizes.
constexpr int kSmallN = 256;

__global__ __launch_bounds__(1024) void hetrd_small_kernel(double2* __restrict__ A, int64_t sA,
                                                           int n, double* __restrict__ D,
                                                           double* __restrict__ E,
                                                           double2* __restrict__ tau) {
  constexpr int kParts = 1024 / kSmallN;
  __shared__ double2 v[kSmallN], w[kSmallN], wp[kParts][kSmallN];
  __shared__ double2 scratch[1024 / 64];
  __shared__ double2 s_tau;
  __shared__ double s_e;
  const int64_t b = blockIdx.x;
  double2* Ab = A + b * sA;
  // M(r, c) = Ab[c n + r]; the stored triangle is r <= c (memory row c,
  // elements 0 .. c); fill r > c with conj(M(c, r)) and make the diagonal real
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int c = e / n, r = e % n;
    if (r > c) {
      const double2 x = Ab[(int64_t)r * n + c];
      Ab[(int64_t)c * n + r] = make_double2(x.x, -x.y);
    } else if (r == c) {
      Ab[(int64_t)c * n + c].y = 0.0;
    }
  }
  __threadfence_block();
  __syncthreads();
  const int r4 = threadIdx.x % kSmallN, part = threadIdx.x / kSmallN;   // kParts x kSmallN threads
  for (int k = n - 1; k >= 1; --k) {
    double2* colk = Ab + (int64_t)k * n;          // M(0 .. n-1, k)
    make_reflector(colk, k, &s_tau, &s_e, scratch);
    const double2 tk = s_tau;
    const double ek = s_e;
    if (tk.x != 0.0 || tk.y != 0.0) {
      for (int r = threadIdx.x; r < k; r += blockDim.x) v[r] = colk[r];
      __syncthreads();
      // w = tau M v over the leading k x k block (kParts column ranges)
      {
        const int c0 = part * k / kParts, c1 = (part + 1) * k / kParts;
        double2 y = make_double2(0.0, 0.0);
        if (r4 < k)
          for (int c = c0; c < c1; ++c) y = cfma(Ab[(int64_t)c * n + r4], v[c], y);
        wp[part][r4] = y;
      }
      __syncthreads();
      double2 dl = make_double2(0.0, 0.0);
      if (threadIdx.x < k) {
        const int r = threadIdx.x;
        double2 y = wp[0][r];
#pragma unroll
        for (int q = 1; q < kParts; ++q) y = cadd(y, wp[q][r]);
        y = cmul(tk, y);
        w[r] = y;
        dl = make_double2(y.x * v[r].x + y.y * v[r].y, y.x * v[r].y - y.y * v[r].x);  // conj(w) v
      }
      const double2 dot = block_sum(dl, scratch);     // w^H v (syncs)
      const double2 a2 = make_double2(-0.5 * (tk.x * dot.x - tk.y * dot.y),
                                      -0.5 * (tk.x * dot.y + tk.y * dot.x));
      if (threadIdx.x < k) w[threadIdx.x] = cfma(a2, v[threadIdx.x], w[threadIdx.x]);
      __syncthreads();
      // M -= v w^H + w v^H on the leading k x k block
      for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
        const int c = e / k, r = e % k;
        const double2 vr = v[r], wc = w[c], wr = w[r], vc = v[c];
        double2 m = Ab[(int64_t)c * n + r];
        m.x -= vr.x * wc.x + vr.y * wc.y + wr.x * vc.x + wr.y * vc.y;
        m.y -= vr.y * wc.x - vr.x * wc.y + wr.y * vc.x - wr.x * vc.y;
        Ab[(int64_t)c * n + r] = m;
      }
      __threadfence_block();
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      colk[k - 1] = make_double2(ek, 0.0);
      E[b * (n - 1) + k - 1] = ek;
      tau[b * (n - 1) + k - 1] = tk;
      D[b * n + k] = colk[k].x;
    }
    __threadfence_block();
    __syncthreads();
  }
  if (threadIdx.x == 0) D[b * n] = Ab[0].x;
}

// ------------------------------------------------------------ wide column kernels
// The per-column work between two matrix-vector launches, spread over many
// work-groups per matrix (each work-group owns 256 rows r):
//   hetrd_finish_kernel   W(:, iw) = tau z + alpha v from the tile partials,
//                         the dots and the quadratic form; with `update`,
//                         column c = ii - 1 takes the panel's reflectors
//                         (rows 0 .. c) and the group writes its partial of
//                         that column's reflector norm
//   hetrd_reflect_kernel  per matrix: D[c] and column c's reflector (tau, E,
//                         scaled vector, unit element)
constexpr int kWide = 256;

__device__ __forceinline__ double2 block_sum256(double2 x, double2* scratch) {
  for (int off = 32; off > 0; off >>= 1) {
    x.x += __shfl_xor(x.x, off);
    x.y += __shfl_xor(x.y, off);
  }
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = x;
  __syncthreads();
  double2 s = scratch[0];
  for (int q = 1; q < kWide / 64; ++q) s = cadd(s, scratch[q]);
  return s;
}

// sum of the nblk + 1 tile partials of row r: slots blk .. nblk-1 (P1) and
// nblk .. nblk + blk (P2), blk = r / 64, one run of consecutive slots
__device__ __forceinline__ double2 row_partials(const double2* __restrict__ part, int n, int nblk,
                                                int r) {
  const double2* pp = part + (int64_t)(r / kHvTile) * n + r;
  const int cnt = nblk + 1;
  double2 y = make_double2(0.0, 0.0);
  int q = 0;
  for (; q + 8 <= cnt; q += 8) {
    double2 x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = pp[(int64_t)(q + u) * n];
#pragma unroll
    for (int u = 0; u < 8; ++u) y = cadd(y, x[u]);
  }
  for (; q < cnt; ++q) y = cadd(y, pp[(int64_t)q * n]);
  return y;
}

// W(:, iw) = tau z + alpha v, z = A_0 v - V_p t1 - W_p t2 (t1 = W_p^H v,
// t2 = V_p^H v over the panel's later columns).  LAPACK's alpha = -1/2 tau
// (tau z)^H v needs a dot product over all rows; here it comes from
// quantities every group holds: z^H v = v^H A_0 v - 2 Re(t1^H t2) (A_0
// Hermitian), so alpha = -1/2 |tau|^2 (q - 2 Re(t1^H t2)) with q summed from
// the tiles' shares.  With `update`, column c = ii - 1 (rows 0 .. c) takes the
// panel's reflectors,
//   A(r, c) -= sum_{j < cnt} A(r, c+1+j) conj(W(c, iw+j)) + W(r, iw+j) conj(A(c, c+1+j)),
// whose terms j >= 1 read the same panel elements as z's correction (one
// pass does both), and each group writes its partial of the reflector norm.
// W(c, iw), the update's j = 0 coefficient, is computed by every group from
// row c's partials (one wave, lanes over the slots and panel columns).
__global__ __launch_bounds__(kWide) void hetrd_finish_kernel(
    double2* __restrict__ A, int64_t sA, int n, int N, int nb, int ii, int nblk, int tiles,
    int update, double2* __restrict__ W, int64_t sW, const double2* __restrict__ part,
    int64_t sPart, const double2* __restrict__ dots, int64_t sDots,
    const double* __restrict__ qpart, int64_t sQ, const double2* __restrict__ tau,
    double* __restrict__ nrm, int64_t sNrm) {
  __shared__ double2 t1[32], t2[32], u1[33], u2[33], dq[kWide / 64][64], scratch[kWide / 64];
  __shared__ double2 s_wc;
  const int64_t b = blockIdx.y;
  double2* Ab = A + b * sA;
  double2* Wb = W + b * sW;
  const double2* Pb = part + b * sPart;
  const int p0 = N - nb, iw = ii - p0, later = N - 1 - ii;
  const int c = ii - 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double2 zero = make_double2(0.0, 0.0);
  const double2* v = Ab + (int64_t)ii * n;
  const double2* ap = Ab + (int64_t)(ii + 1) * n;  // panel reflector j at ap + j n
  const double2* wp = Wb + (int64_t)(iw + 1) * n;  // panel W column j at wp + j n
  // phase 1: every load that does not depend on a reduction
  double2 dacc = zero;
  if (later > 0)
    for (int q = wv; q < nblk; q += kWide / 64) dacc = cadd(dacc, dots[b * sDots + (int64_t)q * 64 + lane]);
  dq[wv][lane] = dacc;
  double qs = 0.0;
  for (int t = threadIdx.x; t < tiles; t += kWide) qs += qpart[b * sQ + t];
  double2 pc = zero, ac = zero, wc = zero;  // row c pieces (wave 1)
  if (update) {
    if (wv == 1) {
      for (int q = lane; q <= nblk; q += 64) pc = cadd(pc, Pb[(int64_t)(c / kHvTile + q) * n + c]);
      if (lane < later) {
        ac = ap[(int64_t)lane * n + c];
        wc = wp[(int64_t)lane * n + c];
      }
    }
    if (threadIdx.x < later) {
      u1[threadIdx.x + 1] = cconj(wp[(int64_t)threadIdx.x * n + c]);
      u2[threadIdx.x + 1] = cconj(Ab[(int64_t)(c + 2 + threadIdx.x) * n + c]);
    }
  }
  const int r = blockIdx.x * kWide + threadIdx.x;
  const bool own = r < ii;
  double2 z = zero, vr = zero, s = zero;
  if (own) {
    vr = v[r];
    if (update) s = Ab[(int64_t)c * n + r];
    z = row_partials(Pb, n, nblk, r);
  }
  // phase 2: t1, t2 and q
  __syncthreads();
  if (threadIdx.x < 64) {
    double2 acc = zero;
#pragma unroll
    for (int w = 0; w < kWide / 64; ++w) acc = cadd(acc, dq[w][threadIdx.x]);
    const int j = threadIdx.x & 31;
    if (j < later) {
      if (threadIdx.x < 32) t1[j] = acc;
      else t2[j] = acc;
    }
  }
  const double q = block_sum256(make_double2(qs, 0.0), scratch).x;  // syncs: t1, t2 ready
  double re12 = 0.0;
  for (int j = 0; j < later; ++j) re12 += t1[j].x * t2[j].x + t1[j].y * t2[j].y;
  const double2 tu = tau[b * (int64_t)(n - 1) + ii - 1];
  const double al = -0.5 * (tu.x * tu.x + tu.y * tu.y) * (q - 2.0 * re12);
  if (update && wv == 1) {
    double2 x = pc;
    if (lane < later) {
      const double2 p = t1[lane], g = t2[lane];
      x.x -= ac.x * p.x - ac.y * p.y + wc.x * g.x - wc.y * g.y;
      x.y -= ac.x * p.y + ac.y * p.x + wc.x * g.y + wc.y * g.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      x.x += __shfl_xor(x.x, off);
      x.y += __shfl_xor(x.y, off);
    }
    if (lane == 0) {
      const double2 vc = v[c];
      const double2 w = make_double2(tu.x * x.x - tu.y * x.y + al * vc.x,
                                     tu.x * x.y + tu.y * x.x + al * vc.y);
      s_wc = w;
      u1[0] = cconj(w);
      u2[0] = cconj(vc);
    }
  }
  // phase 3: the panel pass for z (and the update's terms j >= 1)
  if (own && later > 0) {
    const double2* a = ap + r;
    const double2* w = wp + r;
    int j = 0;
    for (; j + 8 <= later; j += 8) {
      double2 xa[8], xw[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        xa[u] = a[(int64_t)(j + u) * n];
        xw[u] = w[(int64_t)(j + u) * n];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double2 p = t1[j + u], g = t2[j + u];
        z.x -= xa[u].x * p.x - xa[u].y * p.y + xw[u].x * g.x - xw[u].y * g.y;
        z.y -= xa[u].x * p.y + xa[u].y * p.x + xw[u].x * g.y + xw[u].y * g.x;
        if (update) {
          const double2 p2 = u1[j + u + 1], g2 = u2[j + u + 1];
          s.x -= xa[u].x * p2.x - xa[u].y * p2.y + xw[u].x * g2.x - xw[u].y * g2.y;
          s.y -= xa[u].x * p2.y + xa[u].y * p2.x + xw[u].x * g2.y + xw[u].y * g2.x;
        }
      }
    }
    for (; j < later; ++j) {
      const double2 xa = a[(int64_t)j * n], xw = w[(int64_t)j * n];
      const double2 p = t1[j], g = t2[j];
      z.x -= xa.x * p.x - xa.y * p.y + xw.x * g.x - xw.y * g.y;
      z.y -= xa.x * p.y + xa.y * p.x + xw.x * g.y + xw.y * g.x;
      if (update) {
        const double2 p2 = u1[j + 1], g2 = u2[j + 1];
        s.x -= xa.x * p2.x - xa.y * p2.y + xw.x * g2.x - xw.y * g2.y;
        s.y -= xa.x * p2.y + xa.y * p2.x + xw.x * g2.y + xw.y * g2.x;
      }
    }
  }
  if (update) __syncthreads();  // s_wc, u1[0], u2[0]
  double part2 = 0.0;
  if (own) {
    const double2 wfin = (update && r == c)
                             ? s_wc  // the value the update's coefficient used
                             : make_double2(tu.x * z.x - tu.y * z.y + al * vr.x,
                                            tu.x * z.y + tu.y * z.x + al * vr.y);
    Wb[(int64_t)iw * n + r] = wfin;
    if (update) {
      const double2 p = u1[0], g = u2[0];
      s.x -= vr.x * p.x - vr.y * p.y + wfin.x * g.x - wfin.y * g.y;
      s.y -= vr.x * p.y + vr.y * p.x + wfin.x * g.y + wfin.y * g.x;
      if (r == c) s.y = 0.0;
      Ab[(int64_t)c * n + r] = s;
      if (r + 1 < c) part2 = s.x * s.x + s.y * s.y;
    }
  }
  if (update) {
    const double tot = block_sum256(make_double2(part2, 0.0), scratch).x;
    if (threadIdx.x == 0) nrm[b * sNrm + blockIdx.x] = tot;
  }
}

__global__ __launch_bounds__(kWide) void hetrd_reflect_kernel(
    double2* __restrict__ A, int64_t sA, int n, int ii, double2* __restrict__ tau, int nred,
    const double* __restrict__ nrm, int64_t sNrm, double* __restrict__ D,
    double* __restrict__ E) {
  __shared__ double2 s_scale;
  const int64_t b = blockIdx.x;
  double2* Ab = A + b * sA;
  double2* Tb = tau + b * (int64_t)(n - 1);
  const int c = ii - 1;
  double2* col = Ab + (int64_t)c * n;
  if (threadIdx.x == 0) {
    D[b * (int64_t)n + c] = col[c].x;
    s_scale = make_double2(1.0, 0.0);
    if (c > 0) {
      double xn2 = 0.0;
      for (int q = 0; q < nred; ++q) xn2 += nrm[b * sNrm + q];
      const double2 alpha = col[c - 1];
      double2 t;
      double beta;
      double2 scale;
      if (xn2 == 0.0 && alpha.y == 0.0) {
        t = make_double2(0.0, 0.0);
        beta = alpha.x;
        scale = make_double2(1.0, 0.0);
      } else {
        const double nrm2 = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        beta = alpha.x >= 0.0 ? -nrm2 : nrm2;
        t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
        const double dr = alpha.x - beta, di = alpha.y;
        const double den = dr * dr + di * di;
        scale = make_double2(dr / den, -di / den);
      }
      Tb[c - 1] = t;
      E[b * (int64_t)(n - 1) + c - 1] = beta;
      s_scale = scale;
    }
  }
  __syncthreads();
  if (c > 0) {
    const double2 scale = s_scale;
    for (int r = threadIdx.x; r < c - 1; r += blockDim.x) col[r] = cmul(col[r], scale);
    if (threadIdx.x == 0) col[c - 1] = make_double2(1.0, 0.0);
  }
}

// ------------------------------------------------------------ panel head
// A panel's first column ii is already current (the last rank-2nb update
// covered it): record D[ii] and, when `reflect` and ii > 0, form its reflector.
__global__ __launch_bounds__(kStepThreads) void hetrd_head_kernel(
    double2* __restrict__ A, int64_t sA, int n, int ii, int reflect, double* __restrict__ D,
    double* __restrict__ E, double2* __restrict__ tau) {
  __shared__ double2 scratch[kStepThreads / 64];
  const int64_t b = blockIdx.x;
  double2* Ab = A + b * sA;
  if (threadIdx.x == 0) {
    const double2 d = Ab[(int64_t)ii * n + ii];
    D[b * (int64_t)n + ii] = d.x;
    Ab[(int64_t)ii * n + ii] = make_double2(d.x, 0.0);
  }
  __syncthreads();
  if (reflect && ii > 0)
    make_reflector(Ab + (int64_t)ii * n, ii, tau + b * (int64_t)(n - 1) + (ii - 1),
                   E + b * (int64_t)(n - 1) + (ii - 1), scratch);
}

}  // namespace aq

using namespace aq;

extern "C" int64_t aq_hetrd_workspace_elems(int64_t n, int64_t nb, int64_t batch) {
  const int64_t nblk = cdiv(n, kHvTile);
  // [V; W; V] | tile partials | dot partials | (norm partials, quadratic-form
  // shares) as doubles
  return batch * (3 * nb * n + 2 * nblk * n + 64 * nblk +
                  (cdiv(n, kWide) + nblk * (nblk + 1) / 2 + 1) / 2 + 1);
}

// copies the panel's reflector vectors (memory rows p0 .. p0+nb-1, elements
// 0 .. p0-1) into rows 0 .. nb-1 and 2nb .. 3nb-1 of the [V; W; V] block
__global__ void hetrd_copy_v_kernel(const double2* __restrict__ A, int64_t sA, int n, int p0,
                                    int nb, double2* __restrict__ VWV, int64_t sV) {
  const int64_t b = blockIdx.y;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)nb * p0;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / p0), r = (int)(idx % p0);
    const double2 v = A[b * sA + (int64_t)(p0 + j) * n + r];
    VWV[b * sV + (int64_t)j * n + r] = v;
    VWV[b * sV + (int64_t)(2 * nb + j) * n + r] = v;
  }
}

// The reduction itself.  ws: aq_hetrd_workspace_elems(n, nb, batch) complex128.
// bgemm: the library's batched GEMM entry point (the rank-2nb updates).
extern "C" int aq_bgemm_lower(int64_t M, int64_t K, double alpha, const void* A, int64_t a_s0,
                              int64_t a_s1, int64_t a_sb, int conj_a, const void* B,
                              int64_t b_s0, int64_t b_s1, int64_t b_sb, int conj_b, double beta,
                              void* C, int64_t c_s0, int64_t c_s1, int64_t c_sb, int64_t batch,
                              hipStream_t stream);

namespace aq {
// Worker streams of the reduction (default priority), created once per device.
constexpr int kHetrdStreams = 2;
static hipStream_t g_hetrd_streams[64][kHetrdStreams];
static std::mutex g_hetrd_lock;
static int g_hetrd_max_groups = kHetrdStreams;

static int hetrd_streams(hipStream_t* out) {
  int dev = 0;
  AQ_HIP(hipGetDevice(&dev));
  AQ_REQUIRE(dev >= 0 && dev < 64, AQ_ERR_UNSUPPORTED, "aq_hetrd_own: device %d", dev);
  std::lock_guard<std::mutex> g(g_hetrd_lock);
  for (int i = 0; i < kHetrdStreams; ++i) {
    if (!g_hetrd_streams[dev][i]) AQ_HIP(hipStreamCreateWithFlags(&g_hetrd_streams[dev][i],
                                                                  hipStreamNonBlocking));
    out[i] = g_hetrd_streams[dev][i];
  }
  return AQ_OK;
}

// The column loop of one group of matrices on one stream.
struct HetrdGroup {
  double2 *A, *VWV, *W, *part, *dots, *tau;
  double *D, *E, *nrm, *qp;
  int64_t batch;
  hipStream_t s;
};
}  // namespace aq

namespace aq {
// The column loop over up to kHetrdStreams groups of consecutive matrices,
// group q (cnt[q] matrices) on stream st[q], their launches interleaved
// column by column.
static int hetrd_columns(int64_t n, int64_t nb, void* Av, int64_t strideA, double* D, double* E,
                         void* tauv, void* ws, int64_t batch, int ngroups, const int64_t* cnt,
                         const hipStream_t* streams) {
  const int64_t nblk_max = cdiv(n, kHvTile);
  const int64_t sW = 3 * nb * n, sPart = 2 * nblk_max * n, sDots = 64 * nblk_max;
  const int64_t nred_max = cdiv(n, kWide), tiles_max = nblk_max * (nblk_max + 1) / 2;
  const int64_t sR = nred_max + tiles_max;  // doubles per matrix: norms, then quadratic form
  double2* VWV0 = (double2*)ws;
  double2* part0 = VWV0 + batch * sW;
  double2* dots0 = part0 + batch * sPart;
  double* r0 = (double*)(dots0 + batch * sDots);
  HetrdGroup g[kHetrdStreams];
  int64_t start = 0;
  for (int q = 0; q < ngroups; ++q) {
    g[q].A = (double2*)Av + start * strideA;
    g[q].VWV = VWV0 + start * sW;
    g[q].W = g[q].VWV + nb * n;
    g[q].part = part0 + start * sPart;
    g[q].dots = dots0 + start * sDots;
    g[q].nrm = r0 + start * sR;
    g[q].qp = r0 + start * sR + nred_max;
    g[q].tau = (double2*)tauv + start * (n - 1);
    g[q].D = D + start * n;
    g[q].E = E + start * (n - 1);
    g[q].batch = cnt[q];
    g[q].s = streams[q];
    start += cnt[q];
  }
  int64_t N = n;
  for (; N > 1; N -= nb) {
    const int64_t p0 = N - nb;
    for (int q = 0; q < ngroups; ++q) {
      // the panel's first column: diagonal and reflector
      hetrd_head_kernel<<<(unsigned)g[q].batch, kStepThreads, 0, g[q].s>>>(
          g[q].A, strideA, (int)n, (int)(N - 1), 1, g[q].D, g[q].E, g[q].tau);
      AQ_LAUNCH_CHECK();
    }
    const int64_t last = p0 > 0 ? p0 : 1;  // lowest column of the panel with a reflector
    for (int64_t ii = N - 1; ii >= last; --ii) {
      const int64_t nblk = cdiv(ii, kHvTile);
      const int64_t tiles = nblk * (nblk + 1) / 2;
      const int64_t later = N - 1 - ii;
      const int64_t extra = later > 0 ? nblk : 0;  // dot-product waves
      for (int q = 0; q < ngroups; ++q) {
        dim3 grid((unsigned)cdiv(tiles + extra, 4), (unsigned)g[q].batch);
        hetrd_hemv_kernel<<<grid, 256, 0, g[q].s>>>(g[q].A, strideA, (int)n, (int)ii, (int)nblk,
                                                    (int)tiles, g[q].part, sPart, g[q].W, sW,
                                                    (int)(ii - p0), (int)later, g[q].dots, sDots,
                                                    g[q].qp, sR);
        AQ_LAUNCH_CHECK();
        // W(ii) from the partials; then, inside the panel, column ii-1's
        // update and reflector
        const int64_t nred = cdiv(ii, kWide);
        dim3 gw((unsigned)nred, (unsigned)g[q].batch);
        const int update = ii - 1 >= p0 ? 1 : 0;  // column ii-1 lies in the panel
        hetrd_finish_kernel<<<gw, kWide, 0, g[q].s>>>(
            g[q].A, strideA, (int)n, (int)N, (int)nb, (int)ii, (int)nblk, (int)tiles, update,
            g[q].W, sW, g[q].part, sPart, g[q].dots, sDots, g[q].qp, sR, g[q].tau, g[q].nrm, sR);
        AQ_LAUNCH_CHECK();
        if (update) {
          hetrd_reflect_kernel<<<(unsigned)g[q].batch, kWide, 0, g[q].s>>>(
              g[q].A, strideA, (int)n, (int)ii, g[q].tau, (int)nred, g[q].nrm, sR, g[q].D,
              g[q].E);
          AQ_LAUNCH_CHECK();
        }
      }
    }
    if (p0 > 0) {
      // A(0:p0, 0:p0) -= V W^H + W V^H, V = A(0:p0, p0:N) (memory rows p0 .. N-1),
      // in memory coordinates C[c][r] = A(r, c) (only r <= c is ever read):
      //   C -= [conj(W) conj(V)] [V; W]  with [W; V] and [V; W] rows of the block
      for (int q = 0; q < ngroups; ++q) {
        dim3 gr((unsigned)std::min<int64_t>(cdiv(nb * p0, 256), 1024), (unsigned)g[q].batch);
        hetrd_copy_v_kernel<<<gr, 256, 0, g[q].s>>>(g[q].A, strideA, (int)n, (int)p0, (int)nb,
                                                    g[q].VWV, sW);
        AQ_LAUNCH_CHECK();
        const int st = aq_bgemm_lower(p0, 2 * nb, -1.0, g[q].W, 1, n, sW, 1, g[q].VWV, n, 1, sW,
                                      0, 1.0, g[q].A, n, 1, strideA, g[q].batch, g[q].s);
        if (st != AQ_OK) return st;
      }
    }
  }
  if (N == 1) {
    // the last panel ended at column 1: column 0 (current after the last
    // rank-2nb update) has no reflector, only its diagonal
    for (int q = 0; q < ngroups; ++q) {
      hetrd_head_kernel<<<(unsigned)g[q].batch, kStepThreads, 0, g[q].s>>>(
          g[q].A, strideA, (int)n, 0, 0, g[q].D, g[q].E, g[q].tau);
      AQ_LAUNCH_CHECK();
    }
  }
  return AQ_OK;
}
}  // namespace aq

extern "C" int aq_hetrd_own_groups(int64_t n, int64_t nb, void* Av, int64_t strideA,
                                   double* D, double* E, void* tauv, void* ws, int64_t ws_elems,
                                   int64_t batch, int groups, hipStream_t stream) {
  AQ_REQUIRE(Av && D && E && tauv && ws, AQ_ERR_ARG, "aq_hetrd_own: null pointer");
  AQ_REQUIRE(groups >= 1 && groups <= kHetrdStreams, AQ_ERR_ARG,
             "aq_hetrd_own: groups %d out of range", groups);
  AQ_REQUIRE(n >= 2 && n <= 8192 && nb >= 1 && nb <= 32, AQ_ERR_ARG,
             "aq_hetrd_own: n %lld / nb %lld out of range", (long long)n, (long long)nb);
  AQ_REQUIRE(batch >= 1 && batch <= 65535 && strideA >= n * n, AQ_ERR_ARG,
             "aq_hetrd_own: bad batch or stride");
  AQ_REQUIRE(ws_elems >= aq_hetrd_workspace_elems(n, nb, batch), AQ_ERR_ARG,
             "aq_hetrd_own: workspace too small");
  if (n <= kSmallN) {
    hetrd_small_kernel<<<(unsigned)batch, 1024, 0, stream>>>((double2*)Av, strideA, (int)n, D, E,
                                                             (double2*)tauv);
    AQ_LAUNCH_CHECK();
    return AQ_OK;
  }
  // The batch runs as up to two groups on worker streams, their column steps
  // interleaved: one group's small per-matrix step overlaps the other's
  // streaming matrix-vector launch.
  // (from two matrices on: at small batches every launch is latency-bound,
  // and the second group's chain fills the first one's gaps)
  const int ngroups = batch >= 2 ? groups : 1;
  hipStream_t st[kHetrdStreams];
  int64_t cnt[kHetrdStreams];
  if (ngroups > 1) {
    const int rc = hetrd_streams(st);
    if (rc != AQ_OK) return rc;
  } else {
    st[0] = stream;
  }
  for (int q = 0; q < ngroups; ++q) cnt[q] = batch / ngroups + (q < batch % ngroups ? 1 : 0);
  hipEvent_t ready = nullptr;
  if (ngroups > 1) {
    AQ_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    AQ_HIP(hipEventRecord(ready, stream));
    for (int q = 0; q < ngroups; ++q) AQ_HIP(hipStreamWaitEvent(st[q], ready, 0));
  }
  const int rc = hetrd_columns(n, nb, Av, strideA, D, E, tauv, ws, batch, ngroups, cnt, st);
  if (rc != AQ_OK) return rc;
  if (ngroups > 1) {
    for (int q = 0; q < ngroups; ++q) {
      AQ_HIP(hipEventRecord(ready, st[q]));
      AQ_HIP(hipStreamWaitEvent(stream, ready, 0));
    }
    AQ_HIP(hipEventDestroy(ready));
  }
  return AQ_OK;
}

extern "C" int aq_hetrd_own(int64_t n, int64_t nb, void* Av, int64_t strideA, double* D,
                            double* E, void* tauv, void* ws, int64_t ws_elems, int64_t batch,
                            hipStream_t stream) {
  return aq_hetrd_own_groups(n, nb, Av, strideA, D, E, tauv, ws, ws_elems, batch,
                             g_hetrd_max_groups, stream);
}

// Stream groups of the following aq_hetrd_own calls (process-wide): 1 keeps
// the reduction on the caller's stream (for HIP graph capture, where the
// worker streams would have to join the capture), 2 (the default) overlaps
// two groups' column chains.  Returns the previous value.
extern "C" int aq_hetrd_max_groups(int groups) {
  const int old = g_hetrd_max_groups;
  if (groups >= 1 && groups <= kHetrdStreams) g_hetrd_max_groups = groups;
  return old;
}
