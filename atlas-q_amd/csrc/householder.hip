// Householder factorisations in FP64 on the batched MFMA GEMM, built on one
// panel kernel (panel_qr_kernel: zgeqr2 + zlarft on up to 32 columns, one
// work-group per matrix):
//   * the band reduction, first stage of the two-stage tridiagonalisation
//     (aq_he2hb_own, below);
//   * the blocked QR factorisation and the thin Q (aq_geqrf / aq_ungqr, at the
//     end of the file; LAPACK's zgeqrf / zungqr, so R and Q match LAPACK's
//     up to rounding: R's diagonal is real, tau from zlarfg).
//
// Band reduction: batched reduction of Hermitian matrices to band form, the
// first stage of the two-stage tridiagonalisation: A = Q1 B Q1^H with B Hermitian
// of bandwidth 32 (B(r, c) = 0 for |r - c| > 32).
//
// Panels of 32 columns, left to right (LAPACK's zhetrd_he2hb, uplo = lower):
// for the panel at columns j .. j+31, the block below the band, P =
// A(j+32 : n, j : j+32) (m x 32, m = n - j - 32), is factored P = Q_p R by
// Householder reflectors H_c = I - tau_c v_c v_c^H (zgeqr2; Q_p = H_0 .. H_31
// = I - V T V^H, T upper triangular, zlarft), and the trailing block takes
// the two-sided update A22 <- Q_p^H A22 Q_p = A22 - W V^H - V W^H with
// Y = A22 V T, W = Y - 1/2 V (T^H V^H Y).  The panel factorisation runs in
// one work-group per matrix (panel_qr_kernel); every product involving the
// trailing block runs on the batched MFMA GEMM (aq_bgemm), the large ones
// being Y^H = V^H A22 (one read of A22) and the rank-64 update (one read and
// write of A22).  A is held in full (both triangles) so both are plain
// GEMMs.
//
// Storage (row-major A, A(r, c) = mem[r n + c]): after the call the panel
// region A(j+32+r, j+c) holds R (r <= c) and the reflector entries v_c[r]
// (r > c, LAPACK's layout; v_c[c] = 1 implied); T of panel p is written to
// Tall[p]; the band is copied to AB (AB[c][d] = B(c + d, c), d = 0 .. 63,
// zero for d > 32: the second stage needs room for its bulges).
#include <math.h>

#include <algorithm>
#include <mutex>

#include "aq_common.h"

namespace aq {

constexpr int kPB = 32;            // panel width = bandwidth
constexpr int kQRT = 1024;         // panel kernel threads: 32 row groups x 32 columns
constexpr int kQRMaxRows = 4096;   // rows of a panel (the current reflector lives in LDS)
constexpr int kBandLd = 2 * kPB;   // entries per column of the band array

__device__ __forceinline__ double2 dconj(double2 a) { return make_double2(a.x, -a.y); }

// first row >= start in row group g (rows r with r % 32 == g)
__device__ __forceinline__ int first_row(int start, int g) {
  return start + (((g - start) % 32) + 32) % 32;
}

// lower triangle -> upper (A(c, r) = conj(A(r, c)), r > c); real diagonal
__global__ void sym_fill_kernel(double2* __restrict__ A, int64_t sA, int n) {
  const int64_t b = blockIdx.z;
  double2* Ab = A + b * sA;
  // 32 x 32 tiles of the upper triangle: tile (tr, tc), tr <= tc
  __shared__ double2 t[32][33];
  const int tr = blockIdx.y, tc = blockIdx.x;
  if (tr > tc) return;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;  // 256 threads: 8 rows per pass
  // read the mirror tile (rows tc*32.., cols tr*32..) from the lower triangle
  for (int q = ly; q < 32; q += 8) {
    const int r = tc * 32 + q, c = tr * 32 + lx;
    t[q][lx] = (r < n && c < n) ? Ab[(int64_t)r * n + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int q = ly; q < 32; q += 8) {
    const int r = tr * 32 + q, c = tc * 32 + lx;  // upper element (r, c), r <= c in the tile sense
    if (r < n && c < n) {
      if (r < c) Ab[(int64_t)r * n + c] = dconj(t[lx][q]);
      else if (r == c) Ab[(int64_t)r * n + c] = make_double2(t[lx][q].x, 0.0);
    }
  }
}

// Householder QR of one panel of ncol <= 32 columns per matrix (zgeqr2 +
// zlarft).  P(r, c) = P[r ld + c], r < m, c < ncol.  On return: R in P's upper
// triangle, v_c[r] (r > c) below it; V (unit diagonal, zeros above) at
// V[r ldv + c]; its conjugate transpose at Vh[c ldh + r]; T (32 x 32 upper) at
// Tf[c * 32 + c'].  Columns c >= min(m, ncol) get tau = 0 and v = 0 (V, Vh and
// T always have 32 columns).
//
// Thread (g, col) = (tid / 32, tid % 32) owns the rows r = g mod 32 of column
// col.  Per reflector c three passes: dots (v_c^H P(:, c') for c' > c, and
// V(:, c')^H v_c for c' < c, the column of V^H V that T needs), the update
// of the columns c' > c (the threads of column c+1 also sum the norm of its
// new subdiagonal part), and the scaling of column c+1 into v_{c+1} by all
// threads.  Row loops issue 8 loads before using them.
__device__ __forceinline__ void qr_reflector(double2 alpha, double xn2, double2& tau,
                                             double2& scale, double& beta) {
  tau = make_double2(0.0, 0.0);
  scale = make_double2(1.0, 0.0);
  beta = alpha.x;
  if (!(xn2 == 0.0 && alpha.y == 0.0)) {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    const double dr = alpha.x - beta, di = alpha.y, den = dr * dr + di * di;
    scale = make_double2(dr / den, -di / den);
  }
}

__global__ __launch_bounds__(kQRT) void panel_qr_kernel(double2* __restrict__ P, int64_t sP,
                                                        int64_t ld, int m, int ncol,
                                                        double2* __restrict__ V, int64_t sV,
                                                        int64_t ldv, double2* __restrict__ Vh,
                                                        int64_t sVh, int64_t ldh,
                                                        double2* __restrict__ Tf, int64_t sT) {
  __shared__ double2 vcur[kQRMaxRows];
  __shared__ double2 red[32][kPB + 1];
  __shared__ double2 Ts[kPB][kPB + 1];
  __shared__ double2 wv[kPB];
  __shared__ double2 s_tau, s_scale;
  __shared__ double s_beta;
  const int64_t b = blockIdx.x;
  double2* Pb = P + b * sP;
  const int tid = threadIdx.x, col = tid & 31, g = tid >> 5;
  const double2 zero = make_double2(0.0, 0.0);
  for (int i = tid; i < kPB * (kPB + 1); i += kQRT) Ts[i / (kPB + 1)][i % (kPB + 1)] = zero;
  // prologue: norm of column 0 below the diagonal
  {
    double part = 0.0;
    if (col == 0)
      for (int r = first_row(1, g); r < m; r += 32) {
        const double2 x = Pb[(int64_t)r * ld];
        part += x.x * x.x + x.y * x.y;
      }
    if (col == 0) red[g][0] = make_double2(part, 0.0);
  }
  for (int c = 0; c < kPB; ++c) {
    // (1) reflector c from the norm partials in red[*][0]
    __syncthreads();
    if (tid == 0) {
      double xn2 = 0.0;
      for (int q = 0; q < 32; ++q) xn2 += red[q][0].x;
      double2 tau = zero, scale = make_double2(1.0, 0.0);
      double beta = 0.0;
      if (c < m && c < ncol) qr_reflector(Pb[(int64_t)c * ld + c], xn2, tau, scale, beta);
      s_tau = tau;
      s_scale = scale;
      s_beta = beta;
      Ts[c][c] = tau;
    }
    __syncthreads();
    if (c >= m || c >= ncol) continue;  // uniform: nothing left to annihilate
    const double2 tau = s_tau;
    {
      const double2 scale = s_scale;
      for (int r = c + tid; r < m; r += kQRT) {
        double2 v;
        if (r == c) {
          v = make_double2(1.0, 0.0);
          Pb[(int64_t)r * ld + c] = make_double2(s_beta, 0.0);
        } else {
          v = cmul(Pb[(int64_t)r * ld + c], scale);
          Pb[(int64_t)r * ld + c] = v;
        }
        vcur[r] = v;
      }
    }
    __syncthreads();
    // (2) dots over rows r >= c
    {
      double2 s = zero;
      if (col != c && col < ncol) {
        const double sg = col > c ? -1.0 : 1.0;  // conj(v) x  or  conj(x) v
        int r = first_row(c, g);
        for (; r + 7 * 32 < m; r += 8 * 32) {
          double2 x[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) x[u] = Pb[(int64_t)(r + 32 * u) * ld + col];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const double2 v = vcur[r + 32 * u];
            s.x += v.x * x[u].x + v.y * x[u].y;
            s.y += sg * (v.y * x[u].x - v.x * x[u].y);
          }
        }
        for (; r < m; r += 32) {
          const double2 v = vcur[r], x = Pb[(int64_t)r * ld + col];
          s.x += v.x * x.x + v.y * x.y;
          s.y += sg * (v.y * x.x - v.x * x.y);
        }
      }
      red[g][col] = s;
    }
    __syncthreads();
    if (tid < kPB) {
      double2 t = zero;
      for (int q = 0; q < 32; ++q) t = cadd(t, red[q][tid]);
      wv[tid] = t;
    }
    __syncthreads();
    // T(0:c, c) = -tau T(0:c, 0:c) (V(:, 0:c)^H v)
    if (tid < c) {
      double2 t = zero;
      for (int q = tid; q < c; ++q) t = cadd(t, cmul(Ts[tid][q], wv[q]));
      const double2 u = cmul(tau, t);
      Ts[tid][c] = make_double2(-u.x, -u.y);
    }
    // (3) P(:, c') -= conj(tau) v (v^H P(:, c')), c' > c; column c+1 also
    // sums the norm of its rows > c+1
    double part = 0.0;
    if (col > c && col < ncol) {
      const double2 coef = cmul(dconj(tau), wv[col]);
      const bool nxt = col == c + 1;
      int r = first_row(c, g);
      for (; r + 7 * 32 < m; r += 8 * 32) {
        double2 x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = Pb[(int64_t)(r + 32 * u) * ld + col];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const double2 v = vcur[r + 32 * u];
          x[u].x -= v.x * coef.x - v.y * coef.y;
          x[u].y -= v.x * coef.y + v.y * coef.x;
          Pb[(int64_t)(r + 32 * u) * ld + col] = x[u];
          if (nxt && r + 32 * u > c + 1) part += x[u].x * x[u].x + x[u].y * x[u].y;
        }
      }
      for (; r < m; r += 32) {
        const double2 v = vcur[r];
        double2 x = Pb[(int64_t)r * ld + col];
        x.x -= v.x * coef.x - v.y * coef.y;
        x.y -= v.x * coef.y + v.y * coef.x;
        Pb[(int64_t)r * ld + col] = x;
        if (nxt && r > c + 1) part += x.x * x.x + x.y * x.y;
      }
    }
    __syncthreads();  // wv and red[*][*] reads are done
    if (col == c + 1) red[g][0] = make_double2(part, 0.0);
  }
  __syncthreads();
  // clean V, V^H and T
  double2* Vb = V + b * sV;
  double2* Vhb = Vh + b * sVh;
  for (int r = g; r < m; r += 32) {
    double2 x = zero;
    if (col < ncol) {
      if (r == col) x = make_double2(1.0, 0.0);
      else if (r > col) x = Pb[(int64_t)r * ld + col];
    }
    Vb[(int64_t)r * ldv + col] = x;
    Vhb[(int64_t)col * ldh + r] = dconj(x);
  }
  double2* Tb = Tf + b * sT;
  for (int i = tid; i < kPB * kPB; i += kQRT) Tb[i] = Ts[i / kPB][i % kPB];
}

// The panel's W from Yt = V^H A22 (32 x m) and T, in three small kernels:
//   panel_y_kernel  Y = Yt^H T for 64 rows per work-group (to Cat(:, 0:32))
//                   and the group's partial of Z = V^H Y
//   panel_m_kernel  Z from the partials, M = T^H Z (one work-group a matrix)
//   panel_w_kernel  W = Y - 1/2 V M (to Cat(:, 0:32)) and W^H (rows 32 .. 63
//                   of Vh)
// Cat: [m][64] per matrix, ([W | V]); Yt, Vh: row stride ldh.
constexpr int kRowsW = 64;   // rows per work-group of the W kernels

__global__ __launch_bounds__(256) void panel_y_kernel(const double2* __restrict__ Yt, int64_t sYt,
                                                      int64_t ldh, double2* __restrict__ Cat,
                                                      int64_t sC, const double2* __restrict__ Tf,
                                                      int64_t sT, int m,
                                                      double2* __restrict__ Zpart, int64_t sZp) {
  __shared__ double2 ys[kPB][kRowsW + 1];   // Yt(:, rows)
  __shared__ double2 ts[kPB][kPB + 1];
  __shared__ double2 vs[kRowsW][kPB + 1];
  __shared__ double2 yr[kRowsW][kPB + 1];   // Y(rows, :)
  const int64_t b = blockIdx.y;
  const int r0 = blockIdx.x * kRowsW, tid = threadIdx.x;
  const double2 zero = make_double2(0.0, 0.0);
  const double2* Ytb = Yt + b * sYt;
  double2* Cb = Cat + b * sC;
  for (int e = tid; e < kPB * kRowsW; e += 256) {
    const int q = e / kRowsW, rr = e % kRowsW;
    ys[q][rr] = r0 + rr < m ? Ytb[(int64_t)q * ldh + r0 + rr] : zero;
    const int r2 = e / kPB, c2 = e % kPB;
    vs[r2][c2] = r0 + r2 < m ? Cb[(int64_t)(r0 + r2) * (2 * kPB) + kPB + c2] : zero;
  }
  for (int e = tid; e < kPB * kPB; e += 256) ts[e / kPB][e % kPB] = Tf[b * sT + e];
  __syncthreads();
  {
    const int rr = tid >> 2, j0 = (tid & 3) * 8;
    double2 acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = zero;
#pragma unroll 2
    for (int q = 0; q < kPB; ++q) {
      const double2 y = ys[q][rr];  // conj(y) t
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double2 t = ts[q][j0 + u];
        acc[u].x += y.x * t.x + y.y * t.y;
        acc[u].y += y.x * t.y - y.y * t.x;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      yr[rr][j0 + u] = acc[u];
      if (r0 + rr < m) Cb[(int64_t)(r0 + rr) * (2 * kPB) + j0 + u] = acc[u];
    }
  }
  __syncthreads();
  {
    const int i = tid >> 3, j0 = (tid & 7) * 4;
    double2 acc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = zero;
#pragma unroll 2
    for (int rr = 0; rr < kRowsW; ++rr) {
      const double2 v = vs[rr][i];  // conj(v) y
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double2 y = yr[rr][j0 + u];
        acc[u].x += v.x * y.x + v.y * y.y;
        acc[u].y += v.x * y.y - v.y * y.x;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      Zpart[b * sZp + (int64_t)blockIdx.x * kPB * kPB + i * kPB + j0 + u] = acc[u];
  }
}

__global__ __launch_bounds__(1024) void panel_m_kernel(const double2* __restrict__ Zpart,
                                                       int64_t sZp, int nchunk,
                                                       const double2* __restrict__ Tf, int64_t sT,
                                                       double2* __restrict__ Mm, int64_t sM) {
  __shared__ double2 zs[kPB][kPB + 1];
  const int64_t b = blockIdx.x;
  const int i = threadIdx.x >> 5, j = threadIdx.x & 31;
  double2 z = make_double2(0.0, 0.0);
  for (int q = 0; q < nchunk; ++q) z = cadd(z, Zpart[b * sZp + (int64_t)q * kPB * kPB + i * kPB + j]);
  zs[i][j] = z;
  __syncthreads();
  double2 acc = make_double2(0.0, 0.0);
  for (int q = 0; q <= i; ++q) {  // T upper triangular: T(q, i) = 0 for q > i
    const double2 t = Tf[b * sT + q * kPB + i], y = zs[q][j];  // conj(t) y
    acc.x += t.x * y.x + t.y * y.y;
    acc.y += t.x * y.y - t.y * y.x;
  }
  Mm[b * sM + i * kPB + j] = acc;
}

__global__ __launch_bounds__(256) void panel_w_kernel(double2* __restrict__ Cat, int64_t sC,
                                                      const double2* __restrict__ Mm, int64_t sM,
                                                      int m, double2* __restrict__ Vh,
                                                      int64_t sVh, int64_t ldh) {
  __shared__ double2 ms[kPB][kPB + 1];
  __shared__ double2 vs[kRowsW][kPB + 1];
  __shared__ double2 ws[kRowsW][kPB + 1];
  const int64_t b = blockIdx.y;
  const int r0 = blockIdx.x * kRowsW, tid = threadIdx.x;
  const double2 zero = make_double2(0.0, 0.0);
  double2* Cb = Cat + b * sC;
  for (int e = tid; e < kPB * kRowsW; e += 256) {
    const int r2 = e / kPB, c2 = e % kPB;
    vs[r2][c2] = r0 + r2 < m ? Cb[(int64_t)(r0 + r2) * (2 * kPB) + kPB + c2] : zero;
  }
  for (int e = tid; e < kPB * kPB; e += 256) ms[e / kPB][e % kPB] = Mm[b * sM + e];
  __syncthreads();
  {
    const int rr = tid >> 2, j0 = (tid & 3) * 8;
    double2 acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = zero;
#pragma unroll 2
    for (int q = 0; q < kPB; ++q) {
      const double2 v = vs[rr][q];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] = cadd(acc[u], cmul(v, ms[q][j0 + u]));
    }
    if (r0 + rr < m) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        double2* p = Cb + (int64_t)(r0 + rr) * (2 * kPB) + j0 + u;
        const double2 y = *p;
        const double2 w = make_double2(y.x - 0.5 * acc[u].x, y.y - 0.5 * acc[u].y);
        *p = w;
        ws[rr][j0 + u] = w;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < kPB * kRowsW; e += 256) {
    const int j = e / kRowsW, rr = e % kRowsW;
    if (r0 + rr < m) Vh[b * sVh + (int64_t)(kPB + j) * ldh + r0 + rr] = dconj(ws[rr][j]);
  }
}

// AB[c][d] = A(c + d, c) for d <= 32 (real diagonal), 0 for 32 < d < 64
__global__ void band_extract_kernel(const double2* __restrict__ A, int64_t sA, int n,
                                    double2* __restrict__ AB, int64_t sAB) {
  const int64_t b = blockIdx.y;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * kBandLd) return;
  const int c = (int)(idx / kBandLd), d = (int)(idx % kBandLd);
  double2 x = make_double2(0.0, 0.0);
  if (d <= kPB && c + d < n) {
    x = A[b * sA + (int64_t)(c + d) * n + c];
    if (d == 0) x.y = 0.0;
  }
  AB[b * sAB + idx] = x;
}

}  // namespace aq

using namespace aq;

extern "C" int aq_bgemm(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                        int64_t a_s0, int64_t a_s1, int64_t a_sb, int conj_a, const void* B,
                        int64_t b_s0, int64_t b_s1, int64_t b_sb, int conj_b, double beta, void* C,
                        int64_t c_s0, int64_t c_s1, int64_t c_sb, int64_t batch,
                        hipStream_t stream);

extern "C" int64_t aq_he2hb_panels(int64_t n) {
  return n - kPB - 2 >= 0 ? (n - kPB - 2) / kPB + 1 : 0;
}

extern "C" int64_t aq_he2hb_workspace_elems(int64_t n, int64_t batch) {
  // per matrix: Cat [n][64] ([W | V]) | Vh [64][n] | Yt [32][n] | Mm [32][32] |
  // Z partials [cdiv(n, 64)][32][32]
  return batch * (2 * kPB * n + 2 * kPB * n + kPB * n + kPB * kPB + cdiv(n, kRowsW) * kPB * kPB);
}

namespace aq {
// Worker streams of the band reduction (default priority), created once per device.
constexpr int kHe2hbStreams = 2;
static hipStream_t g_he2hb_streams[64][kHe2hbStreams];
static std::mutex g_he2hb_lock;

static int he2hb_streams(hipStream_t* out) {
  int dev = 0;
  AQ_HIP(hipGetDevice(&dev));
  AQ_REQUIRE(dev >= 0 && dev < 64, AQ_ERR_UNSUPPORTED, "aq_he2hb_own: device %d", dev);
  std::lock_guard<std::mutex> g(g_he2hb_lock);
  for (int i = 0; i < kHe2hbStreams; ++i) {
    if (!g_he2hb_streams[dev][i])
      AQ_HIP(hipStreamCreateWithFlags(&g_he2hb_streams[dev][i], hipStreamNonBlocking));
    out[i] = g_he2hb_streams[dev][i];
  }
  return AQ_OK;
}

struct He2hbGroup {
  double2 *A, *AB, *Tall, *Cat, *Vh, *Yt, *Mm, *Zp;
  int64_t batch;
  hipStream_t s;
};
}  // namespace aq

extern "C" int aq_he2hb_own(int64_t n, void* Av, int64_t strideA, void* ABv, void* Tallv,
                            void* ws, int64_t ws_elems, int64_t batch, hipStream_t stream) {
  AQ_REQUIRE(Av && ABv && Tallv && ws, AQ_ERR_ARG, "aq_he2hb_own: null pointer");
  AQ_REQUIRE(n >= 2 && n - kPB <= kQRMaxRows, AQ_ERR_ARG, "aq_he2hb_own: order %lld out of range",
             (long long)n);
  AQ_REQUIRE(batch >= 1 && batch <= 65535 && strideA >= n * n, AQ_ERR_ARG,
             "aq_he2hb_own: bad batch or stride");
  AQ_REQUIRE(ws_elems >= aq_he2hb_workspace_elems(n, batch), AQ_ERR_ARG,
             "aq_he2hb_own: workspace too small");
  const int64_t npan = aq_he2hb_panels(n);
  const int64_t sT = std::max<int64_t>(npan, 1) * kPB * kPB;
  const int64_t sCat = 2 * kPB * n, sVh = 2 * kPB * n, sYt = kPB * n, sM = kPB * kPB;
  const int64_t sZp = cdiv(n, kRowsW) * kPB * kPB;
  double2* Cat0 = (double2*)ws;
  double2* Vh0 = Cat0 + batch * sCat;
  double2* Yt0 = Vh0 + batch * sVh;
  double2* Mm0 = Yt0 + batch * sYt;
  double2* Zp0 = Mm0 + batch * sM;
  // two interleaved groups on worker streams: one group's panel
  // factorisation (a few work-groups) overlaps the other's GEMMs
  const int ngroups = batch >= 8 ? kHe2hbStreams : 1;
  hipStream_t wst[kHe2hbStreams];
  He2hbGroup g[kHe2hbStreams];
  if (ngroups > 1) {
    const int st = he2hb_streams(wst);
    if (st != AQ_OK) return st;
  }
  int64_t start = 0;
  for (int q = 0; q < ngroups; ++q) {
    const int64_t cnt = batch / ngroups + (q < batch % ngroups ? 1 : 0);
    g[q].A = (double2*)Av + start * strideA;
    g[q].AB = (double2*)ABv + start * n * kBandLd;
    g[q].Tall = (double2*)Tallv + start * sT;
    g[q].Cat = Cat0 + start * sCat;
    g[q].Vh = Vh0 + start * sVh;
    g[q].Yt = Yt0 + start * sYt;
    g[q].Mm = Mm0 + start * sM;
    g[q].Zp = Zp0 + start * sZp;
    g[q].batch = cnt;
    g[q].s = ngroups > 1 ? wst[q] : stream;
    start += cnt;
  }
  hipEvent_t ev = nullptr;
  if (ngroups > 1) {
    AQ_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    AQ_HIP(hipEventRecord(ev, stream));
    for (int q = 0; q < ngroups; ++q) AQ_HIP(hipStreamWaitEvent(g[q].s, ev, 0));
  }
  for (int q = 0; q < ngroups; ++q) {
    dim3 grid((unsigned)cdiv(n, 32), (unsigned)cdiv(n, 32), (unsigned)g[q].batch);
    sym_fill_kernel<<<grid, 256, 0, g[q].s>>>(g[q].A, strideA, (int)n);
    AQ_LAUNCH_CHECK();
  }
  const int ldc = 2 * kPB;
  for (int64_t p = 0; p < npan; ++p) {
    const int64_t j = p * kPB, m = n - j - kPB;
    const unsigned nch = (unsigned)cdiv(m, kRowsW);
    for (int q = 0; q < ngroups; ++q) {
      const int64_t bq = g[q].batch;
      hipStream_t s = g[q].s;
      double2* Pp = g[q].A + (j + kPB) * n + j;
      double2* A22 = g[q].A + (j + kPB) * n + (j + kPB);
      double2* Tp = g[q].Tall + p * kPB * kPB;
      panel_qr_kernel<<<(unsigned)bq, kQRT, 0, s>>>(Pp, strideA, n, (int)m, kPB, g[q].Cat + kPB,
                                                    sCat, ldc, g[q].Vh, sVh, n, Tp, sT);
      AQ_LAUNCH_CHECK();
      // Yt = V^H A22  [32 x m]
      int st = aq_bgemm(AQ_C128, kPB, m, m, 1.0, g[q].Vh, n, 1, sVh, 0, A22, n, 1, strideA, 0, 0.0,
                        g[q].Yt, n, 1, sYt, bq, s);
      if (st) return st;
      panel_y_kernel<<<dim3(nch, (unsigned)bq), 256, 0, s>>>(g[q].Yt, sYt, n, g[q].Cat, sCat, Tp,
                                                            sT, (int)m, g[q].Zp, sZp);
      AQ_LAUNCH_CHECK();
      panel_m_kernel<<<(unsigned)bq, 1024, 0, s>>>(g[q].Zp, sZp, (int)nch, Tp, sT, g[q].Mm, sM);
      AQ_LAUNCH_CHECK();
      panel_w_kernel<<<dim3(nch, (unsigned)bq), 256, 0, s>>>(g[q].Cat, sCat, g[q].Mm, sM, (int)m,
                                                            g[q].Vh, sVh, n);
      AQ_LAUNCH_CHECK();
      // A22 -= [W V] [V^H; W^H]
      st = aq_bgemm(AQ_C128, m, m, 2 * kPB, -1.0, g[q].Cat, ldc, 1, sCat, 0, g[q].Vh, n, 1, sVh, 0,
                    1.0, A22, n, 1, strideA, bq, s);
      if (st) return st;
    }
  }
  for (int q = 0; q < ngroups; ++q) {
    dim3 grid((unsigned)cdiv(n * kBandLd, 256), (unsigned)g[q].batch);
    band_extract_kernel<<<grid, 256, 0, g[q].s>>>(g[q].A, strideA, (int)n, g[q].AB, n * kBandLd);
    AQ_LAUNCH_CHECK();
  }
  if (ngroups > 1) {
    for (int q = 0; q < ngroups; ++q) {
      AQ_HIP(hipEventRecord(ev, g[q].s));
      AQ_HIP(hipStreamWaitEvent(stream, ev, 0));
    }
    AQ_HIP(hipEventDestroy(ev));
  }
  return AQ_OK;
}

// ------------------------------------------------------------ QR (zgeqrf, zungqr)
namespace aq {
// V (unit diagonal, zeros above) and V^H of panel columns j .. j+31 from the
// factored A (rows j .. m-1, row stride lda): the form the panel kernel
// leaves, rebuilt for the Q formation.
__global__ void extract_v_kernel(const double2* __restrict__ A, int64_t sA, int64_t lda,
                                 int mp, int ncol, double2* __restrict__ V, int64_t sV,
                                 double2* __restrict__ Vh, int64_t sVh, int64_t ldh) {
  const int64_t b = blockIdx.y;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)mp * kPB) return;
  const int r = (int)(idx / kPB), c = (int)(idx % kPB);
  double2 x = make_double2(0.0, 0.0);
  if (c < ncol && r >= c) x = r == c ? make_double2(1.0, 0.0) : A[b * sA + (int64_t)r * lda + c];
  V[b * sV + idx] = x;
  Vh[b * sVh + (int64_t)c * ldh + r] = dconj(x);
}

// Q = [I_k; 0] (m x k, row stride k)
__global__ void eye_kernel(double2* __restrict__ Q, int64_t sQ, int m, int k) {
  const int64_t b = blockIdx.y;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m * k) return;
  const int r = (int)(idx / k), c = (int)(idx % k);
  Q[b * sQ + idx] = make_double2(r == c ? 1.0 : 0.0, 0.0);
}
}  // namespace aq

extern "C" int64_t aq_geqrf_panels(int64_t m, int64_t n) { return cdiv(std::min(m, n), kPB); }

extern "C" int64_t aq_geqrf_workspace_elems(int64_t m, int64_t n, int64_t batch) {
  // per matrix: V [m][32] | V^H [32][m] | W, W2 [32][max(m, n)]
  return batch * (2 * kPB * m + 2 * kPB * std::max(m, n));
}

extern "C" int aq_geqrf(int64_t m, int64_t n, void* Av, int64_t strideA, void* Tallv, void* ws,
                        int64_t ws_elems, int64_t batch, hipStream_t stream) {
  AQ_REQUIRE(Av && Tallv && ws, AQ_ERR_ARG, "aq_geqrf: null pointer");
  AQ_REQUIRE(m >= 1 && n >= 1 && m <= kQRMaxRows, AQ_ERR_ARG,
             "aq_geqrf: shape %lld x %lld out of range (rows <= %d)", (long long)m, (long long)n,
             kQRMaxRows);
  AQ_REQUIRE(batch >= 1 && batch <= 65535 && strideA >= m * n, AQ_ERR_ARG,
             "aq_geqrf: bad batch or stride");
  AQ_REQUIRE(ws_elems >= aq_geqrf_workspace_elems(m, n, batch), AQ_ERR_ARG,
             "aq_geqrf: workspace too small");
  double2* A = (double2*)Av;
  double2* Tall = (double2*)Tallv;
  const int64_t npan = aq_geqrf_panels(m, n), sT = npan * kPB * kPB;
  const int64_t wmax = std::max(m, n);
  const int64_t sV = kPB * m, sW = kPB * wmax;
  double2* V = (double2*)ws;
  double2* Vh = V + batch * sV;
  double2* W = Vh + batch * sV;
  double2* W2 = W + batch * sW;
  for (int64_t p = 0; p < npan; ++p) {
    const int64_t j = p * kPB, mp = m - j, nc = std::min<int64_t>(kPB, n - j);
    const int64_t ntr = n - j - nc;
    double2* Tp = Tall + p * kPB * kPB;
    panel_qr_kernel<<<(unsigned)batch, kQRT, 0, stream>>>(A + j * n + j, strideA, n, (int)mp,
                                                          (int)nc, V, sV, kPB, Vh, sV, m, Tp, sT);
    AQ_LAUNCH_CHECK();
    if (ntr <= 0) continue;
    double2* Atr = A + j * n + j + nc;
    // W = V^H A_tr ; W2 = T^H W ; A_tr -= V W2
    int st = aq_bgemm(AQ_C128, kPB, ntr, mp, 1.0, Vh, m, 1, sV, 0, Atr, n, 1, strideA, 0, 0.0, W,
                      ntr, 1, sW, batch, stream);
    if (st) return st;
    st = aq_bgemm(AQ_C128, kPB, ntr, kPB, 1.0, Tp, 1, kPB, sT, 1, W, ntr, 1, sW, 0, 0.0, W2, ntr,
                  1, sW, batch, stream);
    if (st) return st;
    st = aq_bgemm(AQ_C128, mp, ntr, kPB, -1.0, V, kPB, 1, sV, 0, W2, ntr, 1, sW, 0, 1.0, Atr, n, 1,
                  strideA, batch, stream);
    if (st) return st;
  }
  return AQ_OK;
}

extern "C" int aq_ungqr(int64_t m, int64_t n, const void* Av, int64_t strideA, const void* Tallv,
                        void* Qv, int64_t strideQ, void* ws, int64_t ws_elems, int64_t batch,
                        hipStream_t stream) {
  AQ_REQUIRE(Av && Tallv && Qv && ws, AQ_ERR_ARG, "aq_ungqr: null pointer");
  AQ_REQUIRE(m >= 1 && n >= 1 && m <= kQRMaxRows, AQ_ERR_ARG, "aq_ungqr: shape out of range");
  const int64_t k = std::min(m, n);
  AQ_REQUIRE(batch >= 1 && batch <= 65535 && strideA >= m * n && strideQ >= m * k, AQ_ERR_ARG,
             "aq_ungqr: bad batch or stride");
  AQ_REQUIRE(ws_elems >= aq_geqrf_workspace_elems(m, n, batch), AQ_ERR_ARG,
             "aq_ungqr: workspace too small");
  const double2* A = (const double2*)Av;
  const double2* Tall = (const double2*)Tallv;
  double2* Q = (double2*)Qv;
  const int64_t npan = aq_geqrf_panels(m, n), sT = npan * kPB * kPB;
  const int64_t wmax = std::max(m, n);
  const int64_t sV = kPB * m, sW = kPB * wmax;
  double2* V = (double2*)ws;
  double2* Vh = V + batch * sV;
  double2* W = Vh + batch * sV;
  double2* W2 = W + batch * sW;
  {
    dim3 grid((unsigned)cdiv(m * k, 256), (unsigned)batch);
    eye_kernel<<<grid, 256, 0, stream>>>(Q, strideQ, (int)m, (int)k);
    AQ_LAUNCH_CHECK();
  }
  // Q = H_0 H_1 ... [I; 0]: panels applied last to first, each to columns j .. k-1
  for (int64_t p = npan - 1; p >= 0; --p) {
    const int64_t j = p * kPB, mp = m - j, nc = std::min<int64_t>(kPB, n - j), kc = k - j;
    const double2* Tp = Tall + p * kPB * kPB;
    {
      dim3 grid((unsigned)cdiv(mp * kPB, 256), (unsigned)batch);
      extract_v_kernel<<<grid, 256, 0, stream>>>(A + j * n + j, strideA, n, (int)mp, (int)nc, V,
                                                 sV, Vh, sV, m);
      AQ_LAUNCH_CHECK();
    }
    double2* Qs = Q + j * k + j;
    // W = V^H Q_s ; W2 = T W ; Q_s -= V W2
    int st = aq_bgemm(AQ_C128, kPB, kc, mp, 1.0, Vh, m, 1, sV, 0, Qs, k, 1, strideQ, 0, 0.0, W, kc,
                      1, sW, batch, stream);
    if (st) return st;
    st = aq_bgemm(AQ_C128, kPB, kc, kPB, 1.0, Tp, kPB, 1, sT, 0, W, kc, 1, sW, 0, 0.0, W2, kc, 1,
                  sW, batch, stream);
    if (st) return st;
    st = aq_bgemm(AQ_C128, mp, kc, kPB, -1.0, V, kPB, 1, sV, 0, W2, kc, 1, sW, 0, 1.0, Qs, k, 1,
                  strideQ, batch, stream);
    if (st) return st;
  }
  return AQ_OK;
}
