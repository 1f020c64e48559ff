// Theta contraction with the two-qubit gate fused into the epilogue, and the
// general complex GEMM entry point.
//
//   Psi[2l+p1, p2*chiR + j] = sum_m A[l,p1,m] B[m,p2,j]     (MFMA mainloop)
//   X[2l+p, q*chiR + j]     = sum_{p1,p2} U[2p+q, 2p1+p2] Psi[2l+p1, p2*chiR+j]
//
// The work-group tile pairs rows (2l, 2l+1) in accumulator registers (r, r+1)
// and the two column halves (p2 = 0, 1) in the two half-blocks, so the 4x4
// gate is applied per lane in registers and X is written once, coalesced
// along j.  Replaces fused_two_qubit_gate_pytorch
// (src/triton_kernels/mps_complex.py:343-368).
#include <algorithm>

#include "gemm_core.h"
#include "bf16x3.h"

namespace aq {

template <typename T> struct Gate4 {
  T re[16];
  T im[16];
};

// WR: waves per column of the work-group (wave w owns rows 32 (w % WR) and
// columns 32 (w / WR) of both halves).
template <typename T, class FR, int WR = 2>
__device__ __forceinline__ void theta_epilogue(const FR& fr, const Gate4<T>& g, c2_t<T>* X,
                                               int64_t chiL, int64_t chiR, int64_t m0,
                                               int64_t j0) {
  using M = Mfma<T>;
  constexpr int E = M::E;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave % WR, wn = wave / WR;
  const int64_t ld = 2 * chiR;
#pragma unroll
  for (int rb = 0; rb < FR::RB; ++rb)
#pragma unroll
    for (int jb = 0; jb < FR::JB; ++jb)
#pragma unroll
      for (int r = 0; r < M::NREG; r += 2) {
        const int64_t row = m0 + 32 * wm + rb * E + M::out_row(lane, r);  // even: 2l
        const int64_t j = j0 + 32 * wn + jb * E + M::out_col(lane);
        if (row >= 2 * chiL || j >= chiR) continue;
        T pr[4], pi[4];  // index 2*p1 + p2
#pragma unroll
        for (int p1 = 0; p1 < 2; ++p1)
#pragma unroll
          for (int p2 = 0; p2 < 2; ++p2) {
            pr[2 * p1 + p2] = fr.re[rb][jb][p2][r + p1];
            pi[2 * p1 + p2] = fr.im[rb][jb][p2][r + p1];
          }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int o = 2 * p + q;
            T xr = 0, xi = 0;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              xr += g.re[4 * o + c] * pr[c] - g.im[4 * o + c] * pi[c];
              xi += g.re[4 * o + c] * pi[c] + g.im[4 * o + c] * pr[c];
            }
            c2_t<T> v;
            v.x = xr;
            v.y = xi;
            X[(row + p) * ld + q * chiR + j] = v;
          }
      }
}

template <typename T>
__global__ __launch_bounds__(kThreads, 2) void theta_kernel(const c2_t<T>* __restrict__ A,
                                                         const c2_t<T>* __restrict__ B,
                                                         Gate4<T> g, c2_t<T>* __restrict__ X,
                                                         int64_t chiL, int64_t chiM,
                                                         int64_t chiR) {
  __shared__ Smem<T> sm;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t j0 = (int64_t)blockIdx.y * kBJ;
  AStrided<T> aop{A, chiM, 1, 2 * chiL, chiM, 0};
  BStrided<T> bop{B, 2 * chiR, 1, chiM, 2 * chiR, 0};
  ColMap cm{1, chiR};
  Frag<T> fr;
  cgemm_mainloop<T, true, true>(aop, bop, cm, chiM, m0, j0, sm, fr);
  theta_epilogue<T>(fr, g, X, chiL, chiR, m0, j0);
}

__global__ __launch_bounds__(kThreads) void theta_bf16x3_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, Gate4<float> g,
    float2* __restrict__ X, int64_t chiL, int64_t chiM, int64_t chiR) {
  __shared__ SmemBf16 sm;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t j0 = (int64_t)blockIdx.y * kBJ;
  AStrided<float> aop{A, chiM, 1, 2 * chiL, chiM, 0};
  BStrided<float> bop{B, 2 * chiR, 1, chiM, 2 * chiR, 0};
  ColMap cm{1, chiR};
  Frag<float> fr;
  cgemm_mainloop_bf16x3<true, true>(aop, bop, cm, chiM, m0, j0, sm, fr);
  theta_epilogue<float>(fr, g, X, chiL, chiR, m0, j0);
}

// A(m, k) = A[m lda + k] (k contiguous) -> planes; one thread per 8 k values
// (16-byte loads of two complex values where the row's 8 values are in range
// and 16-byte aligned, i.e. always for an even lda and aligned A).
__global__ __launch_bounds__(256) void split_rows_kernel(const float2* __restrict__ A, int64_t M,
                                                         int64_t K, int64_t lda, int conj,
                                                         uint16_t* __restrict__ P, int64_t Mp,
                                                         int64_t Kp) {
  const int64_t kq = Kp / 8;
  const int64_t total = Mp * kq;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = idx / kq, k0 = (idx % kq) * 8;
    float2 x[8];
    const float2* src = A + m * lda + k0;
    if (m < M && k0 + 8 <= K && ((uintptr_t)src & 15) == 0) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float4 v = reinterpret_cast<const float4*>(src)[e];
        x[2 * e] = make_float2(v.x, v.y);
        x[2 * e + 1] = make_float2(v.z, v.w);
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e)
        x[e] = (m < M && k0 + e < K) ? src[e] : make_float2(0.f, 0.f);
    }
    if (conj) {
#pragma unroll
      for (int e = 0; e < 8; ++e) x[e].y = -x[e].y;
    }
    uint4 o[4];
    split8(x, o);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      *reinterpret_cast<uint4*>(P + q * Mp * Kp + m * Kp + k0) = o[q];
  }
}

// B(k, n) = B[k ldb + n] (n contiguous) -> transposed planes, through a
// 32 (k) x 64 (j) LDS tile per work-group: reads coalesced along n, writes
// along k.  grid = (Kp / 32, Jp / 64, 2 halves); paired: column h half_n + j,
// else (j / kBJ) kBN + h kBJ + j % kBJ.
__global__ __launch_bounds__(256) void split_cols_kernel(const float2* __restrict__ B, int64_t K,
                                                         int64_t N, int64_t ldb, int paired,
                                                         int64_t half_n, uint16_t* __restrict__ P,
                                                         int64_t Jp, int64_t Kp) {
  __shared__ float2 t[32][64 + 1];
  const int64_t k0 = (int64_t)blockIdx.x * 32, jb = (int64_t)blockIdx.y * 64;
  const int h = blockIdx.z;
  const int tid = threadIdx.x;
  for (int e = tid; e < 32 * 64; e += 256) {
    const int kk = e / 64, jj = e % 64;
    const int64_t j = jb + jj, k = k0 + kk;
    int64_t n;
    bool ok;
    if (paired) {
      n = h * half_n + j;
      ok = j < half_n;
    } else {
      n = (j / kBJ) * kBN + h * kBJ + j % kBJ;
      ok = n < N;
    }
    t[kk][jj] = (ok && k < K) ? B[k * ldb + n] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  // 64 columns x 4 groups of 8 k per thread
  const int jj = tid / 4, kq = tid % 4;
  float2 x[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) x[e] = t[8 * kq + e][jj];
  uint4 o[4];
  split8(x, o);
  const int64_t off = (int64_t)h * Jp * Kp + (jb + jj) * Kp + k0 + 8 * kq;
#pragma unroll
  for (int q = 0; q < 4; ++q) *reinterpret_cast<uint4*>(P + q * 2 * Jp * Kp + off) = o[q];
}

// BF16x3 theta on operands split beforehand (split_rows_kernel /
// split_cols_kernel): the mainloop only copies 16-byte chunks into LDS.
__global__ __launch_bounds__(kThreads, 2) void theta_planes_kernel(PlanesA pa, PlanesB pb,
                                                                 int64_t Kp, Gate4<float> g,
                                                                 float2* __restrict__ X,
                                                                 int64_t chiL, int64_t chiR) {
  __shared__ SmemBf16 sm;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t j0 = (int64_t)blockIdx.y * kBJ;
  Frag<float> fr;
  cgemm_mainloop_planes(pa, pb, Kp, m0, j0, sm, fr);
  theta_epilogue<float>(fr, g, X, chiL, chiR, m0, j0);
}

// The same on a tile of 32 WR rows with 128 WR threads, on a 1-D grid of
// gx x gy tiles.  Work-groups go to the 8 XCDs round-robin (XCD = id mod 8)
// and each XCD has its own L2; with gx % 4 == 0 and gy % 2 == 0 the XCD
// blocks tile the grid as 4 x 2 blocks of (gx/4) x (gy/2) tiles, so an XCD
// reads the A strips of gx/4 row tiles and the B strips of gy/2 column tiles
// (instead of every B strip: a third less operand traffic at 16 x 16 tiles).
template <int WR>
__global__ __launch_bounds__(128 * WR, 2) void theta_planes_w_kernel(PlanesA pa, PlanesB pb,
                                                                   int64_t Kp, Gate4<float> g,
                                                                   float2* __restrict__ X,
                                                                   int64_t chiL, int64_t chiR,
                                                                   int gx, int gy) {
  __shared__ SmemPlanes<WR> sm;
  const int id = blockIdx.x;
  int bx, by;
  if (gx % 4 == 0 && gy % 2 == 0) {
    const int xcd = id % 8, local = id / 8, bw = gx / 4;
    bx = (xcd % 4) * bw + local % bw;
    by = (xcd / 4) * (gy / 2) + local / bw;
  } else {
    bx = id % gx;
    by = id / gx;
  }
  const int64_t m0 = (int64_t)bx * (32 * WR);
  const int64_t j0 = (int64_t)by * kBJ;
  Frag<float> fr;
  cgemm_mainloop_planes_w<WR>(pa, pb, Kp, m0, j0, sm, fr);
  theta_epilogue<float, Frag<float>, WR>(fr, g, X, chiL, chiR, m0, j0);
}

template <typename T, bool KFastA, bool NFastB>
__global__ __launch_bounds__(kThreads, 2) void cgemm_kernel(AStrided<T> aop, BStrided<T> bop,
                                                         c2_t<T>* __restrict__ C, int64_t M,
                                                         int64_t N, int64_t K, int64_t c_s0,
                                                         int64_t c_s1) {
  __shared__ Smem<T> sm;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t j0 = (int64_t)blockIdx.y * kBJ;
  ColMap cm{0, N};
  Frag<T> fr;
  cgemm_mainloop<T, KFastA, NFastB>(aop, bop, cm, K, m0, j0, sm, fr);
  frag_foreach<T>(fr, [&](int row, int h, int jj, T re, T im) {
    const int64_t m = m0 + row;
    const int64_t n = 2 * j0 + h * kBJ + jj;
    if (m < M && n < N) {
      c2_t<T> v;
      v.x = re;
      v.y = im;
      C[m * c_s0 + n * c_s1] = v;
    }
  });
}

template <bool KFastA, bool NFastB>
__global__ __launch_bounds__(kThreads) void cgemm_bf16x3_kernel(AStrided<float> aop,
                                                                BStrided<float> bop,
                                                                float2* __restrict__ C,
                                                                int64_t M, int64_t N, int64_t K,
                                                                int64_t c_s0, int64_t c_s1) {
  __shared__ SmemBf16 sm;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t j0 = (int64_t)blockIdx.y * kBJ;
  ColMap cm{0, N};
  Frag<float> fr;
  cgemm_mainloop_bf16x3<KFastA, NFastB>(aop, bop, cm, K, m0, j0, sm, fr);
  frag_foreach<float>(fr, [&](int row, int h, int jj, float re, float im) {
    const int64_t m = m0 + row;
    const int64_t n = 2 * j0 + h * kBJ + jj;
    if (m < M && n < N) C[m * c_s0 + n * c_s1] = make_float2(re, im);
  });
}

// Split-K: work-group (x, y, z) computes the tile's partial product over the
// reduction slice [z*chunk, (z+1)*chunk) into part[z][M][N]; the reduce kernel
// adds the slices in order (deterministic) into C.
template <typename T, bool KFastA, bool NFastB, bool kBf16x3>
__global__ __launch_bounds__(kThreads, kBf16x3 ? 1 : 2) void cgemm_splitk_kernel(AStrided<T> aop, BStrided<T> bop,
                                                                c2_t<T>* __restrict__ part,
                                                                int64_t M, int64_t N, int64_t K,
                                                                int64_t chunk) {
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int64_t j0 = (int64_t)blockIdx.y * kBJ;
  const int64_t k0 = (int64_t)blockIdx.z * chunk;
  const int64_t klen = (K - k0) < chunk ? (K - k0) : chunk;
  AStrided<T> a = aop;
  a.p += k0 * aop.s1;
  a.K = klen;
  BStrided<T> b = bop;
  b.p += k0 * bop.s0;
  b.K = klen;
  ColMap cm{0, N};
  Frag<T> fr;
  if constexpr (kBf16x3) {
    __shared__ SmemBf16 sm;
    cgemm_mainloop_bf16x3<KFastA, NFastB>(a, b, cm, klen, m0, j0, sm, fr);
  } else {
    __shared__ Smem<T> sm;
    cgemm_mainloop<T, KFastA, NFastB>(a, b, cm, klen, m0, j0, sm, fr);
  }
  c2_t<T>* out = part + (int64_t)blockIdx.z * M * N;
  frag_foreach<T>(fr, [&](int row, int h, int jj, T re, T im) {
    const int64_t m = m0 + row;
    const int64_t n = 2 * j0 + h * kBJ + jj;
    if (m < M && n < N) {
      c2_t<T> v;
      v.x = re;
      v.y = im;
      out[m * N + n] = v;
    }
  });
}

template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const c2_t<T>* __restrict__ part,
                                                            c2_t<T>* __restrict__ C, int64_t M,
                                                            int64_t N, int64_t splits,
                                                            int64_t c_s0, int64_t c_s1) {
  const int64_t total = M * N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    T re = 0, im = 0;
    for (int64_t z = 0; z < splits; ++z) {
      const c2_t<T> v = part[z * total + e];
      re += v.x;
      im += v.y;
    }
    c2_t<T> o;
    o.x = re;
    o.y = im;
    C[(e / N) * c_s0 + (e % N) * c_s1] = o;
  }
}

template <typename T, bool KA, bool NB>
static void launch_splitk(dim3 grid, hipStream_t s, AStrided<T> a, BStrided<T> b, void* part,
                          int64_t M, int64_t N, int64_t K, int64_t chunk, int prec) {
  if constexpr (std::is_same<T, float>::value) {
    if (prec == AQ_PREC_BF16X3) {
      cgemm_splitk_kernel<float, KA, NB, true><<<grid, kThreads, 0, s>>>(
          a, b, (float2*)part, M, N, K, chunk);
      return;
    }
  }
  cgemm_splitk_kernel<T, KA, NB, false><<<grid, kThreads, 0, s>>>(a, b, (c2_t<T>*)part, M, N,
                                                                   K, chunk);
}

template <typename T>
static int cgemm_splitk_impl(int prec, int64_t M, int64_t N, int64_t K, const void* A,
                             int64_t as0, int64_t as1, int ca, const void* B, int64_t bs0,
                             int64_t bs1, int cb, void* C, int64_t cs0, int64_t cs1,
                             int64_t splits, void* ws, int64_t ws_bytes, hipStream_t s) {
  const int64_t chunk = cdiv(cdiv(K, splits), 16) * 16;
  splits = cdiv(K, chunk);
  AQ_REQUIRE(ws_bytes >= splits * M * N * (int64_t)sizeof(c2_t<T>), AQ_ERR_ARG,
             "aq_cgemm_splitk: workspace of %lld bytes < %lld needed", (long long)ws_bytes,
             (long long)(splits * M * N * (int64_t)sizeof(c2_t<T>)));
  AStrided<T> a{(const c2_t<T>*)A, as0, as1, M, K, ca};
  BStrided<T> b{(const c2_t<T>*)B, bs0, bs1, K, N, cb};
  dim3 grid((unsigned)cdiv(M, kBM), (unsigned)cdiv(N, kBN), (unsigned)splits);
  const bool ka = (as1 == 1), nb = (bs1 == 1);
  if (ka && nb) launch_splitk<T, true, true>(grid, s, a, b, ws, M, N, K, chunk, prec);
  else if (ka) launch_splitk<T, true, false>(grid, s, a, b, ws, M, N, K, chunk, prec);
  else if (nb) launch_splitk<T, false, true>(grid, s, a, b, ws, M, N, K, chunk, prec);
  else launch_splitk<T, false, false>(grid, s, a, b, ws, M, N, K, chunk, prec);
  AQ_LAUNCH_CHECK();
  const int64_t blocks = cdiv(M * N, 256) < 4096 ? cdiv(M * N, 256) : 4096;
  splitk_reduce_kernel<T><<<(unsigned)blocks, 256, 0, s>>>((const c2_t<T>*)ws, (c2_t<T>*)C, M,
                                                            N, splits, cs0, cs1);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

template <typename T>
static Gate4<T> make_gate(const double* g) {
  Gate4<T> out;
  for (int i = 0; i < 16; ++i) {
    out.re[i] = (T)g[2 * i];
    out.im[i] = (T)g[2 * i + 1];
  }
  return out;
}

template <typename T, bool KA, bool NB>
static void launch_cgemm(dim3 grid, hipStream_t s, AStrided<T> a, BStrided<T> b, void* C,
                         int64_t M, int64_t N, int64_t K, int64_t cs0, int64_t cs1, int prec) {
  if constexpr (std::is_same<T, float>::value) {
    if (prec == AQ_PREC_BF16X3) {
      cgemm_bf16x3_kernel<KA, NB><<<grid, kThreads, 0, s>>>(a, b, (float2*)C, M, N, K, cs0, cs1);
      return;
    }
  }
  cgemm_kernel<T, KA, NB><<<grid, kThreads, 0, s>>>(a, b, (c2_t<T>*)C, M, N, K, cs0, cs1);
}

template <typename T>
static int cgemm_impl(int prec, int64_t M, int64_t N, int64_t K, const void* A, int64_t as0,
                      int64_t as1, int ca, const void* B, int64_t bs0, int64_t bs1, int cb,
                      void* C, int64_t cs0, int64_t cs1, hipStream_t s) {
  AStrided<T> a{(const c2_t<T>*)A, as0, as1, M, K, ca};
  BStrided<T> b{(const c2_t<T>*)B, bs0, bs1, K, N, cb};
  dim3 grid((unsigned)cdiv(M, kBM), (unsigned)cdiv(N, kBN));
  const bool ka = (as1 == 1), nb = (bs1 == 1);
  if (ka && nb) launch_cgemm<T, true, true>(grid, s, a, b, C, M, N, K, cs0, cs1, prec);
  else if (ka) launch_cgemm<T, true, false>(grid, s, a, b, C, M, N, K, cs0, cs1, prec);
  else if (nb) launch_cgemm<T, false, true>(grid, s, a, b, C, M, N, K, cs0, cs1, prec);
  else launch_cgemm<T, false, false>(grid, s, a, b, C, M, N, K, cs0, cs1, prec);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

}  // namespace aq

using namespace aq;

extern "C" int aq_theta(int dtype, int prec, const void* A, const void* B,
                        const double* gate_host, void* X, int64_t chiL, int64_t chiM,
                        int64_t chiR, hipStream_t stream) {
  AQ_REQUIRE(A && B && X && gate_host, AQ_ERR_ARG, "aq_theta: null pointer");
  AQ_REQUIRE(chiL > 0 && chiM > 0 && chiR > 0, AQ_ERR_ARG,
             "aq_theta: bond sizes must be positive (got %lld, %lld, %lld)", (long long)chiL,
             (long long)chiM, (long long)chiR);
  AQ_REQUIRE(chiL <= (1 << 20) && chiR <= (1 << 20) && chiM <= (1 << 20), AQ_ERR_ARG,
             "aq_theta: bond size too large");
  dim3 grid((unsigned)cdiv(2 * chiL, kBM), (unsigned)cdiv(chiR, kBJ));
  if (dtype == AQ_C64) {
    if (prec == AQ_PREC_BF16X3) {
      theta_bf16x3_kernel<<<grid, kThreads, 0, stream>>>((const float2*)A, (const float2*)B,
                                                         make_gate<float>(gate_host),
                                                         (float2*)X, chiL, chiM, chiR);
    } else {
      AQ_REQUIRE(prec == AQ_PREC_NATIVE, AQ_ERR_ARG, "aq_theta: unknown precision %d", prec);
      theta_kernel<float><<<grid, kThreads, 0, stream>>>((const float2*)A, (const float2*)B,
                                                         make_gate<float>(gate_host),
                                                         (float2*)X, chiL, chiM, chiR);
    }
  } else if (dtype == AQ_C128) {
    AQ_REQUIRE(prec == AQ_PREC_NATIVE, AQ_ERR_ARG,
               "aq_theta: complex128 supports only native FP64 arithmetic");
    theta_kernel<double><<<grid, kThreads, 0, stream>>>((const double2*)A, (const double2*)B,
                                                        make_gate<double>(gate_host),
                                                        (double2*)X, chiL, chiM, chiR);
  } else {
    AQ_REQUIRE(false, AQ_ERR_ARG, "aq_theta: unknown dtype %d", dtype);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

namespace aq {
struct ThetaPlanes {
  int64_t Mp, Jp, Kp, a_elems, b_elems;
};
static ThetaPlanes theta_planes(int64_t chiL, int64_t chiM, int64_t chiR) {
  ThetaPlanes t;
  t.Mp = cdiv(2 * chiL, 2 * kBM) * 2 * kBM;   // whole tiles of the 128-row variants
  t.Jp = cdiv(chiR, kBJ) * kBJ;
  t.Kp = cdiv(chiM, kBKb) * kBKb;
  t.a_elems = 4 * t.Mp * t.Kp;
  t.b_elems = 4 * 2 * t.Jp * t.Kp;
  return t;
}
}  // namespace aq

extern "C" int64_t aq_theta_planes_bytes(int64_t chiL, int64_t chiM, int64_t chiR) {
  if (chiL <= 0 || chiM <= 0 || chiR <= 0) return 0;
  const ThetaPlanes t = theta_planes(chiL, chiM, chiR);
  return 2 * (t.a_elems + t.b_elems);
}

extern "C" int aq_theta_planes(const void* A, const void* B, const double* gate_host, void* X,
                               int64_t chiL, int64_t chiM, int64_t chiR, void* ws,
                               int64_t ws_bytes, hipStream_t stream) {
  AQ_REQUIRE(A && B && X && gate_host && ws, AQ_ERR_ARG, "aq_theta_planes: null pointer");
  AQ_REQUIRE(chiL > 0 && chiM > 0 && chiR > 0, AQ_ERR_ARG,
             "aq_theta_planes: bond sizes must be positive (got %lld, %lld, %lld)",
             (long long)chiL, (long long)chiM, (long long)chiR);
  AQ_REQUIRE(chiL <= (1 << 20) && chiR <= (1 << 20) && chiM <= (1 << 20), AQ_ERR_ARG,
             "aq_theta_planes: bond size too large");
  AQ_REQUIRE(((uintptr_t)ws & 15) == 0, AQ_ERR_ARG, "aq_theta_planes: workspace not 16-byte aligned");
  const ThetaPlanes t = theta_planes(chiL, chiM, chiR);
  AQ_REQUIRE(ws_bytes >= 2 * (t.a_elems + t.b_elems), AQ_ERR_ARG,
             "aq_theta_planes: workspace of %lld bytes, %lld needed", (long long)ws_bytes,
             (long long)(2 * (t.a_elems + t.b_elems)));
  uint16_t* Pa = (uint16_t*)ws;
  uint16_t* Pb = Pa + t.a_elems;
  {
    const int64_t total = t.Mp * t.Kp / 8;
    const unsigned grid = (unsigned)std::min<int64_t>(cdiv(total, 256), 65536);
    split_rows_kernel<<<grid, 256, 0, stream>>>((const float2*)A, 2 * chiL, chiM, chiM, 0, Pa,
                                                t.Mp, t.Kp);
    AQ_LAUNCH_CHECK();
  }
  {
    dim3 grid((unsigned)(t.Kp / 32), (unsigned)(t.Jp / 64), 2);
    split_cols_kernel<<<grid, 256, 0, stream>>>((const float2*)B, chiM, 2 * chiR, 2 * chiR, 1,
                                                chiR, Pb, t.Jp, t.Kp);
    AQ_LAUNCH_CHECK();
  }
  PlanesA pa{Pa, t.Kp, t.Mp * t.Kp};
  PlanesB pb{Pb, t.Kp, 2 * t.Jp * t.Kp, t.Jp * t.Kp};
  const Gate4<float> gg = make_gate<float>(gate_host);
  // 128-row tiles (eight waves, half the operand traffic per MFMA of the
  // 64-row tile) where they still give every CU a work-group; 64-row tiles
  // below that (theta of 2048 x 2048: 81 vs 91 us alone on the GPU)
  const int64_t wide_tiles = (t.Mp / 128) * (t.Jp / kBJ);
  if (wide_tiles >= 256) {
    const int gx = (int)(t.Mp / 128), gy = (int)(t.Jp / kBJ);
    theta_planes_w_kernel<4><<<(unsigned)(gx * gy), 512, 0, stream>>>(pa, pb, t.Kp, gg,
                                                                      (float2*)X, chiL, chiR,
                                                                      gx, gy);
  } else {
    dim3 grid((unsigned)cdiv(2 * chiL, kBM), (unsigned)(t.Jp / kBJ));
    theta_planes_kernel<<<grid, kThreads, 0, stream>>>(pa, pb, t.Kp, gg, (float2*)X, chiL, chiR);
  }
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_cgemm(int dtype, int prec, int64_t M, int64_t N, int64_t K, const void* A,
                        int64_t a_s0, int64_t a_s1, int conj_a, const void* B, int64_t b_s0,
                        int64_t b_s1, int conj_b, void* C, int64_t c_s0, int64_t c_s1,
                        hipStream_t stream) {
  AQ_REQUIRE(A && B && C, AQ_ERR_ARG, "aq_cgemm: null pointer");
  AQ_REQUIRE(M > 0 && N > 0 && K > 0, AQ_ERR_ARG, "aq_cgemm: sizes must be positive");
  AQ_REQUIRE(cdiv(M, kBM) < (1LL << 31) && cdiv(N, kBN) < 65536, AQ_ERR_ARG,
             "aq_cgemm: grid too large");
  if (dtype == AQ_C64)
    return cgemm_impl<float>(prec, M, N, K, A, a_s0, a_s1, conj_a, B, b_s0, b_s1, conj_b, C,
                             c_s0, c_s1, stream);
  AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_cgemm: unknown dtype %d", dtype);
  AQ_REQUIRE(prec == AQ_PREC_NATIVE, AQ_ERR_ARG, "aq_cgemm: complex128 is FP64 only");
  return cgemm_impl<double>(prec, M, N, K, A, a_s0, a_s1, conj_a, B, b_s0, b_s1, conj_b, C,
                            c_s0, c_s1, stream);
}

extern "C" int aq_cgemm_splitk(int dtype, int prec, int64_t M, int64_t N, int64_t K,
                               const void* A, int64_t a_s0, int64_t a_s1, int conj_a,
                               const void* B, int64_t b_s0, int64_t b_s1, int conj_b, void* C,
                               int64_t c_s0, int64_t c_s1, int64_t splits, void* workspace,
                               int64_t workspace_bytes, hipStream_t stream) {
  AQ_REQUIRE(A && B && C && workspace, AQ_ERR_ARG, "aq_cgemm_splitk: null pointer");
  AQ_REQUIRE(M > 0 && N > 0 && K > 0, AQ_ERR_ARG, "aq_cgemm_splitk: sizes must be positive");
  AQ_REQUIRE(splits >= 1 && splits <= 1024, AQ_ERR_ARG, "aq_cgemm_splitk: splits in [1, 1024]");
  AQ_REQUIRE(cdiv(M, kBM) < (1LL << 31) && cdiv(N, kBN) < 65536, AQ_ERR_ARG,
             "aq_cgemm_splitk: grid too large");
  if (dtype == AQ_C64)
    return cgemm_splitk_impl<float>(prec, M, N, K, A, a_s0, a_s1, conj_a, B, b_s0, b_s1, conj_b,
                                    C, c_s0, c_s1, splits, workspace, workspace_bytes, stream);
  AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_cgemm_splitk: unknown dtype %d", dtype);
  AQ_REQUIRE(prec == AQ_PREC_NATIVE, AQ_ERR_ARG, "aq_cgemm_splitk: complex128 is FP64 only");
  return cgemm_splitk_impl<double>(prec, M, N, K, A, a_s0, a_s1, conj_a, B, b_s0, b_s1, conj_b,
                                   C, c_s0, c_s1, splits, workspace, workspace_bytes, stream);
}
