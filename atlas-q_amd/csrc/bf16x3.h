// BF16x3 split-precision mainloop for complex64 GEMMs on gfx950 MFMA.
//
// Each FP32 operand x is split into x_hi = bf16(x) and x_lo = bf16(x - x_hi)
// while it is staged into LDS; a real product is formed as
//   x*y ~ x_hi*y_hi + x_hi*y_lo + x_lo*y_hi        (FP32 accumulation)
// with v_mfma_f32_32x32x16_bf16 (k = 16 per instruction).  The dropped
// x_lo*y_lo term and the rounding of x_lo leave a relative error of about
// 2^-17 per product.  The accumulator layout is that of the FP32 32x32
// instruction, so the FP32 epilogues (theta gate, sampler reductions, plain
// stores) are shared unchanged.
#pragma once

#include "gemm_core.h"

namespace aq {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x8 __attribute__((ext_vector_type(8)));

constexpr int kBKb = 32;      // k per LDS tile
constexpr int kPadB = 8;      // bf16 elements of row padding

struct alignas(16) SmemBf16 {
  // planes: 0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo; rows k-contiguous
  uint16_t a[4][kBM][kBKb + kPadB];
  uint16_t b[4][kBN][kBKb + kPadB];
};

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
  hi = f32_to_bf16_rne(x);
  lo = f32_to_bf16_rne(x - bf16_to_f32(hi));
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  i16x8 v = *reinterpret_cast<const i16x8*>(p);
  return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 neg8(bf16x8 v) {
  i16x8 b = __builtin_bit_cast(i16x8, v);
  b ^= (short)0x8000;
  return __builtin_bit_cast(bf16x8, b);
}

__device__ __forceinline__ f32x16 mma_bf16(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

template <bool KFastA, bool NFastB, class AOp, class BOp>
__device__ __forceinline__ void cgemm_mainloop_bf16x3(const AOp& aop, const BOp& bop,
                                                      const ColMap& cm, int64_t K, int64_t m0,
                                                      int64_t j0, SmemBf16& sm,
                                                      Frag<float>& fr) {
  constexpr int BK = kBKb;
  constexpr int A_PER = kBM * BK / kThreads;   // 8
  constexpr int B_PER = kBN * BK / kThreads;   // 16
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      fr.re[0][0][h][r] = 0;
      fr.im[0][0][h][r] = 0;
    }

  int64_t bcol[B_PER];
  int bkk[B_PER], btc[B_PER];
#pragma unroll
  for (int i = 0; i < B_PER; ++i) {
    int e = tid + i * kThreads;
    int tc, kk;
    if (NFastB) { tc = e % kBN; kk = e / kBN; }
    else        { kk = e % BK;  tc = e / BK; }
    bkk[i] = kk;
    btc[i] = tc;
    bcol[i] = cm.col(j0, tc / kBJ, tc % kBJ);
  }
  int amm[A_PER], akk[A_PER];
#pragma unroll
  for (int i = 0; i < A_PER; ++i) {
    int e = tid + i * kThreads;
    if (KFastA) { akk[i] = e % BK; amm[i] = e / BK; }
    else        { amm[i] = e % kBM; akk[i] = e / kBM; }
  }

  float2 ra[A_PER], rbv[B_PER];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) ra[i] = aop(m0 + amm[i], k0 + akk[i]);
#pragma unroll
    for (int i = 0; i < B_PER; ++i)
      rbv[i] = (bcol[i] != INT64_MAX) ? bop(k0 + bkk[i], bcol[i]) : make_float2(0.f, 0.f);
  };

  fetch(0);
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      uint16_t h, l;
      split_bf16(ra[i].x, h, l);
      sm.a[0][amm[i]][akk[i]] = h;
      sm.a[1][amm[i]][akk[i]] = l;
      split_bf16(ra[i].y, h, l);
      sm.a[2][amm[i]][akk[i]] = h;
      sm.a[3][amm[i]][akk[i]] = l;
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      uint16_t h, l;
      split_bf16(rbv[i].x, h, l);
      sm.b[0][btc[i]][bkk[i]] = h;
      sm.b[1][btc[i]][bkk[i]] = l;
      split_bf16(rbv[i].y, h, l);
      sm.b[2][btc[i]][bkk[i]] = h;
      sm.b[3][btc[i]][bkk[i]] = l;
    }
    __syncthreads();
    if (k0 + BK < K) fetch(k0 + BK);

#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int kof = kk + 8 * (lane >> 5);
      const int row = 32 * wm + (lane & 31);
      const bf16x8 arh = load8(&sm.a[0][row][kof]);
      const bf16x8 arl = load8(&sm.a[1][row][kof]);
      const bf16x8 aih = load8(&sm.a[2][row][kof]);
      const bf16x8 ail = load8(&sm.a[3][row][kof]);
      const bf16x8 naih = neg8(aih), nail = neg8(ail);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int col = h * kBJ + 32 * wn + (lane & 31);
        const bf16x8 brh = load8(&sm.b[0][col][kof]);
        const bf16x8 brl = load8(&sm.b[1][col][kof]);
        const bf16x8 bih = load8(&sm.b[2][col][kof]);
        const bf16x8 bil = load8(&sm.b[3][col][kof]);
        f32x16 cr = fr.re[0][0][h], ci = fr.im[0][0][h];
        // small terms first, then the hi*hi terms
        cr = mma_bf16(arh, brl, cr);
        cr = mma_bf16(arl, brh, cr);
        cr = mma_bf16(naih, bil, cr);
        cr = mma_bf16(nail, bih, cr);
        cr = mma_bf16(arh, brh, cr);
        cr = mma_bf16(naih, bih, cr);
        ci = mma_bf16(arh, bil, ci);
        ci = mma_bf16(arl, bih, ci);
        ci = mma_bf16(aih, brl, ci);
        ci = mma_bf16(ail, brh, ci);
        ci = mma_bf16(arh, bih, ci);
        ci = mma_bf16(aih, brh, ci);
        fr.re[0][0][h] = cr;
        fr.im[0][0][h] = ci;
      }
    }
  }
}

}  // namespace aq

// ---------------------------------------------------------------- pre-split planes
// The operands split once, before the GEMM, into four BF16 planes (re_hi,
// re_lo, im_hi, im_lo), k-contiguous and zero-padded to whole tiles, so the
// mainloop stages each k step with 16-byte copies (no per-tile conversion,
// no bounds checks):
//   A planes  P[plane][m][k],     m < Mp (multiple of kBM), k < Kp (of kBKb)
//   B planes  P[plane][h][j][k],  h < 2 halves, j < Jp (multiple of kBJ)
// For a linear (unpaired) B the two halves are the left and right kBJ
// columns of every kBN-column tile: column n = (n / kBN) kBN + h kBJ + j'.
namespace aq {

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) u32x4* gptr_u32x4;

struct PlanesA {
  const uint16_t* p;
  int64_t Kp, plane;     // row stride (elements), plane stride
};
struct PlanesB {
  const uint16_t* p;
  int64_t Kp, plane, half;  // column stride, plane stride, half stride
};

__device__ __forceinline__ void split8(const float2 (&x)[8], uint4 (&o)[4]) {
  uint16_t h[4][8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    split_bf16(x[e].x, h[0][e], h[1][e]);
    split_bf16(x[e].y, h[2][e], h[3][e]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
    o[q] = make_uint4(h[q][0] | ((uint32_t)h[q][1] << 16), h[q][2] | ((uint32_t)h[q][3] << 16),
                      h[q][4] | ((uint32_t)h[q][5] << 16), h[q][6] | ((uint32_t)h[q][7] << 16));
}

// The mainloop on pre-split planes: tile rows [m0, m0 + kBM), columns j0 ..
// j0 + kBJ of both halves; K padded (Kp a multiple of kBKb, zeros beyond K).
__device__ __forceinline__ void cgemm_mainloop_planes(const PlanesA& pa, const PlanesB& pb,
                                                      int64_t Kp, int64_t m0, int64_t j0,
                                                      SmemBf16& sm, Frag<float>& fr) {
  constexpr int BK = kBKb;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      fr.re[0][0][h][r] = 0;
      fr.im[0][0][h][r] = 0;
    }
  // 16-byte chunks: A 4 planes x 64 rows x 4 = 1024 (4 per thread),
  // B 4 planes x 2 halves x 64 columns x 4 = 2048 (8 per thread)
  gptr_u32x4 ga[4];
  gptr_u32x4 gb[8];
  uint16_t* la[4];
  uint16_t* lb[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int c = tid + 256 * i;
    const int q = c >> 8, row = (c >> 2) & 63, kq = c & 3;
    ga[i] = (gptr_u32x4)(pa.p + q * pa.plane +
                                                            (m0 + row) * pa.Kp + 8 * kq);
    la[i] = &sm.a[q][row][8 * kq];
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int c = tid + 256 * i;
    const int q = c >> 9, h = (c >> 8) & 1, col = (c >> 2) & 63, kq = c & 3;
    gb[i] = (gptr_u32x4)(pb.p + q * pb.plane + h * pb.half +
                                                            (j0 + col) * pb.Kp + 8 * kq);
    lb[i] = &sm.b[q][h * kBJ + col][8 * kq];
  }
  u32x4 ra[4], rb[8];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) ra[i] = ga[i][k0 / 8];
#pragma unroll
    for (int i = 0; i < 8; ++i) rb[i] = gb[i][k0 / 8];
  };
  fetch(0);
  for (int64_t k0 = 0; k0 < Kp; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) *reinterpret_cast<u32x4*>(la[i]) = ra[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) *reinterpret_cast<u32x4*>(lb[i]) = rb[i];
    __syncthreads();
    if (k0 + BK < Kp) fetch(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int kof = kk + 8 * (lane >> 5);
      const int row = 32 * wm + (lane & 31);
      const bf16x8 arh = load8(&sm.a[0][row][kof]);
      const bf16x8 arl = load8(&sm.a[1][row][kof]);
      const bf16x8 aih = load8(&sm.a[2][row][kof]);
      const bf16x8 ail = load8(&sm.a[3][row][kof]);
      const bf16x8 naih = neg8(aih), nail = neg8(ail);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int col = h * kBJ + 32 * wn + (lane & 31);
        const bf16x8 brh = load8(&sm.b[0][col][kof]);
        const bf16x8 brl = load8(&sm.b[1][col][kof]);
        const bf16x8 bih = load8(&sm.b[2][col][kof]);
        const bf16x8 bil = load8(&sm.b[3][col][kof]);
        f32x16 cr = fr.re[0][0][h], ci = fr.im[0][0][h];
        cr = mma_bf16(arh, brl, cr);
        cr = mma_bf16(arl, brh, cr);
        cr = mma_bf16(naih, bil, cr);
        cr = mma_bf16(nail, bih, cr);
        cr = mma_bf16(arh, brh, cr);
        cr = mma_bf16(naih, bih, cr);
        ci = mma_bf16(arh, bil, ci);
        ci = mma_bf16(arl, bih, ci);
        ci = mma_bf16(aih, brl, ci);
        ci = mma_bf16(ail, brh, ci);
        ci = mma_bf16(arh, bih, ci);
        ci = mma_bf16(aih, brh, ci);
        fr.re[0][0][h] = cr;
        fr.im[0][0][h] = ci;
      }
    }
  }
}


// Taller variant: work-group of 128 WR threads (WR waves of 32 rows x two
// column waves), tile 32 WR rows x two halves of 64 columns; the next k step's
// 16-byte chunks are prefetched into registers while this one is computed.
// (Two register sets, loads issued two steps ahead, measured no faster at
// WR = 4: 248 VGPRs, same time.)
template <int WR>
struct alignas(16) SmemPlanes {
  uint16_t a[4][32 * WR][kBKb + kPadB];
  uint16_t b[4][kBN][kBKb + kPadB];
};

template <int WR>
__device__ __forceinline__ void cgemm_mainloop_planes_w(const PlanesA& pa, const PlanesB& pb,
                                                        int64_t Kp, int64_t m0, int64_t j0,
                                                        SmemPlanes<WR>& sm, Frag<float>& fr) {
  constexpr int BK = kBKb;
  constexpr int NT = 128 * WR;
  constexpr int ROWS = 32 * WR;
  constexpr int AP = 4 * ROWS * 4 / NT;        // A chunks per thread (4)
  constexpr int BP = 4 * 2 * 64 * 4 / NT;      // B chunks per thread (16 / WR)
  static_assert(AP * NT == 4 * ROWS * 4 && BP * NT == 2048, "chunk split");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WR, wn = wave / WR;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      fr.re[0][0][h][r] = 0;
      fr.im[0][0][h][r] = 0;
    }
  gptr_u32x4 ga[AP];
  gptr_u32x4 gb[BP];
  uint16_t* la[AP];
  uint16_t* lb[BP];
#pragma unroll
  for (int i = 0; i < AP; ++i) {
    const int c = tid + NT * i;
    const int q = c / (ROWS * 4), row = (c >> 2) % ROWS, kq = c & 3;
    ga[i] = (gptr_u32x4)(pa.p + q * pa.plane + (m0 + row) * pa.Kp + 8 * kq);
    la[i] = &sm.a[q][row][8 * kq];
  }
#pragma unroll
  for (int i = 0; i < BP; ++i) {
    const int c = tid + NT * i;
    const int q = c >> 9, h = (c >> 8) & 1, col = (c >> 2) & 63, kq = c & 3;
    gb[i] = (gptr_u32x4)(pb.p + q * pb.plane + h * pb.half + (j0 + col) * pb.Kp + 8 * kq);
    lb[i] = &sm.b[q][h * kBJ + col][8 * kq];
  }
  u32x4 ra[AP], rb[BP];
  const int64_t steps = Kp / BK;
  auto fetch = [&](int64_t s) {
#pragma unroll
    for (int i = 0; i < AP; ++i) ra[i] = ga[i][s * (BK / 8)];
#pragma unroll
    for (int i = 0; i < BP; ++i) rb[i] = gb[i][s * (BK / 8)];
  };
  fetch(0);
  for (int64_t s = 0; s < steps; ++s) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < AP; ++i) *reinterpret_cast<u32x4*>(la[i]) = ra[i];
#pragma unroll
    for (int i = 0; i < BP; ++i) *reinterpret_cast<u32x4*>(lb[i]) = rb[i];
    __syncthreads();
    if (s + 1 < steps) fetch(s + 1);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int kof = kk + 8 * (lane >> 5);
      const int row = 32 * wm + (lane & 31);
      const bf16x8 arh = load8(&sm.a[0][row][kof]);
      const bf16x8 arl = load8(&sm.a[1][row][kof]);
      const bf16x8 aih = load8(&sm.a[2][row][kof]);
      const bf16x8 ail = load8(&sm.a[3][row][kof]);
      const bf16x8 naih = neg8(aih), nail = neg8(ail);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int col = h * kBJ + 32 * wn + (lane & 31);
        const bf16x8 brh = load8(&sm.b[0][col][kof]);
        const bf16x8 brl = load8(&sm.b[1][col][kof]);
        const bf16x8 bih = load8(&sm.b[2][col][kof]);
        const bf16x8 bil = load8(&sm.b[3][col][kof]);
        f32x16 cr = fr.re[0][0][h], ci = fr.im[0][0][h];
        cr = mma_bf16(arh, brl, cr);
        cr = mma_bf16(arl, brh, cr);
        cr = mma_bf16(naih, bil, cr);
        cr = mma_bf16(nail, bih, cr);
        cr = mma_bf16(arh, brh, cr);
        cr = mma_bf16(naih, bih, cr);
        ci = mma_bf16(arh, bil, ci);
        ci = mma_bf16(arl, bih, ci);
        ci = mma_bf16(aih, brl, ci);
        ci = mma_bf16(ail, brh, ci);
        ci = mma_bf16(arh, bih, ci);
        ci = mma_bf16(aih, brh, ci);
        fr.re[0][0][h] = cr;
        fr.im[0][0][h] = ci;
      }
    }
  }
}

}  // namespace aq
