// Complex GEMM mainloop on CDNA4 MFMA (gfx950), shared by the theta
// contraction, the sweep sampler, the generic C = op(A) op(B) entry point and
// the projections of the truncated-SVD variants.
//
// Work-group tile: 64 rows x 2 "halves" x 64 columns per half, 256 threads =
// 4 waves in a 2 x 2 arrangement; wave (wm, wn) owns rows [32 wm, 32 wm + 32)
// and columns [32 wn, 32 wn + 32) of BOTH halves.  For the MPS kernels the two
// halves are the physical index of the right site (p2 = 0 / 1), so one lane
// holds, for a given (l, j), all four products Psi[2l+p1, p2*chiR + j]: the
// 4x4 gate (theta) or the two outcome probabilities (sampler) are applied in
// registers with no LDS round trip.  For a plain GEMM the halves are just the
// left and right 64 columns of a 128-column tile.
//
// Complex products use the 4-multiplication form on real MFMA:
//   Cr += Ar Br - Ai Bi,  Ci += Ar Bi + Ai Br
// or, where a kernel has the registers for a third accumulator set
// (ThreeM = true), the 3-multiplication form:
//   P1 += Ar Br,  P2 += Ai Bi,  P3 += (Ar + Ai)(Br + Bi),
//   Cr = P1 - P2,  Ci = P3 - P1 - P2   (after the k loop)
// -- a quarter fewer MFMAs, absolute error of the same order as the
// 4-multiplication form (a few units of the FP32 rounding of |A| |B|).
// FP32: v_mfma_f32_32x32x2f32 (one 32x32 block per wave and half);
// FP64: v_mfma_f64_16x16x4f64 (2 x 2 blocks of 16x16 per wave and half).
// Operand tiles are staged through LDS as separate real / imaginary planes,
// k-major, so each MFMA operand fetch is one conflict-free ds_read per lane.
#pragma once

#include "aq_common.h"

namespace aq {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;

template <> struct Mfma<float> {
  static constexpr int E = 32;   // block edge
  static constexpr int TK = 2;   // k per instruction
  static constexpr int NREG = 16;
  using acc_t = f32x16;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  // operand lane layout: lane holds A[op_idx][op_k] / B[op_k][op_idx]
  static __device__ __forceinline__ int op_idx(int lane) { return lane & 31; }
  static __device__ __forceinline__ int op_k(int lane) { return lane >> 5; }
  // tile row fed to MFMA row op_idx(lane): identity (rows 2l, 2l+1 already land
  // in accumulator registers r, r+1)
  static __device__ __forceinline__ int a_row(int lane) { return lane & 31; }
  // accumulator layout: register r of lane -> (out_row, out_col)
  static __device__ __forceinline__ int out_row(int lane, int r) {
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
  }
  static __device__ __forceinline__ int out_col(int lane) { return lane & 31; }
  static constexpr int COL_LANES = 32;  // lanes sharing a row set
};

template <> struct Mfma<double> {
  static constexpr int E = 16;
  static constexpr int TK = 4;
  static constexpr int NREG = 4;
  using acc_t = f64x4;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int op_idx(int lane) { return lane & 15; }
  static __device__ __forceinline__ int op_k(int lane) { return lane >> 4; }
  // The FP64 instruction returns MFMA row 4r + lane/16 in register r.  Feeding
  // tile row pi(i) = 8(r>>1) + 2g + (r&1) (i = 4r + g) to MFMA row i puts tile
  // rows (2l, 2l+1) in registers (r, r+1) of one lane, as the FP32 layout does,
  // so the pairwise epilogues (gate, outcome halves) are shared.
  static __device__ __forceinline__ int a_row(int lane) {
    const int i = lane & 15, r = i >> 2, g = i & 3;
    return 8 * (r >> 1) + 2 * g + (r & 1);
  }
  static __device__ __forceinline__ int out_row(int lane, int r) {
    return 8 * (r >> 1) + 2 * (lane >> 4) + (r & 1);
  }
  static __device__ __forceinline__ int out_col(int lane) { return lane & 15; }
  static constexpr int COL_LANES = 16;
};

constexpr int kBM = 64;        // rows per work-group tile
constexpr int kBJ = 64;        // columns per half
constexpr int kBN = 2 * kBJ;   // columns per tile
constexpr int kThreads = 256;

template <typename T> struct TileK;
template <> struct TileK<float> { static constexpr int value = 16; };
template <> struct TileK<double> { static constexpr int value = 8; };

// ---------------------------------------------------------------- operands
// One complex element through a global (address space 1) pointer: operand
// pointers that come from descriptors in memory are otherwise generic, and
// generic (flat) loads also count against the LDS counter, so every wait for
// an LDS operand would wait for the next tile's global loads as well.
template <typename T> struct Vec2;
template <> struct Vec2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct Vec2<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <typename T> using v2_t = typename Vec2<T>::type;

template <typename T>
__device__ __forceinline__ v2_t<T> gload2(const c2_t<T>* p) {
  return *(const __attribute__((address_space(1))) v2_t<T>*)(p);
}

// Operand functors: operator() is the bounds-checked element read; base(),
// kstep(), conj_flag() and ones() give the mainloop's split access -- a
// per-row (A) or per-column (B) pointer clamped into the operand (base_of),
// advanced by k * kstep(), so every load is in bounds and unconditional.
// A(m, k) = p[m*s0 + k*s1]
template <typename T> struct AStrided {
  const c2_t<T>* p;
  int64_t s0, s1, M, K;
  int conj;
  __device__ __forceinline__ c2_t<T> operator()(int64_t m, int64_t k) const {
    c2_t<T> v;
    if (m < M && k < K) {
      v = p[m * s0 + k * s1];
      if (conj) v.y = -v.y;
    } else {
      v.x = 0; v.y = 0;
    }
    return v;
  }
  __device__ __forceinline__ const c2_t<T>* base_of(int64_t m, bool& ok) const {
    ok = m < M;
    return p + (ok ? m : 0) * s0;
  }
  __device__ __forceinline__ int64_t kstep() const { return s1; }
  __device__ __forceinline__ int conj_flag() const { return conj; }
  __device__ __forceinline__ bool ones() const { return false; }
};

// Sampler left states: A(m, k) = base[m*ld + off[m] + k]; base == nullptr means
// the all-ones [M, 1] boundary vector.
template <typename T> struct AGather {
  const c2_t<T>* base;
  const int32_t* off;
  int64_t ld, M, K;
  __device__ __forceinline__ c2_t<T> operator()(int64_t m, int64_t k) const {
    c2_t<T> v;
    v.x = 0; v.y = 0;
    if (m < M && k < K) {
      if (base == nullptr) {
        v.x = 1;
      } else {
        v = base[m * ld + (off ? off[m] : 0) + k];
      }
    }
    return v;
  }
  // base == nullptr (ones()): no load is made, the element is 1 for k < K
  __device__ __forceinline__ const c2_t<T>* base_of(int64_t m, bool& ok) const {
    ok = m < M;
    if (base == nullptr) return nullptr;
    const int64_t mc = ok ? m : 0;
    return base + mc * ld + (off ? off[mc] : 0);
  }
  __device__ __forceinline__ int64_t kstep() const { return 1; }
  __device__ __forceinline__ int conj_flag() const { return 0; }
  __device__ __forceinline__ bool ones() const { return base == nullptr; }
};

// B(k, n) = p[k*s0 + n*s1]
template <typename T> struct BStrided {
  const c2_t<T>* p;
  int64_t s0, s1, K, N;
  int conj;
  __device__ __forceinline__ c2_t<T> operator()(int64_t k, int64_t n) const {
    c2_t<T> v;
    if (k < K && n < N) {
      v = p[k * s0 + n * s1];
      if (conj) v.y = -v.y;
    } else {
      v.x = 0; v.y = 0;
    }
    return v;
  }
  __device__ __forceinline__ const c2_t<T>* base_of(int64_t n, bool& ok) const {
    ok = n < N;
    return p + (ok ? n : 0) * s1;
  }
  __device__ __forceinline__ int64_t kstep() const { return s0; }
  __device__ __forceinline__ int conj_flag() const { return conj; }
};

// Column map of a tile: PAIRED -> half h, column jj maps to h*half_n + j0 + jj
// (valid when j0 + jj < half_n); linear -> n0 + h*kBJ + jj (valid when < N).
struct ColMap {
  int paired;
  int64_t half_n;  // paired: columns per half (chiR); linear: N
  __device__ __forceinline__ int64_t col(int64_t tile_j0, int h, int jj) const {
    if (paired) {
      int64_t j = tile_j0 + jj;
      return j < half_n ? h * half_n + j : INT64_MAX;
    }
    int64_t n = 2 * tile_j0 + h * kBJ + jj;
    return n < half_n ? n : INT64_MAX;
  }
};

// ---------------------------------------------------------------- mainloop
// Accumulator fragment of one wave: [RB][JB][2 halves] blocks of E x E.
template <typename T> struct Frag {
  static constexpr int E = Mfma<T>::E;
  static constexpr int RB = 32 / E;
  static constexpr int JB = 32 / E;
  typename Mfma<T>::acc_t re[RB][JB][2];
  typename Mfma<T>::acc_t im[RB][JB][2];
  typename Mfma<T>::acc_t p3[RB][JB][2];   // 3M form only
};

// One LDS tile, two barriers per k step (a two-slot ring with one barrier
// measured slower: the barrier is not what bounds the loop).
// BM: rows of the work-group tile (64, or 32 for the half-height tile of the
// grouped kernel's tile-shape option).
template <typename T, int BM = kBM> struct alignas(16) Smem {
  static constexpr int BK = TileK<T>::value;
  static constexpr int PAD = 4;
  T a[2][BK][BM + PAD];
  T b[2][BK][kBN + PAD];
};

// k_fast_a: the A functor is contiguous along k (else along m).
// n_fast_b: the B functor is contiguous along n (else along k).
// K is the reduction length the functors were built with: element k is
// in range when k < K (rows / columns are checked by the functors).
// WR: waves per tile column (2: the 64-row tile of 256 threads; 1: a 32-row
// tile of 128 threads, two waves side by side).
template <typename T, bool KFastA, bool NFastB, bool ThreeM = false, int WR = 2, class AOp,
          class BOp>
__device__ __forceinline__ void cgemm_mainloop(const AOp& aop, const BOp& bop, const ColMap& cm,
                                               int64_t K, int64_t m0, int64_t j0,
                                               Smem<T, 32 * WR>& sm, Frag<T>& fr) {
  using M = Mfma<T>;
  constexpr int BK = Smem<T, 32 * WR>::BK;
  constexpr int E = M::E;
  constexpr int BM = 32 * WR;
  constexpr int NT = 128 * WR;
  constexpr int A_PER = BM * BK / NT;
  constexpr int B_PER = kBN * BK / NT;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WR, wn = wave / WR;

#pragma unroll
  for (int rb = 0; rb < Frag<T>::RB; ++rb)
#pragma unroll
    for (int jb = 0; jb < Frag<T>::JB; ++jb)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < M::NREG; ++r) {
          fr.re[rb][jb][h][r] = 0;
          fr.im[rb][jb][h][r] = 0;
          if (ThreeM) fr.p3[rb][jb][h][r] = 0;
        }

  if (K <= 0) return;

  // Per-thread operand pointers are loop-invariant.  The loads of a tile are
  // unconditional (from addresses clamped into the operands) and the values
  // are used only when the tile is stored to LDS, so the loads of the next
  // tile issue back to back and stay in flight during this tile's MFMAs;
  // out-of-range elements are zeroed and conjugation applied at the store.
  const c2_t<T>* bptr[B_PER];
  bool bok[B_PER];
  int bkk[B_PER], btc[B_PER];
#pragma unroll
  for (int i = 0; i < B_PER; ++i) {
    int e = tid + i * NT;
    int tc, kk;
    if (NFastB) { tc = e % kBN; kk = e / kBN; }
    else        { kk = e % BK;  tc = e / BK; }
    bkk[i] = kk;
    btc[i] = tc;
    bptr[i] = bop.base_of(cm.col(j0, tc / kBJ, tc % kBJ), bok[i]);
  }
  const c2_t<T>* aptr[A_PER];
  bool aok[A_PER];
  int amm[A_PER], akk[A_PER];
#pragma unroll
  for (int i = 0; i < A_PER; ++i) {
    int e = tid + i * NT;
    if (KFastA) { akk[i] = e % BK; amm[i] = e / BK; }
    else        { amm[i] = e % BM; akk[i] = e / BM; }
    aptr[i] = aop.base_of(m0 + amm[i], aok[i]);
  }
  const int64_t a_ks = aop.kstep(), b_ks = bop.kstep();
  const bool a_ones = aop.ones();
  const int a_conj = aop.conj_flag(), b_conj = bop.conj_flag();

  v2_t<T> ra[A_PER], rbv[B_PER];
  unsigned amask = 0, bmask = 0;
  auto fetch = [&](int64_t k0) {
    amask = 0;
    bmask = 0;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int64_t k = k0 + akk[i];
      amask |= (unsigned)(aok[i] && k < K) << i;
      if (a_ones) {
        ra[i].x = 1;
        ra[i].y = 0;
      } else {
        ra[i] = gload2<T>(aptr[i] + (k < K ? k : 0) * a_ks);
      }
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int64_t k = k0 + bkk[i];
      bmask |= (unsigned)(bok[i] && k < K) << i;
      rbv[i] = gload2<T>(bptr[i] + (k < K ? k : 0) * b_ks);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const bool ok = (amask >> i) & 1u;
      sm.a[0][akk[i]][amm[i]] = ok ? ra[i].x : T(0);
      sm.a[1][akk[i]][amm[i]] = ok ? (a_conj ? -ra[i].y : ra[i].y) : T(0);
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const bool ok = (bmask >> i) & 1u;
      sm.b[0][bkk[i]][btc[i]] = ok ? rbv[i].x : T(0);
      sm.b[1][bkk[i]][btc[i]] = ok ? (b_conj ? -rbv[i].y : rbv[i].y) : T(0);
    }
  };
  auto compute = [&]() {
#pragma unroll
    for (int kk = 0; kk < BK; kk += M::TK) {
      const int kr = kk + M::op_k(lane);
      T ar[Frag<T>::RB], ai[Frag<T>::RB], nai[Frag<T>::RB], as[Frag<T>::RB];
#pragma unroll
      for (int rb = 0; rb < Frag<T>::RB; ++rb) {
        const int row = 32 * wm + rb * E + M::a_row(lane);
        ar[rb] = sm.a[0][kr][row];
        ai[rb] = sm.a[1][kr][row];
        nai[rb] = -ai[rb];
        as[rb] = ar[rb] + ai[rb];
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int jb = 0; jb < Frag<T>::JB; ++jb) {
          const int col = h * kBJ + 32 * wn + jb * E + M::op_idx(lane);
          const T br = sm.b[0][kr][col];
          const T bi = sm.b[1][kr][col];
          const T bs = br + bi;
#pragma unroll
          for (int rb = 0; rb < Frag<T>::RB; ++rb) {
            if (ThreeM) {
              fr.re[rb][jb][h] = M::mma(ar[rb], br, fr.re[rb][jb][h]);   // P1
              fr.im[rb][jb][h] = M::mma(ai[rb], bi, fr.im[rb][jb][h]);   // P2
              fr.p3[rb][jb][h] = M::mma(as[rb], bs, fr.p3[rb][jb][h]);   // P3
            } else {
              fr.re[rb][jb][h] = M::mma(ar[rb], br, fr.re[rb][jb][h]);
              fr.re[rb][jb][h] = M::mma(nai[rb], bi, fr.re[rb][jb][h]);
              fr.im[rb][jb][h] = M::mma(ar[rb], bi, fr.im[rb][jb][h]);
              fr.im[rb][jb][h] = M::mma(ai[rb], br, fr.im[rb][jb][h]);
            }
          }
        }
    }
  };

  fetch(0);
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    store();
    __syncthreads();
    if (k0 + BK < K) fetch(k0 + BK);  // in flight during this tile's MFMAs
    compute();
  }
  if (ThreeM) {
    // Cr = P1 - P2, Ci = P3 - P1 - P2
#pragma unroll
    for (int rb = 0; rb < Frag<T>::RB; ++rb)
#pragma unroll
      for (int jb = 0; jb < Frag<T>::JB; ++jb)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int r = 0; r < M::NREG; ++r) {
            const T p1 = fr.re[rb][jb][h][r], p2 = fr.im[rb][jb][h][r];
            fr.re[rb][jb][h][r] = p1 - p2;
            fr.im[rb][jb][h][r] = fr.p3[rb][jb][h][r] - p1 - p2;
          }
  }
}

// Visit every accumulator element of this thread: f(tile_row, h, tile_jj, re, im)
// where tile_row in [0, 32 WR), tile_jj in [0, 64).
template <typename T, int WR = 2, class F>
__device__ __forceinline__ void frag_foreach(const Frag<T>& fr, F&& f) {
  using M = Mfma<T>;
  constexpr int E = M::E;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave % WR, wn = wave / WR;
#pragma unroll
  for (int rb = 0; rb < Frag<T>::RB; ++rb)
#pragma unroll
    for (int jb = 0; jb < Frag<T>::JB; ++jb)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < M::NREG; ++r)
          f(32 * wm + rb * E + M::out_row(lane, r), h, 32 * wn + jb * E + M::out_col(lane),
            fr.re[rb][jb][h][r], fr.im[rb][jb][h][r]);
}

}  // namespace aq
