// Device pieces of the sweep sampler shared by the single-MPS and grouped
// launches (mps_pytorch.py:247-355 semantics).
#pragma once

#include "gemm_core.h"

namespace aq {

template <typename T>
__device__ __forceinline__ void sampler_epilogue(const Frag<T>& fr, c2_t<T>* __restrict__ Tout,
                                                 T* __restrict__ part, int64_t S, int64_t chiR,
                                                 int64_t m0, int64_t j0, int ntiles) {
  using M = Mfma<T>;
  constexpr int E = M::E;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t ld = 2 * chiR;
  const int64_t tile32 = (j0 + 32 * wn) / 32;  // index of this wave's 32-column group
#pragma unroll
  for (int rb = 0; rb < Frag<T>::RB; ++rb)
#pragma unroll
    for (int r = 0; r < M::NREG; ++r) {
      const int64_t row = m0 + 32 * wm + rb * E + M::out_row(lane, r);
      T s0 = 0, s1 = 0;
#pragma unroll
      for (int jb = 0; jb < Frag<T>::JB; ++jb) {
        const int64_t j = j0 + 32 * wn + jb * E + M::out_col(lane);
        const T r0 = fr.re[rb][jb][0][r], i0 = fr.im[rb][jb][0][r];
        const T r1 = fr.re[rb][jb][1][r], i1 = fr.im[rb][jb][1][r];
        s0 += r0 * r0 + i0 * i0;  // zero for columns beyond chiR (zero-filled B)
        s1 += r1 * r1 + i1 * i1;
        if (row < S && j < chiR) {
          c2_t<T> v0, v1;
          v0.x = r0; v0.y = i0;
          v1.x = r1; v1.y = i1;
          Tout[row * ld + j] = v0;
          Tout[row * ld + chiR + j] = v1;
        }
      }
      // reduce over the lanes that hold the other columns of this row
#pragma unroll
      for (int w = M::COL_LANES / 2; w >= 1; w >>= 1) {
        s0 += __shfl_xor(s0, w, 64);
        s1 += __shfl_xor(s1, w, 64);
      }
      if ((lane & (M::COL_LANES - 1)) == 0 && row < S && tile32 < ntiles) {
        part[(row * ntiles + tile32) * 2 + 0] = s0;
        part[(row * ntiles + tile32) * 2 + 1] = s1;
      }
    }
}

// Draw for shot s at `site`: sum the partials in a fixed order, form the
// reference's probabilities and set the outcome bit and the next gather offset.
template <typename T>
__device__ __forceinline__ void select_one(const T* __restrict__ part, int ntiles,
                                           const double* __restrict__ uniforms, int64_t n,
                                           int64_t site, int64_t s, int64_t chiR,
                                           int32_t* __restrict__ off, uint64_t* __restrict__ bits,
                                           int64_t words, double* __restrict__ p0_out,
                                           int single) {
  T a = 0, b = 0;
  for (int t = 0; t < ntiles; ++t) {
    a += part[(s * ntiles + t) * 2 + 0];
    b += part[(s * ntiles + t) * 2 + 1];
  }
  double p0d;
  if (single) {
    // single-shot rule (mps_pytorch.py:217-232): Python-float normalisation,
    // 0.5 when the total is below 1e-15
    const double ad = (double)a, bd = (double)b, tot = ad + bd;
    p0d = tot > 1e-15 ? ad / tot : 0.5;
  } else {
    // mps_pytorch.py:307-314: abs, +1e-15 floor, normalise (in the real dtype)
    a = fabs(a) + (T)1e-15;
    b = fabs(b) + (T)1e-15;
    p0d = (double)(a / (a + b));
  }
  const double u = uniforms[s * n + site];
  const int o = (u >= p0d) ? 1 : 0;
  const int64_t pos = n - 1 - site;
  if (o) bits[s * words + pos / 64] |= (1ull << (pos % 64));
  off[s] = o ? (int32_t)chiR : 0;
  if (p0_out) p0_out[s * n + site] = p0d;
}


}  // namespace aq
