// Grouped (batched, uneven) work executed by persistent launches.
//
// Many independent MPS of different lengths and bond dimensions are advanced
// together: one launch processes every two-site contraction (or plain GEMM,
// or single-qubit gate, or rank rule) of a circuit layer across the whole
// population.  The host uploads a descriptor per problem and the prefix sum of
// their tile counts; a fixed grid of work-groups strides over the global tile
// index and binary-searches the owning problem, so the grid size is a free
// (autotuned) parameter independent of the work mix.
#include "gemm_core.h"
#include "sampler_core.h"

namespace aq {

constexpr int kXcds = 8;  // accelerator complex dies of an MI355X

__device__ __forceinline__ int find_owner(const int64_t* __restrict__ prefix, int P, int64_t t) {
  int lo = 0, hi = P - 1;
  while (lo < hi) {  // largest p with prefix[p] <= t
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------ grouped GEMM
// WithTheta = false: the group holds no AQ_GROUPED_THETA problem and the gate
// epilogue is not compiled in (it doubles the kernel's register footprint
// and halves its occupancy).
// WR: tile height in waves (2: 64-row tiles of 256 threads; 1: 32-row tiles of
// 128 threads, which waste half as much of a tile on problems of few rows).
// The host's tile counts follow the tile height; with WR = 1 an
// AQ_GROUPED_HERK problem is computed as a full product (every tile).
template <typename T, bool KFastA, bool NFastB, bool WithTheta, int WR>
__global__ __launch_bounds__(128 * WR, 2) void grouped_gemm_kernel(const aq_gemm_desc_t* __restrict__ d,
                                                                const int64_t* __restrict__ prefix,
                                                                int P, int64_t total,
                                                                int xcd_ranges) {
  constexpr int BM = 32 * WR;
  __shared__ Smem<T, BM> sm;
  // Work-groups are dispatched to the 8 XCDs round-robin (XCD = blockIdx.x
  // mod 8) and every XCD has its own L2.  With xcd_ranges (and a grid that is
  // a multiple of 8) each XCD takes a contiguous range of tiles, so the tiles
  // that run together on one XCD are neighbours of one problem and share
  // operand strips in its L2; otherwise the tiles are strided over the grid.
  int64_t t0 = blockIdx.x, t_end = total, t_step = gridDim.x;
  if (xcd_ranges && gridDim.x % kXcds == 0) {
    const int64_t chunk = cdiv(total, kXcds);
    t0 = (int64_t)(blockIdx.x % kXcds) * chunk + blockIdx.x / kXcds;
    t_end = min(total, (int64_t)(blockIdx.x % kXcds + 1) * chunk);
    t_step = gridDim.x / kXcds;
  }
  for (int64_t t = t0; t < t_end; t += t_step) {
    const int p = find_owner(prefix, P, t);
    const aq_gemm_desc_t g = d[p];
    const int64_t local = t - prefix[p];
    const int paired = WithTheta && g.kind == AQ_GROUPED_THETA;
    const int64_t half = paired ? g.N / 2 : g.N;
    const int64_t tn = paired ? cdiv(half, kBJ) : cdiv(g.N, kBN);
    int64_t m0, j0;
    if (WR == 2 && g.kind == AQ_GROUPED_HERK) {
      // tiles meeting the lower triangle of a square C: tile row bm holds
      // floor(bm/2) + 1 tiles, so rows (2q, 2q+1) hold 2(q+1) tiles and the
      // tiles before them number q(q+1)
      int64_t q = (int64_t)((sqrt(4.0 * (double)local + 1.0) - 1.0) * 0.5);
      while (q > 0 && q * (q + 1) > local) --q;
      while ((q + 1) * (q + 2) <= local) ++q;
      const int64_t r = local - q * (q + 1);
      m0 = (2 * q + r / (q + 1)) * kBM;
      j0 = (r % (q + 1)) * kBJ;
    } else {
      m0 = (local / tn) * BM;
      j0 = (local % tn) * kBJ;
    }
    // a tile past the problem (device-planned problems are smaller than the
    // bounds their tile counts come from): nothing to compute or store
    if (m0 >= g.M || (paired ? j0 >= half : 2 * j0 >= g.N)) continue;
    AStrided<T> aop{(const c2_t<T>*)g.A, g.a_s0, g.a_s1, g.M, g.K, g.conj_a};
    BStrided<T> bop{(const c2_t<T>*)g.B, g.b_s0, g.b_s1, g.K, g.N, g.conj_b};
    ColMap cm{paired, half};
    Frag<T> fr;
    __syncthreads();  // LDS reuse across tiles
    cgemm_mainloop<T, KFastA, NFastB, false, WR>(aop, bop, cm, g.K, m0, j0, sm, fr);
    c2_t<T>* C = (c2_t<T>*)g.C;
    if constexpr (WithTheta) {
    if (paired) {
      // theta epilogue: X[2l+p, q*chiR + j] = sum U[2p+q, 2p1+p2] Psi[2l+p1, p2*chiR + j]
      using M = Mfma<T>;
      constexpr int E = M::E;
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      const int wm = wave % WR, wn = wave / WR;
      T ur[16], ui[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        ur[i] = (T)g.gate[2 * i];
        ui[i] = (T)g.gate[2 * i + 1];
      }
#pragma unroll
      for (int rb = 0; rb < Frag<T>::RB; ++rb)
#pragma unroll
        for (int jb = 0; jb < Frag<T>::JB; ++jb)
#pragma unroll
          for (int r = 0; r < M::NREG; r += 2) {
            const int64_t row = m0 + 32 * wm + rb * E + M::out_row(lane, r);
            const int64_t j = j0 + 32 * wn + jb * E + M::out_col(lane);
            if (row >= g.M || j >= half) continue;
            T pr[4], pi[4];
#pragma unroll
            for (int p1 = 0; p1 < 2; ++p1)
#pragma unroll
              for (int p2 = 0; p2 < 2; ++p2) {
                pr[2 * p1 + p2] = fr.re[rb][jb][p2][r + p1];
                pi[2 * p1 + p2] = fr.im[rb][jb][p2][r + p1];
              }
#pragma unroll
            for (int pp = 0; pp < 2; ++pp)
#pragma unroll
              for (int q = 0; q < 2; ++q) {
                const int o = 2 * pp + q;
                T xr = 0, xi = 0;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                  xr += ur[4 * o + c] * pr[c] - ui[4 * o + c] * pi[c];
                  xi += ur[4 * o + c] * pi[c] + ui[4 * o + c] * pr[c];
                }
                c2_t<T> v;
                v.x = xr;
                v.y = xi;
                C[(row + pp) * g.c_s0 + (q * half + j) * g.c_s1] = v;
              }
          }
      continue;
    }
    }
    frag_foreach<T, WR>(fr, [&](int row, int h, int jj, T re, T im) {
      const int64_t m = m0 + row;
      const int64_t n = 2 * j0 + h * kBJ + jj;
      if (m < g.M && n < g.N) {
        c2_t<T> v;
        v.x = re;
        v.y = im;
        C[m * g.c_s0 + n * g.c_s1] = v;
      }
    });
  }
}

// ------------------------------------------------------------ grouped 1q gates
// T_out[a,t,b] = sum_s U[s,t] T_in[a,s,b] for every site in the list
__global__ void grouped_1q_kernel(const aq_site1q_desc_t* __restrict__ d,
                                  const int64_t* __restrict__ prefix, int P, int64_t total) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int p = find_owner(prefix, P, e);
    const aq_site1q_desc_t g = d[p];
    const int64_t idx = e - prefix[p];
    if (idx >= g.chiL * g.chiR) continue;   // a device-planned site below its bound
    const int64_t a = idx / g.chiR, b = idx % g.chiR;
    const float2* in = (const float2*)g.T_in;
    float2* out = (float2*)g.T_out;
    const float2 x0 = in[(a * 2 + 0) * g.chiR + b];
    const float2 x1 = in[(a * 2 + 1) * g.chiR + b];
    float2 u[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = make_float2((float)g.gate[2 * i], (float)g.gate[2 * i + 1]);
    out[(a * 2 + 0) * g.chiR + b] = cadd(cmul(u[0], x0), cmul(u[2], x1));
    out[(a * 2 + 1) * g.chiR + b] = cadd(cmul(u[1], x0), cmul(u[3], x1));
  }
}

// ------------------------------------------------------------ grouped sampler
// one site index for many MPS: tiles of every MPS's [S, chiL] x [chiL, 2 chiR]
__global__ __launch_bounds__(kThreads, 2) void grouped_sample_gemm_kernel(
    const aq_sample_desc_t* __restrict__ d, const int64_t* __restrict__ prefix, int P,
    int64_t total) {
  __shared__ Smem<float> sm;
  for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
    const int p = find_owner(prefix, P, t);
    const aq_sample_desc_t g = d[p];
    const int64_t local = t - prefix[p];
    const int64_t tn = cdiv(g.chiR, kBJ);
    const int64_t m0 = (local / tn) * kBM;
    const int64_t j0 = (local % tn) * kBJ;
    AGather<float> aop{(const float2*)g.L, g.off, g.ldL, g.S, g.chiL};
    BStrided<float> bop{(const float2*)g.site, 2 * g.chiR, 1, g.chiL, 2 * g.chiR, 0};
    ColMap cm{1, g.chiR};
    Frag<float> fr;
    __syncthreads();
    cgemm_mainloop<float, true, true>(aop, bop, cm, g.chiL, m0, j0, sm, fr);
    sampler_epilogue<float>(fr, (float2*)g.Tout, (float*)g.part, g.S, g.chiR, m0, j0,
                            (int)cdiv(g.chiR, 32));
  }
}

__global__ void grouped_sample_select_kernel(const aq_sample_desc_t* __restrict__ d,
                                             const int64_t* __restrict__ prefix, int P,
                                             int64_t total) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int p = find_owner(prefix, P, e);
    const aq_sample_desc_t g = d[p];
    const int64_t s = e - prefix[p];
    select_one<float>((const float*)g.part, (int)cdiv(g.chiR, 32), g.uniforms, g.n,
                      g.site_index, s, g.chiR, g.off_next, g.bits, g.words, nullptr, 0);
  }
}

// ------------------------------------------------------------ grouped copy
// dst (complex64) [r, c] = op(src (complex128) at r*s0 + c*s1) * row[r] * col[c]
__global__ void grouped_copy_kernel(const aq_copy_desc_t* __restrict__ d,
                                    const int64_t* __restrict__ prefix, int P, int64_t total) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int p = find_owner(prefix, P, e);
    const aq_copy_desc_t g = d[p];
    const int64_t idx = e - prefix[p];
    if (idx >= g.rows * g.cols) continue;   // a device-planned core below its bound
    const int64_t r = idx / g.cols, c = idx % g.cols;
    double2 v = ((const double2*)g.src)[r * g.s0 + c * g.s1];
    if (g.conj) v.y = -v.y;
    double sc = 1.0;
    if (g.row_scale) sc *= g.row_scale[r];
    if (g.col_scale) sc *= g.col_scale[c];
    ((float2*)g.dst)[r * g.cols + c] = make_float2((float)(v.x * sc), (float)(v.y * sc));
  }
}

__global__ void zero_u64_dev(uint64_t* p, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

static int upload_prefix(const int64_t* prefix_dev, int P) {
  AQ_REQUIRE(prefix_dev != nullptr && P > 0, AQ_ERR_ARG, "grouped: empty problem list");
  return AQ_OK;
}

}  // namespace aq

using namespace aq;

extern "C" int aq_grouped_gemm(int dtype, const aq_gemm_desc_t* descs, const int64_t* prefix,
                               int num_problems, int64_t total_tiles, int layout, int grid,
                               hipStream_t stream) {
  AQ_REQUIRE(descs && prefix, AQ_ERR_ARG, "aq_grouped_gemm: null pointer");
  int rc = upload_prefix(prefix, num_problems);
  if (rc) return rc;
  AQ_REQUIRE(total_tiles > 0, AQ_ERR_ARG, "aq_grouped_gemm: no tiles");
  AQ_REQUIRE(grid >= 1 && grid <= 65536, AQ_ERR_ARG, "aq_grouped_gemm: grid out of range");
  AQ_REQUIRE(layout >= 0 && layout <= 31, AQ_ERR_ARG, "aq_grouped_gemm: layout 0..31");
  const bool ka = layout & 1, nb = layout & 2, no_theta = layout & 4;
  const int xcd = (layout & 8) ? 1 : 0;
  const bool half = layout & 16;  // 32-row tiles
#define AQ_GG1(T, KA, NB, TH, WR)                                                              \
  grouped_gemm_kernel<T, KA, NB, TH, WR><<<grid, 128 * WR, 0, stream>>>(descs, prefix,         \
                                                                        num_problems,          \
                                                                        total_tiles, xcd)
#define AQ_GG(T, KA, NB)                                                                       \
  do {                                                                                         \
    if (half) {                                                                                \
      if (no_theta) AQ_GG1(T, KA, NB, false, 1);                                               \
      else AQ_GG1(T, KA, NB, true, 1);                                                         \
    } else {                                                                                   \
      if (no_theta) AQ_GG1(T, KA, NB, false, 2);                                               \
      else AQ_GG1(T, KA, NB, true, 2);                                                         \
    }                                                                                          \
  } while (0)
  if (dtype == AQ_C64) {
    if (ka && nb) AQ_GG(float, true, true);
    else if (ka) AQ_GG(float, true, false);
    else if (nb) AQ_GG(float, false, true);
    else AQ_GG(float, false, false);
  } else {
    AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_grouped_gemm: unknown dtype");
    if (ka && nb) AQ_GG(double, true, true);
    else if (ka) AQ_GG(double, true, false);
    else if (nb) AQ_GG(double, false, true);
    else AQ_GG(double, false, false);
  }
#undef AQ_GG
#undef AQ_GG1
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_grouped_1q(const aq_site1q_desc_t* descs, const int64_t* prefix,
                             int num_sites, int64_t total_elems, int grid, hipStream_t stream) {
  AQ_REQUIRE(descs && prefix && num_sites > 0 && total_elems > 0, AQ_ERR_ARG,
             "aq_grouped_1q: empty work list");
  AQ_REQUIRE(grid >= 1 && grid <= 1 << 20, AQ_ERR_ARG, "aq_grouped_1q: grid out of range");
  grouped_1q_kernel<<<grid, 256, 0, stream>>>(descs, prefix, num_sites, total_elems);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_sample_grouped_site(const aq_sample_desc_t* descs, const int64_t* tile_prefix,
                                      const int64_t* shot_prefix, int count, int64_t total_tiles,
                                      int64_t total_shots, int grid, hipStream_t stream) {
  AQ_REQUIRE(descs && tile_prefix && shot_prefix && count > 0, AQ_ERR_ARG,
             "aq_sample_grouped_site: empty work list");
  AQ_REQUIRE(total_tiles > 0 && total_shots > 0, AQ_ERR_ARG, "aq_sample_grouped_site: no work");
  AQ_REQUIRE(grid >= 1 && grid <= 65536, AQ_ERR_ARG, "aq_sample_grouped_site: grid out of range");
  grouped_sample_gemm_kernel<<<grid, kThreads, 0, stream>>>(descs, tile_prefix, count,
                                                            total_tiles);
  AQ_LAUNCH_CHECK();
  const int64_t sg = cdiv(total_shots, 256);
  grouped_sample_select_kernel<<<(unsigned)(sg < 65536 ? sg : 65536), 256, 0, stream>>>(
      descs, shot_prefix, count, total_shots);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_zero_u64(uint64_t* p, int64_t count, hipStream_t stream) {
  AQ_REQUIRE(p && count > 0, AQ_ERR_ARG, "aq_zero_u64: empty");
  zero_u64_dev<<<(unsigned)cdiv(count, 256), 256, 0, stream>>>(p, count);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_grouped_copy(const aq_copy_desc_t* descs, const int64_t* prefix, int count,
                               int64_t total_elems, int grid, hipStream_t stream) {
  AQ_REQUIRE(descs && prefix && count > 0 && total_elems > 0, AQ_ERR_ARG,
             "aq_grouped_copy: empty work list");
  AQ_REQUIRE(grid >= 1 && grid <= 1 << 20, AQ_ERR_ARG, "aq_grouped_copy: grid out of range");
  grouped_copy_kernel<<<grid, 256, 0, stream>>>(descs, prefix, count, total_elems);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}
