// Strided-batched complex GEMM with a scaled accumulate epilogue,
//   C_b = alpha op(A_b) op(B_b) + beta C_b,   b = 0 .. batch-1,
// on the MFMA mainloop of gemm_core.h (64 x 128 work-group tiles).  The
// eigensolver's back-transformations and checks run one such launch per step
// over every Gram core of a brickwork sub-layer: the problems of a batch
// share their shapes, so the batch index is the grid's z dimension and no
// descriptors are needed.  alpha and beta are real; beta == 0 never reads C
// (C may hold NaN there).
#include "gemm_core.h"

namespace aq {

template <typename T, bool KFastA, bool NFastB, bool Lower>
__global__ __launch_bounds__(kThreads, 2) void bgemm_kernel(
    const c2_t<T>* __restrict__ A, int64_t a_s0, int64_t a_s1, int64_t a_sb, int conj_a,
    const c2_t<T>* __restrict__ B, int64_t b_s0, int64_t b_s1, int64_t b_sb, int conj_b,
    c2_t<T>* __restrict__ C, int64_t c_s0, int64_t c_s1, int64_t c_sb, int64_t M, int64_t N,
    int64_t K, T alpha, T beta, int tiles_n) {
  __shared__ Smem<T> sm;
  const int64_t b = blockIdx.z;
  const int64_t tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
  const int64_t m0 = tm * kBM, j0 = tn * kBJ;
  if (Lower && 2 * j0 > m0 + kBM - 1) return;  // the tile lies above the diagonal
  AStrided<T> aop{A + b * a_sb, a_s0, a_s1, M, K, conj_a};
  BStrided<T> bop{B + b * b_sb, b_s0, b_s1, K, N, conj_b};
  ColMap cm{0, N};
  Frag<T> fr;
  cgemm_mainloop<T, KFastA, NFastB>(aop, bop, cm, K, m0, j0, sm, fr);
  c2_t<T>* Cb = C + b * c_sb;
  frag_foreach<T>(fr, [&](int row, int h, int jj, T re, T im) {
    const int64_t m = m0 + row;
    const int64_t n = 2 * j0 + h * kBJ + jj;
    if (m < M && n < N && (!Lower || n <= m)) {
      c2_t<T>* p = Cb + m * c_s0 + n * c_s1;
      c2_t<T> v;
      if (beta != T(0)) {
        const c2_t<T> c = *p;
        v.x = alpha * re + beta * c.x;
        v.y = alpha * im + beta * c.y;
      } else {
        v.x = alpha * re;
        v.y = alpha * im;
      }
      *p = v;
    }
  });
}

template <typename T, bool Lower>
static int bgemm_impl(int64_t M, int64_t N, int64_t K, double alpha, const void* A, int64_t a_s0,
                      int64_t a_s1, int64_t a_sb, int conj_a, const void* B, int64_t b_s0,
                      int64_t b_s1, int64_t b_sb, int conj_b, double beta, void* C, int64_t c_s0,
                      int64_t c_s1, int64_t c_sb, int64_t batch, hipStream_t s) {
  const int64_t tiles_m = cdiv(M, kBM), tiles_n = cdiv(N, kBN);
  AQ_REQUIRE(tiles_m * tiles_n < (1LL << 31) && batch <= 65535, AQ_ERR_ARG,
             "aq_bgemm: grid too large (%lld tiles, batch %lld)",
             (long long)(tiles_m * tiles_n), (long long)batch);
  dim3 grid((unsigned)(tiles_m * tiles_n), 1, (unsigned)batch);
  const c2_t<T>* a = (const c2_t<T>*)A;
  const c2_t<T>* b = (const c2_t<T>*)B;
  c2_t<T>* c = (c2_t<T>*)C;
  const bool ka = a_s1 == 1, nb = b_s1 == 1;
#define AQ_BG(KA, NB)                                                                          \
  bgemm_kernel<T, KA, NB, Lower><<<grid, kThreads, 0, s>>>(a, a_s0, a_s1, a_sb, conj_a, b, b_s0, b_s1, \
                                                    b_sb, conj_b, c, c_s0, c_s1, c_sb, M, N, K, \
                                                    (T)alpha, (T)beta, (int)tiles_n)
  if (ka && nb) AQ_BG(true, true);
  else if (ka) AQ_BG(true, false);
  else if (nb) AQ_BG(false, true);
  else AQ_BG(false, false);
#undef AQ_BG
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

}  // namespace aq

using namespace aq;

extern "C" int aq_bgemm(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
                        int64_t a_s0, int64_t a_s1, int64_t a_sb, int conj_a, const void* B,
                        int64_t b_s0, int64_t b_s1, int64_t b_sb, int conj_b, double beta, void* C,
                        int64_t c_s0, int64_t c_s1, int64_t c_sb, int64_t batch,
                        hipStream_t stream) {
  AQ_REQUIRE(A && B && C, AQ_ERR_ARG, "aq_bgemm: null pointer");
  AQ_REQUIRE(M > 0 && N > 0 && K >= 0 && batch > 0, AQ_ERR_ARG,
             "aq_bgemm: sizes must be positive (M %lld N %lld K %lld batch %lld)", (long long)M,
             (long long)N, (long long)K, (long long)batch);
  if (dtype == AQ_C64)
    return bgemm_impl<float, false>(M, N, K, alpha, A, a_s0, a_s1, a_sb, conj_a, B, b_s0, b_s1,
                                    b_sb, conj_b, beta, C, c_s0, c_s1, c_sb, batch, stream);
  AQ_REQUIRE(dtype == AQ_C128, AQ_ERR_ARG, "aq_bgemm: unknown dtype %d", dtype);
  return bgemm_impl<double, false>(M, N, K, alpha, A, a_s0, a_s1, a_sb, conj_a, B, b_s0, b_s1,
                                   b_sb, conj_b, beta, C, c_s0, c_s1, c_sb, batch, stream);
}

extern "C" int aq_bgemm_lower(int64_t M, int64_t K, double alpha, const void* A, int64_t a_s0,
                              int64_t a_s1, int64_t a_sb, int conj_a, const void* B,
                              int64_t b_s0, int64_t b_s1, int64_t b_sb, int conj_b, double beta,
                              void* C, int64_t c_s0, int64_t c_s1, int64_t c_sb, int64_t batch,
                              hipStream_t stream) {
  AQ_REQUIRE(A && B && C, AQ_ERR_ARG, "aq_bgemm_lower: null pointer");
  AQ_REQUIRE(M > 0 && K >= 0 && batch > 0, AQ_ERR_ARG, "aq_bgemm_lower: bad sizes");
  return bgemm_impl<double, true>(M, M, K, alpha, A, a_s0, a_s1, a_sb, conj_a, B, b_s0, b_s1,
                                  b_sb, conj_b, beta, C, c_s0, c_s1, c_sb, batch, stream);
}
