// Batched FP64 Hermitian eigen-decomposition for the Gram cores of the
// truncated SVD: rocSOLVER's strided-batched divide and conquer (zheevd).
//
// One call factors every Gram matrix of a brickwork sub-layer.  rocSOLVER's
// tridiagonalisation is a long chain of small, latency-bound kernels; in the
// strided-batched form each of those kernels covers the whole batch, so the
// chain is paid once per sub-layer instead of once per bond.
//
// rocSOLVER is column-major.  A row-major Hermitian matrix G read column-major
// is G^T = conj(G); its eigenvectors are conj(v) for the eigenvectors v of G,
// and written column-major they read row-major as V^H.  The entry point takes
// and returns the row-major view: on exit row j of matrix b holds the complex
// conjugate of the j-th eigenvector (ascending eigenvalue order), i.e. the
// matrix b is V_b^H.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "aq_common.h"

namespace {

// One rocBLAS handle per host thread (rocBLAS handles are not thread-safe;
// the library's callers may factor from several threads, each on its own
// stream).  The stream is set on every call.
struct Handle {
  rocblas_handle h = nullptr;
  ~Handle() {
    if (h) rocblas_destroy_handle(h);
  }
};
thread_local Handle t_handle;

}  // namespace

extern "C" int aq_heevd_batched(int64_t n, void* A, int64_t strideA, double* W, int64_t strideW,
                                double* E, int32_t* info, int64_t batch, hipStream_t stream) {
  AQ_REQUIRE(A && W && E && info, AQ_ERR_ARG, "aq_heevd_batched: null pointer");
  AQ_REQUIRE(n >= 1 && n <= (1 << 15), AQ_ERR_ARG, "aq_heevd_batched: n out of range (%lld)",
             (long long)n);
  AQ_REQUIRE(batch >= 1 && batch <= (1 << 20), AQ_ERR_ARG, "aq_heevd_batched: bad batch");
  AQ_REQUIRE(strideA >= n * n && strideW >= n, AQ_ERR_ARG,
             "aq_heevd_batched: strides smaller than one matrix / one spectrum");
  if (!t_handle.h) {
    rocblas_status s = rocblas_create_handle(&t_handle.h);
    AQ_REQUIRE(s == rocblas_status_success, AQ_ERR_HIP,
               "aq_heevd_batched: rocblas_create_handle failed (%d)", (int)s);
  }
  rocblas_status s = rocblas_set_stream(t_handle.h, stream);
  AQ_REQUIRE(s == rocblas_status_success, AQ_ERR_HIP,
             "aq_heevd_batched: rocblas_set_stream failed (%d)", (int)s);
  s = rocsolver_zheevd_strided_batched(
      t_handle.h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n,
      reinterpret_cast<rocblas_double_complex*>(A), (rocblas_int)n, (rocblas_stride)strideA, W,
      (rocblas_stride)strideW, E, (rocblas_stride)n, reinterpret_cast<rocblas_int*>(info),
      (rocblas_int)batch);
  AQ_REQUIRE(s == rocblas_status_success, AQ_ERR_HIP,
             "aq_heevd_batched: rocsolver_zheevd_strided_batched failed (%d)", (int)s);
  return AQ_OK;
}
