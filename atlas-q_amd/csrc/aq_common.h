// Shared definitions for the atlas-q MI355X (gfx950) kernels.
//
// Complex data is interleaved (re, im) as float2 / double2, the layout torch
// uses for complex64 / complex128 tensors.  Every entry point is asynchronous
// on the caller's hipStream_t, returns 0 on success and a negative status on
// failure, and leaves a message in a thread-local buffer (aq_last_error).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/atlasq.h"

namespace aq {

void set_error(const char* fmt, ...);

#define AQ_REQUIRE(cond, code, ...)        \
  do {                                     \
    if (!(cond)) {                         \
      ::aq::set_error(__VA_ARGS__);        \
      return (code);                       \
    }                                      \
  } while (0)

#define AQ_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::aq::set_error("%s failed: %s", #call, hipGetErrorString(e_));         \
      return AQ_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

#define AQ_LAUNCH_CHECK()                                                     \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      ::aq::set_error("kernel launch failed: %s", hipGetErrorString(e_));     \
      return AQ_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

// complex helpers --------------------------------------------------------
template <typename T> struct C2;
template <> struct C2<float> { using type = float2; };
template <> struct C2<double> { using type = double2; };
template <typename T> using c2_t = typename C2<T>::type;

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename V>
__device__ __forceinline__ V cadd(V a, V b) {
  V r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace aq
