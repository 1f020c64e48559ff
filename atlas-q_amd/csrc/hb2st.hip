// Second stage of the two-stage tridiagonalisation (FP64): a Hermitian band
// matrix of bandwidth 32 to real tridiagonal form by bulge chasing, one
// work-group per matrix.
//
// Band storage: AB[c][d] = B(c + d, c), d = 0 .. 63 (lower triangle; entries
// with d > 32 hold the bulges while they are chased).
//
// Sweep i (i = 0 .. n-2) annihilates column i below its subdiagonal and chases
// the bulge this creates down the band.  Its task s acts on the 32 indices
// r0 = i + 1 + 32 s .. r0 + L - 1 (L = min(32, n - r0)):
//   s = 0:  the reflector H = I - tau v v^H of LAPACK's zlarfg for
//           x = B(r0 : r0 + L, i) (H^H x = beta e_1, beta real);
//   s >= 1: the block Bk = B(r0 : r0+L, r0-32 : r0) first takes the previous
//           task's reflector from the right (Bk <- Bk H'), its first column
//           gives the new reflector, and the rest of the block takes it from
//           the left (Bk(:, 1:) <- H^H Bk(:, 1:));
//   then the diagonal block D = B(r0 : r0+L, r0 : r0+L) <- H^H D H
//   (D - w v^H - v w^H, w = tau D v - 1/2 |tau|^2 (v^H D v) v).
// Only the first column of each bulge is annihilated; what remains of it is
// removed by the following sweeps (Lang's scheme for one-column sweeps).
//
// Schedule: sweep i runs on wave i mod 8 of the work-group; its task s
// starts once sweep i - 1 has finished its task s + 1 (progress counters in
// LDS, below), which orders every pair of tasks that touch common entries
// as the sequential sweep order does.
//
// Each task's v (32 entries, v[0] = 1) and tau are logged for the
// back-transformation: Vlog[i][s][32], Tlog[i][s].
//
// Wave layout of a 32 x 32 block: lane (q, h) = (lane % 32, lane / 32) holds
// column q, rows 16 h .. 16 h + 15.  Column dot products finish with one
// shuffle; row sums use a reduce-scatter butterfly over the 32 lanes of a
// half (after which lane q holds row 16 h + q / 2).
#include <math.h>

#include "aq_common.h"

namespace aq {

constexpr int kBW = 32;            // bandwidth
constexpr int kABld = 2 * kBW;     // band array entries per column
constexpr int kChaseWaves = 8;   // sweeps in flight per matrix

__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 shx(double2 a, int m) {
  return make_double2(__shfl_xor(a.x, m), __shfl_xor(a.y, m));
}
__device__ __forceinline__ double2 shl(double2 a, int src) {
  return make_double2(__shfl(a.x, src), __shfl(a.y, src));
}
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Row sums of the products x[e] vq of a 32 x 32 block held in column
// layout (x[e] = the lane's element in row 16 h + e).  Returns the sum for
// row 16 h + q / 2.  The products are formed inside the first exchange so
// that they never occupy registers of their own.
__device__ __forceinline__ double2 row_sum_scatter(const double2 (&x)[16], double2 vq, int q) {
  double2 b8[8], b4[4], b2[2];
  const bool k4 = q & 16, k3 = q & 8, k2 = q & 4, k1 = q & 2;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const double2 lo = cmul(x[e], vq), hi = cmul(x[e + 8], vq);
    const double2 keep = k4 ? hi : lo, send = k4 ? lo : hi;
    b8[e] = cadd(keep, shx(send, 16));
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double2 keep = k3 ? b8[e + 4] : b8[e], send = k3 ? b8[e] : b8[e + 4];
    b4[e] = cadd(keep, shx(send, 8));
  }
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const double2 keep = k2 ? b4[e + 2] : b4[e], send = k2 ? b4[e] : b4[e + 2];
    b2[e] = cadd(keep, shx(send, 4));
  }
  const double2 keep = k1 ? b2[1] : b2[0], send = k1 ? b2[0] : b2[1];
  double2 r = cadd(keep, shx(send, 2));
  return cadd(r, shx(r, 1));
}

__device__ __forceinline__ int sweep_tasks(int n, int i) { return (int)cdiv(n - 1 - i, kBW); }

struct ChaseLds {
  double2 v[kChaseWaves][2][kBW];   // current / previous reflector of each wave's sweep
  double2 y[kChaseWaves][kBW];
  double2 cd[kChaseWaves][kBW];
  double2 tau[kChaseWaves][2];
};

// One task.  par: the slot of this task's reflector (the previous one is in
// slot par ^ 1).
__device__ void chase_task(double2* __restrict__ AB, int n, int i, int s, int par,
                           double2* __restrict__ Vlog, double2* __restrict__ Tlog, int smax,
                           ChaseLds& sh, int w) {
  const int lane = threadIdx.x & 63, q = lane & 31, h = lane >> 5;
  const int r0 = i + 1 + kBW * s, L = min(kBW, n - r0);
  const double2 zero = make_double2(0.0, 0.0);
  double2 x[16];
  // ---- the block and its first column
  const int p0 = r0 - kBW;
  if (s == 0) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int p = 16 * h + e;
      x[e] = (q == 0 && p < L) ? AB[(int64_t)i * kABld + 1 + p] : zero;
    }
  } else {
    const double2* col = AB + (int64_t)(p0 + q) * kABld + kBW - q;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int p = 16 * h + e;
      x[e] = p < L ? col[p] : zero;
    }
    // Bk <- Bk H' : y'_p = sum_q Bk(p, q) v'_q, Bk(p, q) -= tau' y'_p conj(v'_q)
    const double2 vq = sh.v[w][par ^ 1][q];
    const double2 ys = row_sum_scatter(x, vq, q);
    if ((q & 1) == 0) sh.y[w][16 * h + (q >> 1)] = ys;
    wave_sync();
    const double2 tp = sh.tau[w][par ^ 1];
    const double2 cvq = cmul(tp, zconj(vq));  // tau' conj(v'_q)
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const double2 yp = sh.y[w][16 * h + e];
      x[e].x -= yp.x * cvq.x - yp.y * cvq.y;
      x[e].y -= yp.x * cvq.y + yp.y * cvq.x;
      if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
  // ---- reflector from column 0 (lanes 0 and 32)
  double part = 0.0;
  if (q == 0) {
#pragma unroll
    for (int e = 0; e < 16; ++e)
      if (16 * h + e >= 1) part += x[e].x * x[e].x + x[e].y * x[e].y;
  }
  const double xn2 = __shfl(part, 0) + __shfl(part, 32);
  const double2 alpha = shl(x[0], 0);
  double2 tau = zero, scale = make_double2(1.0, 0.0);
  double beta = alpha.x;
  if (!(xn2 == 0.0 && alpha.y == 0.0)) {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    const double dr = alpha.x - beta, di = alpha.y, den = dr * dr + di * di;
    scale = make_double2(dr / den, -di / den);
  }
  double2* vlog = Vlog + ((int64_t)i * smax + s) * kBW;
  if (q == 0) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int p = 16 * h + e;
      const double2 v = p == 0 ? make_double2(1.0, 0.0) : (p < L ? cmul(x[e], scale) : zero);
      sh.v[w][par][p] = v;
      vlog[p] = v;
      x[e] = p == 0 ? make_double2(beta, 0.0) : zero;
    }
  }
  if (lane == 0) {
    sh.tau[w][par] = tau;
    Tlog[(int64_t)i * smax + s] = tau;
  }
  wave_sync();
  if (s == 0) {
    if (q == 0) {
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int p = 16 * h + e;
        if (p < L) AB[(int64_t)i * kABld + 1 + p] = x[e];
      }
    }
  } else {
    // Bk(:, q) <- Bk(:, q) - conj(tau) v (v^H Bk(:, q)), q >= 1
    double2 u = zero;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const double2 v = sh.v[w][par][16 * h + e];  // conj(v) x
      u.x += v.x * x[e].x + v.y * x[e].y;
      u.y += v.x * x[e].y - v.y * x[e].x;
      if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    u = cadd(u, shx(u, 32));
    if (q >= 1) {
      const double2 cu = cmul(zconj(tau), u);
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const double2 v = sh.v[w][par][16 * h + e];
        x[e].x -= v.x * cu.x - v.y * cu.y;
        x[e].y -= v.x * cu.y + v.y * cu.x;
        if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
      }
    }
    double2* col = AB + (int64_t)(p0 + q) * kABld + kBW - q;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int p = 16 * h + e;
      if (p < L) col[p] = x[e];
    }
  }
  // ---- diagonal block: D <- H^H D H on its lower triangle
  {
    double2* col = AB + (int64_t)(r0 + q) * kABld - q;  // col[p] = D(p, q), p >= q
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int p = 16 * h + e;
      x[e] = (p >= q && p < L && q < L) ? col[p] : zero;
      if (p == q) x[e].y = 0.0;
    }
    // y = D v: row sums of L(p, q) v_q (q <= p) plus column dots
    // sum_{p' > q} conj(L(p', q)) v_p' (the strictly upper part, by symmetry)
    const double2 vq = sh.v[w][par][q];
    double2 c = zero;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int p = 16 * h + e;
      if (p > q) {
        const double2 vp = sh.v[w][par][p];  // conj(L) v
        c.x += x[e].x * vp.x + x[e].y * vp.y;
        c.y += x[e].x * vp.y - x[e].y * vp.x;
      }
      if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    const double2 ys = row_sum_scatter(x, vq, q);
    c = cadd(c, shx(c, 32));
    if ((q & 1) == 0) sh.y[w][16 * h + (q >> 1)] = ys;
    if (h == 0) sh.cd[w][q] = c;
    wave_sync();
    // rho = v^H y (real), w = tau y - 1/2 |tau|^2 rho v
    const double2 yq = cadd(sh.y[w][q], sh.cd[w][q]);
    double rho = vq.x * yq.x + vq.y * yq.y;  // Re(conj(v_q) y_q), lanes q of both halves
    for (int m = 16; m > 0; m >>= 1) rho += __shfl_xor(rho, m);
    const double hk = -0.5 * (tau.x * tau.x + tau.y * tau.y) * rho;
    const double2 wq = make_double2(tau.x * yq.x - tau.y * yq.y + hk * vq.x,
                                    tau.x * yq.y + tau.y * yq.x + hk * vq.y);
    wave_sync();
    if (h == 0) sh.cd[w][q] = wq;  // w replaces the column dots
    wave_sync();
    const double2 cwq = zconj(wq), cvq = zconj(vq);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int p = 16 * h + e;
      if (p >= q && p < L && q < L) {
        const double2 vp = sh.v[w][par][p], wp = sh.cd[w][p];
        const double2 t1 = cmul(vp, cwq), t2 = cmul(wp, cvq);
        double2 z = make_double2(x[e].x - t1.x - t2.x, x[e].y - t1.y - t2.y);
        if (p == q) z.y = 0.0;
        col[p] = z;
      }
      if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Dataflow between the waves: wave w runs sweeps w, w + 8, ... and
// publishes in pub[w] the position 128 i + (tasks done) of its sweep i (128 i
// + 128 once the sweep is complete).  Task s of sweep i starts when sweep
// i - 1 (on the previous wave) has completed its tasks 0 .. s+1, i.e. pub >=
// 128 (i - 1) + s + 2; the tasks then running on other waves touch disjoint
// entries (sweep i - 1 is at least at task s + 2, sweep i - 2 at s + 4, ...).
// Eight waves of 178 registers each keep the two 32 x 32 blocks of a task in
// registers without spills (sixteen waves at 128 registers spilled them,
// which made every task a chain of scratch round trips).
// A wait that exceeds kChaseSpinLimit polls (a bug, never a slow schedule)
// marks the matrix's result invalid (D[0] = NaN) and stops the sweep.
constexpr int kSweepKey = 128;
constexpr int kChaseSpinLimit = 1 << 22;

// Waits until *p >= need (false after kChaseSpinLimit polls).
__device__ __forceinline__ bool wait_at_least(int* p, int need) {
  for (int polls = 0; polls < kChaseSpinLimit; ++polls) {
    if (__hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >= need) return true;
    __builtin_amdgcn_s_sleep(1);
  }
  return false;
}

__global__ __launch_bounds__(kChaseWaves * 64) void bulge_chase_kernel(
    double2* __restrict__ AB, int64_t sAB, int n, double2* __restrict__ Vlog,
    double2* __restrict__ Tlog, int64_t sLog, int smax, double* __restrict__ D,
    double* __restrict__ E) {
  __shared__ ChaseLds sh;
  __shared__ int pub[kChaseWaves];
  __shared__ int failed;
  const int64_t b = blockIdx.x;
  double2* ABb = AB + b * sAB;
  double2* Vb = Vlog + b * sLog * kBW;
  double2* Tb = Tlog + b * sLog;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x < kChaseWaves) pub[threadIdx.x] = -1;
  if (threadIdx.x == 0) failed = 0;
  __syncthreads();
  const int wprev = (w + kChaseWaves - 1) % kChaseWaves;
  bool ok = true;
  int i = w, s = 0, ns = w <= n - 2 ? sweep_tasks(n, w) : 0;
  while (i <= n - 2) {
    if (i > 0 && !wait_at_least(&pub[wprev], kSweepKey * (i - 1) + s + 2)) {
      ok = false;
      break;
    }
    chase_task(ABb, n, i, s, s & 1, Vb, Tb, smax, sh, w);
    const int done = s + 1 == ns ? kSweepKey * (i + 1) : kSweepKey * i + s + 1;
    if (lane == 0)
      __hip_atomic_store(&pub[w], done, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (++s == ns) {
      i += kChaseWaves;
      s = 0;
      ns = i <= n - 2 ? sweep_tasks(n, i) : 0;
    }
  }
  if (!ok && lane == 0) {
    failed = 1;
    // unblock the waves that wait on this one
    __hip_atomic_store(&pub[w], 1 << 30, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    D[b * n + c] = ABb[(int64_t)c * kABld].x;
    if (c < n - 1) E[b * (n - 1) + c] = ABb[(int64_t)c * kABld + 1].x;
  }
  __syncthreads();
  if (threadIdx.x == 0 && failed) D[b * n] = __builtin_nan("");
}

}  // namespace aq

using namespace aq;

extern "C" int64_t aq_hb2st_log_tasks(int64_t n) {
  // per matrix: (n - 1) sweeps x ceil((n - 1) / 32) task slots
  return n >= 2 ? (n - 1) * cdiv(n - 1, kBW) : 1;
}

extern "C" int aq_hb2st_own(int64_t n, void* ABv, void* Vlogv, void* Tlogv, double* D, double* E,
                            int64_t batch, hipStream_t stream) {
  AQ_REQUIRE(ABv && Vlogv && Tlogv && D && E, AQ_ERR_ARG, "aq_hb2st_own: null pointer");
  AQ_REQUIRE(n >= 2 && n <= 4033, AQ_ERR_ARG, "aq_hb2st_own: order %lld out of range (2 .. 4033)",
             (long long)n);
  AQ_REQUIRE(batch >= 1 && batch <= 65535, AQ_ERR_ARG, "aq_hb2st_own: bad batch");
  const int smax = (int)cdiv(n - 1, kBW);
  bulge_chase_kernel<<<(unsigned)batch, kChaseWaves * 64, 0, stream>>>(
      (double2*)ABv, n * kABld, (int)n, (double2*)Vlogv, (double2*)Tlogv,
      aq_hb2st_log_tasks(n), smax, D, E);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}
