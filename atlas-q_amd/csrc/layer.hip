// Device-side planning of a batched two-qubit sub-layer, so that a whole
// circuit can run (and be captured in a HIP graph) without host round trips.
//
// The host plans every buffer, bucket and launch shape once, from upper
// bounds of the bond dimensions (a gate's new bond is at most
// min(cap, 2 chi_left, 2 chi_right, r chi_middle), r the gate's operator
// Schmidt rank, propagated through the circuit); the
// bond dimensions themselves live on the device, and these kernels write the
// size-dependent fields of the grouped launches' descriptors from them:
//
//   layer_pre_kernel    theta and Gram descriptors of every gate, and the
//                       order of each Gram matrix (the rank rule's length)
//   layer_pad_kernel    padding of the tridiagonal buckets' rows past each
//                       matrix's order (zeros, distinct negative diagonal)
//   layer_post_kernel   from the rank results: projection and rebuild
//                       descriptors, the new bond dimension, the kept ranks
//                       of the tridiagonal rows and the statistics logs
//   layer_1q_kernel     the single-qubit descriptors' bond dimensions
//
// Launch shapes (tile prefix sums, totals, grids) come from the bounds, so a
// problem smaller than its bound leaves some tiles or elements without work;
// the grouped kernels skip them.  The arithmetic of a gate is the one of the
// host-planned path (same kernels, descriptors and bucket conventions).
#include "aq_common.h"

namespace aq {

__global__ void layer_pre_kernel(const aq_layer_gate_t* __restrict__ gates, int count,
                                 aq_gemm_desc_t* __restrict__ theta, aq_gemm_desc_t* __restrict__ gram,
                                 int64_t* __restrict__ order) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  const aq_layer_gate_t q = gates[g];
  const int64_t dl = q.dims[q.bond], dm = q.dims[q.bond + 1], dr = q.dims[q.bond + 2];
  const int64_t m = 2 * dl, n = 2 * dr;
  aq_gemm_desc_t t;
  t.A = q.siteA;
  t.B = q.siteB;
  t.C = q.X;
  t.M = m;
  t.N = n;
  t.K = dm;
  t.a_s0 = dm;
  t.a_s1 = 1;
  t.b_s0 = n;
  t.b_s1 = 1;
  t.c_s0 = n;
  t.c_s1 = 1;
  t.conj_a = 0;
  t.conj_b = 0;
  t.kind = AQ_GROUPED_THETA;
  t.pad = 0;
  for (int i = 0; i < 32; ++i) t.gate[i] = q.gate[i];
  theta[g] = t;
  // Gram of the side the plan chose: left X X^H (order m), right X^H X (order n)
  aq_gemm_desc_t r;
  const int left = q.left != 0;
  const int64_t s = left ? m : n;
  r.A = q.X64;
  r.B = q.X64;
  r.C = q.G;
  r.M = s;
  r.N = s;
  r.K = left ? n : m;
  r.a_s0 = left ? n : 1;
  r.a_s1 = left ? 1 : n;
  r.b_s0 = left ? 1 : n;
  r.b_s1 = left ? n : 1;
  r.c_s0 = q.sz;
  r.c_s1 = 1;
  r.conj_a = left ? 0 : 1;
  r.conj_b = left ? 1 : 0;
  r.kind = AQ_GROUPED_GEMM;
  r.pad = 0;
  for (int i = 0; i < 32; ++i) r.gate[i] = 0.0;
  gram[g] = r;
  order[g] = s;
}

// One work-group per bucket row: rows / columns >= order[row] of the order-sz
// matrix become zero except the diagonal, set to -d (1 + i / sz) with d the
// largest diagonal entry of the leading block (1 when it is zero).
__global__ __launch_bounds__(256) void layer_pad_kernel(double2* __restrict__ T, int64_t sz,
                                                        const int64_t* __restrict__ row_gate,
                                                        const int64_t* __restrict__ order) {
  __shared__ double red[4];
  const int64_t b = blockIdx.x;
  const int64_t s = order[row_gate[b]];
  double2* Tb = T + b * sz * sz;
  double d = 0.0;
  for (int64_t i = threadIdx.x; i < s; i += blockDim.x) d = fmax(d, Tb[i * sz + i].x);
  for (int off = 32; off > 0; off >>= 1) d = fmax(d, __shfl_xor(d, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d;
  __syncthreads();
  d = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  if (!(d > 0.0)) d = 1.0;
  for (int64_t e = threadIdx.x; e < sz * sz; e += blockDim.x) {
    const int64_t r = e / sz, c = e % sz;
    if (r < s && c < s) continue;
    Tb[e] = r == c ? make_double2(-d * (1.0 + (double)r / (double)sz), 0.0)
                   : make_double2(0.0, 0.0);
  }
}

__global__ void layer_post_kernel(const aq_layer_gate_t* __restrict__ gates, int count,
                                  const aq_rank_result_t* __restrict__ rr,
                                  aq_gemm_desc_t* __restrict__ proj, aq_copy_desc_t* __restrict__ cp,
                                  int32_t* __restrict__ kvec, int64_t* __restrict__ log_k,
                                  double* __restrict__ log_eps, double* __restrict__ log_ent,
                                  int64_t* __restrict__ overflow) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  const aq_layer_gate_t q = gates[g];
  const int64_t dl = q.dims[q.bond], dr = q.dims[q.bond + 2];
  const int64_t m = 2 * dl, n = 2 * dr;
  int64_t k = rr[g].k;
  if (k > q.kbound) {
    // a rank past the planned bound would overrun the bound-sized buffers:
    // keep the bound and count the event (the caller reports it)
    k = q.kbound;
    atomicAdd((unsigned long long*)overflow, 1ULL);
  }
  const int left = q.left != 0;
  // eigenvector accessor V[r, c] = (conj) V[r vs0 + c vs1]
  const int64_t vs0 = q.vfmt ? 1 : q.sz, vs1 = q.vfmt ? q.sz : 1;
  const int vcj = q.vfmt ? 1 : 0;
  aq_gemm_desc_t p;
  p.C = q.P;
  p.kind = AQ_GROUPED_GEMM;
  p.pad = 0;
  for (int i = 0; i < 32; ++i) p.gate[i] = 0.0;
  if (left) {  // Vh_k = U_k^H X: A(i, r) = conj V[r, i]
    p.A = q.V;
    p.B = q.X64;
    p.M = k;
    p.N = n;
    p.K = m;
    p.a_s0 = vs1;
    p.a_s1 = vs0;
    p.conj_a = 1 - vcj;
    p.b_s0 = n;
    p.b_s1 = 1;
    p.conj_b = 0;
    p.c_s0 = n;
  } else {     // U_k S_k = X V_k: B(r, c) = V[r, c]
    p.A = q.X64;
    p.B = q.V;
    p.M = m;
    p.N = k;
    p.K = n;
    p.a_s0 = n;
    p.a_s1 = 1;
    p.conj_a = 0;
    p.b_s0 = vs0;
    p.b_s1 = vs1;
    p.conj_b = vcj;
    p.c_s0 = k;
  }
  p.c_s1 = 1;
  proj[g] = p;
  aq_copy_desc_t a, b;
  a.dst = (void*)q.siteA;
  a.rows = m;
  a.cols = k;
  a.row_scale = nullptr;
  a.pad = 0;
  b.dst = (void*)q.siteB;
  b.rows = k;
  b.cols = n;
  b.col_scale = nullptr;
  b.pad = 0;
  if (left) {  // A' = V[:, :k] S, B' = P / S
    a.src = q.V;
    a.s0 = vs0;
    a.s1 = vs1;
    a.conj = vcj;
    a.col_scale = q.S64;
    b.src = q.P;
    b.s0 = n;
    b.s1 = 1;
    b.conj = 0;
    b.row_scale = q.Sinv;
  } else {     // A' = P (= X V_k), B' = V_k^H
    a.src = q.P;
    a.s0 = k;
    a.s1 = 1;
    a.conj = 0;
    a.col_scale = nullptr;
    b.src = q.V;
    b.s0 = vs1;
    b.s1 = vs0;
    b.conj = 1 - vcj;
    b.row_scale = nullptr;
  }
  cp[g] = a;
  cp[count + g] = b;
  if (q.kslot >= 0) kvec[q.kslot] = (int32_t)k;
  q.dims[q.bond + 1] = k;
  if (log_k) {
    log_k[g] = k;
    log_eps[g] = rr[g].eps_local;
    log_ent[g] = rr[g].entropy;
  }
}

__global__ void layer_1q_kernel(aq_site1q_desc_t* __restrict__ d, int64_t* const* __restrict__ dims,
                                const int64_t* __restrict__ site, int count) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  d[g].chiL = dims[g][site[g]];
  d[g].chiR = dims[g][site[g] + 1];
}

}  // namespace aq

using namespace aq;

extern "C" int aq_layer_pre(const aq_layer_gate_t* gates, int count, aq_gemm_desc_t* theta,
                            aq_gemm_desc_t* gram, int64_t* order, hipStream_t stream) {
  AQ_REQUIRE(gates && theta && gram && order && count > 0, AQ_ERR_ARG, "aq_layer_pre: empty");
  layer_pre_kernel<<<(unsigned)cdiv(count, 128), 128, 0, stream>>>(gates, count, theta, gram,
                                                                   order);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_layer_pad(void* T, int64_t sz, int64_t rows, const int64_t* row_gate,
                            const int64_t* order, hipStream_t stream) {
  AQ_REQUIRE(T && row_gate && order && sz >= 1 && rows >= 1 && rows < (1LL << 31), AQ_ERR_ARG,
             "aq_layer_pad: bad arguments");
  layer_pad_kernel<<<(unsigned)rows, 256, 0, stream>>>((double2*)T, sz, row_gate, order);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_layer_post(const aq_layer_gate_t* gates, int count, const aq_rank_result_t* rr,
                             aq_gemm_desc_t* proj, aq_copy_desc_t* copies, int32_t* kvec,
                             int64_t* log_k, double* log_eps, double* log_ent, int64_t* overflow,
                             hipStream_t stream) {
  AQ_REQUIRE(gates && rr && proj && copies && overflow && count > 0, AQ_ERR_ARG,
             "aq_layer_post: empty");
  AQ_REQUIRE((log_k && log_eps && log_ent) || (!log_k && !log_eps && !log_ent), AQ_ERR_ARG,
             "aq_layer_post: logs all or none");
  layer_post_kernel<<<(unsigned)cdiv(count, 128), 128, 0, stream>>>(
      gates, count, rr, proj, copies, kvec, log_k, log_eps, log_ent, overflow);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int aq_layer_1q(aq_site1q_desc_t* descs, int64_t* const* dims, const int64_t* site,
                           int count, hipStream_t stream) {
  AQ_REQUIRE(descs && dims && site && count > 0, AQ_ERR_ARG, "aq_layer_1q: empty");
  layer_1q_kernel<<<(unsigned)cdiv(count, 128), 128, 0, stream>>>(descs, dims, site, count);
  AQ_LAUNCH_CHECK();
  return AQ_OK;
}

extern "C" int64_t aq_layer_gate_size(void) { return (int64_t)sizeof(aq_layer_gate_t); }
