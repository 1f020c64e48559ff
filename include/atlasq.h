/* atlasq.h -- C ABI of the MI355X (gfx950) adaptive-MPS hot path.
 *
 * The reference (ATLAS-Q) has no native ABI on this path: it is Python over
 * torch.  Each entry point below replaces one reference function or inner
 * step; the citation names it (paths relative to the reference checkout).
 *
 * Conventions
 *   - Complex data is interleaved (re, im): complex64 = float2, complex128 =
 *     double2, the torch layout.  Sizes are int64_t element counts.
 *   - The caller allocates every input, output and workspace (torch owns
 *     device memory); the library never frees or keeps caller pointers.
 *   - Every call is asynchronous on `stream`; no call synchronises the
 *     device.  Returns AQ_OK (0) or a negative AQ_ERR_* status; the message
 *     is in aq_last_error() (thread-local).
 *   - Site tensors are row-major [chiL, 2, chiR].  Theta X is row-major
 *     [2*chiL, 2*chiR] with X[2l+p, q*chiR+j].
 */
#ifndef ATLASQ_H
#define ATLASQ_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AQ_OK 0
#define AQ_ERR_ARG (-1)
#define AQ_ERR_HIP (-2)
#define AQ_ERR_UNSUPPORTED (-3)

/* arithmetic of the contractions */
#define AQ_PREC_NATIVE 0 /* FP32 MFMA for complex64, FP64 MFMA for complex128 */
#define AQ_PREC_BF16X3 1 /* complex64 only: hi/lo BF16 split, 3 MFMAs, FP32 acc */

/* dtype codes */
#define AQ_C64 0
#define AQ_C128 1

/* rank-rule arithmetic */
#define AQ_RANK_REF 0 /* reference CPU semantics in S's dtype (see aq_rank_select) */
#define AQ_RANK_TDVP 1 /* the same with threshold (1 - eps) * total instead of
                          (1 - eps^2) * total: the split of TDVP2Site._evolve_two_site
                          (src/atlas_q/tdvp.py:546-562) */

int aq_version(void);
const char* aq_last_error(void);
/* sizeof of aq_gemm_desc_t, aq_site1q_desc_t, aq_sample_desc_t, aq_copy_desc_t,
 * aq_rank_result_t, aq_tn_trunc_t, aq_dense_layer_t (7 int64), so bindings
 * can check their struct layouts. */
int aq_desc_sizes(int64_t* out);

/* Theta: X = gate-applied two-site contraction, "U" convention
 *   X[2l+p, q*chiR+j] = sum_{p1,p2} U[2p+q, 2p1+p2] sum_m A[l,p1,m] B[m,p2,j].
 * Replaces fused_two_qubit_gate_pytorch (src/triton_kernels/mps_complex.py:343-368)
 * and fused_two_qubit_gate_triton (mps_complex.py:221-336), as called from
 * AdaptiveMPS.apply_two_site_gate (src/atlas_q/adaptive_mps.py:225-247).
 * The gate is passed by value from the host: gate_host = 32 doubles,
 * U[r][c] = (gate_host[2*(4r+c)], gate_host[2*(4r+c)+1]).
 * A [chiL,2,chiM], B [chiM,2,chiR], X [2chiL,2chiR], all contiguous, dtype by
 * `dtype` (AQ_C64 / AQ_C128).  prec: AQ_PREC_NATIVE or AQ_PREC_BF16X3 (c64). */
int aq_theta(int dtype, int prec, const void* A, const void* B, const double* gate_host,
             void* X, int64_t chiL, int64_t chiM, int64_t chiR, hipStream_t stream);

/* aq_theta for complex64 in BF16x3 with the operands split once, before the
 * GEMM, into four BF16 planes (re hi / lo, im hi / lo) in caller workspace
 * `ws` (16-byte aligned, >= aq_theta_planes_bytes(chiL, chiM, chiR) bytes),
 * k-contiguous and zero-padded to whole tiles, so the MFMA mainloop only
 * copies them into LDS.  Same contraction, gate and result as aq_theta with
 * AQ_PREC_BF16X3 (replaces the same reference functions:
 * src/triton_kernels/mps_complex.py:343-368, as called at
 * src/atlas_q/adaptive_mps.py:225-247); the planes are valid after the call
 * and overwritten by the next one. */
int64_t aq_theta_planes_bytes(int64_t chiL, int64_t chiM, int64_t chiR);
int aq_theta_planes(const void* A, const void* B, const double* gate_host, void* X,
                    int64_t chiL, int64_t chiM, int64_t chiR, void* ws, int64_t ws_bytes,
                    hipStream_t stream);

/* Rank rule: replaces choose_rank_from_sigma (src/atlas_q/truncation.py:21-83)
 * as called at adaptive_mps.py:266/273.  S: m singular values (float32 when
 * s_dtype == AQ_C64, float64 when AQ_C128), descending.  cap: per-bond cap
 * already reduced by the memory budget (adaptive_mps.py:255-264 is evaluated on
 * the host, where the tensor sizes live).  Writes one aq_rank_result_t to
 * device memory `out`.  AQ_RANK_REF follows the reference's CPU rounding:
 * E = cumsum(S*S) accumulated sequentially in double and rounded to S's dtype,
 * thresh = (1-eps^2 rounded to S's dtype) * total in S's dtype, left
 * searchsorted, k = min(idx+1, m, cap). */
typedef struct {
  int32_t k;
  int32_t flags; /* bit0: total < 1e-30 (degenerate) */
  double eps_local;
  double entropy; /* nats, over the kept values */
  double condS;   /* S[0]/S[k-1], +inf when k <= 1 or S[k-1] == 0 */
  double total_weight;
  double kept_weight;
} aq_rank_result_t;
int aq_rank_select(int s_dtype, const void* S, int64_t m, double eps, int64_t cap, int mode,
                   aq_rank_result_t* out, hipStream_t stream);

/* Core rebuild (adaptive_mps.py:280-290): A' = (U[:, :k] * S[:k]) as
 * [rows/2, 2, k], B' = Vh[:k, :] as [k, 2, cols/2].  U is rows x >=k with
 * element (r,c) at U[r*u_s0 + c*u_s1]; Vh is >=k x cols with (r,c) at
 * Vh[r*v_s0 + c*v_s1] (element strides; torch's SVD factors may be
 * column-major).  S real (float / double per dtype). */
int aq_rebuild(int dtype, const void* U, int64_t u_s0, int64_t u_s1, const void* S,
               const void* Vh, int64_t v_s0, int64_t v_s1, int64_t rows, int64_t cols,
               int64_t k, void* A_out, void* B_out, hipStream_t stream);

/* Single-qubit gate, AdaptiveMPS convention (adaptive_mps.py:159-180):
 * T_out[a,t,b] = sum_s U[s,t] T_in[a,s,b].  gate_host: 8 doubles (U[r][c]
 * interleaved re/im, row-major).  T_in, T_out contiguous [chiL,2,chiR]. */
int aq_apply_1q(int dtype, const void* T_in, void* T_out, int64_t chiL, int64_t chiR,
                const double* gate_host, hipStream_t stream);

/* General complex GEMM  C = op(A) op(B)  on MFMA, element strides.
 * A: M x K, element (m,k) at A[m*a_s0 + k*a_s1], conjugated when conj_a.
 * B: K x N, element (k,n) at B[k*b_s0 + n*b_s1], conjugated when conj_b.
 * C: M x N, element (m,n) at C[m*c_s0 + n*c_s1] (overwritten).
 * Used for the canonical-form absorptions (adaptive_mps.py:333,370,385), the
 * state-vector contraction (adaptive_mps.py:463-481) and the projections of
 * the truncated-SVD variants. */
int aq_cgemm(int dtype, int prec, int64_t M, int64_t N, int64_t K,
             const void* A, int64_t a_s0, int64_t a_s1, int conj_a,
             const void* B, int64_t b_s0, int64_t b_s1, int conj_b,
             void* C, int64_t c_s0, int64_t c_s1, hipStream_t stream);

/* aq_cgemm for outputs too small to fill the GPU over a long reduction
 * (sketches X Omega, Cholesky-QR and core Gram products Y^H Y of the
 * truncated-SVD variants; no reference counterpart beyond the torch matmuls
 * of the plain path).  The reduction is cut into `splits` slices computed by
 * separate work-groups into `workspace` (>= splits*M*N complex elements of
 * dtype) and added in slice order, so the result is deterministic. */
int aq_cgemm_splitk(int dtype, int prec, int64_t M, int64_t N, int64_t K, const void* A,
                    int64_t a_s0, int64_t a_s1, int conj_a, const void* B, int64_t b_s0,
                    int64_t b_s1, int conj_b, void* C, int64_t c_s0, int64_t c_s1,
                    int64_t splits, void* workspace, int64_t workspace_bytes,
                    hipStream_t stream);

/* Batched sweep sampler: replaces MatrixProductStatePyTorch._batch_sweep_sample
 * (src/atlas_q/mps_pytorch.py:247-355) with its multinomial draw expressed as
 * the inverse CDF  outcome = (u >= p0)  over caller-supplied uniforms.
 * sites: host array of n device pointers, site i is [chi[i], 2, chi[i+1]]
 * (chi: host array of n+1 bond sizes, chi[0] = chi[n] = 1), all of `dtype`.
 * uniforms: device float64 [shots, n].  bits: device uint64
 * [shots, ceil(n/64)]: qubit i is bit (n-1-i) of the little-endian word
 * array (qubit 0 = most significant, as mps_pytorch.py:321; word 0 is the
 * reference's int64 value modulo 2^64).  p0_out (nullable): device float64
 * [shots, n], the normalised probability of outcome 0 used at each draw.
 * workspace: device memory of aq_sample_workspace_bytes(...) bytes.
 * flags: AQ_SAMPLE_SINGLE_SHOT_RULES selects the reference's one-shot path
 * (mps_pytorch.py:177-245): first-site probabilities |sum_b A[0,s,b]|^2 and
 * Python-float normalisation without the 1e-15 floor (0.5/0.5 below 1e-15). */
#define AQ_SAMPLE_SINGLE_SHOT_RULES 1
int64_t aq_sample_workspace_bytes(int dtype, int64_t n, const int64_t* chi, int64_t shots);
int aq_sample(int dtype, int prec, int flags, int64_t n, const void* const* sites, const int64_t* chi,
              int64_t shots, const double* uniforms, uint64_t* bits, double* p0_out,
              void* workspace, int64_t workspace_bytes, hipStream_t stream);

/* aq_sample without a caller workspace (no reference counterpart: the
 * reference's sampler allocates freely).  For callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc: this entry point calls no hipMalloc,
 * hipFree or allocating library routine.  Its workspace is a fixed device
 * array of AQ_SAMPLE_SCRATCH_BYTES per GPU, held until the process ends;
 * a sample whose aq_sample_workspace_bytes(...) exceeds it is rejected with
 * AQ_ERR_ARG.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other host threads) race, so callers
 * serialise them.  Results are aq_sample's, bit for bit. */
#define AQ_SAMPLE_SCRATCH_BYTES (64ll << 20)
int aq_sample_noalloc(int dtype, int prec, int flags, int64_t n, const void* const* sites,
                      const int64_t* chi, int64_t shots, const double* uniforms, uint64_t* bits,
                      double* p0_out, hipStream_t stream);

/* Small Hermitian eigensolver (no reference counterpart: the reference calls
 * LAPACK/cuSOLVER SVD, linalg_robust.py:17-52).  Batched two-sided Jacobi in
 * FP64, one work-group per n x n matrix (n <= 64) held in LDS.  A: complex128,
 * matrix b at A + b*strideA, row-major with leading dimension lda (the input
 * is symmetrised).  W: float64 [batch, n] eigenvalues in descending order;
 * V: complex128 eigenvectors as columns, matrix b at V + b*strideV.  Stops
 * when the off-diagonal norm falls below tol*||A||_F or after max_sweeps;
 * sweeps (nullable, int[batch]) receives the sweep count. */
int aq_heev_small(int64_t n, const void* A, int64_t lda, int64_t strideA, double* W, void* V,
                  int64_t ldv, int64_t strideV, int64_t batch, int max_sweeps, double tol,
                  int* sweeps, hipStream_t stream);
/* aq_heev_small with a per-matrix order (no reference counterpart; used for
 * size buckets of Gram matrices of different orders): orders (device,
 * int64[batch]) holds matrix b's order nb <= n; only the leading nb x nb block
 * is read and diagonalised, W[b, nb:n] = 0 and the trailing rows/columns of
 * V's n x n block are those of the identity. */
int aq_heev_small_orders(int64_t n, const int64_t* orders, const void* A, int64_t lda,
                         int64_t strideA, double* W, void* V, int64_t ldv, int64_t strideV,
                         int64_t batch, int max_sweeps, double tol, int* sweeps,
                         hipStream_t stream);

/* Batched FP64 Hermitian eigen-decomposition of the truncated SVD's Gram
 * cores (no reference counterpart: the reference factors theta with a thin
 * SVD, adaptive_mps.py:249 via linalg_robust.py:17-52; the Gram method needs
 * the eigenpairs of X X^H).  rocSOLVER zheevd, strided-batched.  A: complex128
 * row-major n x n Hermitian matrices at A + b*strideA (elements), overwritten:
 * row j of matrix b becomes the conjugate of the j-th eigenvector.  W: float64
 * eigenvalues, ascending, at W + b*strideW.  E: float64 workspace [batch*n].
 * info: int32 [batch] on the device, 0 on success (> 0: not converged). */
int aq_heevd_batched(int64_t n, void* A, int64_t strideA, double* W, int64_t strideW, double* E,
                     int32_t* info, int64_t batch, hipStream_t stream);

/* One-sided (Hestenes) Jacobi SVD of small Gram cores, FP64, orders
 * n <= 64 (no reference counterpart: the factorization of the randomized
 * SVD's projected core, part of the method that replaces
 * linalg_robust.py:17-52 as called at adaptive_mps.py:249-250).  G:
 * complex128 row-major Hermitian positive semidefinite matrices (lower
 * triangle read) at G + b*strideG.  G = L L^H (Cholesky in LDS, stopped at
 * the first pivot below n eps max G_ii: the rest of L is zero), then the
 * columns of L are rotated until orthogonal, L J = U Sigma, with
 * wavefront-level reductions for the inner products.  Outputs: W[b, :]
 * = Sigma^2 descending (the eigenvalues of G), U [b, n, n] row-major with the
 * corresponding unit eigenvectors as columns (zero columns past the numerical
 * rank), err[b] = max |U^H U - I| over the non-zero columns (FP64 re-check).
 * Sweeps stop when every pair has |a_p^H a_q| <= tol ||a_p|| ||a_q||. */
int aq_gram_svd_small(int64_t n, const void* G, int64_t ldg, int64_t strideG, double* W, void* U,
                      int64_t ldu, int64_t strideU, int64_t batch, int max_sweeps, double tol,
                      double* err, hipStream_t stream);

/* One Cholesky QR pass of a tall-skinny complex128 matrix (no reference
 * counterpart: the orthonormalisation of the randomized SVD's sketch, which
 * sizes the factorization that replaces linalg_robust.py:17-52 as called at
 * adaptive_mps.py:249-250): W = Y^H Y + shift_scale trace(W) I, W = L L^H
 * (blocked by 64, diagonal blocks factored and inverted in LDS), Q = Y L^{-H}.
 * Y [m, l] row-major (row stride ldy), Q [m, l] (row stride ldq, must not
 * alias Y), 1 <= l <= min(m, 4096).  ws: complex128 workspace of
 * aq_cholqr_workspace_elems(m, l) elements.  info (device int): 0, or 1 + the
 * index of the first pivot that was not positive and finite (then Q is not
 * orthonormal; the pass still ends with finite numbers). */
int64_t aq_cholqr_workspace_elems(int64_t m, int64_t l);
int aq_cholqr(int64_t m, int64_t l, const void* Y, int64_t ldy, void* Q, int64_t ldq,
              double shift_scale, void* ws, int64_t ws_elems, int* info, hipStream_t stream);

/* Strided-batched complex GEMM with a scaled accumulate (no reference
 * counterpart: the steps of the Gram eigensolver below, which replaces the
 * factorization of adaptive_mps.py:249-250 / linalg_robust.py:17-52 in the
 * truncated-SVD methods):  C_b = alpha op(A_b) op(B_b) + beta C_b for
 * b < batch, matrix b of each operand at its pointer + b * (its batch stride),
 * element strides as aq_cgemm, real alpha / beta (beta == 0 does not read C). */
int aq_bgemm(int dtype, int64_t M, int64_t N, int64_t K, double alpha, const void* A,
             int64_t a_s0, int64_t a_s1, int64_t a_sb, int conj_a, const void* B, int64_t b_s0,
             int64_t b_s1, int64_t b_sb, int conj_b, double beta, void* C, int64_t c_s0,
             int64_t c_s1, int64_t c_sb, int64_t batch, hipStream_t stream);

/* aq_bgemm for complex128 with M == N writing only C's lower triangle (row
 * >= column; tiles strictly above it are skipped): the rank-2nb updates of
 * the tridiagonal reduction below, which read only that triangle. */
int aq_bgemm_lower(int64_t M, int64_t K, double alpha, const void* A, int64_t a_s0,
                   int64_t a_s1, int64_t a_sb, int conj_a, const void* B, int64_t b_s0,
                   int64_t b_s1, int64_t b_sb, int conj_b, double beta, void* C, int64_t c_s0,
                   int64_t c_s1, int64_t c_sb, int64_t batch, hipStream_t stream);

/* Gram eigensolver through the tridiagonal form, in FP64 (no reference
 * counterpart: the reference factors theta with a thin SVD,
 * adaptive_mps.py:249-250 via linalg_robust.py:17-52; the Gram method needs
 * the leading eigenpairs of X X^H, and only those the rank rule keeps).
 *
 * aq_hetrd_batched: complex128 row-major Hermitian n x n matrices (lower
 * triangle read) at A + b*strideA reduced to real tridiagonal form, D
 * [batch, n], E [batch, n-1], tau complex128 [batch, n-1]; the Householder
 * vectors stay in A: with Q = H_{n-2} ... H_0, H_i = I - tau_i v_i v_i^H,
 * v_i[i] = 1, v_i[0:i] = row i+1, columns 0..i-1 of A, v_i zero below i,
 * conj(G) = Q T Q^H.  (rocSOLVER zhetrd, strided batched.)
 * aq_tridiag_eigvals: the nval largest eigenvalues of each tridiagonal,
 * descending, into W (bisection on Sturm counts, absolute accuracy
 * ~ 2 eps ||T||).
 * aq_tridiag_eigvecs: for j < kvec[b] (all j < kmax when kvec is NULL), row
 * j of Zt (complex128 [batch, >= kmax, n], matrix stride sZ) = the unit
 * eigenvector of W[b, j] (twisted inverse iteration; rows kvec[b] .. kmax-1
 * zeroed, so with kvec the padded row count kmax may exceed n as long as
 * every kvec[b] <= n); kept eigenvalues of one unreduced block that chain
 * within 1e-5 ||T|| of each other are a cluster, whose members after the
 * first are recomputed by inverse iteration with Gram-Schmidt against the
 * earlier ones (LAPACK dstein's scheme); scratch:
 * aq_tridiag_eigvecs_scratch_doubles() doubles.
 * aq_hetrd_vblocks / aq_hetrd_tfactors: the vectors v_i in blocks of nb <=
 * 64, or an even nb <= 128 (two half-blocks' factors joined), (V [batch,
 * nblk, nb, n], nblk = ceil((n-1)/nb), zero rows past i = n-2)
 * and each block's forward compact-WY factor T' (nb x nb, from the block's
 * Gram Gv = conj(V) V^T [batch, nblk, nb, nb]) with
 * H_{i0}^H ... H_{i0+nb-1}^H = I - V^T T' conj(V): the back-transformation
 * Et = Zt Q^T is, block by block in increasing order,
 * Et <- Et - (Et conj(V)^T) (conj(T') V).
 * aq_orth_band_reduce: err[b] = max |Q - I| over the diagonal Gram tiles Q
 * [batch, ntile, 64, 128] (Q[t] = rows 64t.. of Et times the conjugate
 * transpose of rows 64t .. 64t+127), restricted to indices < kvec[b]. */
int aq_hetrd_batched(int64_t n, void* A, int64_t strideA, double* D, double* E, void* tau,
                     int64_t batch, hipStream_t stream);
/* Stream groups of the following aq_hetrd_own calls (process-wide, 1 or 2):
 * 1 keeps the whole reduction on the caller's stream (HIP graph capture); 2,
 * the default, overlaps two groups' column chains on worker streams.
 * Returns the previous value. */
int aq_hetrd_max_groups(int groups);
int aq_tridiag_eigvals(int64_t n, const double* D, int64_t sD, const double* E, int64_t sE,
                       double* W, int64_t sW, int64_t nval, int64_t batch, hipStream_t stream);
int64_t aq_tridiag_eigvecs_scratch_doubles(int64_t n, int64_t kmax, int64_t batch);
int aq_tridiag_eigvecs(int64_t n, const double* D, int64_t sD, const double* E, int64_t sE,
                       const double* W, int64_t sW, const int32_t* kvec, int64_t kmax, void* Zt,
                       int64_t sZ, double* scratch, int64_t scratch_doubles, int64_t batch,
                       hipStream_t stream);
int aq_hetrd_vblocks(int64_t n, int64_t nb, const void* A, int64_t strideA, void* V,
                     int64_t batch, hipStream_t stream);
int aq_hetrd_tfactors(int64_t n, int64_t nb, const void* tau, int64_t stau, const void* Gv,
                      void* Tf, int64_t batch, hipStream_t stream);
int aq_orth_band_reduce(const void* Q, int64_t ntile, const int32_t* kvec, int64_t kmax,
                        double* err, int64_t batch, hipStream_t stream);
/* The same reduction by the library's own kernels (same outputs and storage
 * as aq_hetrd_batched; LAPACK's blocked zhetrd/zlatrd with uplo "U"): panels
 * of nb <= 32 columns; per column one batched streaming Hermitian
 * matrix-vector product over the stored triangle (which also yields the
 * quadratic form and the panel dot products), one wide kernel forming the W
 * column and updating the next column, and one per-matrix reflector kernel;
 * per panel a rank-2nb update of the lower triangle on MFMA
 * (aq_bgemm_lower).  Batches of 2 or more run as two interleaved groups on
 * the library's worker streams, joined to `stream` by events.
 * ws: aq_hetrd_workspace_elems(n, nb, batch) complex128. */
int64_t aq_hetrd_workspace_elems(int64_t n, int64_t nb, int64_t batch);
int aq_hetrd_own(int64_t n, int64_t nb, void* A, int64_t strideA, double* D, double* E,
                 void* tau, void* ws, int64_t ws_elems, int64_t batch, hipStream_t stream);
/* aq_hetrd_own with the group count given per call instead of the
 * process-wide setting: groups = 1 runs the whole reduction on `stream` (no
 * worker streams, so several calls from one host thread on different
 * streams proceed independently and overlap), 2 as aq_hetrd_own. */
int aq_hetrd_own_groups(int64_t n, int64_t nb, void* A, int64_t strideA, double* D, double* E,
                        void* tau, void* ws, int64_t ws_elems, int64_t batch, int groups,
                        hipStream_t stream);
/* First stage of the two-stage tridiagonalisation (the library's own; no
 * reference counterpart, as for the reduction above): Hermitian A [batch]
 * [n][n] (row-major, lower triangle read; overwritten) = Q1 B Q1^H with B of
 * bandwidth 32, by panels of 32 Householder reflectors (LAPACK's
 * zhetrd_he2hb, lower) whose trailing updates run on the batched MFMA GEMM.
 * On return AB[b][c][d] = B(c + d, c) for d < 64 (zero for d > 32), the
 * reflectors of panel p (columns 32p .. 32p+31) lie below its R factor in A
 * (A(32p + 32 + r, 32p + c), r > c, unit diagonal implied) and its T factor
 * (Q_p = I - V T V^H) at Tall[b][p][32][32].  aq_he2hb_panels(n) panels;
 * ws: aq_he2hb_workspace_elems(n, batch) complex128; n - 32 <= 4096. */
int64_t aq_he2hb_panels(int64_t n);
int64_t aq_he2hb_workspace_elems(int64_t n, int64_t batch);
int aq_he2hb_own(int64_t n, void* A, int64_t strideA, void* AB, void* Tall, void* ws,
                 int64_t ws_elems, int64_t batch, hipStream_t stream);
/* Second stage: the band AB (as aq_he2hb_own leaves it, bandwidth 32) to real
 * tridiagonal form by bulge chasing, one work-group per matrix (sweeps of
 * one-column Householder annihilations, 8 sweeps in flight).  D [batch][n],
 * E [batch][n-1]; every reflector (v of 32 entries, v[0] = 1, and tau) of
 * sweep i, task s is logged at Vlog[b][i * ceil((n-1)/32) + s][32] and
 * Tlog[b][i * ceil((n-1)/32) + s], aq_hb2st_log_tasks(n) slots per matrix;
 * T = Q2^H B Q2 with Q2 the product of the reflectors in sweep order.
 * 2 <= n <= 4033. */
int64_t aq_hb2st_log_tasks(int64_t n);
int aq_hb2st_own(int64_t n, void* AB, void* Vlog, void* Tlog, double* D, double* E,
                 int64_t batch, hipStream_t stream);
/* Blocked Householder QR of complex128 matrices A [batch][m][n] (row-major,
 * m <= 4096), LAPACK's zgeqrf: panels of 32 columns factored by one
 * work-group each (zgeqr2 + zlarft), trailing updates on the batched GEMM.
 * Replaces the device QR behind the reference's robust_qr
 * (linalg_robust.py:55-76; torch.linalg.qr there) and the canonical forms'
 * QRs (adaptive_mps.py:310-385, mps_pytorch.py:91-152).  On return A holds R
 * in its upper triangle and the reflectors below it; Tall[b][p][32][32] the
 * T factor of panel p (aq_geqrf_panels(m, n) panels).  aq_ungqr forms the
 * thin Q [batch][m][min(m, n)] (row stride min(m, n)) from them (zungqr).
 * ws: aq_geqrf_workspace_elems(m, n, batch) complex128, for either call. */
int64_t aq_geqrf_panels(int64_t m, int64_t n);
int64_t aq_geqrf_workspace_elems(int64_t m, int64_t n, int64_t batch);
int aq_geqrf(int64_t m, int64_t n, void* A, int64_t strideA, void* Tall, void* ws,
             int64_t ws_elems, int64_t batch, hipStream_t stream);
int aq_ungqr(int64_t m, int64_t n, const void* A, int64_t strideA, const void* Tall, void* Q,
             int64_t strideQ, void* ws, int64_t ws_elems, int64_t batch, hipStream_t stream);

/* Rank-predictor MLP 32 -> 128 -> 128 -> 1 with GELU and BF16 weights (no
 * reference counterpart for pre-SVD prediction; the reference's predictors,
 * src/atlas_q/tools_qih/ai_rank_predictor.py:14-89 and
 * finetuned_predictor.py:8-23, run after a full SVD at tn_core.py:294, and its
 * training recipe, ai_rank_predictor.py:117-195, is AdamW on an MSE loss).
 * Parameters: one FP32 master vector of aq_mlp_num_params() floats laid out
 * W1[128][32] | b1[128] | W2[128][128] | b2[128] | W3[128] | b3, and a BF16
 * copy (uint16) of the same length used by forward/backward for the weights.
 * X: float32 [batch, 32] (row-major), y: float32 [batch].  The layer products
 * run on BF16 MFMA (v_mfma_f32_32x32x16_bf16, FP32 accumulation) with the
 * BF16 weights exact and FP32 activations / gradients split into BF16 hi +
 * lo (relative error ~2^-16 per product). */
int aq_mlp_num_params(void);
int aq_mlp_sync_bf16(const float* master, uint16_t* wb, hipStream_t stream);
int aq_mlp_forward(const float* X, int64_t batch, const uint16_t* wb, const float* master,
                   float* y, hipStream_t stream);
/* Gradient of sum_s dy[s] * y[s]; workspace: aq_mlp_workspace_floats(batch)
 * floats (per-block partials, summed in a fixed order). */
int64_t aq_mlp_workspace_floats(int64_t batch);
int aq_mlp_backward(const float* X, const float* dy, int64_t batch, const uint16_t* wb,
                    const float* master, float* grad, float* workspace, int64_t workspace_floats,
                    hipStream_t stream);
/* AdamW step on the master (decoupled decay on weights only), refreshing wb. */
int aq_mlp_adamw(float* master, const float* grad, float* m, float* v, uint16_t* wb, double lr,
                 double beta1, double beta2, double eps, double wd, double grad_scale, int step,
                 hipStream_t stream);

/* ---------------------------------------------------------------- grouped
 * Batched, uneven work for many independent MPS in one persistent launch (no
 * reference counterpart: the reference advances one MPS at a time).  The
 * caller uploads one descriptor per problem and `prefix`, the exclusive prefix
 * sum of the problems' work counts (tiles or elements, int64); a grid of
 * `grid` work-groups strides over the global work index. */
#define AQ_GROUPED_GEMM 0  /* C = op(A) op(B), linear columns */
#define AQ_GROUPED_THETA 1 /* theta: A [chiL,2,chiM], B [chiM,2,chiR], C = X,
                              M = 2chiL, N = 2chiR, K = chiM, gate in `gate` */
#define AQ_GROUPED_HERK 2  /* C = op(A) op(B) with M == N, only the tiles that meet
                              the lower triangle (row >= col) are computed: the
                              Gram matrix X X^H where the eigensolver reads one
                              triangle (aq_heevd_batched reads row >= col) */
typedef struct {
  const void* A;
  const void* B;
  void* C;
  int64_t M, N, K;
  int64_t a_s0, a_s1, b_s0, b_s1, c_s0, c_s1; /* element strides */
  int32_t conj_a, conj_b, kind, pad;
  double gate[32]; /* AQ_GROUPED_THETA: U[4][4] interleaved re/im, row-major */
} aq_gemm_desc_t;
/* Tiles of a problem: ceil(M/64) * ceil(N/128) (GEMM), ceil(M/64) *
 * ceil(N/2/64) (theta) or sum_{bm < ceil(M/64)} (floor(bm/2) + 1) (HERK).
 * layout: bit0 = A unit stride along k (a_s1 == 1), bit1 = B unit stride
 * along n (b_s1 == 1), uniform over the group; bit2 = the group holds no
 * AQ_GROUPED_THETA problem (selects a kernel without the gate epilogue, at
 * twice the occupancy; a THETA problem in such a group is computed as a
 * plain GEMM); bit3 = XCD-contiguous tile ranges: with a grid that is a
 * multiple of 8 each of the 8 XCDs (round-robin dispatch) takes a contiguous
 * range of tiles (fewer L2 misses; measured slightly slower for the FP64
 * Gram / projection groups, so off by default); bit4 = 32-row tiles of 128
 * threads instead of 64-row tiles of 256 (the tile counts above with
 * ceil(M/32) rows of tiles; an AQ_GROUPED_HERK problem is then computed as
 * a full product, ceil(M/32) * ceil(N/128) tiles). */
int aq_grouped_gemm(int dtype, const aq_gemm_desc_t* descs, const int64_t* prefix,
                    int num_problems, int64_t total_tiles, int layout, int grid,
                    hipStream_t stream);

/* Single-qubit gates on many complex64 sites (AdaptiveMPS convention,
 * adaptive_mps.py:177): work count of a site = chiL * chiR. */
typedef struct {
  const void* T_in;
  void* T_out;
  int64_t chiL, chiR;
  double gate[8];
} aq_site1q_desc_t;
int aq_grouped_1q(const aq_site1q_desc_t* descs, const int64_t* prefix, int num_sites,
                  int64_t total_elems, int grid, hipStream_t stream);

/* Rank rule (aq_rank_select semantics) on `count` spectra at once, one
 * work-group each: S_ptrs / m / cap are device arrays of length count. */
int aq_rank_select_batched(int s_dtype, const void* const* S_ptrs, const int64_t* m,
                           const int64_t* cap, double eps, int count, aq_rank_result_t* out,
                           hipStream_t stream);

/* Sweep sampler over many MPS: one grouped GEMM launch and one grouped select
 * launch per site index (aq_sample semantics per MPS).  For site index i the
 * caller passes the descriptors of every MPS with more than i sites. */
typedef struct {
  const void* L;       /* previous site's T rows (NULL at site 0) */
  const int32_t* off;  /* per-shot column offset into L (NULL at site 0) */
  const void* site;    /* [chiL, 2, chiR] */
  void* Tout;          /* [S, 2 chiR] */
  void* part;          /* [S, ceil(chiR/32), 2] real */
  const double* uniforms; /* [S, n] */
  uint64_t* bits;      /* [S, words] */
  int32_t* off_next;   /* [S] */
  int64_t ldL, S, chiL, chiR, n, site_index, words;
} aq_sample_desc_t;
int aq_sample_grouped_site(const aq_sample_desc_t* descs, const int64_t* tile_prefix,
                           const int64_t* shot_prefix, int count, int64_t total_tiles,
                           int64_t total_shots, int grid, hipStream_t stream);
/* Zero `count` uint64 words on `stream` (no reference counterpart): the
 * sampler's bit words, and a one-element launch that brackets a timed region
 * in a kernel trace. */
int aq_zero_u64(uint64_t* p, int64_t count, hipStream_t stream);

/* Grouped core rebuild (adaptive_mps.py:280-290 for many updates): complex64
 * dst[r, c] = op(src[r*s0 + c*s1]) * row_scale[r] * col_scale[c] for r < rows,
 * c < cols, src complex128 (FP64 factors), scales float64 (nullable); work
 * count of a descriptor = rows * cols. */
typedef struct {
  const void* src;
  void* dst;
  const double* row_scale;
  const double* col_scale;
  int64_t rows, cols, s0, s1;
  int32_t conj, pad;
} aq_copy_desc_t;
int aq_grouped_copy(const aq_copy_desc_t* descs, const int64_t* prefix, int count,
                    int64_t total_elems, int grid, hipStream_t stream);

/* ---------------------------------------------------------------- device-planned layers
 * A batched two-qubit sub-layer with its bond dimensions on the device (no
 * reference counterpart: the grouped form of adaptive_mps.py:182-308 for many
 * MPS, as BatchedMPS runs it, without host round trips so that a whole
 * circuit can be captured in a HIP graph).  The host fixes every buffer,
 * bucket and launch shape from upper bounds of the bond dimensions; these
 * entry points write the size-dependent descriptor fields from the actual
 * dimensions (dims, int64 per bond of the gate's MPS, updated in place). */
typedef struct {
  const void* siteA;      /* complex64 slot of site `bond` */
  const void* siteB;      /* complex64 slot of site `bond` + 1 */
  void* X;                /* theta (complex64, bound-sized slot) */
  void* X64;              /* its complex128 copy */
  void* G;                /* Gram matrix: row of a size bucket (complex128, sz x sz) */
  const void* V;          /* eigenvectors: vfmt 0 columns of a row-major sz x sz block,
                             vfmt 1 conjugate rows Et (kpad x sz) */
  const double* S64;      /* square roots of the row's eigenvalues, float64 [sz] */
  const double* Sinv;     /* their reciprocals (0 for 0) */
  void* P;                /* projection output (complex128, bound-sized slot) */
  int64_t* dims;          /* the MPS's bond dimensions (device) */
  int64_t bond;
  int64_t sz;             /* bucket order */
  int64_t left;           /* 1: Gram X X^H, 0: X^H X (fixed by the bounds) */
  int64_t vfmt;
  int64_t kslot;          /* index into the tridiagonal buckets' kept-rank array, or -1 */
  int64_t kbound;         /* planned bound of the kept rank (buffers are sized by it) */
  double gate[32];        /* U[4][4] interleaved re/im, "U" convention */
} aq_layer_gate_t;

int64_t aq_layer_gate_size(void);   /* sizeof(aq_layer_gate_t), for bindings */

/* Theta descriptors (AQ_GROUPED_THETA, gate in the epilogue), Gram
 * descriptors (into G, row stride sz) and each Gram matrix's order. */
int aq_layer_pre(const aq_layer_gate_t* gates, int count, aq_gemm_desc_t* theta,
                 aq_gemm_desc_t* gram, int64_t* order, hipStream_t stream);
/* Bucket rows b < rows of T (complex128 [rows, sz, sz]): entries with row or
 * column >= order[row_gate[b]] set to zero, the diagonal there to
 * -d (1 + i/sz), d the largest diagonal entry of the leading block. */
int aq_layer_pad(void* T, int64_t sz, int64_t rows, const int64_t* row_gate,
                 const int64_t* order, hipStream_t stream);
/* From the rank results rr[g] (aq_rank_select_batched): the projection
 * descriptors (Vh_k = U_k^H X or U_k S_k = X V_k, into P), the rebuild
 * descriptors (copies[g] site `bond`, copies[count + g] site `bond` + 1, as
 * BatchedMPS), dims[bond + 1] = k, kvec[kslot] = k, and (nullable) the
 * logs of k, eps_local and entropy.  A rank above the gate's kbound is
 * lowered to it and counted in *overflow (device int64). */
int aq_layer_post(const aq_layer_gate_t* gates, int count, const aq_rank_result_t* rr,
                  aq_gemm_desc_t* proj, aq_copy_desc_t* copies, int32_t* kvec, int64_t* log_k,
                  double* log_eps, double* log_ent, int64_t* overflow, hipStream_t stream);
/* Single-qubit descriptors: chiL = dims[g][site[g]], chiR = dims[g][site[g] + 1]. */
int aq_layer_1q(aq_site1q_desc_t* descs, int64_t* const* dims, const int64_t* site, int count,
                hipStream_t stream);

/* ---------------------------------------------------------------- tn_core
 * The reference's second two-site update, tools_qih/tn_core.py:mps_apply_2q
 * (tn_core.py:186-384): theta in the U^T convention (aq_theta with the gate
 * transposed), SVD, then the truncation below and the sqrt(S) split. */

/* AI-policy codes (QIH_AI_POLICY, tn_core.py:282-333; unknown names map to
 * guardrail on the host, as the reference's fallback branch does). */
#define AQ_TN_POLICY_BASELINE 0
#define AQ_TN_POLICY_GUARDRAIL 1
#define AQ_TN_POLICY_BLEND 2
#define AQ_TN_POLICY_CLAMP 3
#define AQ_TN_POLICY_CLAMP_BLEND 4

typedef struct {
  int64_t keep_count;    /* singular values kept before the S > eps filter */
  int64_t chi;           /* resulting bond dimension (#S[:keep_count] > eps) */
  int64_t adaptive_keep; /* the adaptive cap (the SVD event's adaptive_keep) */
  int32_t prefix;        /* 1 when the S > eps survivors are a prefix */
  int32_t pad;
  double total_weight;   /* sum S^2 */
  double kept_weight;    /* sum over the kept values of S^2 */
  double truncation_error; /* 1 - kept/total (0 when total == 0) */
} aq_tn_trunc_t;

/* Truncation of tn_core.py:233-372 on a descending spectrum S of m values
 * (float32 for AQ_C64, float64 for AQ_C128), one result written to device
 * memory `out`.  adaptive/tol: the adaptive rule keep = #(cumsum(S^2) <=
 * (1-tol) total) + 1 (at most m; 1 when none), applied when adaptive and tol
 * > 0.  chi_max < 0 means None.  ai_keep >= 0: the predictor's rank (already
 * capped by chi_max) combined with the adaptive cap by `policy` (blend weight
 * alpha); ai_keep < 0: no predictor (adaptive rule, else the fixed cap).
 * Arithmetic follows torch's CPU rounding (cumsum accumulated in double,
 * threshold float(1-tol) * total in S's type, S > eps with eps in S's type). */
int aq_tn_truncate(int s_dtype, const void* S, int64_t m, double tol, int adaptive,
                   int64_t chi_max, int64_t ai_keep, int policy, double alpha, double eps,
                   aq_tn_trunc_t* out, hipStream_t stream);

/* Core split of tn_core.py:374-382: A'[r, c] = U[r, c] * sqrt(S[c]) as
 * [rows/2, 2, k], B'[r, c] = sqrt(S[r]) * Vh[r, c] as [k, 2, cols/2]; strides
 * and types as aq_rebuild. */
int aq_rebuild_sqrt(int dtype, const void* U, int64_t u_s0, int64_t u_s1, const void* S,
                    const void* Vh, int64_t v_s0, int64_t v_s1, int64_t rows, int64_t cols,
                    int64_t k, void* A_out, void* B_out, hipStream_t stream);

/* Inputs of the reference's learned rank predictors from a spectrum S (m
 * values, descending, float32/float64 by s_dtype), on the device.
 * AQ_FEAT_LOG1P64: RankPredictor.forward's input (ai_rank_predictor.py:55-69):
 *   64 floats log1p(S[i]/S[0]) for i < min(limit, m, 64), zero elsewhere; limit
 *   is forward's max_ranks, or #(S > 1e-10) when negative (as predict passes).
 * AQ_FEAT_FT135: FinetunedPredictorWrapper._features (finetuned_predictor.py:
 *   51-75): 128 shifted log singular values, spectral entropy, cumulative mass
 *   at 8, 16, 32, 64, 96, 128; 135 floats.
 * n_sv (device int64): #(S > 1e-10), the rank both predictors scale by. */
#define AQ_FEAT_LOG1P64 0
#define AQ_FEAT_FT135 1
#define AQ_FEAT_SPEC32 2 /* no reference counterpart: the inputs of this library's
                            spectral predictor for the predict(S) seam of
                            tn_core.py:294 -- with p_i = S_i^2 / sum S^2 over the
                            leading <= 128 values: log10 p_i (i < 24), log10 of the
                            weight beyond 8,16,32,64,96,128 values, entropy,
                            log2 #(S > 1e-10); 32 floats */
int aq_spectrum_features(int kind, int s_dtype, const void* S, int64_t m, int64_t limit,
                         float* feat, int64_t* n_sv, hipStream_t stream);

/* Forward pass of a stack of dense layers in FP32, the architectures of the
 * reference's RankPredictor (ai_rank_predictor.py:24-42: Linear, LayerNorm,
 * ReLU; Dropout is the identity at inference) and FinetunedMLP
 * (finetuned_predictor.py:11-23: Linear, ReLU, Sigmoid).  Layer l computes
 * z = W z_prev + b (W row-major [d_out, d_in]), then LayerNorm when ln_g is
 * set (biased variance, eps ln_eps, affine ln_g / ln_b), then `act`.
 * layers: host array of n_layers (1..8) descriptors of device pointers; dims
 * 1..512.  x: [batch, d_in of layer 0], y: [batch, d_out of the last layer]. */
#define AQ_ACT_NONE 0
#define AQ_ACT_RELU 1
#define AQ_ACT_GELU 2
#define AQ_ACT_SIGMOID 3
typedef struct {
  const float* W;
  const float* b;    /* nullable */
  const float* ln_g; /* nullable: no LayerNorm */
  const float* ln_b; /* nullable */
  int32_t d_in, d_out, act, pad;
  double ln_eps;
} aq_dense_layer_t;
int aq_dense_forward(int n_layers, const aq_dense_layer_t* layers, const float* x, int64_t batch,
                     float* y, hipStream_t stream);

/* ---------------------------------------------------------------- MPO / envs
 * Callers next to the two-site path.  dtype AQ_C64 / AQ_C128, contiguous. */

/* MPO x MPS site contraction of apply_mpo_to_mps (src/atlas_q/mpo_ops.py:
 * 1034-1082): M[(l*chiL + a), s, (r*chiR + b)] = sum_t W[l,s,t,r] A[a,t,b],
 * W [Dl, 2, 2, Dr], A [chiL, 2, chiR], M [Dl*chiL, 2, Dr*chiR]. */
int aq_mpo_apply_site(int dtype, const void* W, int64_t Dl, int64_t Dr, const void* A,
                      int64_t chiL, int64_t chiR, void* M, hipStream_t stream);

/* Combine step of the MPO expectation-value transfer (mpo_ops.py:1083-1150,
 * E' = einsum("Lab,atr,LstR,bsB->RrB", E, conj(A), W, A)): given
 * Z[L, t] = conj(A[:, t, :])^T E[L] A  (Z [DL, 2, chir, 2*chiB], the two
 * ket outcomes s side by side), E'[R, r, B] = sum_{L,s,t} W[L,s,t,R]
 * Z[L, t, r, s*chiB + B].  W [DL, 2, 2, DR], E [DR, chir, chiB]. */
int aq_mpo_env_combine(int dtype, const void* Z, const void* W, int64_t DL, int64_t DR,
                       int64_t chir, int64_t chiB, void* E, hipStream_t stream);

/* Tile-boundary renormalisation of build_left_envs / build_right_envs
 * (src/atlas_q/sim_env_scan.py:53-130): X_b <- X_b / max(||X_b||_F, floor)
 * for `batch` blocks of `count` contiguous elements at X + b*stride. */
int aq_fro_normalize_batched(int dtype, void* X, int64_t batch, int64_t stride, int64_t count,
                             double floor_, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* ATLASQ_H */
