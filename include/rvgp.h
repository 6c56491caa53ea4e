/*
 * rvgp.h -- C ABI of librvgp.so, the MI355X (gfx950) kernels behind the
 * RVGP connection-Laplacian Gaussian-process path.
 *
 * Conventions (all entry points):
 *   - Every pointer argument is DEVICE memory unless its comment says "host".
 *     The caller allocates every buffer; the library never frees caller
 *     memory.  Scratch comes from caller-provided workspace pointers whose
 *     sizes are given by the matching *_ws_bytes() query.
 *   - `stream` is a hipStream_t passed as void*; work is enqueued on it and
 *     the call returns without synchronising unless the comment says so.
 *   - Return value: RVGP_OK (0) or a negative status.  The Python layer maps
 *     RVGP_EINVAL -> ValueError, RVGP_ERUNTIME / RVGP_EHIP -> RuntimeError,
 *     RVGP_ENOMEM -> MemoryError, the exception types the reference raises
 *     (RVGP/lib/ptu_dijkstra.pyx:68-82, 119-123, 358-359).
 *   - Kernels that detect a data-dependent failure (a rank-deficient
 *     neighbourhood, a non-positive-definite matrix) write a code into a
 *     caller-provided device int `status`; the caller reads it after the
 *     stream completes.
 *   - Matrices are row-major; `ld*` is the row stride in elements.
 *   - No global mutable state other than the thread-local error string.
 *
 * Each entry point names the reference interface it replaces (file:line,
 * relative to the agosztolai/RVGP repository).
 */
#ifndef RVGP_H
#define RVGP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RVGP_OK 0
#define RVGP_EINVAL (-1)
#define RVGP_ERUNTIME (-2)
#define RVGP_ENOMEM (-3)
#define RVGP_EHIP (-4)

/* Device-side status codes written into `status` words. */
#define RVGP_ST_OK 0
#define RVGP_ST_RANK_DEFICIENT 1   /* a neighbourhood singular value < 1e-10 */
#define RVGP_ST_SHORT_COMPONENT 2  /* fewer than K+1 nodes reachable */
#define RVGP_ST_POOL_OVERFLOW 3    /* heap node pool too small (host sizing bug) */
#define RVGP_ST_NOT_SPD 4          /* Cholesky met a non-positive pivot */
#define RVGP_ST_TIMEOUT 5          /* a bounded cross-workgroup wait gave up */

/* Library version and last error message (thread-local, host). */
int rvgp_version(void);
const char* rvgp_last_error(void);

/* ---------------------------------------------------------------------
 * a1  Furthest point sampling  -- replaces geometry.furthest_point_sampling
 *     (RVGP/geometry.py:126-162).
 * ------------------------------------------------------------------- */

/* diam2[0] = max over pairs of |x_i - x_j|^2 (the square of D.max(),
 * geometry.py:142-144).  x: (n, D) f64.  diam2 must be zeroed by the
 * caller (it is accumulated with an atomic max). */
int rvgp_fps_diam2(const double* x, int n, int D, double* diam2, void* stream);

/* Greedy farthest-point loop (geometry.py:146-160) in one persistent
 * workgroup.  n_iter = N, or n when N is None.  If use_spacing != 0 the loop
 * stops at the first i with lambdas[i]/sqrt(diam2) < spacing and writes i to
 * n_sel[0]; otherwise n_sel[0] = n_iter.  ds: workspace of n doubles.
 * perm: (n_iter,) int32, lambdas: (n_iter,) f64.  1 <= D <= 8192 (the
 * spacing mode's rvgp_fps_diam2 takes D <= 64); main.py:59-60 runs it on
 * the spectral rows (D = n_eigenpairs) to pick SGPR inducing points. */
int rvgp_fps(const double* x, int n, int D, int n_iter, int start_idx, double spacing,
             int use_spacing, const double* diam2, double* ds, int* perm, double* lambdas,
             int* n_sel, void* stream);

/* ---------------------------------------------------------------------
 * a2  kNN graph  -- replaces geometry.manifold_graph(typ='knn')
 *     (RVGP/geometry.py:100-123: sklearn kneighbors_graph(include_self=False)
 *     + I -> networkx) and the symmetrisation of ptu_dijkstra.pyx:84-103.
 * ------------------------------------------------------------------- */

/* nbr (n, k) int32: the k nearest other points of every point, ordered by
 * (squared distance, index).  Requires 0 < k < n, k <= 64, D <= 64.
 * qperm: NULL or a permutation of [0, n) giving the order in which queries
 * are packed into waves (points close on the manifold should be close in it);
 * the result does not depend on it.  ws:
 * rvgp_knn_ws_bytes(n, k) bytes -- per-slice top-k lists of the candidate
 * slices scanned in parallel, and for n > 8192 a bounding pass's per-point
 * k-th distance over the first 8192 points, which lets the full scan skip
 * candidates that cannot end in the top k (the result does not depend on
 * either). */
size_t rvgp_knn_ws_bytes(int n, int k);
int rvgp_knn(const double* x, int n, int D, int k, const int* qperm, int* nbr, void* ws,
             size_t ws_bytes, void* stream);

/* Dense Gaussian affinity of manifold_graph(typ='affinity')
 * (geometry.py:114-118): A (n x n, row stride lda) with
 * A[i][j] = exp(-|x_i - x_j|^2 / (2 sigma^2)) (the reference fixes
 * sigma = 0.1); x (n, D) f64, D <= 64.  The non-zero entries are the graph's
 * weighted edges (self loops of weight 1 included), as nx.from_numpy_array
 * reads them. */
int rvgp_affinity(const double* x, int n, int D, double sigma, double* A, int lda, void* stream);

/* Workspace bytes for rvgp_graph_symmetrize. */
size_t rvgp_graph_symmetrize_ws_bytes(int n, int k);

/* Sorted, de-duplicated CSR of max(A, A^T) with A = kNN connectivity + I:
 * row i = {i} u nbr(i) u {j : i in nbr(j)}.  rowptr: (n+1) int32;
 * cols: capacity n*(2k+1) int32.  nnz = rowptr[n]. */
int rvgp_graph_symmetrize(const int* nbr, int n, int k, int* rowptr, int* cols, void* ws,
                          size_t ws_bytes, void* stream);

/* ---------------------------------------------------------------------
 * a3  Geodesic neighbourhoods + tangent frames -- replaces
 *     ptu_dijkstra.tangent_frames (RVGP/lib/ptu_dijkstra.pyx:33-125, 300-437).
 * ------------------------------------------------------------------- */

/* Workspace bytes for rvgp_geo_nbh on n sources with `batch` in flight and
 * a worst-case node pool of `pool` slots per source (pool >= 1 + K * max row
 * length).  A first pass gives each source a 512-node pool; the sources that
 * outgrow it are listed and redone with `pool` (same result). */
size_t rvgp_geo_nbh_ws_bytes(int n, int batch, int pool);

/* Truncated Dijkstra driven by the Fibonacci heap of ptu_dijkstra.pyx:
 * 440-690, one source per lane: nbh (n, K+1) int32 in pop order
 * (pyx:361-394).  w: (nnz,) f64 edge weights (all 1.0 on the RVGP path).
 * status[0] is set to RVGP_ST_SHORT_COMPONENT / RVGP_ST_POOL_OVERFLOW. */
int rvgp_geo_nbh(const int* rowptr, const int* cols, const double* w, int n, int K, int batch,
                 int pool, int* nbh, void* ws, size_t ws_bytes, int* status, void* stream);

/* Centre each neighbourhood (pyx:397-404) and take its local PCA -- the left
 * singular vectors and singular values of the D x (K+1) centred matrix
 * (pyx:406-434) -- via an LDS-staged covariance and a Jacobi eigensolve,
 * one wavefront per node.  tangents (n, D, ncols) f64 (first ncols left
 * singular vectors, ncols <= min(D, K+1)); sigma (n, nsv) f64.  Any of the
 * first n_check singular values below 1e-10 sets status[0] =
 * RVGP_ST_RANK_DEFICIENT (pyx:426-428).  D <= 64, K+1 <= 64. */
int rvgp_gauges(const double* x, int n, int D, const int* nbh, int Kp1, int ncols, int nsv,
                int n_check, double* tangents, double* sigma, int* status, void* stream);

/* ---------------------------------------------------------------------
 * a5/a6/a7  Connections and Laplacians -- replaces ptu_dijkstra.connections
 *     (pyx:128-293), geometry.compute_laplacian (geometry.py:55-63) and
 *     geometry.compute_connection_laplacian (geometry.py:14-52).
 * ------------------------------------------------------------------- */

/* For every CSR entry e = (i, j): R_e = polar factor of T_i^T T_j (= U V^T
 * of its SVD, pyx:259-293), T_i = tangents[i][:, :d] with row stride ldt;
 * nnz = rowptr[n].
 * Writes R (nnz, d, d) f64, and if non-NULL the connection-Laplacian blocks
 * L_ij R_e (L_ii = degree, L_ij = -1) as f64 lc64 and f32 lc32, and the
 * scalar Laplacian values as f64 l64 / f32 l32 (nnz,).  d <= 8. */
int rvgp_connections(const double* tangents, int n, int D, int ldt, int d, const int* rowptr,
                     const int* cols, int nnz, double* R, double* lc64, float* lc32, double* l64,
                     float* l32, void* stream);

/* ---------------------------------------------------------------------
 * Block-sparse (BSR, b x b blocks, b in {1,2,3}) times dense, fused:
 *   Y = alpha * (A X) + beta * X + gamma * Z ;  if acc: acc += delta * Y
 * X, Z, Y, acc: (nb*b, r) with row strides ldx/ldz/ldy/ldacc.  Z and acc may
 * be NULL.  Y must not alias X.  Used by the eigensolver (Chebyshev filter,
 * Rayleigh-Ritz) and by the heat-kernel recurrence, in place of scipy sparse
 * products inside eigsh (geometry.py:73) and dense expm (smoothing.py:15).
 * ------------------------------------------------------------------- */
int rvgp_bsr_spmm_f64(int nb, int b, const int* rowptr, const int* cols, const double* vals,
                      const double* X, int ldx, int r, double alpha, double beta,
                      const double* Z, int ldz, double gamma, double* Y, int ldy,
                      double* acc, int ldacc, double delta, void* stream);
int rvgp_bsr_spmm_f32(int nb, int b, const int* rowptr, const int* cols, const float* vals,
                      const float* X, int ldx, int r, float alpha, float beta, const float* Z,
                      int ldz, float gamma, float* Y, int ldy, float* acc, int ldacc,
                      float delta, void* stream);
/* Half-storage form for the eigensolver's opening Chebyshev passes (the
 * sparse products inside eigsh, geometry.py:73): X, Z, Y are IEEE binary16
 * (uint16 storage), vals f32, all arithmetic f32, Y rounded to nearest even.
 * r and ldx/ldy/ldz multiples of 8, X/Y/Z 16-byte aligned, X spanning
 * < 2 GiB; no acc term. */
int rvgp_bsr_spmm_f16(int nb, int b, const int* rowptr, const int* cols, const float* vals,
                      const void* X, int ldx, int r, float alpha, float beta, const void* Z,
                      int ldz, float gamma, void* Y, int ldy, void* stream);

/* ---------------------------------------------------------------------
 * a10/a11  Vector-field frame algebra -- replaces geometry.project_to_manifold
 *     and express_in_local_frame (RVGP/geometry.py:165-175, used at
 *     dataclass.py:100, 110, 116) and the renormalisation of
 *     smoothing.vector_diffusion (RVGP/smoothing.py:57-62).
 * ------------------------------------------------------------------- */

/* F fields over n nodes, frames T (n, D, ldt) f64 (first d columns used):
 *   mode 0: out (F,n,d) = T^T in (F,n,D);  mode 1: out (F,n,D) = T in (F,n,d);
 *   mode 2: out (F,n,D) = T T^T in (F,n,D)  (in == out allowed). */
int rvgp_frame_apply(int mode, int F, int n, int D, int d, const double* T, int ldt,
                     const double* in, double* out, void* stream);

/* out (rows, d) with row stride ldo, in place:
 *   out_i <- out_i * a_i / (b_i * |out_i|)   (a, b with strides lda, ldb). */
int rvgp_diffusion_normalise(size_t rows, int d, double* out, int ldo, const double* a, int lda,
                             const double* b, int ldb, void* stream);

/* a9  Ambient eigenvector rows Phi = T_i U_i for the listed nodes
 * (dataclass.py:57-59): out ((m*D), k) rows r*D+p, f32 if out_f32 else f64;
 * U (n*d, k) f64 with row stride ldu.  nodes[r] must lie in [0, n): a node
 * outside it reads no memory and its D rows are NaN (the host wrappers raise
 * IndexError before launching, as numpy indexing does). */
int rvgp_phi_rows(int out_f32, int m, int n, int D, int d, const double* T, int ldt, int k,
                  const double* U, int ldu, const int* nodes, void* out, int ldo, void* stream);

/* ---------------------------------------------------------------------
 * Dense linear algebra on the matrix cores.  Replaces the dense BLAS/LAPACK
 * inside ARPACK's Rayleigh-Ritz (geometry.py:73), the GPflow Gram
 * K = (X * S) X^T (kernels.py:55-61) and TensorFlow's Cholesky / triangular
 * solves behind GPR.log_marginal_likelihood and predict_f (main.py:98-116).
 * ------------------------------------------------------------------- */
#define RVGP_GEMM_F64 0       /* f64 in, f64 MFMA, f64 out */
#define RVGP_GEMM_F32 1       /* f32 in, f32 MFMA, f32 out */
#define RVGP_GEMM_F32_ACC64 2 /* f32 in, widened, f64 MFMA, f64 out */
#define RVGP_GEMM_BF16 3      /* f32 in, rounded to bf16, f32 MFMA accumulate, f32 out */

/* C = alpha * op(A) * diag(kscale) * op(B) + beta * C, row-major operands,
 * op = transpose when trans* != 0; kscale (K,) f64 or NULL.  split_k > 1
 * splits the K range over that many workgroup slices whose partials are
 * reduced in a fixed order (deterministic); it needs ws_bytes >=
 * split_k * M * N * sizeof(out element). */
int rvgp_gemm(int mode, int transA, int transB, int M, int N, int K, double alpha, const void* A,
              int lda, const void* B, int ldb, const double* kscale, double beta, void* C, int ldc,
              int split_k, void* ws, size_t ws_bytes, void* stream);

/* Workspace bytes rvgp_gemm needs for this shape; split_k <= 0 lets the
 * library choose the split (from the shape only, so results are
 * reproducible).  0 when no split is used. */
/* Symmetric product C = A^T B (M x M, A and B K x M row-major, f64 out) for
 * Grams known to be symmetric (A^T A, or A^T (S A) with S symmetric): only
 * the upper 128 x 128 tiles are computed, the lower triangle is mirrored from
 * the upper.  mode F64, F32_ACC64, or F32: f32 in, FP32 MFMA over K-chunks of
 * at most 4096 rows, chunk partials summed in f64 (f64 out); split_k / ws as
 * rvgp_gemm (workspace from rvgp_gemm_ws_bytes(mode, M, M, K, split_k)).  The Rayleigh-Ritz Grams
 * of the eigensolver that replaces eigsh (geometry.py:73). */
int rvgp_gemm_sym(int mode, int M, int K, const void* A, int lda, const void* B, int ldb,
                  double* C, int ldc, int split_k, void* ws, size_t ws_bytes, void* stream);
size_t rvgp_gemm_ws_bytes(int mode, int M, int N, int K, int split_k);

/* In-place lower Cholesky A = L L^T of an n x n row-major matrix (f64 or
 * f32 by `f64` flag); the strictly upper triangle is left unspecified.  A
 * non-positive pivot sets status[0] = RVGP_ST_NOT_SPD (the condition that
 * makes tf.linalg.cholesky fail inside GPflow). */
int rvgp_potrf(int f64, int n, void* A, int lda, int* status, void* stream);

/* Solve with the Cholesky factor in place on B (n x r):
 *   trans == 0: L X = B ;  trans != 0: L^T X = B. */
int rvgp_trsm(int f64, int trans, int n, int r, const void* L, int lda, void* B, int ldb,
              void* stream);

/* out[0] = sum_i log(L_ii) accumulated in f64 (half the log-determinant). */
int rvgp_chol_logdet(int f64, int n, const void* L, int lda, double* out, void* stream);

/* Symmetric eigendecomposition A = V diag(w) V^T in FP64, w ascending, the
 * eigenvectors in the columns of V (n x n, row stride ldv): the dense
 * eigenproblem of the Rayleigh-Ritz step that ARPACK solves inside eigsh
 * (geometry.py:73).  Householder tridiagonalisation, multisection bisection,
 * inverse iteration (Gram-Schmidt inside eigenvalue clusters), blocked
 * back-transformation.  A is not modified.  1 <= n <= 2048; workspace
 * rvgp_syev_ws_bytes(n) bytes. */
size_t rvgp_syev_ws_bytes(int n);
int rvgp_syev(int n, double* A, int lda, double* w, double* V, int ldv, void* ws,
              size_t ws_bytes, void* stream);

/* Ritz residuals of a block: res[j] = |AX[:, j] - theta[j] X[:, j]| / |X[:, j]|
 * (f64 sums in a fixed order, the residual block never formed); X and AX
 * (M, k), f64 or f32 by `f64`, row strides ldx / ldax; theta (k) f64 on the
 * device.  The convergence test ARPACK applies to its Ritz pairs inside eigsh
 * (geometry.py:73).  ws: rvgp_ritz_residuals_ws_bytes(M, k) bytes. */
size_t rvgp_ritz_residuals_ws_bytes(int M, int k);
int rvgp_ritz_residuals(int f64, int M, int k, const void* X, int ldx, const void* AX, int ldax,
                        const double* theta, double* res, void* ws, size_t ws_bytes,
                        void* stream);

/* Eigenvector output of compute_spectrum (geometry.py:72-75):
 * out[perm[i] * b + a][j] = scale * X[i * b + a][j] / sqrt(nrm2[j]) for block
 * rows i < nb, a < b, columns j < k -- unit-norm columns times
 * scale = sqrt(rows), mapped back from a renumbering of the operator's block
 * rows (perm[new] = old, int64; NULL = identity).  X f64 or f32 (`f64`), row
 * stride ldx; out f64, row stride ldo, must not alias X. */
int rvgp_evecs_finish(int f64, int nb, int b, int k, const void* X, int ldx, const double* nrm2,
                      const long long* perm, double scale, double* out, int ldo, void* stream);

/* out[j] = sum_i Z[i][j]^2 (f64 accumulation, fixed order); Z (M, k) f64/f32. */
size_t rvgp_colsumsq_ws_bytes(int M, int k);
int rvgp_colsumsq(int f64, int M, int k, const void* Z, int ldz, double* out, void* ws,
                  size_t ws_bytes, void* stream);

/* rvgp_colsumsq without a workspace argument; no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * partial sums live in a fixed 8 MiB device array held by the library (one
 * per GPU, until the process ends), so the call allocates nothing.  Limit:
 * min(1024, ceil(M/32)) * k <= 2^20, else RVGP_EINVAL.  Results equal
 * rvgp_colsumsq's bit for bit.  Calls that overlap on one GPU share the array
 * and race; the caller orders them. */
int rvgp_colsumsq_noalloc(int f64, int M, int k, const void* Z, int ldz, double* out,
                          void* stream);

/* ---------------------------------------------------------------------
 * a12/a14/a15  Spectral kernel, NLML and adjoint -- replaces
 *     ManifoldKernel.eval_S / K / K_diag (RVGP/kernels.py:41-67) and GPflow's
 *     GPR.training_loss with TF autodiff under the scipy optimiser
 *     (RVGP/main.py:87-95).  Parameters theta = (nu, kappa, sigma_f, s2),
 *     theta = softplus(u) + lower.  All scalars live on the device so an
 *     optimiser step can be captured in a HIP graph.
 * ------------------------------------------------------------------- */

/* S (k) and dS (3, k) = dS/d(nu, kappa, sigma_f) of the spectral filter,
 * typ 0 = matern, 1 = se; theta (3,) device; n_rows = rows of evecs_Lc. */
int rvgp_spectral_filter(int typ, int k, const double* evals, const double* theta, double n_rows,
                         double* S, double* dS, void* stream);

/* A = G + diag(s2 / S)  (weight-space system matrix), s2 a device scalar. */
int rvgp_ws_form(int k, const double* G, int ldg, const double* S, const double* s2, double* A,
                 int lda, void* stream);

/* Weight-space NLML and gradients, summed over nrhs right-hand sides that
 * share the hyperparameters (GPflow's GPR with a Y of nrhs columns sums the
 * columns' log-densities).  b = Phi^T Y, w = A^-1 b, Gw = G w are k x nrhs
 * row-major (leading dimension nrhs); yy = |Y|_F^2.  out (9,) = [nlml,
 * d/d(nu,kappa,sigma_f,s2), d/du(...)]; gS: scratch (k,).  M = observations
 * per right-hand side. */
int rvgp_ws_finish(int k, double M, int nrhs, const double* L, int ldl, const double* ainv_diag,
                   const double* b, const double* w, const double* Gw, const double* yy,
                   const double* S, const double* dS, const double* s2, const double* u, double* gS,
                   double* out, void* stream);

/* Function-space NLML and gradients over nrhs right-hand sides from
 * q_j = |L^-1 phi_j|^2, p = Phi^T a (k x nrhs row-major), ya = sum_r y_r^T a_r,
 * aa = |a|_F^2 and logdet_half = sum log L_ii (all device). */
int rvgp_fs_finish(int k, double M, int nrhs, const double* q, const double* p, const double* ya,
                   const double* aa, const double* logdet_half, const double* S, const double* dS,
                   const double* s2, const double* u, double* gS, double* out, void* stream);

/* out[0] = sum_i x_i y_i (f64, fixed summation order); ws of
 * rvgp_dot_ws_bytes(n) bytes.  The Frobenius products y^T a and |a|^2 of the
 * multi-right-hand-side NLML. */
size_t rvgp_dot_ws_bytes(size_t n);
int rvgp_dot(size_t n, const double* x, const double* y, double* out, void* ws, size_t ws_bytes,
             void* stream);

/* Determinant-lemma terms for the mixed-precision function-space form, from
 * the FP64 factor LA of A = s2 S^-1 + Phi^T Phi and diag(A^-1):
 * q_j = phi_j^T K_y^-1 phi_j and logdet_half = log det(K_y) / 2. */
int rvgp_lemma_terms(int k, double M, const double* S, const double* s2, const double* ainv_diag,
                     const double* LA, int lda, double* q, double* logdet_half, void* stream);

/* theta = softplus(u) + lower: kernel_theta (3,) and s2 (1,). */
int rvgp_gp_params(const double* u, const double* lower, double* kernel_theta, double* s2,
                   void* stream);

/* One Adam step on the 4 unconstrained parameters; state m, v (4,), t (1,);
 * trainable (4,) int flags. */
int rvgp_adam_step(double* u, const double* g, double* m, double* v, double* t,
                   const int* trainable, double lr, double b1, double b2, double eps,
                   void* stream);

/* One whole weight-space optimiser step of a single fit in one launch
 * (replaces, per step of the loop behind RVGP/main.py:87-95 -- GPflow's
 * GPR.training_loss, its TF gradient and the optimiser update -- the
 * sequence rvgp_gp_params, rvgp_spectral_filter, rvgp_ws_form, rvgp_potrf,
 * the triangular inverse, rvgp_colsumsq, GEMVs, rvgp_ws_finish and
 * rvgp_adam_step): from theta = softplus(u) + lower, the filter S of evals,
 * A = G + s2 diag(S)^-1 (G: k x k, leading dimension ldg), its Cholesky
 * factor and inverse, diag(A^-1), w = A^-1 b, and out (9) exactly as
 * rvgp_ws_finish for one right-hand side (yy = y^T y, M observations);
 * when do_adam, then the Adam update of u with state m1, m2 (4), t (1).
 * A tile dataflow of ceil(k/32)^2 + 2 ceil(k/32) + 3 workgroups that wait on each other
 * through flags in `sync` (rvgp_ws_step_sync_ints(k) ints): zero `sync`
 * once before the first call and keep it for later calls on the same
 * stream; calls that share a `sync` array must not overlap.  ws:
 * rvgp_ws_step_ws_bytes(k) bytes.  1 <= k <= RVGP_WS_STEP_MAX_K.
 * status[0]: RVGP_ST_OK, RVGP_ST_NOT_SPD, or RVGP_ST_TIMEOUT when a
 * bounded wait gave up (the results of that call are then invalid). */
#define RVGP_WS_STEP_MAX_K 2048
size_t rvgp_ws_step_ws_bytes(int k);
size_t rvgp_ws_step_sync_ints(int k);
int rvgp_ws_step(int k, int typ, const double* evals, double n_rows, const double* G, int ldg,
                 const double* b, const double* yy, double M, const double* lower, double* u,
                 const int* trainable, double* m1, double* m2, double* t, int do_adam, double lr,
                 double b1, double b2, double eps, void* ws, size_t ws_bytes, int* sync,
                 size_t sync_ints, double* out, int* status, void* stream);

/* ---------------------------------------------------------------------
 * Batched fits: F independent fields on one manifold, each with its own
 * training subset and hyperparameters.  The reference fits one field per
 * RVGP.fit call (RVGP/main.py:11-84 -> GPflow GPR + scipy optimiser,
 * main.py:87-95); these entry points run every fit of a step in one launch.
 * ------------------------------------------------------------------- */

/* G (F, k, k) f64: G_f = sum over the fit's nodes i and p < d of
 * U[i*d+p]^T U[i*d+p] (the weight-space Gram Phi_f^T Phi_f; the gauges have
 * orthonormal columns, so the intrinsic rows give the ambient Gram).
 * U (n*d, k) f64 (f64 = 1) or f32 (widened to f64 on the matrix cores);
 * the nodes of fit f are nodes[node_off[f] .. node_off[f+1]);
 * tile = 32, 64 or 128 (the autotuner's choice). */
int rvgp_batch_gram(int f64, int F, int k, int d, const void* U, int ldu, const int* node_off,
                    const int* nodes, int tile, double* G, void* stream);

/* b (F, k): b_f = sum_i U_i^T c_i with the intrinsic targets c (one d-row per
 * listed node, c_i = T_i^T y_i); yy (F): sum of ynorm2 (|y_i|^2 per listed
 * node) over the fit's nodes. */
int rvgp_batch_rhs(int F, int k, int d, const double* U, int ldu, const int* node_off,
                   const int* nodes, const double* c, const double* ynorm2, double* b, double* yy,
                   void* stream);

/* One optimiser step of all F fits in one persistent launch (a workgroup
 * per fit, grid = min(grid, F)): theta_f = softplus(u_f) + lower, spectral
 * filter of the shared evals, A_f = G_f + s2 diag(S)^-1, Cholesky of A_f with
 * L^-1 formed alongside, the weight-space NLML and adjoint of
 * rvgp_ws_finish into out (F, 9), and -- when do_adam -- one Adam update of
 * u_f with state m1, m2 (F, 4) and step counters t (F).  M (F): observations
 * per fit (ambient rows).  panel = 16 or 32 (Cholesky block width; needs
 * 8*(panel+1)*(k+panel) bytes of LDS <= 159 KiB), threads = 512 or 1024 per
 * fit workgroup (the autotuner's choice).  ws: rvgp_batch_step_ws_bytes
 * bytes; per fit it holds [L | L^-1] (k x 2k) then S, dS, gS, diag(A^-1),
 * L^-1 b, w = A^-1 b, G w.  status (F): RVGP_ST_NOT_SPD for a fit whose A
 * is not positive definite. */
size_t rvgp_batch_step_ws_bytes(int F, int k);
int rvgp_batch_step(int F, int k, int typ, const double* evals, double n_rows, const double* G,
                    const double* b, const double* yy, const double* M, const double* lower,
                    double* u, const int* trainable, double* m1, double* m2, double* t,
                    int do_adam, double lr, double b1, double b2, double eps, int panel,
                    int threads, int grid, double* ws, size_t ws_bytes, double* out, int* status,
                    void* stream);

/* ---------------------------------------------------------------------
 * kernel='rbf' branch of RVGP.fit (RVGP/main.py:33-37): gpflow.kernels.RBF
 * on evecs_L rows, the D vector components as D output columns of one GPR.
 * theta = (variance, lengthscale, noise) = softplus(u) + lower.
 * ------------------------------------------------------------------- */
int rvgp_rbf_params(const double* u, const double* lower, double* theta, void* stream);

/* K (M x N) = variance exp(-r^2 / (2 lengthscale^2)), r^2 = max(0, rn_i + cn_j
 * - 2 G_ij) from the Gram G = X X2^T and squared row norms (GPflow's
 * square_distance); sym: X2 = X (r_ii = 0); add_noise: + noise on the
 * diagonal. */
int rvgp_rbf_kernel(int M, int N, const double* G, int ldg, const double* rn, const double* cn,
                    const double* theta, int sym, int add_noise, double* K, int ldk,
                    void* stream);

/* Summed NLML over nrhs outputs and its adjoint, out (7) = [nlml,
 * d/d(variance, lengthscale, noise), d/du(...)], from Kinv = K_y^-1, AAt =
 * A A^T (A = K_y^-1 Y), the Gram G, squared norms sqn, ya = sum y^T K_y^-1 y
 * and logdet_half = sum log L_ii.  ws: rvgp_rbf_grad_ws_bytes(M) bytes. */
size_t rvgp_rbf_grad_ws_bytes(int M);
int rvgp_rbf_grad(int M, int nrhs, const double* Kinv, int ldk, const double* AAt, int lda,
                  const double* G, int ldg, const double* sqn, const double* theta, const double* ya,
                  const double* logdet_half, const double* u, double* ws, size_t ws_bytes,
                  double* out, void* stream);

/* ---------------------------------------------------------------------
 * n_inducing_points branch of RVGP.fit (RVGP/main.py:54-67): GPflow SGPR
 * with the manifold kernel.  -ELBO and its adjoint from the weight-space
 * statistics G (k x k), b (k), yy and GPflow's factors of the inducing rows
 * Z (m x k): Kuu = Z S Z^T + jitter I = L L^T, Q = L^-1 Z, R1 = Q S G,
 * B = I + R1 S Q^T / s2 = L_B L_B^T, Qb = L_B^-1 Q, R1b = L_B^-1 R1,
 * W1 = R1 S Q^T Q, cc = L_B^-1 Q S b, zw = Qb^T cc, hw = R1b^T cc; scal =
 * [sum log diag L_B, unused, |cc|^2]; S, dS of rvgp_spectral_filter; out (9)
 * as rvgp_ws_finish.  N = training rows. */
int rvgp_sgpr_finish(int m, int k, double N, const double* Q, const double* Qb, const double* R1,
                     const double* R1b, const double* W1, const double* G, const double* b,
                     const double* zw, const double* hw, const double* S, const double* dS,
                     const double* scal, const double* yy, double s2, const double* u, double* gS,
                     double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RVGP_H */
