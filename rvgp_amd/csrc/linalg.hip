// Blocked Cholesky and triangular solves (the linear algebra behind GPflow's
// GPR: tf.linalg.cholesky + triangular_solve, reached from
// RVGP/main.py:89-91 and 110-111).
//
// potrf: right-looking, NB = 64.  Per block column: (1) the diagonal block
// is factorised by one wavefront with its rows in registers, (2) the panel
// below is solved against it (one wavefront per panel row), (3) the lower
// triangle of the trailing matrix gets A22 -= A21 A21^T on the matrix cores.
// trsm: blocked right-looking substitution; the off-diagonal updates are
// GEMMs, the diagonal-block solves one wavefront per right-hand side.
#include "common.h"

// The library builds with -ffp-contract=off (the distance kernels need a
// fixed operation order); the substitutions accumulate with FMAs; results are compared with
// tolerances, never bit for bit.
#pragma clang fp contract(fast)

extern "C" int rvgp_gemm(int mode, int transA, int transB, int M, int N, int K, double alpha,
                         const void* A, int lda, const void* B, int ldb, const double* kscale,
                         double beta, void* C, int ldc, int split_k, void* ws, size_t ws_bytes,
                         void* stream);

namespace rvgp {

int gemm_syrk_lower(int f64, int n, int k, double alpha, const void* A, int lda, double beta,
                    void* C, int ldc, void* stream);

constexpr int NB = 64;

// Diagonal block, one wavefront: lane r holds row r of the block in
// registers; the column loop is fully unrolled, so every register index is a
// constant.  The entries of the scaled column that other lanes need are read
// from LDS as broadcasts (two per 16-byte read), except the look-ahead entry,
// which arrives by v_readlane.  The chain
// that bounds the kernel is pivot -> 1/sqrt -> scaled column -> next pivot,
// so column c+1 is updated first and its pivot's reciprocal square root is
// in flight while the rest of the rank-1 update issues.  Entries right of a
// lane's diagonal are scratch: a lane clears its entry c when column c is
// scaled, so nothing it accumulates there reaches the factor, and the update
// runs on every lane without a per-entry predicate.  Rows past nb are
// identity rows.
template <typename T>
__global__ __launch_bounds__(64) void potrf_diag(T* __restrict__ A, int lda, int j0, int nb,
                                                 int* __restrict__ status) {
    __shared__ T col[NB];
    const int r = threadIdx.x;
    T a[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        a[c] = (r < nb && c <= r) ? A[(size_t)(j0 + r) * lda + j0 + c] : (c == r ? T(1) : T(0));
    T piv = bcast(a[0], 0);
    bool bad = false;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        bad |= !(piv > T(0));
        const T inv = rsqrt(piv);
        a[c] = r == c ? piv * inv : (r > c ? a[c] * inv : T(0));
        if (c + 1 < NB) {
            // publish the scaled column; the workgroup is one wavefront, whose
            // LDS operations complete in program order, so the reads below see
            // this write and the next column's write follows these reads
            col[r] = a[c];
            a[c + 1] -= a[c] * bcast(a[c], c + 1);
            piv = bcast(a[c + 1], c + 1);
#pragma unroll
            for (int s = c + 2; s < NB; ++s) a[s] -= a[c] * col[s];
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (bad) {  // wave-uniform: every lane saw the same pivots
        if (r == 0) atomicMax(status, RVGP_ST_NOT_SPD);
        return;
    }
    if (r < nb) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c <= r) A[(size_t)(j0 + r) * lda + j0 + c] = a[c];
    }
}

// Solve the nb x nb diagonal block L11 (at Lm[i0][i0], lower triangular) for
// r right-hand sides: X = L11^-1 B (trans = 0) or L11^-T B (trans = 1), where
// element (a, c) of B is B[a * rs + c * cs] -- rs = ldb, cs = 1 for the rows
// [i0, i0+nb) of a row-major B, and rs = 1, cs = lda for the Cholesky panel,
// whose rows are right-hand sides (x L11^T = a  <=>  L11 x^T = a^T).
// One wavefront per right-hand side, lane a holding x_a: step a broadcasts
// x_a = b_a / L_aa (v_readlane) and the lanes below (above, transposed)
// subtract L_ja x_a -- a 64-step chain of readlane, multiply and FMA instead
// of a serial 2016-FMA substitution per lane.  Each lane keeps the entries of
// L11 it multiplies (its row, or its column when transposed) in registers,
// so the unrolled chain touches no memory; entries outside the triangle are
// zero, which leaves finished lanes unchanged without a predicate.  L11 is
// staged once per workgroup in LDS; 4 right-hand sides per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void trsm_diag(int trans, const T* __restrict__ Lm, int lda,
                                                 int i0, int nb, int r, T* __restrict__ B,
                                                 long long rs, long long cs,
                                                 const int* __restrict__ status) {
    if (status != nullptr && *status) return;
    __shared__ T L[NB][NB + 1];
    __shared__ T dinv[NB];
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * NB; e += 256) {  // whole tile: zeros outside the block
        int a = e / NB, b = e - a * NB;
        L[a][b] = (a < nb && b <= a) ? Lm[(size_t)(i0 + a) * lda + i0 + b] : T(0);
    }
    if (tid < NB) dinv[tid] = tid < nb ? T(1) / Lm[(size_t)(i0 + tid) * lda + i0 + tid] : T(1);
    __syncthreads();
    const int lane = tid & 63;
    const int c = blockIdx.x * 4 + (tid >> 6);
    if (c >= r) return;
    T* bc = B + (size_t)c * cs;
    T x = lane < nb ? bc[(size_t)lane * rs] : T(0);
    const T dv = dinv[lane];
    T l[NB];
    if (!trans) {
#pragma unroll
        for (int a = 0; a < NB; ++a) l[a] = L[lane][a];
#pragma unroll
        for (int a = 0; a < NB; ++a) {
            if (a < nb) {
                const T xa = bcast(x * dv, a);
                x = lane == a ? xa : x - l[a] * xa;
            }
        }
    } else {
#pragma unroll
        for (int a = 0; a < NB; ++a) l[a] = L[a][lane];
#pragma unroll
        for (int a = NB - 1; a >= 0; --a) {
            if (a < nb) {
                const T xa = bcast(x * dv, a);
                x = lane == a ? xa : x - l[a] * xa;
            }
        }
    }
    if (lane < nb) bc[(size_t)lane * rs] = x;
}

template <typename T>
__global__ void logdet_kernel(int n, const T* __restrict__ L, int lda, double* __restrict__ out) {
    __shared__ double part[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += log((double)L[(size_t)i * lda + i]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

template <typename T>
static int potrf_impl(int n, T* A, int lda, int* status, void* stream) {
    RVGP_CHECK_ARG(n >= 0 && lda >= n, "rvgp_potrf: bad n/lda");
    hipStream_t s = as_stream(stream);
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nb = min(NB, n - j0);
        hipLaunchKernelGGL(potrf_diag<T>, dim3(1), dim3(64), 0, s, A, lda, j0, nb, status);
        RVGP_LAUNCH_CHECK();
        const int rest = n - j0 - nb;
        if (rest <= 0) break;
        // panel: rows [j0+nb, n) of block column j0 solve x L11^T = a
        hipLaunchKernelGGL(trsm_diag<T>, dim3(cdiv(rest, 4)), dim3(256), 0, s, 0, A, lda, j0, nb,
                           rest, A + (size_t)(j0 + nb) * lda + j0, 1LL, (long long)lda, status);
        RVGP_LAUNCH_CHECK();
        T* A21 = A + (size_t)(j0 + nb) * lda + j0;
        T* A22 = A + (size_t)(j0 + nb) * lda + j0 + nb;
        int rc = gemm_syrk_lower(sizeof(T) == 8, rest, nb, -1.0, A21, lda, 1.0, A22, lda, stream);
        if (rc) return rc;
    }
    return RVGP_OK;
}

template <typename T>
static int trsm_impl(int trans, int n, int r, const T* L, int lda, T* B, int ldb, void* stream) {
    RVGP_CHECK_ARG(n >= 0 && r >= 0 && lda >= n && ldb >= r, "rvgp_trsm: bad sizes");
    if (n == 0 || r == 0) return RVGP_OK;
    hipStream_t s = as_stream(stream);
    const int mode = sizeof(T) == 8 ? RVGP_GEMM_F64 : RVGP_GEMM_F32;
    const int nblk = (n + NB - 1) / NB;
    for (int t = 0; t < nblk; ++t) {
        const int bi = trans ? nblk - 1 - t : t;
        const int i0 = bi * NB, nb = min(NB, n - i0);
        hipLaunchKernelGGL(trsm_diag<T>, dim3(cdiv(r, 4)), dim3(256), 0, s, trans, L, lda, i0, nb, r,
                           B + (size_t)i0 * ldb, (long long)ldb, 1LL, (const int*)nullptr);
        RVGP_LAUNCH_CHECK();
        // Right-looking update with the block just solved: every update has
        // K = NB and spans all remaining rows, so it fills many tiles.
        if (!trans && i0 + nb < n) {
            // B[i+nb:] -= L[i+nb:, i] X_i
            const int rest = n - i0 - nb;
            int rc = rvgp_gemm(mode, 0, 0, rest, r, nb, -1.0, L + (size_t)(i0 + nb) * lda + i0, lda,
                               B + (size_t)i0 * ldb, ldb, nullptr, 1.0,
                               B + (size_t)(i0 + nb) * ldb, ldb, 1, nullptr, 0, stream);
            if (rc) return rc;
        } else if (trans && i0 > 0) {
            // B[:i] -= L[i, :i]^T X_i
            int rc = rvgp_gemm(mode, 1, 0, i0, r, nb, -1.0, L + (size_t)i0 * lda, lda,
                               B + (size_t)i0 * ldb, ldb, nullptr, 1.0, B, ldb, 1, nullptr, 0, stream);
            if (rc) return rc;
        }
    }
    return RVGP_OK;
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_potrf(int f64, int n, void* A, int lda, int* status, void* stream) {
    return f64 ? potrf_impl<double>(n, static_cast<double*>(A), lda, status, stream)
               : potrf_impl<float>(n, static_cast<float*>(A), lda, status, stream);
}

extern "C" int rvgp_trsm(int f64, int trans, int n, int r, const void* L, int lda, void* B,
                         int ldb, void* stream) {
    return f64 ? trsm_impl<double>(trans, n, r, static_cast<const double*>(L), lda,
                                   static_cast<double*>(B), ldb, stream)
               : trsm_impl<float>(trans, n, r, static_cast<const float*>(L), lda,
                                  static_cast<float*>(B), ldb, stream);
}

extern "C" int rvgp_chol_logdet(int f64, int n, const void* L, int lda, double* out,
                                void* stream) {
    RVGP_CHECK_ARG(n >= 0 && lda >= n, "rvgp_chol_logdet: bad sizes");
    if (f64)
        hipLaunchKernelGGL(logdet_kernel<double>, dim3(1), dim3(256), 0, as_stream(stream), n,
                           static_cast<const double*>(L), lda, out);
    else
        hipLaunchKernelGGL(logdet_kernel<float>, dim3(1), dim3(256), 0, as_stream(stream), n,
                           static_cast<const float*>(L), lda, out);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
