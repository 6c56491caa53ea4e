// Blocked Cholesky and triangular solves (the linear algebra behind GPflow's
// GPR: tf.linalg.cholesky + triangular_solve, reached from
// RVGP/main.py:89-91 and 110-111).
//
// potrf: right-looking, NB = 64.  Per block column: (1) the diagonal block
// is factorised by one workgroup in LDS, (2) the panel below is solved
// against it (one workgroup per 64 rows, L11 in LDS), (3) the trailing
// matrix gets A22 -= A21 A21^T on the matrix cores (rvgp_gemm).
// trsm: blocked substitution; the off-diagonal updates are GEMMs, the
// diagonal-block solves one lane per right-hand-side column.
#include "common.h"

extern "C" int rvgp_gemm(int mode, int transA, int transB, int M, int N, int K, double alpha,
                         const void* A, int lda, const void* B, int ldb, const double* kscale,
                         double beta, void* C, int ldc, int split_k, void* ws, size_t ws_bytes,
                         void* stream);

namespace rvgp {

constexpr int NB = 64;

template <typename T>
__global__ __launch_bounds__(256) void potrf_diag(T* __restrict__ A, int lda, int j0, int nb,
                                                  int* __restrict__ status) {
    __shared__ T L[NB][NB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < nb * nb; e += 256) {
        int r = e / nb, c = e - r * nb;
        L[r][c] = A[(size_t)(j0 + r) * lda + j0 + c];
    }
    __syncthreads();
    __shared__ int bad;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int c = 0; c < nb; ++c) {
        const T piv = L[c][c];
        if (!(piv > T(0))) {
            if (tid == 0) bad = 1;
        }
        __syncthreads();
        if (bad) break;
        const T lcc = sqrt(piv);
        __syncthreads();
        if (tid == 0) L[c][c] = lcc;
        for (int r = c + 1 + tid; r < nb; r += 256) L[r][c] /= lcc;
        __syncthreads();
        const int rem = nb - c - 1;
        for (int e = tid; e < rem * rem; e += 256) {
            int r = c + 1 + e / rem, s = c + 1 + e % rem;
            if (s <= r) L[r][s] -= L[r][c] * L[s][c];
        }
        __syncthreads();
    }
    if (bad) {
        if (tid == 0) atomicMax(status, RVGP_ST_NOT_SPD);
        return;
    }
    for (int e = tid; e < nb * nb; e += 256) {
        int r = e / nb, c = e - r * nb;
        if (c <= r) A[(size_t)(j0 + r) * lda + j0 + c] = L[r][c];
    }
}

// Rows i in [j0+nb, n): x L11^T = a  (forward substitution along the row).
template <typename T>
__global__ __launch_bounds__(64) void potrf_panel(T* __restrict__ A, int lda, int n, int j0,
                                                  int nb, const int* __restrict__ status) {
    if (*status) return;
    __shared__ T L[NB][NB + 1];
    __shared__ T X[64][NB + 1];
    const int tid = threadIdx.x;
    const int r0 = j0 + nb + blockIdx.x * 64;
    for (int e = tid; e < nb * nb; e += 64) {
        int r = e / nb, c = e - r * nb;
        L[r][c] = A[(size_t)(j0 + r) * lda + j0 + c];
    }
    for (int e = tid; e < 64 * nb; e += 64) {
        int r = e / nb, c = e - r * nb;
        X[r][c] = (r0 + r < n) ? A[(size_t)(r0 + r) * lda + j0 + c] : T(0);
    }
    __syncthreads();
    if (r0 + tid < n) {
        for (int c = 0; c < nb; ++c) {
            T s = X[tid][c];
            for (int t = 0; t < c; ++t) s -= X[tid][t] * L[c][t];
            X[tid][c] = s / L[c][c];
        }
    }
    __syncthreads();
    for (int e = tid; e < 64 * nb; e += 64) {
        int r = e / nb, c = e - r * nb;
        if (r0 + r < n) A[(size_t)(r0 + r) * lda + j0 + c] = X[r][c];
    }
}

// Solve the nb x nb diagonal block for all r columns of B rows [i0, i0+nb).
template <typename T>
__global__ __launch_bounds__(64) void trsm_diag(int trans, const T* __restrict__ Lm, int lda,
                                                int i0, int nb, int r, T* __restrict__ B,
                                                int ldb) {
    __shared__ T L[NB][NB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < nb * nb; e += 64) {
        int a = e / nb, b = e - a * nb;
        L[a][b] = (b <= a) ? Lm[(size_t)(i0 + a) * lda + i0 + b] : T(0);
    }
    __syncthreads();
    const int c = blockIdx.x * 64 + tid;
    if (c >= r) return;
    T x[NB];
    for (int a = 0; a < nb; ++a) x[a] = B[(size_t)(i0 + a) * ldb + c];
    if (!trans) {
        for (int a = 0; a < nb; ++a) {
            T s = x[a];
            for (int b = 0; b < a; ++b) s -= L[a][b] * x[b];
            x[a] = s / L[a][a];
        }
    } else {
        for (int a = nb - 1; a >= 0; --a) {
            T s = x[a];
            for (int b = a + 1; b < nb; ++b) s -= L[b][a] * x[b];
            x[a] = s / L[a][a];
        }
    }
    for (int a = 0; a < nb; ++a) B[(size_t)(i0 + a) * ldb + c] = x[a];
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
    const int mode = sizeof(T) == 8 ? RVGP_GEMM_F64 : RVGP_GEMM_F32;
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nb = min(NB, n - j0);
        hipLaunchKernelGGL(potrf_diag<T>, dim3(1), dim3(256), 0, s, A, lda, j0, nb, status);
        RVGP_LAUNCH_CHECK();
        const int rest = n - j0 - nb;
        if (rest <= 0) break;
        hipLaunchKernelGGL(potrf_panel<T>, dim3(cdiv(rest, 64)), dim3(64), 0, s, A, lda, n, j0, nb,
                           status);
        RVGP_LAUNCH_CHECK();
        T* A21 = A + (size_t)(j0 + nb) * lda + j0;
        T* A22 = A + (size_t)(j0 + nb) * lda + j0 + nb;
        int rc = rvgp_gemm(mode, 0, 1, rest, rest, nb, -1.0, A21, lda, A21, lda, nullptr, 1.0, A22,
                           lda, 1, nullptr, 0, stream);
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
        if (!trans && i0 > 0) {
            // B_i -= L[i, :i] X[:i]
            int rc = rvgp_gemm(mode, 0, 0, nb, r, i0, -1.0, L + (size_t)i0 * lda, lda, B, ldb,
                               nullptr, 1.0, B + (size_t)i0 * ldb, ldb, 1, nullptr, 0, stream);
            if (rc) return rc;
        } else if (trans && i0 + nb < n) {
            // B_i -= L[i+nb:, i]^T X[i+nb:]
            const int rest = n - i0 - nb;
            int rc = rvgp_gemm(mode, 1, 0, nb, r, rest, -1.0, L + (size_t)(i0 + nb) * lda + i0, lda,
                               B + (size_t)(i0 + nb) * ldb, ldb, nullptr, 1.0,
                               B + (size_t)i0 * ldb, ldb, 1, nullptr, 0, stream);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(trsm_diag<T>, dim3(cdiv(r, 64)), dim3(64), 0, s, trans, L, lda, i0, nb, r,
                           B, ldb);
        RVGP_LAUNCH_CHECK();
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
