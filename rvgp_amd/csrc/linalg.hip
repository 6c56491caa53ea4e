// Blocked Cholesky and triangular solves (the linear algebra behind GPflow's
// GPR: tf.linalg.cholesky + triangular_solve, reached from
// RVGP/main.py:89-91 and 110-111).
//
// potrf: right-looking, NB = 64.  Per block column: (1) the diagonal block
// is factorised by four wavefronts with its rows in registers, (2) the panel
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

// Diagonal block, one workgroup of four wavefronts: lane r holds row r of
// the block and wave w holds columns [16w, 16w+16) of it in registers, every
// register index a compile-time constant (the panel loop is unrolled).  The
// block is factored in panels of four columns: the wave owning the panel
// factors it in registers -- a four-step pivot chain, the entries other
// lanes need arriving by v_readlane -- and publishes it through LDS (double
// buffered); after one barrier every wave applies the rank-4 update to its
// own columns right of the panel (the owner's next panel first).  Per panel
// the chain is four pivots plus one LDS round trip, instead of one LDS round
// trip per column.  Entries right of a lane's diagonal are scratch: a lane
// clears its entry c when column c is scaled, so nothing accumulated there
// reaches the factor.  Rows past nb are identity rows.
template <typename T>
__global__ __launch_bounds__(256) void potrf_diag(T* __restrict__ A, int lda, int j0, int nb,
                                                  int* __restrict__ status) {
    __shared__ T P[2][NB][4];
    __shared__ int bad_any;
    const int r = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (threadIdx.x == 0) bad_any = 0;
    T a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int c = 16 * w + i;
        a[i] = (r < nb && c <= r) ? A[(size_t)(j0 + r) * lda + j0 + c] : (c == r ? T(1) : T(0));
    }
    bool bad = false;
#pragma unroll
    for (int p = 0; p < NB / 4; ++p) {
        const int lc = (p & 3) * 4;  // the panel's first column in its owner's slice
        if (w == (p >> 2)) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = 4 * p + j;
                const T piv = bcast(a[lc + j], c);
                bad |= !(piv > T(0));
                const T inv = rsqrt(piv);
                a[lc + j] = r == c ? piv * inv : (r > c ? a[lc + j] * inv : T(0));
#pragma unroll
                for (int k = j + 1; k < 4; ++k) a[lc + k] -= a[lc + j] * bcast(a[lc + j], 4 * p + k);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) P[p & 1][r][j] = a[lc + j];
            if (bad && r == 0) bad_any = 1;
        }
        __syncthreads();
        if (p + 1 < NB / 4) {
            const T pr0 = P[p & 1][r][0], pr1 = P[p & 1][r][1];
            const T pr2 = P[p & 1][r][2], pr3 = P[p & 1][r][3];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                if (16 * w + i > 4 * p + 3) {  // wave-uniform
                    const T* ps = P[p & 1][16 * w + i];
                    a[i] -= pr0 * ps[0] + pr1 * ps[1] + pr2 * ps[2] + pr3 * ps[3];
                }
            }
        }
    }
    __syncthreads();
    if (bad_any) {
        if (threadIdx.x == 0) atomicMax(status, RVGP_ST_NOT_SPD);
        return;
    }
    if (r < nb) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int c = 16 * w + i;
            if (c <= r) A[(size_t)(j0 + r) * lda + j0 + c] = a[i];
        }
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
    __shared__ T L[NB][NB + 1];
    __shared__ T dinv[NB];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int c = blockIdx.x * 4 + (tid >> 6);
    // every global load (status word, tile, this wave's right-hand side) is
    // issued before the first wait, so they cost one memory latency together
    const int st = status != nullptr ? *status : 0;
    T* bc = B + (size_t)c * cs;
    const T x0 = c < r && lane < nb ? bc[(size_t)lane * rs] : T(0);
    T t[NB * NB / 256];
#pragma unroll
    for (int q = 0; q < NB * NB / 256; ++q) {  // whole tile: zeros outside the block
        const int e = tid + 256 * q, a = e / NB, b = e - a * NB;
        t[q] = (a < nb && b <= a) ? Lm[(size_t)(i0 + a) * lda + i0 + b] : T(0);
    }
    const T ld = tid < nb ? Lm[(size_t)(i0 + tid) * lda + i0 + tid] : T(1);
    if (st) return;  // the same word for the whole grid
#pragma unroll
    for (int q = 0; q < NB * NB / 256; ++q) {
        const int e = tid + 256 * q, a = e / NB, b = e - a * NB;
        L[a][b] = t[q];
    }
    if (tid < NB) dinv[tid] = T(1) / ld;
    __syncthreads();
    if (c >= r) return;
    T x = x0;
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
        hipLaunchKernelGGL(potrf_diag<T>, dim3(1), dim3(256), 0, s, A, lda, j0, nb, status);
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
