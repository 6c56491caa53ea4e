// Rayleigh-Ritz bookkeeping of the eigensolver that stands in for
// scipy.sparse.linalg.eigsh in compute_spectrum (RVGP/geometry.py:66-80):
//  * Ritz residuals |A x_j - theta_j x_j| / |x_j| of a block, read once
//    (the residual block is never formed), sums in f64 in a fixed order;
//  * the eigenvector output pass: unit-norm columns scaled by sqrt(rows)
//    (geometry.py:75) and block rows mapped back from the locality
//    numbering of the operator, in one read and one write of the block.
#include "common.h"

namespace rvgp {
namespace {

constexpr int kMaxParts = 1024;

// 32-row slices, as the column sums of squares: every thread's loads of a
// slice are in flight together.
inline int res_parts(int M) {
    const int p = (M + 31) / 32;
    return p < 1 ? 1 : (p > kMaxParts ? kMaxParts : p);
}

// part[(slice * k + j) * 2 + {0, 1}] = sums over the slice's rows of
// (AX_ij - theta_j X_ij)^2 and X_ij^2.  Lane = column: every row is read
// coalesced.
template <typename T>
__global__ __launch_bounds__(256) void ritz_res_part_kernel(int M, int k, int parts,
                                                            const T* __restrict__ X, int ldx,
                                                            const T* __restrict__ AX, int ldax,
                                                            const double* __restrict__ theta,
                                                            double* __restrict__ part) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const int per = (M + parts - 1) / parts;
    const int i0 = blockIdx.y * per, i1 = min(M, i0 + per);
    const double th = theta[j];
    double sr[4] = {0.0, 0.0, 0.0, 0.0}, sx[4] = {0.0, 0.0, 0.0, 0.0};
    int i = i0;
    for (; i + 3 < i1; i += 4) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const double x = (double)X[(size_t)(i + t) * ldx + j];
            const double r = (double)AX[(size_t)(i + t) * ldax + j] - th * x;
            sr[t] += r * r;
            sx[t] += x * x;
        }
    }
    for (int t = 0; i < i1; ++i, ++t) {
        const double x = (double)X[(size_t)i * ldx + j];
        const double r = (double)AX[(size_t)i * ldax + j] - th * x;
        sr[t] += r * r;
        sx[t] += x * x;
    }
    double* o = part + ((size_t)blockIdx.y * k + j) * 2;
    o[0] = (sr[0] + sr[1]) + (sr[2] + sr[3]);
    o[1] = (sx[0] + sx[1]) + (sx[2] + sx[3]);
}

// res[j] = sqrt(sum_p r / sum_p x) over the slices: 16 wavefronts per 64
// columns, wavefront w summing the slices p = w (mod 16), the 16 partial
// sums then added in wavefront order (a fixed order).
__global__ __launch_bounds__(1024) void ritz_res_reduce_kernel(int k, int parts,
                                                               const double* __restrict__ part,
                                                               double* __restrict__ res) {
    __shared__ double sr[16][64], sx[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    // four slices in flight per step: the loads do not wait on each other
    double ar[4] = {0.0, 0.0, 0.0, 0.0}, ax[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < k) {
        int q = w;
        for (; q + 48 < parts; q += 64) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                ar[t] += part[((size_t)(q + 16 * t) * k + j) * 2];
                ax[t] += part[((size_t)(q + 16 * t) * k + j) * 2 + 1];
            }
        }
        for (int t = 0; q < parts; q += 16, ++t) {
            ar[t] += part[((size_t)q * k + j) * 2];
            ax[t] += part[((size_t)q * k + j) * 2 + 1];
        }
    }
    sr[w][lane] = (ar[0] + ar[1]) + (ar[2] + ar[3]);
    sx[w][lane] = (ax[0] + ax[1]) + (ax[2] + ax[3]);
    __syncthreads();
    if (w == 0 && j < k) {
        double tr = 0.0, tx = 0.0;
        for (int q = 0; q < 16; ++q) {
            tr += sr[q][lane];
            tx += sx[q][lane];
        }
        res[j] = tx > 0.0 ? sqrt(tr / tx) : 0.0;
    }
}

// out[perm[i] * b + a][j] = scale * X[i * b + a][j] / sqrt(nrm2[j]): one
// wavefront per source row, lanes over the columns.
template <typename T>
__global__ __launch_bounds__(256) void evecs_finish_kernel(int nrows, int b, int k,
                                                           const T* __restrict__ X, int ldx,
                                                           const double* __restrict__ nrm2,
                                                           const long long* __restrict__ perm,
                                                           double scale, double* __restrict__ out,
                                                           int ldo) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrows) return;
    const int lane = threadIdx.x & 63;
    const int blk = row / b, a = row - blk * b;
    const long long dst = perm != nullptr ? perm[blk] * b + a : (long long)row;
    const T* src = X + (size_t)row * ldx;
    double* o = out + (size_t)dst * ldo;
    for (int j = lane; j < k; j += 64) o[j] = (double)src[j] * (scale / sqrt(nrm2[j]));
}

}  // namespace
}  // namespace rvgp

using namespace rvgp;

extern "C" size_t rvgp_ritz_residuals_ws_bytes(int M, int k) {
    return 2 * sizeof(double) * (size_t)res_parts(M) * (size_t)(k > 0 ? k : 0);
}

extern "C" int rvgp_ritz_residuals(int f64, int M, int k, const void* X, int ldx, const void* AX,
                                   int ldax, const double* theta, double* res, void* ws,
                                   size_t ws_bytes, void* stream) {
    RVGP_CHECK_ARG(M >= 1 && k >= 1 && ldx >= k && ldax >= k, "rvgp_ritz_residuals: bad sizes");
    RVGP_CHECK_ARG(X != nullptr && AX != nullptr && theta != nullptr && res != nullptr,
                   "rvgp_ritz_residuals: null pointer");
    RVGP_CHECK_ARG(ws != nullptr && ws_bytes >= rvgp_ritz_residuals_ws_bytes(M, k),
                   "rvgp_ritz_residuals: workspace too small");
    const int parts = res_parts(M);
    double* part = static_cast<double*>(ws);
    dim3 g(cdiv(k, 256), parts);
    if (f64)
        hipLaunchKernelGGL(ritz_res_part_kernel<double>, g, dim3(256), 0, as_stream(stream), M, k,
                           parts, static_cast<const double*>(X), ldx,
                           static_cast<const double*>(AX), ldax, theta, part);
    else
        hipLaunchKernelGGL(ritz_res_part_kernel<float>, g, dim3(256), 0, as_stream(stream), M, k,
                           parts, static_cast<const float*>(X), ldx,
                           static_cast<const float*>(AX), ldax, theta, part);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(ritz_res_reduce_kernel, dim3(cdiv(k, 64)), dim3(1024), 0, as_stream(stream),
                       k, parts, part, res);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_evecs_finish(int f64, int nb, int b, int k, const void* X, int ldx,
                                 const double* nrm2, const long long* perm, double scale,
                                 double* out, int ldo, void* stream) {
    RVGP_CHECK_ARG(nb >= 1 && b >= 1 && b <= 3 && k >= 1 && ldx >= k && ldo >= k,
                   "rvgp_evecs_finish: bad sizes");
    RVGP_CHECK_ARG((long long)nb * b <= 2147483647LL, "rvgp_evecs_finish: too many rows");
    RVGP_CHECK_ARG(X != nullptr && nrm2 != nullptr && out != nullptr,
                   "rvgp_evecs_finish: null pointer");
    RVGP_CHECK_ARG(X != static_cast<const void*>(out), "rvgp_evecs_finish: out must not alias X");
    const int nrows = nb * b;
    const dim3 g(cdiv(nrows, 4));
    if (f64)
        hipLaunchKernelGGL(evecs_finish_kernel<double>, g, dim3(256), 0, as_stream(stream), nrows,
                           b, k, static_cast<const double*>(X), ldx, nrm2, perm, scale, out, ldo);
    else
        hipLaunchKernelGGL(evecs_finish_kernel<float>, g, dim3(256), 0, as_stream(stream), nrows, b,
                           k, static_cast<const float*>(X), ldx, nrm2, perm, scale, out, ldo);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
