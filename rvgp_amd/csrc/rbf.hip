// The kernel='rbf' branch of RVGP.fit (RVGP/main.py:33-37): a squared-
// exponential kernel (gpflow.kernels.RBF) on the rows of the scalar
// Laplacian's eigenvectors, the D vector components as D output columns of
// one GPR with shared hyperparameters (variance, lengthscale, noise).
//
//   rvgp_rbf_params   theta = softplus(u) + lower for the 3 parameters
//   rvgp_rbf_kernel   K = variance exp(-r^2 / (2 l^2)) (+ noise on the
//                     diagonal) from the Gram G = X X2^T (the matrix cores,
//                     rvgp_gemm) and the squared row norms:
//                     r^2 = max(0, |x|^2 + |x'|^2 - 2 x.x'), as GPflow's
//                     square_distance forms it
//   rvgp_rbf_grad     the adjoint 1/2 tr((R K^-1 - A A^T) dK/dtheta) of the
//                     summed NLML, A = K^-1 Y, with dK/dvariance = K_f /
//                     variance, dK/dl = K_f r^2 / l^3, dK/dnoise = I --
//                     K_f and r^2 recomputed from G, never stored
//   rvgp_rbf_finish   NLML and the gradient w.r.t. theta and u
#include "common.h"

namespace rvgp {

constexpr int RT = 256;

__global__ void rbf_params_kernel(const double* __restrict__ u, const double* __restrict__ lower,
                                  double* __restrict__ theta) {
    const int p = threadIdx.x;
    if (p >= 3) return;
    const double x = u[p];
    const double sp = x > 0 ? x + log1p(exp(-x)) : log1p(exp(x));
    theta[p] = sp + lower[p];
}

__global__ void rbf_kernel_kernel(int M, int N, const double* __restrict__ G, int ldg,
                                  const double* __restrict__ rn, const double* __restrict__ cn,
                                  const double* __restrict__ theta, int sym, int add_noise,
                                  double* __restrict__ K, int ldk) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)M * N) return;
    const int i = (int)(e / N), j = (int)(e - (size_t)i * N);
    const double var = theta[0], ls = theta[1];
    double r2 = rn[i] + cn[j] - 2.0 * G[(size_t)i * ldg + j];
    if (r2 < 0.0 || (sym && i == j)) r2 = 0.0;
    double k = var * exp(-0.5 * r2 / (ls * ls));
    if (add_noise && i == j) k += theta[2];
    K[(size_t)i * ldk + j] = k;
}

// Partial sums of W K_f, W K_f r^2 and diag(W), W = R Kinv - AAt: one
// workgroup per row i writes its row's triple; rbf_finish_kernel sums the
// rows in a fixed order.
__global__ __launch_bounds__(RT) void rbf_grad_part_kernel(int M, double R,
                                                           const double* __restrict__ Kinv, int ldk,
                                                           const double* __restrict__ AAt, int lda,
                                                           const double* __restrict__ G, int ldg,
                                                           const double* __restrict__ sqn,
                                                           const double* __restrict__ theta,
                                                           double* __restrict__ part) {
    __shared__ double sh[3][RT / 64];
    const int i = blockIdx.x;
    const double var = theta[0], ls = theta[1];
    double sv = 0.0, sl = 0.0, st = 0.0;
    for (int j = threadIdx.x; j < M; j += RT) {
        double r2 = sqn[i] + sqn[j] - 2.0 * G[(size_t)i * ldg + j];
        if (r2 < 0.0 || i == j) r2 = 0.0;
        const double kf = var * exp(-0.5 * r2 / (ls * ls));
        const double w = R * Kinv[(size_t)i * ldk + j] - AAt[(size_t)i * lda + j];
        sv += w * kf;
        sl += w * kf * r2;
        if (i == j) st += w;
    }
    sv = wave_sum(sv);
    sl = wave_sum(sl);
    st = wave_sum(st);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = sv;
        sh[1][w] = sl;
        sh[2][w] = st;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0.0;
        for (int q = 0; q < RT / 64; ++q) s += sh[threadIdx.x][q];
        part[(size_t)3 * i + threadIdx.x] = s;
    }
}

// out (7) = [nlml, d/d(variance, lengthscale, noise), d/du(...)]
__global__ __launch_bounds__(RT) void rbf_finish_kernel(int M, double R,
                                                        const double* __restrict__ part,
                                                        const double* __restrict__ ya,
                                                        const double* __restrict__ ldh,
                                                        const double* __restrict__ theta,
                                                        const double* __restrict__ u,
                                                        double* __restrict__ out) {
    __shared__ double sh[3][RT / 64];
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < M; i += RT)
        for (int q = 0; q < 3; ++q) s[q] += part[(size_t)3 * i + q];
    for (int q = 0; q < 3; ++q) {
        const double v = wave_sum(s[q]);
        if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[3];
        for (int q = 0; q < 3; ++q) {
            t[q] = 0.0;
            for (int w = 0; w < RT / 64; ++w) t[q] += sh[q][w];
        }
        const double var = theta[0], ls = theta[1];
        const double g[3] = {0.5 * t[0] / var, 0.5 * t[1] / (ls * ls * ls), 0.5 * t[2]};
        out[0] = 0.5 * ya[0] + R * ldh[0] + 0.5 * R * M * log(6.283185307179586);
        for (int p = 0; p < 3; ++p) {
            out[1 + p] = g[p];
            out[4 + p] = g[p] / (1.0 + exp(-u[p]));  // dtheta/du = sigmoid(u)
        }
    }
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_rbf_params(const double* u, const double* lower, double* theta, void* stream) {
    hipLaunchKernelGGL(rbf_params_kernel, dim3(1), dim3(64), 0, as_stream(stream), u, lower, theta);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_rbf_kernel(int M, int N, const double* G, int ldg, const double* rn,
                               const double* cn, const double* theta, int sym, int add_noise,
                               double* K, int ldk, void* stream) {
    RVGP_CHECK_ARG(M >= 1 && N >= 1 && ldg >= N && ldk >= N, "rvgp_rbf_kernel: bad sizes");
    RVGP_CHECK_ARG(!sym || M == N, "rvgp_rbf_kernel: a symmetric kernel matrix must be square");
    hipLaunchKernelGGL(rbf_kernel_kernel, dim3(cdiv((size_t)M * N, 256)), dim3(256), 0,
                       as_stream(stream), M, N, G, ldg, rn, cn, theta, sym, add_noise, K, ldk);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" size_t rvgp_rbf_grad_ws_bytes(int M) { return sizeof(double) * 3 * (size_t)(M > 0 ? M : 0); }

extern "C" int rvgp_rbf_grad(int M, int nrhs, const double* Kinv, int ldk, const double* AAt,
                             int lda, const double* G, int ldg, const double* sqn,
                             const double* theta, const double* ya, const double* logdet_half,
                             const double* u, double* ws, size_t ws_bytes, double* out,
                             void* stream) {
    RVGP_CHECK_ARG(M >= 1 && nrhs >= 1 && ldk >= M && lda >= M && ldg >= M,
                   "rvgp_rbf_grad: bad sizes");
    RVGP_CHECK_ARG(ws_bytes >= rvgp_rbf_grad_ws_bytes(M), "rvgp_rbf_grad: workspace too small");
    hipLaunchKernelGGL(rbf_grad_part_kernel, dim3(M), dim3(RT), 0, as_stream(stream), M,
                       (double)nrhs, Kinv, ldk, AAt, lda, G, ldg, sqn, theta, ws);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(rbf_finish_kernel, dim3(1), dim3(RT), 0, as_stream(stream), M, (double)nrhs,
                       ws, ya, logdet_half, theta, u, out);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
