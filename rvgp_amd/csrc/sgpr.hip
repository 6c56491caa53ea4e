// The n_inducing_points branch of RVGP.fit (RVGP/main.py:54-67): GPflow's
// SGPR (Titsias' collapsed bound) with the spectral kernel
// k(x, x') = x diag(S) x'^T, its negative bound and hand-written adjoint.
//
// With m inducing rows Z (m x k), N training rows X (N x k) reduced to
// G = X^T X, b = X^T y, yy = y^T y, s2 the noise variance, and GPflow's own
// well-conditioned factors (SGPR.elbo):
//     Kuu = Z S Z^T + jitter I = L L^T,  Q = L^-1 Z,  R1 = Q S G  (= L^-1 H,
//     H = Z S G),  A A^T = R1 S Q^T / s2,  B = I + A A^T = L_B L_B^T,
//     cc = L_B^-1 Q S b  (= sigma c),
// the bound is
//     ELBO = -N/2 log 2pi - sum log diag L_B - N/2 log s2 - yy/(2 s2)
//            + |cc|^2/(2 s2^2) - tr(S G)/(2 s2) + tr(A A^T)/2.
// Writing C = Kuu + Z S G S Z^T / s2 = L B L^T and w = C^-1 Z S b, its
// derivatives need, per feature j (columns of Q, R1, Qb = L_B^-1 Q,
// R1b = L_B^-1 R1, W1 = R1 S Q^T Q):
//   a1 = |Qb_j|^2 = z_j^T C^-1 z_j        a2 = Qb_j.R1b_j = z_j^T C^-1 h_j
//   a3 = |Q_j|^2  = z_j^T Kuu^-1 z_j      a4 = Q_j.R1_j   = z_j^T Kuu^-1 h_j
//   a5 = Q_j.W1_j = z_j^T Kuu^-1 M Kuu^-1 z_j   (M = H S Z^T)
//   zw_j = z_j^T w = Qb_j.cc,  hw_j = h_j^T w = R1b_j.cc
// and then
//   dELBO/dS_j = -(a1 + 2 a2 / s2)/2 + a3/2
//                + (2 b_j zw_j - zw_j^2 - 2 zw_j hw_j / s2) / (2 s2^2)
//                - G_jj / (2 s2) + (2 a4 - a5) / (2 s2)
//   dELBO/ds2  = tr(C^-1 M)/(2 s2^2) - N/(2 s2) + yy/(2 s2^2)
//                + (w^T M w / s2^4 - 2 |cc|^2 / s2^3)/2 + tr(S G)/(2 s2^2)
//                - tr(Kuu^-1 M)/(2 s2^2)
// with tr(C^-1 M) = sum_j S_j a2, tr(Kuu^-1 M) = s2 tr(A A^T) = sum_j S_j a4,
// w^T M w = sum_j S_j zw_j hw_j.  The NLML is -ELBO; gradients are mapped to
// (nu, kappa, sigma_f) through dS and to the unconstrained parameters as in
// gp.hip.
#include "common.h"

namespace rvgp {

constexpr int QT = 256;

__device__ double qt_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < QT / 64; ++i) s += sh[i];
    return s;
}

// Q, Qb, R1, R1b, W1: (m, k) row-major (ld k); G (k, k); b, zw, hw, S (k);
// dS (3, k); scal = [sum log diag L_B, unused, |cc|^2]; u (4); out (9) as
// gp.hip's finish kernels.
__global__ __launch_bounds__(QT) void sgpr_finish_kernel(
    int m, int k, double N, const double* __restrict__ Q, const double* __restrict__ Qb,
    const double* __restrict__ R1, const double* __restrict__ R1b, const double* __restrict__ W1,
    const double* __restrict__ G, const double* __restrict__ b, const double* __restrict__ zw,
    const double* __restrict__ hw, const double* __restrict__ S, const double* __restrict__ dS,
    const double* __restrict__ scal, const double* __restrict__ yy, double s2,
    const double* __restrict__ u, double* __restrict__ gS, double* __restrict__ out) {
    __shared__ double sh[QT / 64];
    double trSG = 0.0, trKM = 0.0, trCM = 0.0, wMw = 0.0;
    for (int j = threadIdx.x; j < k; j += QT) {
        double a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, a5 = 0.0;
        for (int i = 0; i < m; ++i) {
            const size_t e = (size_t)i * k + j;
            const double q = Q[e], qb = Qb[e], r = R1[e];
            a1 += qb * qb;
            a2 += qb * R1b[e];
            a3 += q * q;
            a4 += q * r;
            a5 += q * W1[e];
        }
        const double sj = S[j], gjj = G[(size_t)j * k + j];
        trSG += sj * gjj;
        trKM += sj * a4;
        trCM += sj * a2;
        wMw += sj * zw[j] * hw[j];
        const double zwj = zw[j];
        const double gel = -0.5 * (a1 + 2.0 * a2 / s2) + 0.5 * a3 +
                           0.5 * (2.0 * b[j] * zwj - zwj * zwj - 2.0 * zwj * hw[j] / s2) / (s2 * s2) -
                           0.5 * gjj / s2 + 0.5 * (2.0 * a4 - a5) / s2;
        gS[j] = -gel;  // d NLML / d S_j
    }
    trSG = qt_sum(trSG, sh);
    trKM = qt_sum(trKM, sh);
    trCM = qt_sum(trCM, sh);
    wMw = qt_sum(wMw, sh);
    const double ldB = scal[0], vw = scal[2], y2 = yy[0];
    const double elbo = -0.5 * N * log(6.283185307179586) - ldB - 0.5 * N * log(s2) -
                        0.5 * y2 / s2 + 0.5 * vw / (s2 * s2) - 0.5 * trSG / s2 + 0.5 * trKM / s2;
    const double ds2 = 0.5 * trCM / (s2 * s2) - 0.5 * N / s2 + 0.5 * y2 / (s2 * s2) +
                       0.5 * (wMw / (s2 * s2 * s2 * s2) - 2.0 * vw / (s2 * s2 * s2)) +
                       0.5 * trSG / (s2 * s2) - 0.5 * trKM / (s2 * s2);
    __syncthreads();
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int j = threadIdx.x; j < k; j += QT) {
        g0 += gS[j] * dS[j];
        g1 += gS[j] * dS[k + j];
        g2 += gS[j] * dS[2 * k + j];
    }
    g0 = qt_sum(g0, sh);
    g1 = qt_sum(g1, sh);
    g2 = qt_sum(g2, sh);
    if (threadIdx.x == 0) {
        out[0] = -elbo;
        const double g[4] = {g0, g1, g2, -ds2};
        for (int p = 0; p < 4; ++p) {
            out[1 + p] = g[p];
            out[5 + p] = g[p] / (1.0 + exp(-u[p]));
        }
    }
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_sgpr_finish(int m, int k, double N, const double* Q, const double* Qb,
                                const double* R1, const double* R1b, const double* W1,
                                const double* G, const double* b, const double* zw,
                                const double* hw, const double* S, const double* dS,
                                const double* scal, const double* yy, double s2, const double* u,
                                double* gS, double* out, void* stream) {
    RVGP_CHECK_ARG(m >= 1 && k >= 1, "rvgp_sgpr_finish: bad sizes");
    RVGP_CHECK_ARG(s2 > 0.0, "rvgp_sgpr_finish: noise variance must be positive");
    hipLaunchKernelGGL(sgpr_finish_kernel, dim3(1), dim3(QT), 0, as_stream(stream), m, k, N, Q, Qb,
                       R1, R1b, W1, G, b, zw, hw, S, dS, scal, yy, s2, u, gS, out);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
