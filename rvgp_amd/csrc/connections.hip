// a5/a6/a7  Connections (RVGP/lib/ptu_dijkstra.pyx:128-293) and the blocks of
// the scalar / connection Laplacians (RVGP/geometry.py:14-63).
//
// One lane per CSR entry e = (i, j), self loops included.  M = T_i^T T_j is
// a d x d product of two gathered gauge rows; R_e is its orthogonal polar
// factor, which is what U V^T of LAPACK's dgesvd is for a nonsingular M:
//   d = 1: sign(m);
//   d = 2: closed form -- the rotation (det M > 0) or reflection (det M < 0)
//          maximising tr(R^T M), normalised without trigonometry;
//   d >= 3: Higham's scaled Newton iteration X <- (g X + X^{-T}/g)/2, with
//          the inverse by cofactors for d = 3 and Newton-Schulz for d > 3.
// The lane then writes R_e, and the Laplacian entries L_ij R_e
// (L_ii = degree, L_ij = -1) in f64 and f32 block-CSR form, so the
// connection Laplacian is assembled without a COO round trip.
#include "common.h"

namespace rvgp {

template <int DD>
__device__ void polar_newton_schulz(double (&X)[DD][DD]) {
    // scale to spectral norm < sqrt(3) via the Frobenius norm
    double f = 0.0;
#pragma unroll
    for (int a = 0; a < DD; ++a)
#pragma unroll
        for (int b = 0; b < DD; ++b) f += X[a][b] * X[a][b];
    f = sqrt(f);
    if (f == 0.0) return;
#pragma unroll
    for (int a = 0; a < DD; ++a)
#pragma unroll
        for (int b = 0; b < DD; ++b) X[a][b] /= f;
    for (int it = 0; it < 200; ++it) {
        double P[DD][DD], Y[DD][DD];
#pragma unroll
        for (int a = 0; a < DD; ++a)
#pragma unroll
            for (int b = 0; b < DD; ++b) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < DD; ++k) s += X[k][a] * X[k][b];
                P[a][b] = (a == b ? 3.0 : 0.0) - s;
            }
        double diff = 0.0;
#pragma unroll
        for (int a = 0; a < DD; ++a)
#pragma unroll
            for (int b = 0; b < DD; ++b) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < DD; ++k) s += X[a][k] * P[k][b];
                Y[a][b] = 0.5 * s;
                diff += (Y[a][b] - X[a][b]) * (Y[a][b] - X[a][b]);
            }
#pragma unroll
        for (int a = 0; a < DD; ++a)
#pragma unroll
            for (int b = 0; b < DD; ++b) X[a][b] = Y[a][b];
        if (diff < 1e-30) break;
    }
}

__device__ void polar3(double (&X)[3][3]) {
    for (int it = 0; it < 60; ++it) {
        // cofactor matrix C; X^{-T} = C / det
        double C[3][3];
        C[0][0] = X[1][1] * X[2][2] - X[1][2] * X[2][1];
        C[0][1] = X[1][2] * X[2][0] - X[1][0] * X[2][2];
        C[0][2] = X[1][0] * X[2][1] - X[1][1] * X[2][0];
        C[1][0] = X[0][2] * X[2][1] - X[0][1] * X[2][2];
        C[1][1] = X[0][0] * X[2][2] - X[0][2] * X[2][0];
        C[1][2] = X[0][1] * X[2][0] - X[0][0] * X[2][1];
        C[2][0] = X[0][1] * X[1][2] - X[0][2] * X[1][1];
        C[2][1] = X[0][2] * X[1][0] - X[0][0] * X[1][2];
        C[2][2] = X[0][0] * X[1][1] - X[0][1] * X[1][0];
        double det = X[0][0] * C[0][0] + X[0][1] * C[0][1] + X[0][2] * C[0][2];
        if (det == 0.0) return;  // singular: leave as is (measure-zero case)
        double nx = 0.0, ni = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                nx += X[a][b] * X[a][b];
                ni += C[a][b] * C[a][b];
            }
        ni /= det * det;
        // Higham's scaling g = (|X^-1|_F / |X|_F)^(1/2), dropped near convergence
        double g = sqrt(sqrt(ni / nx));
        if (fabs(g - 1.0) < 1e-3) g = 1.0;
        double diff = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double y = 0.5 * (g * X[a][b] + C[a][b] / (g * det));
                diff += (y - X[a][b]) * (y - X[a][b]);
                X[a][b] = y;
            }
        if (diff < 1e-30 && g == 1.0) break;
    }
}

template <int DD>
__global__ __launch_bounds__(256) void connections_kernel(
    const double* __restrict__ T, int n, int D, int ldt, const int* __restrict__ rowptr,
    const int* __restrict__ cols, int nnz, double* __restrict__ R, double* __restrict__ lc64,
    float* __restrict__ lc32, double* __restrict__ l64, float* __restrict__ l32) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    // row of e: last i with rowptr[i] <= e
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (rowptr[mid] <= e) lo = mid; else hi = mid;
    }
    const int i = lo, j = cols[e];
    const double* Ti = T + (size_t)i * D * ldt;
    const double* Tj = T + (size_t)j * D * ldt;
    double X[DD][DD];
#pragma unroll
    for (int p = 0; p < DD; ++p)
#pragma unroll
        for (int q = 0; q < DD; ++q) X[p][q] = 0.0;
    for (int k = 0; k < D; ++k) {
        double a[DD], b[DD];
#pragma unroll
        for (int p = 0; p < DD; ++p) {
            a[p] = Ti[k * ldt + p];
            b[p] = Tj[k * ldt + p];
        }
#pragma unroll
        for (int p = 0; p < DD; ++p)
#pragma unroll
            for (int q = 0; q < DD; ++q) X[p][q] = __fma_rn(a[p], b[q], X[p][q]);
    }
    if constexpr (DD == 1) {
        X[0][0] = X[0][0] < 0.0 ? -1.0 : 1.0;
    } else if constexpr (DD == 2) {
        const double a = X[0][0], b = X[0][1], c = X[1][0], d = X[1][1];
        if (a * d - b * c >= 0.0) {  // rotation [[c0,-s0],[s0,c0]]
            double u = a + d, v = c - b, h = hypot(u, v);
            double c0 = h > 0.0 ? u / h : 1.0, s0 = h > 0.0 ? v / h : 0.0;
            X[0][0] = c0; X[0][1] = -s0; X[1][0] = s0; X[1][1] = c0;
        } else {  // reflection [[c0,s0],[s0,-c0]]
            double u = a - d, v = b + c, h = hypot(u, v);
            double c0 = u / h, s0 = v / h;
            X[0][0] = c0; X[0][1] = s0; X[1][0] = s0; X[1][1] = -c0;
        }
    } else if constexpr (DD == 3) {
        polar3(X);
    } else {
        polar_newton_schulz<DD>(X);
    }
    const int deg = rowptr[i + 1] - rowptr[i] - 1;
    const double lv = (i == j) ? (double)deg : -1.0;
#pragma unroll
    for (int p = 0; p < DD; ++p)
#pragma unroll
        for (int q = 0; q < DD; ++q) {
            const size_t o = (size_t)e * DD * DD + p * DD + q;
            if (R) R[o] = X[p][q];
            const double v = lv * X[p][q];
            if (lc64) lc64[o] = v;
            if (lc32) lc32[o] = (float)v;
        }
    if (l64) l64[e] = lv;
    if (l32) l32[e] = (float)lv;
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_connections(const double* tangents, int n, int D, int ldt, int d,
                                const int* rowptr, const int* cols, int nnz, double* R,
                                double* lc64, float* lc32, double* l64, float* l32,
                                void* stream) {
    RVGP_CHECK_ARG(d >= 1 && d <= 8, "rvgp_connections: 1 <= d <= 8 supported");
    RVGP_CHECK_ARG(D >= d, "Embedding dimension must be less or equal to the ambient dimension of input data");
    RVGP_CHECK_ARG(ldt >= d, "rvgp_connections: ldt < d");
    RVGP_CHECK_ARG(nnz >= 0, "rvgp_connections: nnz < 0");
    if (nnz == 0) return RVGP_OK;
    dim3 g(cdiv(nnz, 256)), b(256);
    hipStream_t s = as_stream(stream);
#define RVGP_CONN(DD) \
    hipLaunchKernelGGL(connections_kernel<DD>, g, b, 0, s, tangents, n, D, ldt, rowptr, cols, nnz, R, lc64, lc32, l64, l32)
    switch (d) {
        case 1: RVGP_CONN(1); break;
        case 2: RVGP_CONN(2); break;
        case 3: RVGP_CONN(3); break;
        case 4: RVGP_CONN(4); break;
        case 5: RVGP_CONN(5); break;
        case 6: RVGP_CONN(6); break;
        case 7: RVGP_CONN(7); break;
        default: RVGP_CONN(8); break;
    }
#undef RVGP_CONN
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
