// Per-node frame algebra of the vector-field path:
//   * express_in_local_frame / project_to_manifold (RVGP/geometry.py:165-175,
//     used by dataclass.py:100, 110, 116): c_i = T_i^T v_i, v_i = T_i c_i,
//     v_i <- T_i T_i^T v_i, for F fields at once;
//   * the magnitude renormalisation of vector_diffusion
//     (RVGP/smoothing.py:57-62): out_i <- out_i * a_i / (b_i * |out_i|);
//   * ambient eigenvector rows Phi_i = T_i U_i (dataclass.py:57-59) for a
//     list of nodes, written straight into the GP design matrix.
#include "common.h"

namespace rvgp {

// mode 0: c = T^T v  (in: (F, n, D) -> out: (F, n, d))
// mode 1: v = T c    (in: (F, n, d) -> out: (F, n, D))
// mode 2: v = T T^T v (in/out: (F, n, D))
__global__ void frame_kernel(int mode, int F, int n, int D, int d, int ldt,
                             const double* __restrict__ T, const double* in, double* out) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)F * n) return;
    const int i = (int)(g % n);
    const double* Ti = T + (size_t)i * D * ldt;
    if (mode == 0) {
        const double* v = in + g * D;
        for (int q = 0; q < d; ++q) {
            double s = 0.0;
            for (int p = 0; p < D; ++p) s += Ti[p * ldt + q] * v[p];
            out[g * d + q] = s;
        }
    } else if (mode == 1) {
        const double* c = in + g * d;
        for (int p = 0; p < D; ++p) {
            double s = 0.0;
            for (int q = 0; q < d; ++q) s += Ti[p * ldt + q] * c[q];
            out[g * D + p] = s;
        }
    } else {
        const double* v = in + g * D;
        double c[8];
        for (int q = 0; q < d; ++q) {
            double s = 0.0;
            for (int p = 0; p < D; ++p) s += Ti[p * ldt + q] * v[p];
            c[q] = s;
        }
        for (int p = 0; p < D; ++p) {
            double s = 0.0;
            for (int q = 0; q < d; ++q) s += Ti[p * ldt + q] * c[q];
            out[g * D + p] = s;
        }
    }
}

// out (F*n, d) scaled in place; a, b: (F*n,) magnitudes from the two scalar
// diffusions exp(-tL)|x| and exp(-tL)1.
__global__ void diffusion_normalise_kernel(size_t rows, int d, double* __restrict__ out,
                                           int ldo, const double* __restrict__ a, int lda,
                                           const double* __restrict__ b, int ldb) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rows) return;
    double* o = out + g * ldo;
    double s = 0.0;
    for (int q = 0; q < d; ++q) s += o[q] * o[q];
    const double scale = a[g * lda] / (b[g * ldb] * sqrt(s));
    for (int q = 0; q < d; ++q) o[q] = o[q] * scale;
}

// Phi rows: out[(r*D + p), j] = sum_q T[node_r][p][q] * U[node_r*d + q][j]
template <typename TO>
__global__ void phi_rows_kernel(int m, int n, int D, int d, int ldt, int k,
                                const double* __restrict__ T, const double* __restrict__ U,
                                int ldu, const int* __restrict__ nodes, TO* __restrict__ out,
                                int ldo) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    for (int r = blockIdx.y; r < m; r += gridDim.y) {
        const int i = nodes[r];
        if ((unsigned)i >= (unsigned)n) {
            // a node outside [0, n): the row reads nothing and is NaN
            for (int p = 0; p < D; ++p) out[(size_t)(r * D + p) * ldo + j] = (TO)__builtin_nan("");
            continue;
        }
        const double* Ti = T + (size_t)i * D * ldt;
        double u[8];
        for (int q = 0; q < d; ++q) u[q] = U[(size_t)(i * d + q) * ldu + j];
        for (int p = 0; p < D; ++p) {
            double s = 0.0;
            for (int q = 0; q < d; ++q) s += Ti[p * ldt + q] * u[q];
            out[(size_t)(r * D + p) * ldo + j] = (TO)s;
        }
    }
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_frame_apply(int mode, int F, int n, int D, int d, const double* T, int ldt,
                                const double* in, double* out, void* stream) {
    RVGP_CHECK_ARG(mode >= 0 && mode <= 2, "rvgp_frame_apply: mode must be 0, 1 or 2");
    RVGP_CHECK_ARG(d >= 1 && d <= 8 && d <= D && ldt >= d, "rvgp_frame_apply: need 1 <= d <= min(D, 8)");
    const size_t rows = (size_t)F * n;
    if (rows == 0) return RVGP_OK;
    hipLaunchKernelGGL(frame_kernel, dim3(cdiv(rows, 256)), dim3(256), 0, as_stream(stream), mode, F,
                       n, D, d, ldt, T, in, out);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_diffusion_normalise(size_t rows, int d, double* out, int ldo, const double* a,
                                        int lda, const double* b, int ldb, void* stream) {
    RVGP_CHECK_ARG(d >= 1 && ldo >= d, "rvgp_diffusion_normalise: bad d/ldo");
    if (rows == 0) return RVGP_OK;
    hipLaunchKernelGGL(diffusion_normalise_kernel, dim3(cdiv(rows, 256)), dim3(256), 0,
                       as_stream(stream), rows, d, out, ldo, a, lda, b, ldb);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_phi_rows(int out_f32, int m, int n, int D, int d, const double* T, int ldt,
                             int k, const double* U, int ldu, const int* nodes, void* out, int ldo,
                             void* stream) {
    RVGP_CHECK_ARG(d >= 1 && d <= 8 && ldt >= d && ldu >= k && ldo >= k && n >= 0,
                   "rvgp_phi_rows: bad sizes");
    if (m == 0 || k == 0) return RVGP_OK;
    dim3 g(cdiv(k, 256), m < 65535 ? m : 65535);
    if (out_f32)
        hipLaunchKernelGGL(phi_rows_kernel<float>, g, dim3(256), 0, as_stream(stream), m, n, D, d, ldt, k,
                           T, U, ldu, nodes, static_cast<float*>(out), ldo);
    else
        hipLaunchKernelGGL(phi_rows_kernel<double>, g, dim3(256), 0, as_stream(stream), m, n, D, d, ldt,
                           k, T, U, ldu, nodes, static_cast<double*>(out), ldo);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
