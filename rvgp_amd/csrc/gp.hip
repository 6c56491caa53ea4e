// a12/a14  Spectral Matern filter, GP negative log marginal likelihood and its
// hand-written adjoint (replaces ManifoldKernel.eval_S / K, RVGP/kernels.py:
// 41-67, and GPflow's GPR.training_loss + TF autodiff reached from
// RVGP/main.py:87-95).
//
// Spectral filter (kernels.py:41-53), with N = number of rows of the ambient
// eigenvector matrix:
//   r_j = (lambda_j + 2 nu / kappa^2)^(-nu)      (matern)
//   r_j = exp(-lambda_j kappa^2 / 2)              (se)
//   S_j = sigma_f * N * r_j / sum_i r_i
// and the derivatives of S through the normalisation:
//   dS_j/dth = S_j (dlog r_j/dth - sum_i w_i dlog r_i/dth),  w = r / sum r,
//   dS_j/dsigma_f = S_j / sigma_f.
//
// NLML of y ~ N(0, Phi diag(S) Phi^T + s2 I) (M observations, k features):
//  function space (Cholesky L of K_y):
//    NLML = y^T a / 2 + sum log L_ii + M/2 log 2pi,   a = K_y^-1 y
//    dNLML/dS_j = (phi_j^T K_y^-1 phi_j - (phi_j^T a)^2) / 2
//    dNLML/ds2  = (tr K_y^-1 - a^T a) / 2,  tr K_y^-1 = (M - sum_j S_j q_j)/s2
//  weight space (A = s2 S^-1 + G, G = Phi^T Phi, b = Phi^T y, w = A^-1 b):
//    NLML = (y^T y - b^T w)/(2 s2) + ((M-k) log s2 + sum log S_j + log det A)/2
//           + M/2 log 2pi
//    dNLML/dS_j = (1/S_j - s2 Ainv_jj / S_j^2 - w_j^2 / S_j^2) / 2
//    dNLML/ds2  = ((M-k)/s2 + sum_j Ainv_jj/S_j - |y - Phi w|^2 / s2^2) / 2,
//                 |y - Phi w|^2 = y^T y - 2 b^T w + w^T G w
// Both give the same numbers (Woodbury / determinant lemma); the weight-space
// form never materialises an M x M matrix.  Gradients are finally mapped to
// the unconstrained parameters u of theta = softplus(u) + lower.
#include "common.h"

#include <mutex>

namespace rvgp {

constexpr int FT = 256;

__device__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < FT / 64; ++i) s += sh[i];
    return s;
}

// theta = [nu, kappa, sigma_f]; S (k), dS (3, k)
__global__ __launch_bounds__(FT) void spectral_filter_kernel(int typ, int k,
                                                             const double* __restrict__ lam,
                                                             const double* __restrict__ theta,
                                                             double n_rows, double* __restrict__ S,
                                                             double* __restrict__ dS) {
    __shared__ double sh[FT / 64];
    const double nu = theta[0], kappa = theta[1], sf = theta[2];
    const double sh2 = 2.0 * nu / (kappa * kappa);
    double sr = 0.0, snu = 0.0, ska = 0.0;
    for (int j = threadIdx.x; j < k; j += FT) {
        double r, dn, dk;
        if (typ == 0) {
            const double base = lam[j] + sh2;
            r = pow(base, -nu);
            dn = -log(base) - nu * (2.0 / (kappa * kappa)) / base;
            dk = 4.0 * nu * nu / (kappa * kappa * kappa * base);
        } else {
            r = exp(-0.5 * lam[j] * kappa * kappa);
            dn = 0.0;
            dk = -lam[j] * kappa;
        }
        sr += r;
        snu += r * dn;
        ska += r * dk;
    }
    sr = block_sum(sr, sh);
    snu = block_sum(snu, sh) / sr;
    ska = block_sum(ska, sh) / sr;
    for (int j = threadIdx.x; j < k; j += FT) {
        double r, dn, dk;
        if (typ == 0) {
            const double base = lam[j] + sh2;
            r = pow(base, -nu);
            dn = -log(base) - nu * (2.0 / (kappa * kappa)) / base;
            dk = 4.0 * nu * nu / (kappa * kappa * kappa * base);
        } else {
            r = exp(-0.5 * lam[j] * kappa * kappa);
            dn = 0.0;
            dk = -lam[j] * kappa;
        }
        const double s = sf * n_rows * r / sr;
        S[j] = s;
        if (dS) {
            dS[j] = s * (dn - snu);
            dS[k + j] = s * (dk - ska);
            dS[2 * k + j] = s / sf;
        }
    }
}

// A = G + diag(s2 / S)   (k x k, row-major)
__global__ void ws_form_kernel(int k, const double* __restrict__ G, int ldg,
                               const double* __restrict__ S, const double* __restrict__ s2p,
                               double* __restrict__ A, int lda) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)k * k) return;
    const int i = (int)(e / k), j = (int)(e % k);
    double v = G[(size_t)i * ldg + j];
    if (i == j) v += s2p[0] / S[i];
    A[(size_t)i * lda + j] = v;
}

// Column sums of squares: out[j] = sum_i Z[i][j]^2, f64 accumulation, in a
// fixed order: P(M) row slices of >= 256 rows (P <= 1024) write partials --
// lane = column, so every row is read coalesced, four interleaved
// accumulators per lane -- and a second pass sums them in slice order.
constexpr int kColMaxParts = 1024;
static inline int colsumsq_parts(int M) {
    // 32-row slices: every thread's loads of a slice are in flight together,
    // so a column sum costs one memory latency, not M/4 of them
    const int p = (M + 31) / 32;
    return p < 1 ? 1 : (p > kColMaxParts ? kColMaxParts : p);
}

template <typename T>
__global__ void colsumsq_part_kernel(int M, int k, int parts, const T* __restrict__ Z, int ldz,
                                     double* __restrict__ part) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const int per = (M + parts - 1) / parts;
    const int i0 = blockIdx.y * per, i1 = min(M, i0 + per);
    double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int i = i0;
    for (; i + 7 < i1; i += 8) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const double z = (double)Z[(size_t)(i + t) * ldz + j];
            s[t] += z * z;
        }
    }
    for (int t = 0; i < i1; ++i, ++t) {
        const double z = (double)Z[(size_t)i * ldz + j];
        s[t] += z * z;
    }
    part[(size_t)blockIdx.y * k + j] =
        ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// out[j] = sum of the slice partials: 16 wavefronts per 64 columns,
// wavefront w summing the slices p = w (mod 16), the 16 partial sums then
// added in wavefront order (a fixed order).
__global__ __launch_bounds__(1024) void colsumsq_reduce_kernel(int k, int parts,
                                                               const double* __restrict__ part,
                                                               double* __restrict__ out) {
    __shared__ double sp[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    // four slices in flight per step: the loads do not wait on each other
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < k) {
        int q = w;
        for (; q + 48 < parts; q += 64) {
#pragma unroll
            for (int t = 0; t < 4; ++t) a[t] += part[(size_t)(q + 16 * t) * k + j];
        }
        for (int t = 0; q < parts; q += 16, ++t) a[t] += part[(size_t)q * k + j];
    }
    sp[w][lane] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (w == 0 && j < k) {
        double t = 0.0;
        for (int q = 0; q < 16; ++q) t += sp[q][lane];
        out[j] = t;
    }
}

// Dot product out = sum_i x_i y_i in f64, fixed order: P slices of 2048
// elements (P <= 1024) write partial sums, one workgroup adds them in slice
// order.  Used for y^T a and |a|^2 over all R right-hand sides at once.
constexpr int kDotSlice = 2048, kDotMaxParts = 1024;
static inline int dot_parts(size_t n) {
    const size_t p = (n + kDotSlice - 1) / kDotSlice;
    return p < 1 ? 1 : (p > (size_t)kDotMaxParts ? kDotMaxParts : (int)p);
}

__global__ __launch_bounds__(FT) void dot_part_kernel(size_t n, int parts,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      double* __restrict__ part) {
    __shared__ double sh[FT / 64];
    const size_t per = (n + parts - 1) / parts;
    const size_t i0 = (size_t)blockIdx.x * per;
    const size_t i1 = i0 + per < n ? i0 + per : n;
    double s = 0.0;
    for (size_t i = i0 + threadIdx.x; i < i1; i += FT) s += x[i] * y[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(FT) void dot_reduce_kernel(int parts, const double* __restrict__ part,
                                                        double* __restrict__ out) {
    __shared__ double sh[FT / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < parts; i += FT) s += part[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[0] = s;
}

// Map dNLML/dS, dNLML/ds2 to theta = (nu, kappa, sigma_f, s2) and to the
// unconstrained u (theta = softplus(u) + lower => dtheta/du = sigmoid(u)).
__device__ void finish_grads(int k, const double* gS, double gs2, const double* dS,
                             const double* u, double* out, double* sh) {
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int j = threadIdx.x; j < k; j += FT) {
        g0 += gS[j] * dS[j];
        g1 += gS[j] * dS[k + j];
        g2 += gS[j] * dS[2 * k + j];
    }
    g0 = block_sum(g0, sh);
    g1 = block_sum(g1, sh);
    g2 = block_sum(g2, sh);
    if (threadIdx.x == 0) {
        out[1] = g0;
        out[2] = g1;
        out[3] = g2;
        out[4] = gs2;
        const double g[4] = {g0, g1, g2, gs2};
        for (int p = 0; p < 4; ++p) {
            const double sig = 1.0 / (1.0 + exp(-u[p]));
            out[5 + p] = g[p] * sig;
        }
    }
}

// Weight-space finish over R right-hand sides sharing the hyperparameters
// (the NLML is the sum of the R columns' NLMLs, as GPflow's GPR sums the
// log-density over the columns of Y).  Inputs: L = chol(A) (lower),
// ainv_diag = diag(A^-1), b = Phi^T Y, w = A^-1 b, Gw = G w (all k x R,
// row-major, leading dimension R), yy = |Y|_F^2.  out: [nlml, dnu, dkappa,
// dsigma_f, ds2, du_nu, du_kappa, du_sigma_f, du_s2]; scratch gS (k).
__global__ __launch_bounds__(FT) void ws_finish_kernel(int k, double M, int R,
                                                       const double* __restrict__ L, int ldl,
                                                       const double* __restrict__ ainv_diag,
                                                       const double* __restrict__ b,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ Gw,
                                                       const double* __restrict__ yy,
                                                       const double* __restrict__ S,
                                                       const double* __restrict__ dS,
                                                       const double* __restrict__ s2p,
                                                       const double* __restrict__ u,
                                                       double* __restrict__ gS,
                                                       double* __restrict__ out) {
    __shared__ double sh[FT / 64];
    const double s2 = s2p[0];
    const double Rd = (double)R;
    double bw = 0.0, wGw = 0.0, ld = 0.0, lS = 0.0, tr = 0.0;
    for (int j = threadIdx.x; j < k; j += FT) {
        double ww = 0.0;
        for (int r = 0; r < R; ++r) {
            const size_t e = (size_t)j * R + r;
            bw += b[e] * w[e];
            wGw += w[e] * Gw[e];
            ww += w[e] * w[e];
        }
        ld += log(L[(size_t)j * ldl + j]);
        lS += log(S[j]);
        const double ainv = ainv_diag[j];  // (A^-1)_jj = sum_i (L^-1)_ij^2
        tr += ainv / S[j];
        const double sj = S[j];
        gS[j] = 0.5 * (Rd / sj - Rd * s2 * ainv / (sj * sj) - ww / (sj * sj));
    }
    bw = block_sum(bw, sh);
    wGw = block_sum(wGw, sh);
    ld = block_sum(ld, sh);
    lS = block_sum(lS, sh);
    tr = block_sum(tr, sh);
    const double y2 = yy[0];
    const double two_pi = 6.283185307179586;
    const double nlml = 0.5 * (y2 - bw) / s2 + 0.5 * Rd * ((M - k) * log(s2) + lS + 2.0 * ld) +
                        0.5 * Rd * M * log(two_pi);
    const double res2 = y2 - 2.0 * bw + wGw;
    const double gs2 = 0.5 * (Rd * (M - k) / s2 + Rd * tr - res2 / (s2 * s2));
    if (threadIdx.x == 0) out[0] = nlml;
    __syncthreads();
    finish_grads(k, gS, gs2, dS, u, out, sh);
}

// Function-space finish over R right-hand sides.  Inputs: q_j =
// |L^-1 phi_j|^2, p = Phi^T a (k x R, row-major, leading dimension R),
// ya = sum_r y_r^T a_r, aa = |a|_F^2, logdet_half = sum log L_ii.
__global__ __launch_bounds__(FT) void fs_finish_kernel(int k, double M, int R,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ ya,
                                                       const double* __restrict__ aa,
                                                       const double* __restrict__ ldh,
                                                       const double* __restrict__ S,
                                                       const double* __restrict__ dS,
                                                       const double* __restrict__ s2p,
                                                       const double* __restrict__ u,
                                                       double* __restrict__ gS,
                                                       double* __restrict__ out) {
    __shared__ double sh[FT / 64];
    const double s2 = s2p[0];
    const double Rd = (double)R;
    double sq = 0.0;
    for (int j = threadIdx.x; j < k; j += FT) {
        sq += S[j] * q[j];
        double pp = 0.0;
        for (int r = 0; r < R; ++r) pp += p[(size_t)j * R + r] * p[(size_t)j * R + r];
        gS[j] = 0.5 * (Rd * q[j] - pp);
    }
    sq = block_sum(sq, sh);
    const double two_pi = 6.283185307179586;
    const double nlml = 0.5 * ya[0] + Rd * ldh[0] + 0.5 * Rd * M * log(two_pi);
    const double trinv = (M - sq) / s2;
    const double gs2 = 0.5 * (Rd * trinv - aa[0]);
    if (threadIdx.x == 0) out[0] = nlml;
    __syncthreads();
    finish_grads(k, gS, gs2, dS, u, out, sh);
}

// theta = softplus(u) + lower for the 4 parameters; writes theta[0..2] to
// kth (nu, kappa, sigma_f) and theta[3] to s2.
__global__ void params_kernel(const double* __restrict__ u, const double* __restrict__ lower,
                              double* __restrict__ kth, double* __restrict__ s2) {
    const int p = threadIdx.x;
    if (p >= 4) return;
    const double x = u[p];
    const double sp = x > 0 ? x + log1p(exp(-x)) : log1p(exp(x));
    const double th = sp + lower[p];
    if (p < 3) kth[p] = th; else s2[0] = th;
}

// One Adam step on the 4 unconstrained parameters (state m, v, t on device).
__global__ void adam_kernel(double* __restrict__ u, const double* __restrict__ g,
                            double* __restrict__ m, double* __restrict__ v,
                            double* __restrict__ t, const int* __restrict__ trainable, double lr,
                            double b1, double b2, double eps) {
    // launched as one wavefront of 4 lanes: t[0] is read by every lane
    // before lane 0 stores the incremented count
    const int p = threadIdx.x;
    const double step = t[0] + 1.0;
    if (trainable[p]) {
        const double gp = g[p];
        m[p] = b1 * m[p] + (1.0 - b1) * gp;
        v[p] = b2 * v[p] + (1.0 - b2) * gp * gp;
        const double mh = m[p] / (1.0 - pow(b1, step));
        const double vh = v[p] / (1.0 - pow(b2, step));
        u[p] -= lr * mh / (sqrt(vh) + eps);
    }
    if (p == 0) t[0] = step;
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_spectral_filter(int typ, int k, const double* evals, const double* theta,
                                    double n_rows, double* S, double* dS, void* stream) {
    RVGP_CHECK_ARG(typ == 0 || typ == 1, "rvgp_spectral_filter: typ must be 0 (matern) or 1 (se)");
    RVGP_CHECK_ARG(k >= 1, "rvgp_spectral_filter: k >= 1");
    hipLaunchKernelGGL(spectral_filter_kernel, dim3(1), dim3(FT), 0, as_stream(stream), typ, k, evals,
                       theta, n_rows, S, dS);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_ws_form(int k, const double* G, int ldg, const double* S, const double* s2,
                            double* A, int lda, void* stream) {
    RVGP_CHECK_ARG(k >= 1 && ldg >= k && lda >= k, "rvgp_ws_form: bad sizes");
    hipLaunchKernelGGL(ws_form_kernel, dim3(cdiv((size_t)k * k, 256)), dim3(256), 0,
                       as_stream(stream), k, G, ldg, S, s2, A, lda);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" size_t rvgp_colsumsq_ws_bytes(int M, int k) {
    return sizeof(double) * (size_t)colsumsq_parts(M) * (k > 0 ? k : 0);
}

static int colsumsq_launch(int f64, int M, int k, const void* Z, int ldz, double* out,
                           double* part, void* stream) {
    const int parts = colsumsq_parts(M);
    dim3 g(cdiv(k, 256), parts);
    if (f64)
        hipLaunchKernelGGL(colsumsq_part_kernel<double>, g, dim3(256), 0, as_stream(stream), M, k,
                           parts, static_cast<const double*>(Z), ldz, part);
    else
        hipLaunchKernelGGL(colsumsq_part_kernel<float>, g, dim3(256), 0, as_stream(stream), M, k,
                           parts, static_cast<const float*>(Z), ldz, part);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(colsumsq_reduce_kernel, dim3(cdiv(k, 64)), dim3(1024), 0, as_stream(stream), k,
                       parts, part, out);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_colsumsq(int f64, int M, int k, const void* Z, int ldz, double* out, void* ws,
                             size_t ws_bytes, void* stream) {
    RVGP_CHECK_ARG(M >= 0 && k >= 1 && ldz >= k, "rvgp_colsumsq: bad sizes");
    RVGP_CHECK_ARG(ws_bytes >= rvgp_colsumsq_ws_bytes(M, k), "rvgp_colsumsq: workspace too small");
    return colsumsq_launch(f64, M, k, Z, ldz, out, static_cast<double*>(ws), stream);
}

// ---------------------------------------------------------------------------
// rvgp_colsumsq without a workspace argument, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops).  The
// per-slice partial sums live in a fixed device array, `scratch`, of
// kNoAllocScratchDoubles doubles (8 MiB), one copy per GPU that the library
// runs on, held until the process ends.  Host code never passes the array's
// name to a kernel (in host code it is only a stand-in for the symbol): the
// device address of each GPU's copy is looked up once per device with
// hipGetSymbolAddress and cached.  Every call on one GPU shares that GPU's
// copy, so calls that overlap on one GPU (different streams, or host threads)
// race and must be ordered by the caller.  The partial-sum kernel writes
// every element the reduction reads, so nothing needs clearing between calls.
constexpr size_t kNoAllocScratchDoubles = (size_t)1 << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

static int scratch_address(double** out) {
    int dev = 0;
    RVGP_CHECK_HIP(hipGetDevice(&dev));
    RVGP_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "scratch: device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        RVGP_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RVGP_CHECK_ARG(p != nullptr, "scratch: no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return RVGP_OK;
}

extern "C" int rvgp_colsumsq_noalloc(int f64, int M, int k, const void* Z, int ldz, double* out,
                                     void* stream) {
    RVGP_CHECK_ARG(M >= 0 && k >= 1 && ldz >= k, "rvgp_colsumsq_noalloc: bad sizes");
    const size_t need = (size_t)colsumsq_parts(M) * (size_t)k;
    RVGP_CHECK_ARG(need <= kNoAllocScratchDoubles,
                   "rvgp_colsumsq_noalloc: %d columns of %d rows need %zu scratch doubles, more "
                   "than the %zu the library holds",
                   k, M, need, kNoAllocScratchDoubles);
    double* part = nullptr;
    const int rc = scratch_address(&part);
    if (rc != RVGP_OK) return rc;
    return colsumsq_launch(f64, M, k, Z, ldz, out, part, stream);
}

extern "C" int rvgp_ws_finish(int k, double M, int nrhs, const double* L, int ldl,
                              const double* ainv_diag, const double* b, const double* w,
                              const double* Gw, const double* yy, const double* S,
                              const double* dS, const double* s2, const double* u, double* gS,
                              double* out, void* stream) {
    RVGP_CHECK_ARG(k >= 1 && ldl >= k && nrhs >= 1, "rvgp_ws_finish: bad sizes");
    hipLaunchKernelGGL(ws_finish_kernel, dim3(1), dim3(FT), 0, as_stream(stream), k, M, nrhs, L, ldl,
                       ainv_diag,
                       b, w, Gw, yy, S, dS, s2, u, gS, out);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_fs_finish(int k, double M, int nrhs, const double* q, const double* p,
                              const double* ya, const double* aa, const double* logdet_half,
                              const double* S, const double* dS, const double* s2, const double* u,
                              double* gS, double* out, void* stream) {
    RVGP_CHECK_ARG(k >= 1 && nrhs >= 1, "rvgp_fs_finish: bad sizes");
    hipLaunchKernelGGL(fs_finish_kernel, dim3(1), dim3(FT), 0, as_stream(stream), k, M, nrhs, q, p, ya,
                       aa,
                       logdet_half, S, dS, s2, u, gS, out);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" size_t rvgp_dot_ws_bytes(size_t n) { return sizeof(double) * (size_t)dot_parts(n); }

extern "C" int rvgp_dot(size_t n, const double* x, const double* y, double* out, void* ws,
                        size_t ws_bytes, void* stream) {
    RVGP_CHECK_ARG(ws_bytes >= rvgp_dot_ws_bytes(n), "rvgp_dot: workspace too small");
    const int parts = dot_parts(n);
    double* part = static_cast<double*>(ws);
    hipLaunchKernelGGL(dot_part_kernel, dim3(parts), dim3(FT), 0, as_stream(stream), n, parts, x, y,
                       part);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(dot_reduce_kernel, dim3(1), dim3(FT), 0, as_stream(stream), parts, part, out);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_gp_params(const double* u, const double* lower, double* kernel_theta,
                              double* s2, void* stream) {
    hipLaunchKernelGGL(params_kernel, dim3(1), dim3(64), 0, as_stream(stream), u, lower, kernel_theta,
                       s2);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_adam_step(double* u, const double* g, double* m, double* v, double* t,
                              const int* trainable, double lr, double b1, double b2, double eps,
                              void* stream) {
    hipLaunchKernelGGL(adam_kernel, dim3(1), dim3(4), 0, as_stream(stream), u, g, m, v, t, trainable,
                       lr, b1, b2, eps);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

// Determinant-lemma terms for the mixed-precision function-space form: from
// the FP64 Cholesky factor LA of A = s2 S^-1 + Phi^T Phi and diag(A^-1),
//   q_j = phi_j^T K_y^-1 phi_j = 1/S_j - s2 (A^-1)_jj / S_j^2
//   logdet_half = ((M-k) log s2 + sum_j log S_j) / 2 + sum_j log LA_jj
// so neither depends on the FP32 factor of the M x M system.
__global__ __launch_bounds__(256) void lemma_terms_kernel(int k, double M,
                                                          const double* __restrict__ S,
                                                          const double* __restrict__ s2p,
                                                          const double* __restrict__ ainv,
                                                          const double* __restrict__ LA, int lda,
                                                          double* __restrict__ q,
                                                          double* __restrict__ logdet_half) {
    __shared__ double sh[4];
    const double s2 = s2p[0];
    double acc = 0.0;
    for (int j = threadIdx.x; j < k; j += 256) {
        const double sj = S[j];
        q[j] = 1.0 / sj - s2 * ainv[j] / (sj * sj);
        acc += 0.5 * log(sj) + log(LA[(size_t)j * lda + j]);
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) logdet_half[0] = sh[0] + sh[1] + sh[2] + sh[3] + 0.5 * (M - k) * log(s2);
}

extern "C" int rvgp_lemma_terms(int k, double M, const double* S, const double* s2,
                                const double* ainv_diag, const double* LA, int lda, double* q,
                                double* logdet_half, void* stream) {
    RVGP_CHECK_ARG(k >= 1 && lda >= k, "rvgp_lemma_terms: bad sizes");
    hipLaunchKernelGGL(lemma_terms_kernel, dim3(1), dim3(256), 0, as_stream(stream), k, M, S, s2,
                       ainv_diag, LA, lda, q, logdet_half);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
