// a3 (part 2)  Tangent frames: centring (RVGP/lib/ptu_dijkstra.pyx:397-404)
// and the SVD of the D x (K+1) neighbourhood matrix (pyx:406-434).
//
// One wavefront per node.  The neighbourhood is gathered into LDS (lanes
// read consecutive coordinates of one point, so each point row is one
// coalesced segment), centred with the mean summed in pop order exactly as
// the reference does, and reduced to the smaller of its two symmetric Gram
// matrices (D x D covariance when D <= K+1, else (K+1) x (K+1)).  A cyclic
// Jacobi eigensolve with the round-robin pair ordering rotates m/2 disjoint
// pairs per step; the two-sided update J^T A J of one step is one pass over
// independent 2 x 2 blocks, the lanes splitting the blocks and the vector
// update.
// Singular values are the square roots of the eigenvalues; the left singular
// vectors are the eigenvectors (covariance case) or G v / s (Gram case).
// Singular-vector signs are free, as they are for LAPACK's dgesvd.
#include "common.h"

// The library builds with -ffp-contract=off; the Jacobi rotations here may
// fuse multiply-adds (the centring and the Gram keep their explicit
// __fma_rn order, and frames / singular values are compared with
// tolerances).
#pragma clang fp contract(fast)

namespace rvgp {

__global__ __launch_bounds__(64) void gauges_kernel(const double* __restrict__ x, int n, int D,
                                                    const int* __restrict__ nbh, int Kp1,
                                                    int ncols, int nsv, int n_check,
                                                    double* __restrict__ tangents,
                                                    double* __restrict__ sigma,
                                                    int* __restrict__ status) {
    extern __shared__ double smem[];
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const bool cov = D <= Kp1;
    const int m = cov ? D : Kp1;
    const int m2 = m + (m & 1);
    double* G = smem;                  // D x Kp1
    double* A = G + D * Kp1;           // m2 x m2
    const int lda = m2 + 1;            // padded row stride of A and V
    double* V = A + m2 * lda;          // m2 x m2
    double* cs = V + m2 * lda;         // m2/2 cosines
    double* sn = cs + m2 / 2;          // m2/2 sines
    int* pp = reinterpret_cast<int*>(sn + m2 / 2);  // m2/2
    int* qq = pp + m2 / 2;                          // m2/2
    int* idx = qq + m2 / 2;                         // Kp1
    int* order = idx + Kp1;                         // m

    for (int q = lane; q < Kp1; q += 64) idx[q] = nbh[(size_t)i * Kp1 + q];
    __syncthreads();
    for (int e = lane; e < D * Kp1; e += 64) {
        int q = e / D, p = e - q * D;
        G[p * Kp1 + q] = x[(size_t)idx[q] * D + p];
    }
    __syncthreads();
    for (int p = lane; p < D; p += 64) {  // pyx:397-404
        double mean = 0.0;
        for (int q = 0; q < Kp1; ++q) mean += G[p * Kp1 + q];
        mean = mean / Kp1;
        for (int q = 0; q < Kp1; ++q) G[p * Kp1 + q] -= mean;
    }
    __syncthreads();
    for (int e = lane; e < m2 * m2; e += 64) {
        int a = e / m2, b = e - a * m2;
        double s = 0.0;
        if (a < m && b < m) {
            if (cov) {
                for (int q = 0; q < Kp1; ++q) s = __fma_rn(G[a * Kp1 + q], G[b * Kp1 + q], s);
            } else {
                for (int p = 0; p < D; ++p) s = __fma_rn(G[p * Kp1 + a], G[p * Kp1 + b], s);
            }
        }
        A[a * lda + b] = s;
        V[a * lda + b] = (a == b) ? 1.0 : 0.0;
    }
    __syncthreads();

    const int half = m2 / 2;
    // t / half and t / m2 for t < 4096 by multiply-shift (exact: the rounding
    // error t (ceil(2^20/d) - 2^20/d) / 2^20 < 1/64 <= 1/d, and t * r < 2^32)
    const unsigned rh = ((1u << 20) + half - 1) / half, rm = ((1u << 20) + m2 - 1) / m2;
    for (int sweep = 0; sweep < 40; ++sweep) {
        // threshold Jacobi: a pair is rotated only while |a_pq| exceeds the
        // rounding level of the largest diagonal entry, and the iteration
        // stops after a sweep that rotates nothing.  (A test on the total
        // off-diagonal mass never passes for these rank-deficient Grams:
        // rotations among the near-zero eigen-directions keep leaving
        // rounding noise of that size.)
        double dmax = lane < m2 ? fabs(A[lane * lda + lane]) : 0.0;
        dmax = wave_max(dmax);
        const double thr = 1e-15 * dmax;
        bool any_rot = false;
        for (int step = 0; step < m2 - 1; ++step) {
            bool rot = false;
            if (lane < half) {
                // round robin: position 0 fixed, positions 1..m2-1 rotate
                int a = (lane == 0) ? 0 : 1 + (lane - 1 + step) % (m2 - 1);
                int bpos = m2 - 1 - lane;
                int b = 1 + (bpos - 1 + step) % (m2 - 1);
                int p = a < b ? a : b, q = a < b ? b : a;
                double apq = A[p * lda + q];
                double c = 1.0, s = 0.0;
                if (fabs(apq) > thr) {
                    rot = true;
                    double tau = (A[q * lda + q] - A[p * lda + p]) / (2.0 * apq);
                    double t;
                    if (fabs(tau) > 1e150) t = 0.5 / tau;
                    else t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + t * t);
                    s = t * c;
                }
                pp[lane] = p;
                qq[lane] = q;
                cs[lane] = c;
                sn[lane] = s;
            }
            // the workgroup is this one wavefront: skip a step that rotates
            // nothing (identity rotations)
            if (!__any(rot)) continue;
            any_rot = true;
            __syncthreads();
            // A <- J^T A J in one pass: the pairs partition the indices, so
            // each entry lies in exactly one 2 x 2 block (rows of pair k,
            // columns of pair k2) and the blocks update independently.
            for (int t = lane; t < half * half; t += 64) {
                const int k = (int)(((unsigned)t * rh) >> 20), k2 = t - k * half;
                const int p = pp[k], q = qq[k], p2 = pp[k2], q2 = qq[k2];
                const double c = cs[k], s = sn[k], c2 = cs[k2], s2 = sn[k2];
                const double a00 = A[p * lda + p2], a01 = A[p * lda + q2];
                const double a10 = A[q * lda + p2], a11 = A[q * lda + q2];
                // rows: [p; q] <- [c -s; s c]^T ... applied as J^T on the left
                const double b00 = c * a00 - s * a10, b01 = c * a01 - s * a11;
                const double b10 = s * a00 + c * a10, b11 = s * a01 + c * a11;
                // columns: [p2 q2] <- [p2 q2] [c2 s2; -s2 c2]
                A[p * lda + p2] = c2 * b00 - s2 * b01;
                A[p * lda + q2] = s2 * b00 + c2 * b01;
                A[q * lda + p2] = c2 * b10 - s2 * b11;
                A[q * lda + q2] = s2 * b10 + c2 * b11;
            }
            for (int t = lane; t < half * m2; t += 64) {  // V <- V J
                const int k = (int)(((unsigned)t * rm) >> 20), r = t - k * m2;
                const int p = pp[k], q = qq[k];
                const double c = cs[k], s = sn[k];
                const double vrp = V[r * lda + p], vrq = V[r * lda + q];
                V[r * lda + p] = c * vrp - s * vrq;
                V[r * lda + q] = s * vrp + c * vrq;
            }
            __syncthreads();
        }
        if (!any_rot) break;
    }
    // eigenvalues descending (ties: lower index first)
    for (int a = lane; a < m; a += 64) {
        double la = A[a * lda + a];
        int rank = 0;
        for (int b = 0; b < m; ++b) {
            double lb = A[b * lda + b];
            rank += (lb > la) || (lb == la && b < a);
        }
        order[rank] = a;
    }
    __syncthreads();
    int bad = 0;
    for (int r = lane; r < m; r += 64) {
        double lam = A[order[r] * lda + order[r]];
        double s = sqrt(lam > 0.0 ? lam : 0.0);
        if (r < nsv) sigma[(size_t)i * nsv + r] = s;
        if (r < n_check && s < 1e-10) bad = 1;
    }
    for (int e = lane; e < D * ncols; e += 64) {
        int p = e / ncols, r = e - p * ncols;
        int c = order[r];
        double u;
        if (cov) {
            u = V[p * lda + c];
        } else {
            double lam = A[c * lda + c];
            double s = sqrt(lam > 0.0 ? lam : 0.0);
            double acc = 0.0;
            for (int q = 0; q < Kp1; ++q) acc = __fma_rn(G[p * Kp1 + q], V[q * lda + c], acc);
            u = s > 0.0 ? acc / s : 0.0;
        }
        tangents[(size_t)i * D * ncols + e] = u;
    }
    if (__any(bad) && lane == 0) atomicMax(status, RVGP_ST_RANK_DEFICIENT);
}

static size_t gauges_smem(int D, int Kp1) {
    int m = D <= Kp1 ? D : Kp1;
    int m2 = m + (m & 1);
    size_t dbl = (size_t)D * Kp1 + 2 * (size_t)m2 * (m2 + 1) + (size_t)m2;
    size_t ints = 2 * (size_t)(m2 / 2) + Kp1 + m;
    return dbl * sizeof(double) + ints * sizeof(int);
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_gauges(const double* x, int n, int D, const int* nbh, int Kp1, int ncols,
                           int nsv, int n_check, double* tangents, double* sigma, int* status,
                           void* stream) {
    RVGP_CHECK_ARG(n >= 1 && D >= 1 && D <= 64 && Kp1 >= 2 && Kp1 <= 64,
                   "rvgp_gauges: need D <= 64 and 2 <= K+1 <= 64");
    const int m = D <= Kp1 ? D : Kp1;
    RVGP_CHECK_ARG(ncols >= 0 && ncols <= m && nsv >= 0 && nsv <= m && n_check <= m,
                   "rvgp_gauges: ncols/nsv/n_check must not exceed min(D, K+1)");
    const size_t sm = gauges_smem(D, Kp1);
    RVGP_CHECK_ARG(sm <= 65536, "rvgp_gauges: D=%d, K+1=%d needs %zu bytes of LDS (> 64 KiB)", D, Kp1, sm);
    hipLaunchKernelGGL(gauges_kernel, dim3(n), dim3(64), sm, as_stream(stream), x, n, D, nbh, Kp1,
                       ncols, nsv, n_check, tangents, sigma, status);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
