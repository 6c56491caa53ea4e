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

// The neighbourhood matrix G (D x Kp1, point q in column q) gathered from x
// and centred (pyx:397-404): the mean of each coordinate summed in pop order.
// G[p * rs + q * cs] (rs = Kp1, cs = 1: row-major).
__device__ void gauges_load_centred(const double* __restrict__ x, int D, const int* idx,
                                    int Kp1, double* G, int lane, int rs, int cs, int nl = 64) {
    for (int e = lane; e < D * Kp1; e += nl) {
        int q = e / D, p = e - q * D;
        G[p * rs + q * cs] = x[(size_t)idx[q] * D + p];
    }
    __syncthreads();
    for (int p = lane; p < D; p += nl) {
        double mean = 0.0;
        for (int q = 0; q < Kp1; ++q) mean += G[p * rs + q * cs];
        mean = mean / Kp1;
        for (int q = 0; q < Kp1; ++q) G[p * rs + q * cs] -= mean;
    }
    __syncthreads();
}

// LDS plan, in doubles (regions rounded to even counts so the 16-byte
// rotation records stay aligned):
//   R1: G while the Gram matrix is formed, then V;
//   R2: A; in the Gram case G again once the eigenvalues are copied out;
//   then the rotation records (c, s), the eigenvalues, the pairs (p, q), the
//   neighbourhood ids and the eigenvalue order.
// Reusing G's space for V and A's for the second G keeps one wave's
// footprint at ~17 KiB for D = 32, K + 1 = 31, so nine waves share a CU.
struct GaugesLds {
    int m, m2, lda, half, r1, r2;
    __host__ __device__ GaugesLds(int D, int Kp1) {
        const bool cov = D <= Kp1;
        m = cov ? D : Kp1;
        m2 = m + (m & 1);
        lda = m2 + 1;
        half = m2 / 2;
        const int g = D * Kp1, a = m2 * lda;
        r1 = g > a ? g : a;
        r2 = (cov || a >= g) ? a : g;
        r1 += r1 & 1;
        r2 += r2 & 1;
    }
    __host__ __device__ size_t bytes(int Kp1) const {
        return sizeof(double) * ((size_t)r1 + r2 + 2 * (size_t)half + m2) +
               sizeof(int) * (2 * (size_t)half + Kp1 + m);
    }
};

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
    const GaugesLds L(D, Kp1);
    const int m = L.m, m2 = L.m2, lda = L.lda, half = L.half;
    double* G = smem;                                        // R1
    double* V = smem;                                        // R1, once A is formed
    double* A = smem + L.r1;                                 // R2
    double2* csn = reinterpret_cast<double2*>(A + L.r2);     // half (c, s)
    double* lam = reinterpret_cast<double*>(csn + half);     // m2
    int2* pqv = reinterpret_cast<int2*>(lam + m2);           // half (p, q)
    int* idx = reinterpret_cast<int*>(pqv + half);           // Kp1
    int* order = idx + Kp1;                                  // m

    for (int q = lane; q < Kp1; q += 64) idx[q] = nbh[(size_t)i * Kp1 + q];
    __syncthreads();
    gauges_load_centred(x, D, idx, Kp1, G, lane, Kp1, 1);
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
    }
    __syncthreads();  // G is dead: V takes its space
    for (int e = lane; e < m2 * m2; e += 64) {
        int a = e / m2, b = e - a * m2;
        V[a * lda + b] = (a == b) ? 1.0 : 0.0;
    }
    __syncthreads();

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
                pqv[lane] = make_int2(p, q);
                csn[lane] = make_double2(c, s);
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
                const int2 r = pqv[k], cc = pqv[k2];
                const double2 w = csn[k], w2 = csn[k2];
                const int p = r.x, q = r.y, p2 = cc.x, q2 = cc.y;
                const double c = w.x, s = w.y, c2 = w2.x, s2 = w2.y;
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
                const int2 pq = pqv[k];
                const double2 w = csn[k];
                const double vrp = V[r * lda + pq.x], vrq = V[r * lda + pq.y];
                V[r * lda + pq.x] = w.x * vrp - w.y * vrq;
                V[r * lda + pq.y] = w.y * vrp + w.x * vrq;
            }
            __syncthreads();
        }
        if (!any_rot) break;
    }
    for (int a = lane; a < m; a += 64) lam[a] = A[a * lda + a];
    __syncthreads();
    // eigenvalues descending (ties: lower index first)
    for (int a = lane; a < m; a += 64) {
        double la = lam[a];
        int rank = 0;
        for (int b = 0; b < m; ++b) {
            double lb = lam[b];
            rank += (lb > la) || (lb == la && b < a);
        }
        order[rank] = a;
    }
    __syncthreads();
    int bad = 0;
    for (int r = lane; r < m; r += 64) {
        double lv = lam[order[r]];
        double s = sqrt(lv > 0.0 ? lv : 0.0);
        if (r < nsv) sigma[(size_t)i * nsv + r] = s;
        if (r < n_check && s < 1e-10) bad = 1;
    }
    // Gram case: the left singular vectors are G v / s, with G rebuilt in
    // A's space (the same gather and centring, so the same values)
    double* G2 = A;
    if (!cov && ncols > 0) gauges_load_centred(x, D, idx, Kp1, G2, lane, Kp1, 1);
    for (int e = lane; e < D * ncols; e += 64) {
        int p = e / ncols, r = e - p * ncols;
        int c = order[r];
        double u;
        if (cov) {
            u = V[p * lda + c];
        } else {
            double lv = lam[c];
            double s = sqrt(lv > 0.0 ? lv : 0.0);
            double acc = 0.0;
            for (int q = 0; q < Kp1; ++q) acc = __fma_rn(G2[p * Kp1 + q], V[q * lda + c], acc);
            u = s > 0.0 ? acc / s : 0.0;
        }
        tangents[(size_t)i * D * ncols + e] = u;
    }
    if (__any(bad) && lane == 0) atomicMax(status, RVGP_ST_RANK_DEFICIENT);
}

static size_t gauges_smem(int D, int Kp1) { return GaugesLds(D, Kp1).bytes(Kp1); }

// Wide ambient space (D > K+1): one-sided (Hestenes) Jacobi on the columns
// of the centred neighbourhood matrix itself.  Rotating column pairs until
// every pair is orthogonal leaves G V = U Sigma: the column norms are the
// singular values and the normalised columns the left singular vectors, so
// neither the Gram matrix nor V is formed, and small singular values keep
// their relative accuracy (the Gram matrix squares them).  G is held column
// by column in LDS; a step rotates the m2/2 disjoint pairs of the
// round-robin ordering, LP lanes per pair, each lane holding RPL rows of
// both columns in registers, with the three dot products summed across the
// pair's lanes by an xor butterfly (identical sums in every lane, so the
// lanes agree on every rotation).  Padding columns and rows are zero.
struct OneSidedPlan {
    int m2, half, lp, rpl, ldg;
    // nl: the lanes one neighbourhood gets (64, or 32 when a wave takes two)
    __host__ __device__ OneSidedPlan(int D, int Kp1, int nl = 64) {
        m2 = Kp1 + (Kp1 & 1);
        half = m2 / 2;
        int hp = 1;
        while (hp < half) hp <<= 1;
        lp = nl / hp;                     // lanes per pair (>= 2 for nl = 64: K+1 < D <= 64)
        const int r = (D + lp - 1) / lp;  // rows per lane
        rpl = r <= 4 ? 4 : r <= 8 ? 8 : r <= 16 ? 16 : 32;
        // +2 doubles: consecutive columns start 16 bytes x (odd) apart, so
        // the lanes' 16-byte reads of different columns spread over the banks
        ldg = lp * rpl + 2;
    }
    __host__ __device__ size_t bytes(int Kp1) const {
        return sizeof(double) * ((size_t)m2 * ldg + m2) + sizeof(int) * ((size_t)Kp1 + m2);
    }
    // one neighbourhood's LDS in doubles, rounded up to 16 bytes (the next
    // neighbourhood's columns stay 16-byte aligned)
    __host__ __device__ size_t node_doubles(int Kp1) const { return (bytes(Kp1) + 15) / 16 * 2; }
    // two neighbourhoods per wave: every pair keeps >= 2 lanes and <= 32 rows a lane
    __host__ __device__ bool fits(int D) const { return lp >= 2 && (D + lp - 1) / lp <= 32; }
};

// sum over the nl lanes of one neighbourhood (nl = 64: the whole wave)
template <int NL>
__device__ __forceinline__ double node_sum(double v) {
    if (NL == 64) return wave_sum(v);
#pragma unroll
    for (int o = NL / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// NPW neighbourhoods per wave (1, or 2 when the plan for 32 lanes fits):
// each gets NL = 64 / NPW lanes and its own LDS region; a step's fixed cost
// (the rotation from the three sums, the shuffles, the barrier) is then
// shared by NPW neighbourhoods.  With two, the wave sweeps until neither
// rotates: a converged neighbourhood's pairs stay below the rotation
// threshold, so its columns do not change meanwhile.
template <int RPL, int NPW>
__global__ __launch_bounds__(64) void gauges_onesided_kernel(
    const double* __restrict__ x, int n, int D, const int* __restrict__ nbh, int Kp1, int ncols,
    int nsv, int n_check, double* __restrict__ tangents, double* __restrict__ sigma,
    int* __restrict__ status) {
    extern __shared__ double smem[];
    constexpr int NL = 64 / NPW;
    const int nd = threadIdx.x / NL, lane = threadIdx.x - nd * NL;
    const int i = blockIdx.x * NPW + nd;
    const bool live = i < n;  // a missing second neighbourhood reads node 0's, writes nothing
    const OneSidedPlan P(D, Kp1, NL);
    const int m2 = P.m2, half = P.half, lp = P.lp, ldg = P.ldg;
    double* Gc = smem + nd * P.node_doubles(Kp1);            // column c at Gc + c * ldg
    double* sig = Gc + (size_t)m2 * ldg;                     // m2 column norms
    int* idx = reinterpret_cast<int*>(sig + m2);             // Kp1
    int* order = idx + Kp1;                                  // m2

    for (int e = lane; e < m2 * ldg; e += NL) Gc[e] = 0.0;
    for (int q = lane; q < Kp1; q += NL) idx[q] = nbh[(size_t)(live ? i : 0) * Kp1 + q];
    __syncthreads();
    gauges_load_centred(x, D, idx, Kp1, Gc, lane, 1, ldg, NL);  // pyx:397-404

    // columns of norm^2 below 1e-28 |G|_F^2 are numerically zero (the
    // centring leaves at least one): their rotations would only stir noise
    double fro = 0.0;
    for (int e = lane; e < m2 * ldg; e += NL) fro = __fma_rn(Gc[e], Gc[e], fro);
    fro = node_sum<NL>(fro);
    const double floor2 = 1e-28 * fro;

    const int k = lane / lp, j = lane - k * lp;   // pair, row block
    const bool act = k < half;
    const int r0 = j * RPL;
    for (int sweep = 0; sweep < 40; ++sweep) {
        bool any_rot = false;
        for (int step = 0; step < m2 - 1; ++step) {
            int p = 0, q = 1;
            if (act) {
                // round robin: position 0 fixed, positions 1..m2-1 rotate
                const int a = (k == 0) ? 0 : 1 + (k - 1 + step) % (m2 - 1);
                const int b = 1 + (m2 - 1 - k - 1 + step) % (m2 - 1);
                p = a < b ? a : b;
                q = a < b ? b : a;
            }
            double xp[RPL], xq[RPL];
            double al = 0.0, be = 0.0, ga = 0.0;
            if (act) {
#pragma unroll
                for (int t = 0; t < RPL; t += 2) {
                    const double2 u = *reinterpret_cast<const double2*>(Gc + p * ldg + r0 + t);
                    const double2 v = *reinterpret_cast<const double2*>(Gc + q * ldg + r0 + t);
                    xp[t] = u.x; xp[t + 1] = u.y;
                    xq[t] = v.x; xq[t + 1] = v.y;
                }
#pragma unroll
                for (int t = 0; t < RPL; ++t) {
                    al = __fma_rn(xp[t], xp[t], al);
                    be = __fma_rn(xq[t], xq[t], be);
                    ga = __fma_rn(xp[t], xq[t], ga);
                }
            }
            for (int o = 1; o < lp; o <<= 1) {
                al += __shfl_xor(al, o, 64);
                be += __shfl_xor(be, o, 64);
                ga += __shfl_xor(ga, o, 64);
            }
            const double nab = sqrt(al * be);
            const bool rot = act && al > floor2 && be > floor2 && fabs(ga) > 1e-14 * nab;
            if (!__any(rot)) continue;
            any_rot = true;
            if (rot) {
                // [p q] <- [p q] [c s; -s c] with p.q = 0 afterwards
                const double zeta = (be - al) / (2.0 * ga);
                double t;
                if (fabs(zeta) > 1e150) t = 0.5 / zeta;
                else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
#pragma unroll
                for (int u = 0; u < RPL; u += 2) {
                    const double2 np = make_double2(c * xp[u] - s * xq[u], c * xp[u + 1] - s * xq[u + 1]);
                    const double2 nq = make_double2(s * xp[u] + c * xq[u], s * xp[u + 1] + c * xq[u + 1]);
                    *reinterpret_cast<double2*>(Gc + p * ldg + r0 + u) = np;
                    *reinterpret_cast<double2*>(Gc + q * ldg + r0 + u) = nq;
                }
            }
            __syncthreads();  // the next step pairs columns written by other lanes
        }
        // Stop only after a sweep that rotated nothing: every pair was then
        // within 1e-14 of orthogonal when visited and nothing moved after.
        // A sweep whose rotated pairs were all nearly orthogonal is not
        // enough: between columns of nearly equal norm a small cosine still
        // means a large rotation angle, which mixes a pair orthogonalised
        // earlier in the sweep with one not yet visited, leaving residual
        // cosines of the size of the last sweep's (not its square), and the
        // weight-space statistics rely on T^T T = I.
        if (!any_rot) break;
    }
    // singular values: column norms; descending, ties lower index first
    for (int c = lane; c < m2; c += NL) {
        double s2 = 0.0;
        for (int r = 0; r < D; ++r) s2 = __fma_rn(Gc[c * ldg + r], Gc[c * ldg + r], s2);
        sig[c] = sqrt(s2);
    }
    __syncthreads();
    for (int a = lane; a < Kp1; a += NL) {
        const double la = sig[a];
        int rank = 0;
        for (int b = 0; b < Kp1; ++b) {
            const double lb = sig[b];
            rank += (lb > la) || (lb == la && b < a);
        }
        order[rank] = a;
    }
    __syncthreads();
    int bad = 0;
    if (live) {
        for (int r = lane; r < Kp1; r += NL) {
            const double sv = sig[order[r]];
            if (r < nsv) sigma[(size_t)i * nsv + r] = sv;
            if (r < n_check && sv < 1e-10) bad = 1;
        }
        for (int e = lane; e < D * ncols; e += NL) {
            const int p = e / ncols, r = e - p * ncols;
            const int c = order[r];
            const double sv = sig[c];
            tangents[(size_t)i * D * ncols + e] = sv > 0.0 ? Gc[c * ldg + p] / sv : 0.0;
        }
    }
    if (__any(bad) && threadIdx.x == 0) atomicMax(status, RVGP_ST_RANK_DEFICIENT);
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
    hipStream_t st = as_stream(stream);
    if (D > Kp1) {
        // two neighbourhoods per wave where the 32-lane plan fits (K+1 <= 32)
        const OneSidedPlan P2(D, Kp1, 32);
        const int npw = P2.fits(D) ? 2 : 1;
        const OneSidedPlan P = npw == 2 ? P2 : OneSidedPlan(D, Kp1);
        const size_t sm = npw == 2 ? 2 * sizeof(double) * P.node_doubles(Kp1) : P.bytes(Kp1);
        RVGP_CHECK_ARG(sm <= 65536, "rvgp_gauges: D=%d, K+1=%d needs %zu bytes of LDS", D, Kp1, sm);
        const dim3 grid((unsigned)cdiv(n, npw));
#define RVGP_ONESIDED(R, W)                                                                    \
    hipLaunchKernelGGL((gauges_onesided_kernel<R, W>), grid, dim3(64), sm, st, x, n, D, nbh, Kp1, \
                       ncols, nsv, n_check, tangents, sigma, status)
#define RVGP_ONESIDED_W(W)                       \
    if (P.rpl == 4) RVGP_ONESIDED(4, W);         \
    else if (P.rpl == 8) RVGP_ONESIDED(8, W);    \
    else if (P.rpl == 16) RVGP_ONESIDED(16, W);  \
    else RVGP_ONESIDED(32, W);
        if (npw == 2) { RVGP_ONESIDED_W(2) }
        else { RVGP_ONESIDED_W(1) }
#undef RVGP_ONESIDED_W
#undef RVGP_ONESIDED
        RVGP_LAUNCH_CHECK();
        return RVGP_OK;
    }
    const size_t sm = gauges_smem(D, Kp1);
    RVGP_CHECK_ARG(sm <= 65536, "rvgp_gauges: D=%d, K+1=%d needs %zu bytes of LDS (> 64 KiB)", D, Kp1, sm);
    hipLaunchKernelGGL(gauges_kernel, dim3(n), dim3(64), sm, st, x, n, D, nbh, Kp1,
                       ncols, nsv, n_check, tangents, sigma, status);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
