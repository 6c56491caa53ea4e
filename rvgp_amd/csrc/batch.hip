// Batched independent GP fits on one shared manifold: F vector fields, each
// with its own training subset (of any size) and its own hyperparameters,
// fitted together.  The single-fit path (GPflow's GPR reached from
// RVGP/main.py:87-95) runs F times in the reference; here every fit of a
// step runs inside one kernel launch.
//
// Weight-space form (exact for any subset size): with the intrinsic
// eigenvector rows U_i (d x k) of the training nodes of fit f,
//     G_f = sum_i U_i^T U_i      b_f = sum_i U_i^T c_i,   c_i = T_i^T y_i
// equal Phi^T Phi and Phi^T y over the ambient rows Phi_i = T_i U_i (the
// gauge T_i has orthonormal columns), so the ambient design matrix is never
// formed.  Per optimiser step and fit:
//     A = G + s2 diag(S)^-1 = L L^T,  X = L^-1,  diag(A^-1)_j = |X e_j|^2,
//     w = X^T (X b),  then the NLML / adjoint of gp.hip's weight-space
//     finish and one Adam update of the fit's 4 unconstrained parameters.
//
// Kernels:
//   batch_gram_kernel<BT, WT>  upper BT x BT tiles of every G_f on the matrix
//                              cores (v_mfma_f64_16x16x4f64), U rows gathered
//                              through the fit's node list into LDS (FP32
//                              input is widened to FP64 in LDS); XCD-aware
//                              block order keeps one fit's tiles on one XCD.
//   batch_rhs_kernel           b_f and y^T y per fit.
//   batch_step_kernel<NB>      one workgroup per fit (persistent over fits):
//                              filter, A, right-looking blocked Cholesky of
//                              [A | I] that yields L and L^-1 together
//                              (diagonal block in registers of one wave,
//                              panel staged in LDS, trailing update of both
//                              halves on the matrix cores), solves, NLML,
//                              adjoint and Adam.
#include "common.h"

#pragma clang fp contract(fast)

namespace rvgp {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ Gram
template <typename T>
__device__ __forceinline__ double as_f64(T v) { return (double)v; }

constexpr int GBK = 16;  // gathered rows per LDS stage

template <int BT, int WT, typename TIn>
__global__ __launch_bounds__((BT / WT) * (BT / WT) * 64) void batch_gram_kernel(
    int F, int k, int d, const TIn* __restrict__ U, int ldu, const int* __restrict__ off,
    const int* __restrict__ nodes, int ntiles, int T, double* __restrict__ G) {
    constexpr int NW = (BT / WT) * (BT / WT);
    constexpr int NTH = NW * 64;
    constexpr int MT = WT / 16;  // MFMA tiles per wave side
    __shared__ double As[GBK][BT + 1];
    __shared__ double Bs[GBK][BT + 1];
    __shared__ int rows[GBK];

    // XCD-aware order: blocks are dealt round-robin to the 8 XCDs, so logical
    // block q = (L mod 8) * (total / 8) + L / 8 keeps consecutive q -- the
    // tiles of one fit -- on one XCD and its L2
    const int total = ntiles * F;
    int q = blockIdx.x;
    if (total % 8 == 0) q = (blockIdx.x % 8) * (total / 8) + blockIdx.x / 8;
    const int f = q / ntiles;
    int t = q - f * ntiles;
    // tile t -> (ti, tj), ti <= tj, row-major over the upper triangle
    int ti = 0;
    while (t >= T - ti) { t -= T - ti; ++ti; }
    const int tj = ti + t;
    const bool diag = ti == tj;
    const int m0 = ti * BT, n0 = tj * BT;
    const int r0 = off[f] * d, r1 = off[f + 1] * d;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w / (BT / WT)) * WT, wn = (w % (BT / WT)) * WT;
    dbl4 acc[MT][MT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) acc[i][j] = dbl4{0, 0, 0, 0};

    for (int k0 = r0; k0 < r1; k0 += GBK) {
        if (tid < GBK) {
            const int r = k0 + tid;
            rows[tid] = r < r1 ? nodes[r / d] * d + (r % d) : -1;
        }
        __syncthreads();
        for (int e = tid; e < GBK * BT; e += NTH) {
            const int kk = e / BT, mm = e - kk * BT;
            const int ur = rows[kk];
            double va = 0.0, vb = 0.0;
            if (ur >= 0) {
                const TIn* up = U + (size_t)ur * ldu;
                if (m0 + mm < k) va = as_f64(up[m0 + mm]);
                if (!diag && n0 + mm < k) vb = as_f64(up[n0 + mm]);
            }
            As[kk][mm] = va;
            if (!diag) Bs[kk][mm] = vb;
        }
        __syncthreads();
        const double(*Bp)[BT + 1] = diag ? As : Bs;
#pragma unroll
        for (int ks = 0; ks < GBK; ks += 4) {
            const int kr = ks + (lane >> 4);
            double a[MT], b[MT];
#pragma unroll
            for (int i = 0; i < MT; ++i) {
                a[i] = As[kr][wm + i * 16 + (lane & 15)];
                b[i] = Bp[kr][wn + i * 16 + (lane & 15)];
            }
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < MT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // f64 MFMA output: lane holds column lane%16, rows lane/16 + 4 r
    double* Gf = G + (size_t)f * k * k;
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + i * 16 + (lane >> 4) + 4 * r;
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < k && gn < k) {
                    Gf[(size_t)gm * k + gn] = acc[i][j][r];
                    if (!diag) Gf[(size_t)gn * k + gm] = acc[i][j][r];
                }
            }
}

// b_f[j] = sum over the fit's rows of U[row][j] c[row]; yy_f = sum of the
// fit's per-node |y_i|^2.  grid (ceil(k / 256), F).
__global__ __launch_bounds__(256) void batch_rhs_kernel(int F, int k, int d,
                                                        const double* __restrict__ U, int ldu,
                                                        const int* __restrict__ off,
                                                        const int* __restrict__ nodes,
                                                        const double* __restrict__ c,
                                                        const double* __restrict__ ynorm2,
                                                        double* __restrict__ b,
                                                        double* __restrict__ yy) {
    const int f = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int t0 = off[f], t1 = off[f + 1];
    if (j < k) {
        double s0 = 0.0, s1 = 0.0;
        for (int t = t0; t < t1; ++t) {
            const double* up = U + (size_t)nodes[t] * d * ldu + j;
            const double* cp = c + (size_t)t * d;
            for (int p = 0; p < d; ++p) {
                if (p & 1) s1 += up[(size_t)p * ldu] * cp[p];
                else s0 += up[(size_t)p * ldu] * cp[p];
            }
        }
        b[(size_t)f * k + j] = s0 + s1;
    }
    if (blockIdx.x == 0) {
        __shared__ double sh[4];
        double s = 0.0;
        for (int t = t0 + threadIdx.x; t < t1; t += 256) s += ynorm2[t];
        s = wave_sum(s);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) yy[f] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    }
}

// ------------------------------------------------------------------ step
// Threads per fit workgroup (ST = 512 or 1024, a template parameter chosen
// by the autotuner): more waves hide more of each panel step's memory
// latency; fewer leave more registers per thread.
template <int ST>
__device__ double wg_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < ST / 64; ++i) s += sh[i];
    return s;
}

// Per-fit workspace: W = [A | X] (k x 2k, row-major) then vectors
// S, dS (3k), gS, ainv, t, w, Gw (9k doubles).
__host__ __device__ inline size_t step_ws_doubles(int k) { return 2 * (size_t)k * k + 9 * (size_t)k; }

template <int NB, int ST>
__global__ __launch_bounds__(ST) void batch_step_kernel(
    int F, int k, int typ, const double* __restrict__ lam, double n_rows,
    const double* __restrict__ G, const double* __restrict__ bvec, const double* __restrict__ yy,
    const double* __restrict__ Mrows, const double* __restrict__ lower, double* __restrict__ u,
    const int* __restrict__ trainable, double* __restrict__ m1, double* __restrict__ m2,
    double* __restrict__ tstep, int do_adam, double lr, double b1, double b2, double eps,
    double* __restrict__ ws, double* __restrict__ out, int* __restrict__ status) {
    extern __shared__ double lds[];
    // LDS: Ld[NB][NB + 1] (diagonal factor), P[k][NB + 1] (panel rows)
    double* Ld = lds;
    double* P = lds + NB * (NB + 1);
    __shared__ double sh[ST / 64];
    __shared__ double th[8];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int ldw = 2 * k;

    for (int f = blockIdx.x; f < F; f += gridDim.x) {
        const double* Gf = G + (size_t)f * k * k;
        const double* bf = bvec + (size_t)f * k;
        double* W = ws + (size_t)f * step_ws_doubles(k);
        double* S = W + 2 * (size_t)k * k;
        double* dS = S + k;
        double* gS = dS + 3 * k;
        double* ainv = gS + k;
        double* tv = ainv + k;
        double* wv_ = tv + k;
        double* Gw = wv_ + k;
        double* uf = u + 4 * (size_t)f;

        // ---- theta = softplus(u) + lower
        if (tid < 4) {
            const double x = uf[tid];
            const double sp = x > 0 ? x + log1p(exp(-x)) : log1p(exp(x));
            th[tid] = sp + lower[tid];
        }
        __syncthreads();
        const double nu = th[0], kappa = th[1], sf = th[2], s2 = th[3];

        // ---- spectral filter S and dS/d(nu, kappa, sigma_f) (kernels.py:41-53)
        const double shf = 2.0 * nu / (kappa * kappa);
        double sr = 0.0, snu = 0.0, ska = 0.0;
        for (int j = tid; j < k; j += ST) {
            double r, dn, dk;
            if (typ == 0) {
                const double base = lam[j] + shf;
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
        sr = wg_sum<ST>(sr, sh);
        snu = wg_sum<ST>(snu, sh) / sr;
        ska = wg_sum<ST>(ska, sh) / sr;
        for (int j = tid; j < k; j += ST) {
            double r, dn, dk;
            if (typ == 0) {
                const double base = lam[j] + shf;
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
            dS[j] = s * (dn - snu);
            dS[k + j] = s * (dk - ska);
            dS[2 * k + j] = s / sf;
        }
        __syncthreads();

        // ---- W = [G + diag(s2 / S) | I]
        for (size_t e = tid; e < (size_t)k * ldw; e += ST) {
            const int i = (int)(e / ldw), j = (int)(e - (size_t)i * ldw);
            double v;
            if (j < k) v = Gf[(size_t)i * k + j] + (i == j ? s2 / S[i] : 0.0);
            else v = (j - k == i) ? 1.0 : 0.0;
            W[e] = v;
        }
        __syncthreads();

        // ---- blocked right-looking Cholesky of A, applied to X at once
        int bad = 0;
        for (int j0 = 0; j0 < k; j0 += NB) {
            const int jb = min(NB, k - j0);
            const int t0 = j0 + jb;  // first trailing row
            // (1) diagonal block: lane i holds row i in registers
            if (wv == 0) {
                double row[NB];
                const int i = lane;
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    row[c] = (i < jb && c <= i) ? W[(size_t)(j0 + i) * ldw + j0 + c] : 0.0;
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    if (c < jb) {
                        double piv = bcast(row[c], c);
                        if (!(piv > 0.0)) {  // not positive definite (or NaN)
                            bad = 1;
                            piv = 1.0;
                        }
                        const double dc = sqrt(piv);
                        const double lic = i > c ? row[c] / dc : (i == c ? dc : 0.0);
                        row[c] = lic;
#pragma unroll
                        for (int l = c + 1; l < NB; ++l) {
                            const double llc = __shfl(lic, l, 64);  // L[l][c] (VGPR)
                            if (l <= i && l < jb) row[l] -= lic * llc;
                        }
                    }
                }
                if (i < NB) {
#pragma unroll
                    for (int c = 0; c < NB; ++c) {
                        Ld[i * (NB + 1) + c] = (c <= i && i < jb) ? row[c] : 0.0;
                        if (c <= i && i < jb) W[(size_t)(j0 + i) * ldw + j0 + c] = row[c];
                    }
                }
            }
            __syncthreads();
            // (2) panel rows i >= t0: L[i, J] = A[i, J] L_JJ^-T (row solve),
            //     and rows J of X: X[J, :t0] = L_JJ^-1 X[J, :t0] (column solve)
            const int nrow = k - t0, ncol = t0;
            for (int e = tid; e < nrow + ncol; e += ST) {
                double x[NB];
                if (e < nrow) {
                    const int i = t0 + e;
                    double* wr = W + (size_t)i * ldw + j0;
#pragma unroll
                    for (int c = 0; c < NB; ++c) x[c] = c < jb ? wr[c] : 0.0;
#pragma unroll
                    for (int c = 0; c < NB; ++c) {
                        if (c < jb) {
                            double s = x[c];
#pragma unroll
                            for (int l = 0; l < c; ++l) s -= x[l] * Ld[c * (NB + 1) + l];
                            x[c] = s / Ld[c * (NB + 1) + c];
                        }
                    }
#pragma unroll
                    for (int c = 0; c < NB; ++c) {
                        const double v = c < jb ? x[c] : 0.0;
                        if (c < jb) wr[c] = v;
                        P[e * (NB + 1) + c] = v;
                    }
                } else {
                    const int cc = e - nrow;  // column of X
                    double* xc = W + (size_t)j0 * ldw + k + cc;
#pragma unroll
                    for (int r = 0; r < NB; ++r) x[r] = r < jb ? xc[(size_t)r * ldw] : 0.0;
#pragma unroll
                    for (int r = 0; r < NB; ++r) {
                        if (r < jb) {
                            double s = x[r];
#pragma unroll
                            for (int l = 0; l < r; ++l) s -= Ld[r * (NB + 1) + l] * x[l];
                            x[r] = s / Ld[r * (NB + 1) + r];
                        }
                    }
#pragma unroll
                    for (int r = 0; r < NB; ++r)
                        if (r < jb) xc[(size_t)r * ldw] = x[r];
                }
            }
            __syncthreads();
            // (3) trailing update on the matrix cores, 16 x 16 tiles:
            //     A[i, j] -= P_i . P_j   (t0 <= j <= i: lower tiles)
            //     X[i, c] -= P_i . X[J, c]  (c < t0)
            // A wave takes TG tiles at a time and issues all their loads of
            // the old values before the products, so TG round trips to the
            // cache overlap instead of following one another.
            const int TR = (nrow + 15) / 16, TC = (ncol + 15) / 16;
            const int nA = TR * (TR + 1) / 2, ntile = nA + TR * TC;
            constexpr int TG = 4;
            for (int base = wv * TG; base < ntile; base += (ST / 64) * TG) {
                int tis[TG], tjs[TG];
                dbl4 cv[TG];
#pragma unroll
                for (int g = 0; g < TG; ++g) {
                    const int tI = base + g;
                    int ti = -1, tj = 0;
                    if (tI < nA) {
                        // lower-triangle tile index -> (ti, tj), tj <= ti
                        ti = (int)((sqrt(8.0 * tI + 1.0) - 1.0) * 0.5);
                        while ((ti + 1) * (ti + 2) / 2 <= tI) ++ti;
                        while (ti * (ti + 1) / 2 > tI) --ti;
                        tj = tI - ti * (ti + 1) / 2;
                    } else if (tI < ntile) {
                        const int q = tI - nA;
                        ti = q / TC;
                        tj = TR + q - (q / TC) * TC;  // X tiles: tj >= TR marks them
                    }
                    tis[g] = ti;
                    tjs[g] = tj;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = ti * 16 + (lane >> 4) + 4 * r;
                        double v = 0.0;
                        if (ti >= 0 && row < nrow) {
                            if (tj < TR) {
                                const int col = tj * 16 + (lane & 15);
                                if (col <= row) v = W[(size_t)(t0 + row) * ldw + t0 + col];
                            } else {
                                const int col = (tj - TR) * 16 + (lane & 15);
                                if (col < ncol) v = W[(size_t)(t0 + row) * ldw + k + col];
                            }
                        }
                        cv[g][r] = v;
                    }
                }
#pragma unroll
                for (int g = 0; g < TG; ++g) {
                    const int ti = tis[g], tj = tjs[g];
                    if (ti < 0) continue;
                    dbl4 acc = dbl4{0, 0, 0, 0};
                    const int ar = ti * 16 + (lane & 15);  // panel row of the A fragment
#pragma unroll
                    for (int ks = 0; ks < NB; ks += 4) {
                        const int kc = ks + (lane >> 4);
                        const double a = (ar < nrow && kc < jb) ? P[ar * (NB + 1) + kc] : 0.0;
                        double b;
                        if (tj < TR) {
                            const int br = tj * 16 + (lane & 15);
                            b = (br < nrow && kc < jb) ? P[br * (NB + 1) + kc] : 0.0;
                        } else {
                            const int bc = (tj - TR) * 16 + (lane & 15);
                            b = (bc < ncol && kc < jb) ? W[(size_t)(j0 + kc) * ldw + k + bc] : 0.0;
                        }
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = ti * 16 + (lane >> 4) + 4 * r;
                        if (row >= nrow) continue;
                        const double v = cv[g][r] - acc[r];
                        if (tj < TR) {
                            const int col = tj * 16 + (lane & 15);
                            if (col <= row) W[(size_t)(t0 + row) * ldw + t0 + col] = v;
                        } else {
                            const int col = (tj - TR) * 16 + (lane & 15);
                            if (col < ncol) W[(size_t)(t0 + row) * ldw + k + col] = v;
                        }
                    }
                }
            }
            __syncthreads();
        }
        bad = (int)wave_max((double)bad);
        if (wv == 0 && lane == 0 && status) status[f] = bad ? RVGP_ST_NOT_SPD : RVGP_ST_OK;

        // ---- one coalesced pass over the rows of X = L^-1 (a wave per row):
        //      t = X b, and the column norms diag(A^-1)_c = sum_i X[i][c]^2
        //      accumulated per wave in LDS (the panel's space), then summed
        //      in wave order
        double* part = P;  // (ST / 64) x k partial sums
        for (int e = tid; e < (ST / 64) * k; e += ST) part[e] = 0.0;
        __syncthreads();
        for (int i = wv; i < k; i += ST / 64) {
            double s = 0.0;
            const double* xr = W + (size_t)i * ldw + k;
            for (int c = lane; c <= i; c += 64) {
                const double x = xr[c];
                s += x * bf[c];
                part[wv * k + c] += x * x;
            }
            s = wave_sum(s);
            if (lane == 0) tv[i] = s;
        }
        __syncthreads();
        for (int c = tid; c < k; c += ST) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < ST / 64; ++q) {
                s += part[q * k + c];
                part[q * k + c] = 0.0;
            }
            ainv[c] = s;
        }
        __syncthreads();
        // ---- w = X^T t, the same row pass
        for (int i = wv; i < k; i += ST / 64) {
            const double ti = tv[i];
            const double* xr = W + (size_t)i * ldw + k;
            for (int c = lane; c <= i; c += 64) part[wv * k + c] += xr[c] * ti;
        }
        __syncthreads();
        for (int c = tid; c < k; c += ST) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < ST / 64; ++q) s += part[q * k + c];
            wv_[c] = s;
        }
        __syncthreads();
        // ---- Gw = G w
        for (int i = wv; i < k; i += ST / 64) {
            double s = 0.0;
            for (int c = lane; c < k; c += 64) s += Gf[(size_t)i * k + c] * wv_[c];
            s = wave_sum(s);
            if (lane == 0) Gw[i] = s;
        }
        __syncthreads();

        // ---- NLML and adjoint (gp.hip weight-space finish, one right-hand side)
        const double Mf = Mrows[f];
        double bw = 0.0, wGw = 0.0, ld = 0.0, lS = 0.0, tr = 0.0;
        for (int j = tid; j < k; j += ST) {
            const double sj = S[j], wj = wv_[j];
            bw += bf[j] * wj;
            wGw += wj * Gw[j];
            ld += log(W[(size_t)j * ldw + j]);
            lS += log(sj);
            tr += ainv[j] / sj;
            gS[j] = 0.5 * (1.0 / sj - s2 * ainv[j] / (sj * sj) - wj * wj / (sj * sj));
        }
        bw = wg_sum<ST>(bw, sh);
        wGw = wg_sum<ST>(wGw, sh);
        ld = wg_sum<ST>(ld, sh);
        lS = wg_sum<ST>(lS, sh);
        tr = wg_sum<ST>(tr, sh);
        const double y2 = yy[f];
        const double nlml = 0.5 * (y2 - bw) / s2 + 0.5 * ((Mf - k) * log(s2) + lS + 2.0 * ld) +
                            0.5 * Mf * log(6.283185307179586);
        const double res2 = y2 - 2.0 * bw + wGw;
        const double gs2 = 0.5 * ((Mf - k) / s2 + tr - res2 / (s2 * s2));
        __syncthreads();
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        for (int j = tid; j < k; j += ST) {
            g0 += gS[j] * dS[j];
            g1 += gS[j] * dS[k + j];
            g2 += gS[j] * dS[2 * k + j];
        }
        g0 = wg_sum<ST>(g0, sh);
        g1 = wg_sum<ST>(g1, sh);
        g2 = wg_sum<ST>(g2, sh);
        if (tid < 4) {
            const double gth = tid == 0 ? g0 : tid == 1 ? g1 : tid == 2 ? g2 : gs2;
            const double x = uf[tid];
            const double gu = gth / (1.0 + exp(-x));  // dtheta/du = sigmoid(u)
            double* o = out + 9 * (size_t)f;
            if (tid == 0) o[0] = nlml;
            o[1 + tid] = gth;
            o[5 + tid] = gu;
            if (do_adam) {
                // Adam on the fit's own state; its step counter is read by
                // lanes 0-3 before lane 0 stores the increment
                const double step = tstep[f] + 1.0;
                if (trainable[tid]) {
                    double* mf = m1 + 4 * (size_t)f;
                    double* vf = m2 + 4 * (size_t)f;
                    mf[tid] = b1 * mf[tid] + (1.0 - b1) * gu;
                    vf[tid] = b2 * vf[tid] + (1.0 - b2) * gu * gu;
                    const double mh = mf[tid] / (1.0 - pow(b1, step));
                    const double vh = vf[tid] / (1.0 - pow(b2, step));
                    uf[tid] = x - lr * mh / (sqrt(vh) + eps);
                }
                __builtin_amdgcn_wave_barrier();
                if (tid == 0) tstep[f] = step;
            }
        }
        __syncthreads();
    }
}

template <int NB, int ST>
static size_t step_lds_bytes(int k) {
    // Ld, then the panel P (k x (NB + 1)), whose space also holds the
    // per-wave partial sums (ST / 64) x k of the final row passes
    const size_t pn = (size_t)k * (NB + 1) > (size_t)(ST / 64) * k ? (size_t)k * (NB + 1)
                                                                    : (size_t)(ST / 64) * k;
    return sizeof(double) * ((size_t)NB * (NB + 1) + pn);
}

template <int NB, int ST>
static int launch_step(int grid, int F, int k, int typ, const double* lam, double n_rows,
                       const double* G, const double* b, const double* yy, const double* M,
                       const double* lower, double* u, const int* trainable, double* m1,
                       double* m2, double* t, int do_adam, double lr, double b1, double b2,
                       double eps, double* ws, double* out, int* status, hipStream_t s) {
    const size_t lds = step_lds_bytes<NB, ST>(k);
    // 160 KiB of LDS per workgroup, less 1 KiB kept for the static arrays
    constexpr size_t kMaxDyn = 160 * 1024 - 1024;
    RVGP_CHECK_ARG(lds <= kMaxDyn, "rvgp_batch_step: k = %d needs %zu bytes of LDS with panel "
                                   "width %d (at most %zu); use a narrower panel",
                   k, lds, NB, kMaxDyn);
    // the LDS opt-in is raised to the 160 KiB maximum once per panel width
    // (not per launch: the launch may be under HIP graph capture)
    static bool opted_in = false;
    if (!opted_in) {
        RVGP_CHECK_HIP(hipFuncSetAttribute((const void*)batch_step_kernel<NB, ST>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxDyn));
        opted_in = true;
    }
    hipLaunchKernelGGL((batch_step_kernel<NB, ST>), dim3(grid), dim3(ST), lds, s, F, k, typ, lam, n_rows,
                       G, b, yy, M, lower, u, trainable, m1, m2, t, do_adam, lr, b1, b2, eps, ws,
                       out, status);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

template <int BT, int WT, typename TIn>
static int launch_gram(int F, int k, int d, const TIn* U, int ldu, const int* off,
                       const int* nodes, double* G, hipStream_t s) {
    const int T = (k + BT - 1) / BT;
    const int ntiles = T * (T + 1) / 2;
    const size_t total = (size_t)ntiles * F;
    RVGP_CHECK_ARG(total < (1u << 31), "rvgp_batch_gram: too many tiles");
    hipLaunchKernelGGL((batch_gram_kernel<BT, WT, TIn>), dim3((unsigned)total),
                       dim3((BT / WT) * (BT / WT) * 64), 0, s, F, k, d, U, ldu, off, nodes, ntiles,
                       T, G);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_batch_gram(int f64, int F, int k, int d, const void* U, int ldu,
                               const int* node_off, const int* nodes, int tile, double* G,
                               void* stream) {
    RVGP_CHECK_ARG(F >= 1 && k >= 1 && d >= 1 && ldu >= k, "rvgp_batch_gram: bad sizes");
    hipStream_t s = as_stream(stream);
    if (f64) {
        const double* u = static_cast<const double*>(U);
        if (tile == 32) return launch_gram<32, 16, double>(F, k, d, u, ldu, node_off, nodes, G, s);
        if (tile == 64) return launch_gram<64, 32, double>(F, k, d, u, ldu, node_off, nodes, G, s);
        if (tile == 128) return launch_gram<128, 64, double>(F, k, d, u, ldu, node_off, nodes, G, s);
    } else {
        const float* u = static_cast<const float*>(U);
        if (tile == 32) return launch_gram<32, 16, float>(F, k, d, u, ldu, node_off, nodes, G, s);
        if (tile == 64) return launch_gram<64, 32, float>(F, k, d, u, ldu, node_off, nodes, G, s);
        if (tile == 128) return launch_gram<128, 64, float>(F, k, d, u, ldu, node_off, nodes, G, s);
    }
    RVGP_CHECK_ARG(false, "rvgp_batch_gram: tile must be 32, 64 or 128 (got %d)", tile);
    return RVGP_EINVAL;
}

extern "C" int rvgp_batch_rhs(int F, int k, int d, const double* U, int ldu, const int* node_off,
                              const int* nodes, const double* c, const double* ynorm2, double* b,
                              double* yy, void* stream) {
    RVGP_CHECK_ARG(F >= 1 && k >= 1 && d >= 1 && ldu >= k, "rvgp_batch_rhs: bad sizes");
    hipLaunchKernelGGL(batch_rhs_kernel, dim3(cdiv(k, 256), F), dim3(256), 0, as_stream(stream), F, k,
                       d, U, ldu, node_off, nodes, c, ynorm2, b, yy);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" size_t rvgp_batch_step_ws_bytes(int F, int k) {
    return sizeof(double) * (size_t)(F > 0 ? F : 0) * step_ws_doubles(k > 0 ? k : 0);
}

extern "C" int rvgp_batch_step(int F, int k, int typ, const double* evals, double n_rows,
                               const double* G, const double* b, const double* yy,
                               const double* M, const double* lower, double* u,
                               const int* trainable, double* m1, double* m2, double* t,
                               int do_adam, double lr, double b1, double b2, double eps,
                               int panel, int threads, int grid, double* ws, size_t ws_bytes,
                               double* out, int* status, void* stream) {
    RVGP_CHECK_ARG(F >= 1 && k >= 1, "rvgp_batch_step: bad sizes");
    RVGP_CHECK_ARG(typ == 0 || typ == 1, "rvgp_batch_step: typ must be 0 (matern) or 1 (se)");
    RVGP_CHECK_ARG(ws_bytes >= rvgp_batch_step_ws_bytes(F, k), "rvgp_batch_step: workspace too small");
    RVGP_CHECK_ARG(grid >= 1, "rvgp_batch_step: grid >= 1");
    const int g = grid < F ? grid : F;
    hipStream_t s = as_stream(stream);
#define RVGP_STEP(NB_, ST_)                                                                        \
    return launch_step<NB_, ST_>(g, F, k, typ, evals, n_rows, G, b, yy, M, lower, u, trainable, m1, \
                                 m2, t, do_adam, lr, b1, b2, eps, ws, out, status, s)
    if (panel == 16 && threads == 512) RVGP_STEP(16, 512);
    if (panel == 16 && threads == 1024) RVGP_STEP(16, 1024);
    if (panel == 32 && threads == 512) RVGP_STEP(32, 512);
    if (panel == 32 && threads == 1024) RVGP_STEP(32, 1024);
#undef RVGP_STEP
    RVGP_CHECK_ARG(false, "rvgp_batch_step: (panel, threads) must be 16 or 32 and 512 or 1024 "
                          "(got %d, %d)", panel, threads);
    return RVGP_EINVAL;
}
