// Block-sparse (BSR) times dense with a fused affine epilogue:
//     Y = alpha (A X) + beta X + gamma Z ;   acc += delta Y   (optional)
// This one kernel family carries the eigensolver (Chebyshev filter steps
// Y_{k+1} = (2/e)(A - c) Y_k - Y_{k-1} and Rayleigh-Ritz products) and the
// heat-kernel recurrence exp(-tA) x (the Chebyshev sum accumulates into
// `acc`), replacing the sparse products inside ARPACK (RVGP/geometry.py:73)
// and the dense scipy expm (RVGP/smoothing.py:13-15).
//
// Layout: block rows i = 0..nb-1, blocks of b x b values row-major, stored in
// CSR order (rowptr / cols over block indices).  Dense operands are
// (nb*b, r) row-major with explicit row strides.
//
// Three mappings:
//  * narrow (r <= 4): LPR lanes share one block row and split its blocks
//    (nnz-parallel), reducing with cross-lane shuffles -- consecutive lanes
//    read consecutive blocks, so block values stream coalesced.
//  * vector (r > 4, 16-byte aligned rows): one wavefront per (block row,
//    column group of <= 64 16-byte vectors); each X row segment is one
//    coalesced load of up to 1 KiB.
//  * wide (r > 4, otherwise): one wavefront per (block row, 64*CPL-column
//    chunk), one 4/8-byte column per lane per load.
// In the wide forms the block row index is wave-uniform: the row's column
// indices and block values are fetched coalesced and broadcast with
// v_readlane.  Column chunks form the slow grid dimension so that all block
// rows of one chunk run together and the chunk of X they gather stays in
// cache.
#include "common.h"

// The library builds with -ffp-contract=off (the distance kernels need a
// fixed operation order); products here are accumulated as FMAs (one VALU op per multiply-add);
// nothing in this file is compared bit for bit with the oracle.
#pragma clang fp contract(fast)

namespace rvgp {

template <typename T, int B, int LPR, int RMAX>
__global__ __launch_bounds__(256) void spmm_narrow(
    int nb, const int* __restrict__ rowptr, const int* __restrict__ cols,
    const T* __restrict__ vals, const T* __restrict__ X, int ldx, int r, T alpha, T beta,
    const T* __restrict__ Z, int ldz, T gamma, T* __restrict__ Y, int ldy, T* __restrict__ acc,
    int ldacc, T delta) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = gid / LPR;
    const int l = gid - row * LPR;
    const bool live = row < nb;
    T y[B][RMAX];
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
        for (int c = 0; c < RMAX; ++c) y[a][c] = T(0);
    if (live) {
        const int e1 = rowptr[row + 1];
        for (int e = rowptr[row] + l; e < e1; e += LPR) {
            const int j = cols[e];
            const T* blk = vals + (size_t)e * B * B;
            T v[B][B];
#pragma unroll
            for (int a = 0; a < B; ++a)
#pragma unroll
                for (int bb = 0; bb < B; ++bb) v[a][bb] = blk[a * B + bb];
#pragma unroll
            for (int c = 0; c < RMAX; ++c) {
                if (c < r) {
                    T xb[B];
#pragma unroll
                    for (int bb = 0; bb < B; ++bb) xb[bb] = X[(size_t)(j * B + bb) * ldx + c];
#pragma unroll
                    for (int a = 0; a < B; ++a)
#pragma unroll
                        for (int bb = 0; bb < B; ++bb) y[a][c] += v[a][bb] * xb[bb];
                }
            }
        }
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1)
#pragma unroll
        for (int a = 0; a < B; ++a)
#pragma unroll
            for (int c = 0; c < RMAX; ++c) y[a][c] += __shfl_xor(y[a][c], o, LPR);
    if (!live) return;
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
        for (int c = 0; c < RMAX; ++c) {
            if (c < r && ((a * RMAX + c) % LPR) == l) {
                const size_t rr = (size_t)(row * B + a);
                T out = alpha * y[a][c];
                if (beta != T(0)) out += beta * X[rr * ldx + c];
                if (Z != nullptr && gamma != T(0)) out += gamma * Z[rr * ldz + c];
                Y[rr * ldy + c] = out;
                if (acc != nullptr) acc[rr * ldacc + c] += delta * out;
            }
        }
}

// Wide SpMM, vector form (16-byte rows).  When r, every leading dimension and
// every base pointer are multiples of the vector width VW (4 f32 / 2 f64),
// the r/VW vectors of a row are cut into column groups of gv <= 64 vectors;
// lane l of the wave for (block row, group g) owns vector g*gv + l, so one
// gather instruction moves up to 1 KiB of an X row.  Groups are the slow grid
// dimension: one group of every block row is swept before the next, which
// keeps the neighbour rows the in-flight rows of an XCD gather within its
// 4 MB L2 (a full-width sweep of r = 576 f32 columns needs ~6 KB per
// neighbour row and overflowed it, re-fetching X from HBM ~4x).  Block values
// and column indices are loaded and broadcast as in spmm_wide2.
template <typename T, int VW>
struct VecOf;
template <>
struct VecOf<float, 4> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct VecOf<double, 2> {
    typedef double type __attribute__((ext_vector_type(2)));
};

template <typename T, int B, int VW>
__global__ __launch_bounds__(256) void spmm_vec(
    int nb, const int* __restrict__ rowptr, const int* __restrict__ cols,
    const T* __restrict__ vals, const T* __restrict__ X, int ldx, int r, int gv, T alpha, T beta,
    const T* __restrict__ Z, int ldz, T gamma, T* __restrict__ Y, int ldy, T* __restrict__ acc,
    int ldacc, T delta) {
    using V = typename VecOf<T, VW>::type;
    constexpr int WPB = 4;
    constexpr int BB = B * B;
    const int lane = threadIdx.x & 63;
    // XCD-aware mapping: workgroups are dealt round-robin to the 8 XCDs, so
    // workgroup w runs on XCD w % 8; give each XCD one contiguous range of
    // block rows (gridDim.x is a multiple of 8) so that its L2 serves the
    // neighbouring X rows of consecutive rows.
    const int per = gridDim.x >> 3;
    const int lw = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int row = __builtin_amdgcn_readfirstlane(lw * WPB + (threadIdx.x >> 6));
    if (row >= nb) return;
    const int c = (blockIdx.y * gv + lane) * VW;
    const bool on = lane < gv && c < r;
    V y[B];
#pragma unroll
    for (int a = 0; a < B; ++a) y[a] = V(0);
    const int e0 = rowptr[row], e1 = rowptr[row + 1];
    for (int base = e0; base < e1; base += 64) {
        const int cnt = min(64, e1 - base);
        const int mycol = lane < cnt ? cols[base + lane] : 0;
        T mv[BB];
#pragma unroll
        for (int q = 0; q < BB; ++q) {
            const int f = lane + 64 * q;
            mv[q] = f < cnt * BB ? vals[(size_t)base * BB + f] : T(0);
        }
#pragma unroll
        for (int e = 0; e < 64; ++e) {
            if (e < cnt) {
                const int j = __builtin_amdgcn_readlane(mycol, e);
                T v[B][B];
#pragma unroll
                for (int a = 0; a < B; ++a)
#pragma unroll
                    for (int bb = 0; bb < B; ++bb) {
                        const int f = e * BB + a * B + bb;  // constant after unrolling
                        v[a][bb] = bcast(mv[f >> 6], f & 63);
                    }
                if (on) {
                    const T* xr = X + (size_t)j * B * ldx + c;
                    V xb[B];
#pragma unroll
                    for (int bb = 0; bb < B; ++bb)
                        xb[bb] = *reinterpret_cast<const V*>(xr + (size_t)bb * ldx);
#pragma unroll
                    for (int a = 0; a < B; ++a)
#pragma unroll
                        for (int bb = 0; bb < B; ++bb) y[a] += v[a][bb] * xb[bb];
                }
            }
        }
    }
    if (!on) return;
#pragma unroll
    for (int a = 0; a < B; ++a) {
        const size_t rr = (size_t)(row * B + a);
        V out = alpha * y[a];
        if (beta != T(0)) out += beta * *reinterpret_cast<const V*>(X + rr * ldx + c);
        if (Z != nullptr && gamma != T(0))
            out += gamma * *reinterpret_cast<const V*>(Z + rr * ldz + c);
        *reinterpret_cast<V*>(Y + rr * ldy + c) = out;
        if (acc != nullptr) {
            V* ap = reinterpret_cast<V*>(acc + rr * ldacc + c);
            *ap = *ap + delta * out;
        }
    }
}

// Wide SpMM, buffer form: the vector form with the gather address path moved
// off the vector ALU.  Column indices and block values are read with scalar
// loads (wave-uniform addresses) straight into SGPRs that feed the FMAs, and
// X is gathered with raw buffer loads: the per-lane column offset is a fixed
// VGPR, the neighbour row's byte offset a scalar -- so per stored block a
// lane issues B 16-byte loads and B*B*VW FMAs and nothing else.  Needs X to
// span < 2 GiB (32-bit offsets).
constexpr int kBufRsrcWord3 = 0x00020000;  // raw 32-bit dwords, gfx9 family

template <typename T, int B, int VW>
__global__ __launch_bounds__(256) void spmm_buf(
    int nb, const int* __restrict__ rowptr, const int* __restrict__ cols,
    const T* __restrict__ vals, const T* __restrict__ X, int ldx, int r, int gv, int xbytes,
    T alpha, T beta, const T* __restrict__ Z, int ldz, T gamma, T* __restrict__ Y, int ldy,
    T* __restrict__ acc, int ldacc, T delta) {
    using V = typename VecOf<T, VW>::type;
    constexpr int WPB = 4;
    constexpr int BB = B * B;
    const int lane = threadIdx.x & 63;
    const int per = gridDim.x >> 3;  // XCD-aware row ranges, as in spmm_vec
    const int lw = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int row = __builtin_amdgcn_readfirstlane(lw * WPB + (threadIdx.x >> 6));
    if (row >= nb) return;
    const int c = (blockIdx.y * gv + lane) * VW;
    // lanes past the row's columns leave before the gathers: an exec-masked
    // lane issues no address, so the wave's loads carry only live vectors
    if (!(lane < gv && c < r)) return;
    const __amdgpu_buffer_rsrc_t xr =
        __builtin_amdgcn_make_buffer_rsrc((void*)X, (short)0, xbytes, kBufRsrcWord3);
    const int voff = c * (int)sizeof(T);
    const int rowbytes = ldx * (int)sizeof(T);
    V y[B];
#pragma unroll
    for (int a = 0; a < B; ++a) y[a] = V(0);
    const int e0 = rowptr[row], e1 = rowptr[row + 1];
#pragma unroll 4
    for (int e = e0; e < e1; ++e) {
        const int j = cols[e];
        const T* blk = vals + (size_t)e * BB;
        V xb[B];
#pragma unroll
        for (int bb = 0; bb < B; ++bb)
            xb[bb] = __builtin_bit_cast(
                V, __builtin_amdgcn_raw_buffer_load_b128(xr, voff, (j * B + bb) * rowbytes, 0));
#pragma unroll
        for (int a = 0; a < B; ++a)
#pragma unroll
            for (int bb = 0; bb < B; ++bb) y[a] += blk[a * B + bb] * xb[bb];
    }
#pragma unroll
    for (int a = 0; a < B; ++a) {
        const size_t rr = (size_t)(row * B + a);
        V out = alpha * y[a];
        if (beta != T(0)) out += beta * *reinterpret_cast<const V*>(X + rr * ldx + c);
        if (Z != nullptr && gamma != T(0))
            out += gamma * *reinterpret_cast<const V*>(Z + rr * ldz + c);
        *reinterpret_cast<V*>(Y + rr * ldy + c) = out;
        if (acc != nullptr) {
            V* ap = reinterpret_cast<V*>(acc + rr * ldacc + c);
            *ap = *ap + delta * out;
        }
    }
}

// Half-storage form of spmm_buf: X, Z and Y are stored as IEEE binary16,
// block values stay f32 and every product and sum is f32; only the output is
// rounded to binary16 (round to nearest even).  A 16-byte gather moves 8
// columns instead of 4, so a sweep issues half the gathers and half the
// L1->L2 line requests of the f32 form for the same columns.  The
// eigensolver runs its opening Chebyshev passes on this form, where the
// residual targets (>= 1e-3) sit above binary16's 2^-11 rounding.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float8 __attribute__((ext_vector_type(8)));

template <int B>
__global__ __launch_bounds__(256) void spmm_buf_h(
    int nb, const int* __restrict__ rowptr, const int* __restrict__ cols,
    const float* __restrict__ vals, const _Float16* __restrict__ X, int ldx, int r, int gv,
    int xbytes, float alpha, float beta, const _Float16* __restrict__ Z, int ldz, float gamma,
    _Float16* __restrict__ Y, int ldy) {
    constexpr int WPB = 4;
    constexpr int BB = B * B;
    const int lane = threadIdx.x & 63;
    const int per = gridDim.x >> 3;  // XCD-aware row ranges, as in spmm_vec
    const int lw = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int row = __builtin_amdgcn_readfirstlane(lw * WPB + (threadIdx.x >> 6));
    if (row >= nb) return;
    const int c = (blockIdx.y * gv + lane) * 8;
    if (!(lane < gv && c < r)) return;
    const __amdgpu_buffer_rsrc_t xr =
        __builtin_amdgcn_make_buffer_rsrc((void*)X, (short)0, xbytes, kBufRsrcWord3);
    const int voff = c * 2;
    const int rowbytes = ldx * 2;
    float8 y[B];
#pragma unroll
    for (int a = 0; a < B; ++a) y[a] = float8(0.f);
    const int e0 = rowptr[row], e1 = rowptr[row + 1];
#pragma unroll 4
    for (int e = e0; e < e1; ++e) {
        const int j = cols[e];
        const float* blk = vals + (size_t)e * BB;
        float8 xb[B];
#pragma unroll
        for (int bb = 0; bb < B; ++bb)
            xb[bb] = __builtin_convertvector(
                __builtin_bit_cast(half8, __builtin_amdgcn_raw_buffer_load_b128(
                                              xr, voff, (j * B + bb) * rowbytes, 0)),
                float8);
#pragma unroll
        for (int a = 0; a < B; ++a)
#pragma unroll
            for (int bb = 0; bb < B; ++bb) y[a] += blk[a * B + bb] * xb[bb];
    }
#pragma unroll
    for (int a = 0; a < B; ++a) {
        const size_t rr = (size_t)(row * B + a);
        float8 out = alpha * y[a];
        if (beta != 0.f)
            out += beta * __builtin_convertvector(
                              *reinterpret_cast<const half8*>(X + rr * ldx + c), float8);
        if (Z != nullptr && gamma != 0.f)
            out += gamma * __builtin_convertvector(
                               *reinterpret_cast<const half8*>(Z + rr * ldz + c), float8);
        *reinterpret_cast<half8*>(Y + rr * ldy + c) = __builtin_convertvector(out, half8);
    }
}

// Wide SpMM, latency-hiding form.  One wavefront per (block row, column
// chunk).  The row's column indices and block values (up to 64 blocks at a
// time) are fetched with ONE coalesced load per register (lane l holds entry
// l of the flattened list) and broadcast with v_readlane; the block loop is
// fully unrolled, so every lane/register index is a compile-time constant
// and the X gathers of different blocks are independent loads the hardware
// keeps in flight together, instead of a dependent index->address->load
// chain per block.
template <typename T, int B, int CPL>
__global__ __launch_bounds__(256) void spmm_wide2(
    int nb, const int* __restrict__ rowptr, const int* __restrict__ cols,
    const T* __restrict__ vals, const T* __restrict__ X, int ldx, int r, T alpha, T beta,
    const T* __restrict__ Z, int ldz, T gamma, T* __restrict__ Y, int ldy, T* __restrict__ acc,
    int ldacc, T delta) {
    constexpr int WPB = 4;
    constexpr int BB = B * B;
    const int lane = threadIdx.x & 63;
    const int per = gridDim.x >> 3;
    const int lw = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int row = __builtin_amdgcn_readfirstlane(lw * WPB + (threadIdx.x >> 6));
    if (row >= nb) return;
    const int c0 = blockIdx.y * (64 * CPL) + lane;
    T y[B][CPL];
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
        for (int t = 0; t < CPL; ++t) y[a][t] = T(0);
    const int e0 = rowptr[row], e1 = rowptr[row + 1];
    for (int base = e0; base < e1; base += 64) {
        const int cnt = min(64, e1 - base);
        const int mycol = lane < cnt ? cols[base + lane] : 0;
        T mv[BB];
#pragma unroll
        for (int q = 0; q < BB; ++q) {
            const int f = lane + 64 * q;
            mv[q] = f < cnt * BB ? vals[(size_t)base * BB + f] : T(0);
        }
#pragma unroll
        for (int e = 0; e < 64; ++e) {
            if (e < cnt) {
                const int j = __builtin_amdgcn_readlane(mycol, e);
                T v[B][B];
#pragma unroll
                for (int a = 0; a < B; ++a)
#pragma unroll
                    for (int bb = 0; bb < B; ++bb) {
                        const int f = e * BB + a * B + bb;  // constant after unrolling
                        v[a][bb] = bcast(mv[f >> 6], f & 63);
                    }
                const T* xr = X + (size_t)j * B * ldx;
#pragma unroll
                for (int t = 0; t < CPL; ++t) {
                    const int c = c0 + 64 * t;
                    if (c < r) {
                        T xb[B];
#pragma unroll
                        for (int bb = 0; bb < B; ++bb) xb[bb] = xr[(size_t)bb * ldx + c];
#pragma unroll
                        for (int a = 0; a < B; ++a)
#pragma unroll
                            for (int bb = 0; bb < B; ++bb) y[a][t] += v[a][bb] * xb[bb];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < B; ++a) {
        const size_t rr = (size_t)(row * B + a);
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int c = c0 + 64 * t;
            if (c < r) {
                T out = alpha * y[a][t];
                if (beta != T(0)) out += beta * X[rr * ldx + c];
                if (Z != nullptr && gamma != T(0)) out += gamma * Z[rr * ldz + c];
                Y[rr * ldy + c] = out;
                if (acc != nullptr) acc[rr * ldacc + c] += delta * out;
            }
        }
    }
}

// The vector form needs r, every leading dimension and every base address to
// be multiples of the 16-byte vector.
template <typename T>
static bool vec_ok(int r, const T* X, int ldx, const T* Y, int ldy, const T* Z, int ldz,
                   const T* acc, int ldacc) {
    constexpr int VW = 16 / sizeof(T);
    auto al = [](const T* p, int ld) { return ((uintptr_t)p % 16) == 0 && ld % VW == 0; };
    return r % VW == 0 && al(X, ldx) && al(Y, ldy) && (Z == nullptr || al(Z, ldz)) &&
           (acc == nullptr || al(acc, ldacc));
}

template <typename T>
static int spmm_dispatch(int nb, int b, const int* rowptr, const int* cols, const T* vals,
                         const T* X, int ldx, int r, T alpha, T beta, const T* Z, int ldz,
                         T gamma, T* Y, int ldy, T* acc, int ldacc, T delta, void* stream) {
    RVGP_CHECK_ARG(nb >= 0 && r >= 1, "rvgp_bsr_spmm: need nb >= 0 and r >= 1");
    RVGP_CHECK_ARG(b >= 1 && b <= 3, "rvgp_bsr_spmm: block size must be 1, 2 or 3");
    RVGP_CHECK_ARG(ldx >= r && ldy >= r, "rvgp_bsr_spmm: leading dimension < r");
    RVGP_CHECK_ARG(Z == nullptr || ldz >= r, "rvgp_bsr_spmm: ldz < r");
    RVGP_CHECK_ARG(acc == nullptr || ldacc >= r, "rvgp_bsr_spmm: ldacc < r");
    RVGP_CHECK_ARG(X != Y, "rvgp_bsr_spmm: Y must not alias X");
    if (nb == 0) return RVGP_OK;
    hipStream_t s = as_stream(stream);
    if (r <= 4) {
        constexpr int LPR = 8;
        dim3 g(cdiv((size_t)nb * LPR, 256));
#define RVGP_NARROW(BB)                                                                         \
    hipLaunchKernelGGL((spmm_narrow<T, BB, LPR, 4>), g, dim3(256), 0, s, nb, rowptr, cols, vals, \
                       X, ldx, r, alpha, beta, Z, ldz, gamma, Y, ldy, acc, ldacc, delta)
        if (b == 1) RVGP_NARROW(1);
        else if (b == 2) RVGP_NARROW(2);
        else RVGP_NARROW(3);
#undef RVGP_NARROW
    } else if (vec_ok<T>(r, X, ldx, Y, ldy, Z, ldz, acc, ldacc)) {
        constexpr int VW = 16 / sizeof(T);
        const int nv = r / VW;                        // 16-byte vectors per row
        // column groups of <= 64 vectors, balanced, each a whole number of
        // 128-byte lines (a multiple of 8 vectors) so no group starts or ends
        // inside a cache line of a gathered X row
        int groups = (int)cdiv(nv, 64);
        int gv = (nv + groups - 1) / groups;
        gv = gv > 8 ? (gv + 7) / 8 * 8 : gv;
        groups = (int)cdiv(nv, gv);
        dim3 g(8 * cdiv(cdiv(nb, 4), 8), groups);
        const size_t xbytes = (size_t)nb * b * ldx * sizeof(T);
        if (xbytes < ((size_t)1 << 31)) {
#define RVGP_BUF(BB)                                                                          \
    hipLaunchKernelGGL((spmm_buf<T, BB, VW>), g, dim3(256), 0, s, nb, rowptr, cols, vals, X, \
                       ldx, r, gv, (int)xbytes, alpha, beta, Z, ldz, gamma, Y, ldy, acc,    \
                       ldacc, delta)
            if (b == 1) RVGP_BUF(1);
            else if (b == 2) RVGP_BUF(2);
            else RVGP_BUF(3);
#undef RVGP_BUF
            RVGP_LAUNCH_CHECK();
            return RVGP_OK;
        }
#define RVGP_VEC(BB)                                                                          \
    hipLaunchKernelGGL((spmm_vec<T, BB, VW>), g, dim3(256), 0, s, nb, rowptr, cols, vals, X, \
                       ldx, r, gv, alpha, beta, Z, ldz, gamma, Y, ldy, acc, ldacc, delta)
        if (b == 1) RVGP_VEC(1);
        else if (b == 2) RVGP_VEC(2);
        else RVGP_VEC(3);
#undef RVGP_VEC
    } else {
        // columns per lane: balance the column chunks (at most 4 f32 / 2 f64)
        const int cmax = sizeof(T) == 4 ? 4 : 2;
        const int nchunks = (r + 64 * cmax - 1) / (64 * cmax);
        const int cpl = (r + 64 * nchunks - 1) / (64 * nchunks);
        dim3 g(8 * cdiv(cdiv(nb, 4), 8), nchunks);
#define RVGP_WIDE2(BB, CC)                                                                      \
    hipLaunchKernelGGL((spmm_wide2<T, BB, CC>), g, dim3(256), 0, s, nb, rowptr, cols, vals, X, \
                       ldx, r, alpha, beta, Z, ldz, gamma, Y, ldy, acc, ldacc, delta)
#define RVGP_WIDE2_B(CC)                 \
    if (b == 1) RVGP_WIDE2(1, CC);       \
    else if (b == 2) RVGP_WIDE2(2, CC);  \
    else RVGP_WIDE2(3, CC);
        if (cpl == 1) { RVGP_WIDE2_B(1) }
        else if (cpl == 2) { RVGP_WIDE2_B(2) }
        else if constexpr (sizeof(T) == 4) {
            if (cpl == 3) { RVGP_WIDE2_B(3) }
            else { RVGP_WIDE2_B(4) }
        }
#undef RVGP_WIDE2_B
#undef RVGP_WIDE2
    }
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_bsr_spmm_f64(int nb, int b, const int* rowptr, const int* cols,
                                 const double* vals, const double* X, int ldx, int r,
                                 double alpha, double beta, const double* Z, int ldz,
                                 double gamma, double* Y, int ldy, double* acc, int ldacc,
                                 double delta, void* stream) {
    return spmm_dispatch<double>(nb, b, rowptr, cols, vals, X, ldx, r, alpha, beta, Z, ldz, gamma,
                                 Y, ldy, acc, ldacc, delta, stream);
}

extern "C" int rvgp_bsr_spmm_f16(int nb, int b, const int* rowptr, const int* cols,
                                 const float* vals, const void* X, int ldx, int r, float alpha,
                                 float beta, const void* Z, int ldz, float gamma, void* Y, int ldy,
                                 void* stream) {
    RVGP_CHECK_ARG(nb >= 0 && r >= 8 && r % 8 == 0,
                   "rvgp_bsr_spmm_f16: r must be a positive multiple of 8");
    RVGP_CHECK_ARG(b >= 1 && b <= 3, "rvgp_bsr_spmm_f16: block size must be 1, 2 or 3");
    RVGP_CHECK_ARG(ldx >= r && ldy >= r && ldx % 8 == 0 && ldy % 8 == 0 &&
                       (Z == nullptr || (ldz >= r && ldz % 8 == 0)),
                   "rvgp_bsr_spmm_f16: leading dimensions must be >= r and multiples of 8");
    RVGP_CHECK_ARG(((uintptr_t)X | (uintptr_t)Y | (uintptr_t)Z) % 16 == 0,
                   "rvgp_bsr_spmm_f16: X, Y and Z must be 16-byte aligned");
    RVGP_CHECK_ARG(X != Y, "rvgp_bsr_spmm_f16: Y must not alias X");
    const size_t xbytes = (size_t)nb * b * ldx * 2;
    RVGP_CHECK_ARG(xbytes < ((size_t)1 << 31), "rvgp_bsr_spmm_f16: X must span < 2 GiB");
    if (nb == 0) return RVGP_OK;
    RVGP_CHECK_ARG(rowptr != nullptr && cols != nullptr && vals != nullptr && X != nullptr &&
                       Y != nullptr,
                   "rvgp_bsr_spmm_f16: null pointer");
    // column groups of <= 64 8-column vectors, balanced, whole 128-byte lines
    const int nv = r / 8;
    int groups = (int)cdiv(nv, 64);
    int gv = (nv + groups - 1) / groups;
    gv = gv > 8 ? (gv + 7) / 8 * 8 : gv;
    groups = (int)cdiv(nv, gv);
    const dim3 g(8 * cdiv(cdiv(nb, 4), 8), groups);
    hipStream_t s = as_stream(stream);
    const _Float16* Xh = static_cast<const _Float16*>(X);
    const _Float16* Zh = static_cast<const _Float16*>(Z);
    _Float16* Yh = static_cast<_Float16*>(Y);
#define RVGP_BUFH(BB)                                                                        \
    hipLaunchKernelGGL((spmm_buf_h<BB>), g, dim3(256), 0, s, nb, rowptr, cols, vals, Xh, ldx, \
                       r, gv, (int)xbytes, alpha, beta, Zh, ldz, gamma, Yh, ldy)
    if (b == 1) RVGP_BUFH(1);
    else if (b == 2) RVGP_BUFH(2);
    else RVGP_BUFH(3);
#undef RVGP_BUFH
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_bsr_spmm_f32(int nb, int b, const int* rowptr, const int* cols,
                                 const float* vals, const float* X, int ldx, int r, float alpha,
                                 float beta, const float* Z, int ldz, float gamma, float* Y,
                                 int ldy, float* acc, int ldacc, float delta, void* stream) {
    return spmm_dispatch<float>(nb, b, rowptr, cols, vals, X, ldx, r, alpha, beta, Z, ldz, gamma,
                                Y, ldy, acc, ldacc, delta, stream);
}
