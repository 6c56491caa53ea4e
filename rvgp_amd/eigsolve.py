"""Smallest eigenpairs of the (connection) Laplacian on the GPU.

Replaces ``scipy.sparse.linalg.eigsh(L, k, which="SM")`` of
RVGP/geometry.py:66-80 (ARPACK Lanczos without shift-invert).

Method: a block Chebyshev-filtered subspace iteration (Zhou & Saad 2007,
"scaled" filter) with a Rayleigh-Ritz step.  Each iteration

  1. filters the current block X through a Chebyshev polynomial of A that
     is ~1 on [0, cut] and damped on [cut, ub] -- fused BSR SpMMs
     Y <- (2s/e)(A - c) Y - s s' Y_prev (rvgp_bsr_spmm), the only work that
     touches the n*d-sized operator.  Each column gets the degree its own
     Ritz value and residual need, so the later steps run on a shrinking
     trailing slice of the block;
  2. Rayleigh-Ritz in FP64: G = Y^T Y and H = Y^T (A Y) on the matrix cores
     (rvgp_gemm_sym, split-K), B = L^-T from the Cholesky factor of G (an
     eigendecomposition of G instead when G is numerically singular, which
     drops the dependent directions), eig(B^T H B) = Q diag(theta) Q^T,
     X = Y B Q and AX = A X (one more SpMM);
  3. residuals |AX - X theta| per column (rvgp_ritz_residuals) decide
     convergence; converged leading pairs are locked.

precision="fp64": everything in FP64 (the plain FP64 GPU form).
precision="mixed": the filter and the tall products run on FP32 blocks with
the Grams accumulated in FP64 (f32 -> f64 MFMA) -- the opening filter passes,
whose residual targets lie above binary16 rounding, on binary16 blocks (f32
arithmetic, half the gathered bytes); once the FP32 residuals
stall at FP32 level, FP64 polish iterations (FP64 SpMM + FP64 RR) bring the
wanted pairs to the FP64 tolerance.  The small projected problems (p x p,
p = k + guard) are solved on the GPU in FP64 by the library's own symmetric
eigensolver (dense.syev: tridiagonalisation, bisection, inverse iteration),
and the p x p products run on its matrix-core GEMM.

Returns eigenvalues ascending and eigenvectors with unit 2-norm times the
caller's scale (sqrt(rows) as geometry.py:75), in the operator's original
block-row numbering (rvgp_evecs_finish).
"""
import math
import sys
import time

import numpy as np
import torch

from . import dense
from .sparse import BSR, spmm


# block sizes whose FP32 phase filters twice (Cholesky QR in between) while the
# degree is capped by conditioning
DOUBLE_FILTER_BLOCKS = (1,)
# filter degrees chosen per column (shrinking trailing slices of the block)
# rather than one degree for the whole block
ADAPTIVE_DEGREES = True
# largest amplification spread a filter may put across the block: FP32
# blocks keep ~300 (the weakest directions then carry ~2e-5 relative
# precision); FP64 blocks FP64_MAX_RATIO (the Rayleigh-Ritz Gram's condition
# is its square, which the Cholesky basis must still resolve)
FP32_MAX_RATIO = 300.0
FP64_MAX_RATIO = 1e4
# the same cap for the mixed form's FP64 polish (its block is already
# accurate to the FP32 floor)
POLISH_MAX_RATIO = 1e4
# the mixed form's FP32 phase hands over to FP64 at this residual (relative to
# the Gershgorin bound), a few times above the FP32 floor
MIXED_HANDOVER = 1e-5
# FP32-phase filter passes that start from a worst residual above this run
# on binary16 blocks (rvgp_bsr_spmm_f16: f32 values and arithmetic, each
# output rounded to 2^-11): on the 3-torus operators (n = 1e5, k = 500) the
# last such pass still ends at the FP32 pass's residual (~5e-4); a pass
# started below ~2e-3 on binary16 ends higher and costs later passes more
# than it saved.  The FP32 passes and the FP64 polish after them set the
# final accuracy, so this threshold moves only the time.
HALF_ABOVE = 2e-3
# a filter is given the degree that would cut the worst wanted residual by
# NEED_SAFETY x the factor still needed
NEED_SAFETY = 10.0


class EigStats(dict):
    pass


def _rayleigh_ritz(A, Y, work_dtype):
    """Returns X (N x p'), theta (p',) f64, AX (N x p') in work dtype."""
    W = spmm(A, Y)
    # both Grams are symmetric (A is): upper tiles only, mirrored
    G = dense.gemm_sym(Y, Y)
    H = dense.gemm_sym(Y, W)
    drop = 1e-13 if work_dtype == torch.float64 else 1e-10
    Bm = _gram_basis_chol(G, drop)
    if Bm is None:
        # numerically dependent block: eigendecomposition of G drops the
        # directions below the rounding level instead of failing
        g, V = dense.syev(G)
        gmax = float(g[-1].item())
        keep = g > gmax * drop
        Bm = (V[:, keep] / torch.sqrt(g[keep])).contiguous()
    Hb = dense.gemm(Bm, dense.gemm(H, Bm), transA=True)
    Hb = 0.5 * (Hb + Hb.T)
    theta, Q = dense.syev(Hb)
    Cm = dense.gemm(Bm, Q).to(work_dtype).contiguous()
    X = dense.gemm(Y, Cm)
    # A X by one more block SpMM rather than W Cm: at N = 3e5, p = 576 the
    # SpMM reads ~2 GB (1.1 ms FP32, 2.1 ms FP64 for b = 3) where the GEMM
    # costs 2 N p^2 flops (3.3 ms FP32, ~10 ms FP64)
    AX = spmm(A, X)
    return X, theta, AX


def _gram_basis_chol(G, drop):
    """B with B^T G B = I from the Cholesky factor G = L L^T (B = L^-T), on the
    matrix cores; None when G is not numerically positive definite at the
    level `drop` (smallest pivot^2 below drop * largest diagonal)."""
    p = G.shape[0]
    L = G.clone()
    _, st = dense.potrf(L, check=False)
    d = torch.diagonal(L)
    ok = torch.logical_and(st[0] == 0, (d.min() ** 2) > drop * torch.diagonal(G).max())
    if not bool(ok.item()):
        return None
    Linv = torch.eye(p, dtype=G.dtype, device=G.device)
    dense.trsm(torch.tril(L).contiguous(), Linv)
    return Linv.T.contiguous()


def _cholqr(Y):
    """Y R^-1 with R^T R = Y^T Y (Cholesky QR; FP64 Gram and factor, result in
    Y's dtype), or None when the Gram is not numerically positive definite."""
    G = dense.gemm_sym(Y, Y)
    B = _gram_basis_chol(G, 1e-13 if Y.dtype == torch.float64 else 1e-10)
    if B is None:
        return None
    return dense.gemm(Y, B.to(Y.dtype).contiguous())


def _residuals(X, theta, AX):
    """|A x_j - theta_j x_j| / |x_j| per column (one pass, sums in FP64)."""
    return dense.ritz_residuals(X, AX, theta)


def _cheb_filter(A, X, m, a, b, a0):
    """Scaled Chebyshev filter of degree m (X is overwritten)."""
    e = (b - a) / 2.0
    c = (b + a) / 2.0
    sigma = e / (a0 - c)
    tau = 2.0 / sigma
    Y = spmm(A, X, alpha=sigma / e, beta=-sigma * c / e)
    prev, cur = X, Y
    s1 = sigma
    for _ in range(2, m + 1):
        s2 = 1.0 / (tau - s1)
        spmm(A, cur, Y=prev, alpha=2.0 * s2 / e, beta=-2.0 * s2 * c / e, Z=prev, gamma=-s1 * s2)
        prev, cur = cur, prev
        s1 = s2
    return cur


def _cheb_filter_cols(A, X, ms, a, b, a0):
    """Scaled Chebyshev filter with a degree per column.  ms (numpy ints) is
    non-decreasing along the block and constant over groups of 8 columns;
    the three-term recurrence acts column by column, so after step s only the
    columns whose degree exceeds s go on -- a shrinking trailing slice of the
    block -- and the others keep their last iterate.  X is overwritten; the
    result is returned (X's storage or a new block)."""
    m = int(ms[-1])
    if int(ms[0]) == m:
        return _cheb_filter(A, X, m, a, b, a0)
    e = (b - a) / 2.0
    c = (b + a) / 2.0
    sigma = e / (a0 - c)
    tau = 2.0 / sigma
    Y = spmm(A, X, alpha=sigma / e, beta=-sigma * c / e)
    bufs = (X, Y)  # the iterate of step s is in bufs[s % 2]
    # starts[s] = first column whose degree is at least s
    starts = np.searchsorted(ms, np.arange(m + 2), side="left")
    s1 = sigma
    for step in range(2, m + 1):
        c0 = int(starts[step])
        cur, prev = bufs[(step - 1) % 2], bufs[step % 2]
        s2 = 1.0 / (tau - s1)
        spmm(A, cur[:, c0:], Y=prev[:, c0:], alpha=2.0 * s2 / e, beta=-2.0 * s2 * c / e,
             Z=prev[:, c0:], gamma=-s1 * s2)
        s1 = s2
    out = bufs[m % 2]
    for step in range(1, m):  # columns that stopped on a step of the other parity
        c0, c1 = int(starts[step]), int(starts[step + 1])
        if step % 2 != m % 2 and c1 > c0:
            out[:, c0:c1].copy_(bufs[step % 2][:, c0:c1])
    return out


def _half_ok(X):
    """Whether block X can be filtered on binary16 storage: whole 16-byte
    rows of 8 columns, X spanning < 2 GiB at 2 bytes an entry."""
    return X.shape[1] % 8 == 0 and X.shape[0] * X.shape[1] * 2 < 2 ** 31


def _filter_pass(A, X, ms, m, a, b, a0, half=False):
    """One Chebyshev filter pass over block X (degrees ms per column, or m
    for every column when ms is None); returns the filtered block in X's
    dtype.  half=True stores the block as binary16 for the pass, each column
    first scaled to unit RMS: the filtered subspace is what the Rayleigh-Ritz
    step uses, not the column scales.  The binary16 pass normalises the
    filter at 0 instead of a0: the polynomial differs only by a constant
    factor, and on a PSD operator no eigen-direction is then amplified at any
    step (a random block's a0 is an estimate well inside the spectrum, below
    which directions grow by up to ~1e4 in a pass -- past binary16's 65504)."""
    if half:
        a0 = min(a0, 0.0)
        nrm = dense.colsumsq(X)
        f = torch.rsqrt(torch.clamp(nrm / X.shape[0], min=1e-300)).to(X.dtype)
        Xh = (X * f).to(torch.float16)
        Yh = _filter_pass(A, Xh, ms, m, a, b, a0)
        return Yh.to(X.dtype)
    if ms is not None:
        return _cheb_filter_cols(A, X, ms, a, b, a0)
    return _cheb_filter(A, X, m, a, b, a0)


def _column_degrees(theta, res, worst, ncols, kk, cut, ub, m, ptol, m_min=4):
    """Per-column filter degrees for a block whose first min(kk, len(theta))
    columns are wanted Ritz vectors (Ritz values theta, residuals res, or the
    worst residual for all when res is None): column i gets the degree that
    would cut its own residual by NEED_SAFETY x the factor still needed,
    ln(need_i) / acosh(x_i) with x_i its Ritz value mapped to the damped
    interval, at most m; guard and fresh columns get m.  The degrees are made
    non-decreasing along the block and constant over groups of 8 columns
    (16-byte SpMM rows), so the columns still filtering at any step form a
    trailing slice."""
    ms = np.full(ncols, m, dtype=np.int64)
    nt = min(len(theta), kk, ncols)
    if nt > 0:
        e = (ub - cut) / 2.0
        c = (ub + cut) / 2.0
        x = (c - np.asarray(theta[:nt], dtype=np.float64)) / e
        a = np.arccosh(np.maximum(x, 1.0 + 1e-15))
        r = np.full(nt, worst) if res is None else np.asarray(res[:nt], dtype=np.float64)
        need = np.maximum(NEED_SAFETY * r / ptol, 1.0 + 1e-12)
        ms[:nt] = np.clip(np.ceil(np.log(need) / a), m_min, m).astype(np.int64)
    ms = np.maximum.accumulate(ms)
    g = -(-ncols // 8)
    grp = np.concatenate([ms, np.full(g * 8 - ncols, ms[-1])]).reshape(g, 8).max(1)
    return np.repeat(grp, 8)[:ncols]


def _degree(theta, k, cut, ub, max_ratio, m_min=4, m_max=60, need=None):
    """Largest degree whose amplification spread over the block stays below
    max_ratio (keeps the filtered block well enough conditioned), and no
    larger than the Chebyshev estimate of what the worst wanted residual still
    needs: a degree-m filter damps the unwanted end against the k-th Ritz
    value by ~exp(-m acosh(x_k)), so m = ln(need) / acosh(x_k) (need = the
    residual reduction still required, with a safety factor applied by the
    caller)."""
    e = (ub - cut) / 2.0
    c = (ub + cut) / 2.0
    x0 = (c - float(theta[0])) / e
    xk = (c - float(theta[min(k, len(theta)) - 1])) / e
    a0 = math.acosh(max(x0, 1.0 + 1e-15))
    ak = math.acosh(max(xk, 1.0 + 1e-15))
    m = m_max if a0 - ak <= 1e-12 else int(math.log(max_ratio) / (a0 - ak))
    m_need = None
    if need is not None and need > 1.0 and ak > 1e-12:
        m_need = int(math.ceil(math.log(need) / ak))
        m = min(m, m_need)
    return max(m_min, min(m_max, m)), m_need


def _random_block_bracket(A, p):
    """Where the Ritz values of a random p-column block of A lie, without
    forming it: around the mean eigenvalue mu = tr(A) / N, within
    +-2 sigma sqrt(p / N), sigma^2 = |A|_F^2 / N - mu^2 the eigenvalue variance
    (the edge of the Marchenko-Pastur law for a random projection).  The
    opening filter of a random block takes its bounds and degree from this
    bracket in place of a Rayleigh-Ritz step on the unfiltered block."""
    N = A.shape[0]
    b = A.b
    rows = torch.repeat_interleave(torch.arange(A.nb, device=A.cols.device),
                                   torch.diff(A.rowptr.long()))
    blocks = A.vals64.reshape(A.nnzb, b, b)
    diag = blocks[A.cols.long() == rows]
    tr = float(torch.diagonal(diag, dim1=1, dim2=2).sum().item())
    fro2 = float((A.vals64 * A.vals64).sum().item())
    mu = tr / N
    sig = math.sqrt(max(fro2 / N - mu * mu, 0.0))
    half = 2.0 * sig * math.sqrt(p / N)
    lo, hi = max(mu - half, 0.0), mu + half
    return torch.tensor([lo, hi], dtype=torch.float64, device=A.cols.device)


def _deflate(Y, Xl):
    """Y <- Y - Xl (Xl^T Y), twice (classical Gram-Schmidt with reorthogonalisation)."""
    if Xl is None or Xl.shape[1] == 0:
        return Y
    for _ in range(2):
        C = dense.gemm(Xl, Y, transA=True, out_dtype=torch.float64).to(Y.dtype).contiguous()
        dense.gemm(Xl, C, Y, alpha=-1.0, beta=1.0)
    return Y


def _run_phase(A, X, k, ub, wdt, ptol, max_iter, max_ratio, lock, fresh, stats, verbose, tag,
               stall_exit=False, m_max=60, ritz=None, scale=None, half_above=None):
    """Filter + Rayleigh-Ritz iterations on block X (work dtype wdt) until the
    k wanted pairs have residual/ub < ptol.  With ``lock``, converged leading
    pairs are frozen and deflated out of the filtered block.  ``ritz`` =
    (theta, worst residual, per-column residuals or None) of X when X
    already is a Ritz basis (the previous phase's result; residuals relative
    to ``scale``): the phase then opens with a filter pass instead of a
    Rayleigh-Ritz step, which would only re-derive those values.
    ``ub`` bounds the spectrum for the filters; residuals are measured
    relative to ``scale`` (default ub).  Filter passes that start from a
    worst residual above ``half_above`` run on binary16 blocks."""
    scale = ub if scale is None else scale
    p = X.shape[1]
    history = []
    locked_res = np.zeros(0)  # residual / scale of each locked pair when it was locked
    Xl = X[:, :0]
    thl = torch.zeros(0, dtype=torch.float64, device=X.device)
    Y = X
    it = 0
    if ritz is not None:
        th0, worst0, res0 = ritz
        tf = time.perf_counter()
        cut, low = float(th0.max().item()), float(th0.min().item())
        m, m_need = _degree(torch.sort(th0).values, min(k, th0.numel()), cut, ub, max_ratio,
                            m_max=m_max, need=NEED_SAFETY * worst0 / ptol)
        half = half_above is not None and worst0 > half_above and _half_ok(X)
        if verbose:
            print(f"[chfsi {tag}] opening filter m={m} m_need={m_need} worst={worst0:.3e}"
                  f"{' binary16' if half else ''}", file=sys.stderr, flush=True)
        ms = None
        if ADAPTIVE_DEGREES:
            th_h = th0.cpu().numpy()
            ms = _column_degrees(th_h, res0, worst0, X.shape[1], min(k, th_h.size), cut, ub, m, ptol)
        Y = _filter_pass(A, X, ms, m, cut, ub, low, half)
        stats["spmm_cols"] += int(ms.sum()) if ms is not None else m * X.shape[1]
        stats["spmm"] += m
        stats["half_passes"] += int(half)
        torch.cuda.current_stream().synchronize()
        stats["t_filter"] += time.perf_counter() - tf
        stats["iters"] += 1
    while True:
        tr = time.perf_counter()
        Y = _deflate(Y, Xl)
        Xa, tha, AXa = _rayleigh_ritz(A, Y, wdt)
        stats["rr"] += 1
        stats["spmm"] += 2
        stats["t_rr"] += time.perf_counter() - tr
        res = _residuals(Xa, tha, AXa) / scale
        need = k - Xl.shape[1]
        kk = min(need, res.numel())
        worst = float(res[:kk].max().item()) if kk > 0 else 0.0
        if verbose:
            print(f"[chfsi {tag}] it={it} locked={Xl.shape[1]} active={Xa.shape[1]} "
                  f"worst_res={worst:.3e} spmm={stats['spmm']}", file=sys.stderr, flush=True)
        stats["last_worst"] = worst
        stats["phase_residual"] = max([worst, *locked_res])
        res_h = res.cpu().numpy()
        # per-column residuals of the block this phase returns (locked pairs
        # as locked), for the next phase's per-column filter degrees
        stats["last_res"] = np.concatenate([locked_res, res_h])
        if kk >= need and worst < ptol:
            X = torch.cat([Xl, Xa], 1)
            return X, torch.cat([thl, tha]), it, True
        # a phase whose residuals have stopped falling (the FP32 rounding
        # floor of the mixed form) hands over to the next phase
        history.append(worst)
        stalled = stall_exit and len(history) > 3 and min(history[-2:]) > 0.5 * min(history[:-2])
        if it >= max_iter or stalled:
            if stalled:
                stats["stalled_" + tag] = worst
            X = torch.cat([Xl, Xa], 1)
            stats["unconverged"] = worst
            return X, torch.cat([thl, tha]), it, False
        c = 0
        if lock:
            while c < kk and res_h[c] < ptol:
                c += 1
            c = min(c, max(0, Xa.shape[1] - 8))  # keep a working block
            c -= c % 8  # block widths stay multiples of 8: 16-byte SpMM rows
            if c > 0:
                locked_res = np.concatenate([locked_res, res_h[:c]])
                Xl = torch.cat([Xl, Xa[:, :c]], 1).contiguous()
                thl = torch.cat([thl, tha[:c]])
                Xa, tha = Xa[:, c:].contiguous(), tha[c:]
        # keep the total width p: top the active block up with random columns
        if Xl.shape[1] + Xa.shape[1] < p:
            Xa = torch.cat([Xa, fresh(p - Xl.shape[1] - Xa.shape[1], wdt)], 1).contiguous()
        cut = float(tha[min(tha.numel(), Xa.shape[1]) - 1].item())
        low = float(tha[0].item())
        kk = min(k - Xl.shape[1], tha.numel())
        m, m_need = _degree(tha, max(kk, 1), cut, ub, max_ratio, m_max=m_max,
                            need=NEED_SAFETY * worst / ptol if kk > 0 else None)
        half = half_above is not None and worst > half_above and _half_ok(Xa)
        if verbose:
            print(f"[chfsi {tag}] filter m={m} m_need={m_need} cut={cut:.4e} low={low:.4e}"
                  f"{' binary16' if half else ''}", file=sys.stderr, flush=True)
        tf = time.perf_counter()
        ms = None
        if ADAPTIVE_DEGREES:
            ms = _column_degrees(tha.cpu().numpy(), res_h[c:], worst, Xa.shape[1], max(kk, 0), cut,
                                 ub, m, ptol)
        Y = _filter_pass(A, Xa, ms, m, cut, ub, low, half)
        stats["spmm_cols"] += int(ms.sum()) if ms is not None else m * Xa.shape[1]
        stats["spmm"] += m
        stats["half_passes"] += int(half)
        if (A.b in DOUBLE_FILTER_BLOCKS and wdt == torch.float32 and Xl.shape[1] == 0
                and m_need is not None
                and m_need > m):
            # FP32 phase of a scalar operator, degree capped by the block's
            # conditioning rather than by the residual target:
            # re-orthonormalise (Cholesky QR; no locked vectors to deflate
            # yet) and filter again with the same bounds, in place of a full
            # Rayleigh-Ritz between the two passes.  Only where a block SpMM
            # is cheap against a Rayleigh-Ritz pass (b = 1: ~48 SpMMs per RR
            # at p = 576 on MI355X; b = 3: ~18): elsewhere the second
            # full-width pass costs more than the RR it saves, because it
            # also delays locking.
            Q = _cholqr(Y)
            if Q is not None:
                m2 = max(4, min(m, m_need - m))
                Y = _filter_pass(A, Q, None, m2, cut, ub, low, half)
                stats["spmm"] += m2
                stats["spmm_cols"] += m2 * Q.shape[1]
                stats["cholqr"] = stats.get("cholqr", 0) + 1
        torch.cuda.current_stream().synchronize()  # this solve's stream only
        stats["t_filter"] += time.perf_counter() - tf
        stats["iters"] += 1
        it += 1


def chfsi(A: BSR, k, precision="fp64", tol=1e-9, max_iter=300, guard=None, seed=0,
          verbose=False, X0=None, perm=None, scale=1.0):
    """k smallest eigenpairs of the symmetric PSD operator A: eigenvalues
    ascending, eigenvectors as unit-norm columns times ``scale``, with block
    rows mapped back through ``perm`` (perm[new] = old block row) when A is a
    renumbered operator.  stats["residual"] bounds |A x - theta x| / |x| over
    the returned pairs, relative to the Gershgorin bound."""
    N = A.shape[0]
    if k >= N - 1:
        raise TypeError("Cannot use scipy.linalg.eigh for sparse A with k >= N - 1. "
                        "Use scipy.linalg.eigh(A.toarray()) or reduce k.")
    if k < 1:
        raise ValueError("k must be positive")
    dev = A.rowptr.device
    if guard is None:
        guard = max(8, int(math.ceil(0.15 * k)))
    p = min(N, -(-(k + guard) // 8) * 8)
    # residuals are measured against the Gershgorin bound (a norm scale that
    # does not depend on a random start); the filters use the tighter
    # Lanczos bound of lambda_max
    scale_g = A.gershgorin_bound() * 1.0001 + 1e-12
    ub = min(scale_g, A.spectral_upper_bound() * 1.0001 + 1e-12)
    stats_ub = ub
    stats = EigStats(spmm=0, spmm_cols=0, rr=0, iters=0, t_filter=0.0, t_rr=0.0, half_passes=0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))

    def fresh(cols, dtype):
        return torch.randn((N, cols), generator=gen, device=dev, dtype=torch.float64).to(dtype)

    t0 = time.perf_counter()
    X = X0.contiguous() if X0 is not None else fresh(p, torch.float64)
    # a random block opens with a filter pass bounded by where its Ritz
    # values would lie, instead of a Rayleigh-Ritz step on the unfiltered block
    opening = None if X0 is not None else (_random_block_bracket(A, p), 1.0, None)
    if precision == "mixed":
        # FP32 filter body with locking; hands over at 1e-5, a few times above
        # the FP32 residual floor (a phase that reaches the floor first
        # stalls there)
        X, theta, it, _ = _run_phase(A, X.to(torch.float32).contiguous(), k, ub, torch.float32,
                                     MIXED_HANDOVER, max_iter, FP32_MAX_RATIO, True, fresh, stats,
                                     verbose, "fp32", stall_exit=True, scale=scale_g,
                                     ritz=opening, half_above=HALF_ABOVE)
        stats["iters_fp32"] = it
        stats["t_fp32"] = time.perf_counter() - t0
        # FP64 polish with locking: the FP32 block is already accurate to
        # ~1e-5, so a few filter passes finish the pairs
        X, theta, it, done = _run_phase(A, X.double().contiguous(), k, ub, torch.float64, tol, 50,
                                        POLISH_MAX_RATIO, True, fresh, stats, verbose, "fp64-polish",
                                        ritz=(theta, stats["last_worst"], stats["last_res"]),
                                        scale=scale_g)
        if not done:
            # an unlocked Rayleigh-Ritz over the whole block orders and refines
            X, theta, _, _ = _run_phase(A, X, k, ub, torch.float64, tol, 3, FP64_MAX_RATIO, False, fresh,
                                        stats, verbose, "fp64-final", m_max=16, scale=scale_g)
        stats["iters_polish"] = it
    elif precision == "fp64":
        X, theta, it, done = _run_phase(A, X.to(torch.float64).contiguous(), k, ub, torch.float64,
                                        tol, max_iter, FP64_MAX_RATIO, True, fresh, stats, verbose, "fp64",
                                        scale=scale_g, ritz=opening)
        if not done:
            X, theta, _, _ = _run_phase(A, X, k, ub, torch.float64, tol, 3, FP64_MAX_RATIO, False, fresh,
                                        stats, verbose, "fp64-final", m_max=16, scale=scale_g)
    else:
        raise ValueError("precision must be 'fp64' or 'mixed'")
    o = torch.argsort(theta)
    theta = theta[o]
    in_order = bool(torch.equal(o[:k], torch.arange(k, device=o.device)))
    Xk = X[:, :k] if in_order else X[:, o[:k]].contiguous()
    # the last phase's residuals (relative to the Gershgorin bound): the
    # active pairs' at its final Rayleigh-Ritz step, the locked pairs' when
    # they were locked
    stats["residual"] = float(stats.pop("phase_residual"))
    stats.pop("last_res", None)
    stats["ub_filter"] = stats_ub
    stats["ub_gershgorin"] = scale_g
    evals = theta[:k].clone()
    evecs = dense.evecs_finish(Xk, A.b, scale, perm)
    stats["t_total"] = time.perf_counter() - t0
    return evals, evecs, stats


def spectrum_of_scipy(laplacian, n_eigenpairs=None):
    """compute_spectrum drop-in for a host scipy matrix (geometry.py:66-80)."""
    N = laplacian.shape[0]
    k = N if n_eigenpairs is None else min(int(n_eigenpairs), N)
    A = BSR.from_scipy(laplacian, 1)
    evals, evecs, _ = chfsi(A, k, scale=math.sqrt(N))  # geometry.py:75
    return evals.cpu().numpy(), evecs.cpu().numpy()
