"""Heat-kernel smoothing on the device -- drop-in for RVGP/smoothing.py.

The reference forms exp(-t L) densely with scipy.linalg.expm
(smoothing.py:13-15), O((n d)^3) time and O((n d)^2) memory.  Here
exp(-t A) X is a Chebyshev expansion on the spectral interval [0, ub]
(ub >= lambda_max: the Lanczos bound of BSR.spectral_upper_bound, at most
the Gershgorin bound):

    exp(-t lambda) = sum_j c_j T_j(2 lambda / ub - 1),
    c_j = (2 - [j = 0]) (-1)^j I_j(t ub / 2) exp(-t ub / 2),

with I_j the modified Bessel functions (scipy.special.ive on the host, a
handful of scalars).  The three-term recurrence runs as fused BSR SpMMs that
also accumulate c_j T_j X into the result (rvgp_bsr_spmm's `acc` epilogue),
so the operator is applied once per term with no extra passes.  The series is
truncated when the remaining coefficients sum below 1e-17 of the first.
"""
import numpy as np
import scipy.special
import torch

from . import _lib
from .sparse import BSR, spmm


def cheb_coeffs(t, ub, rtol=1e-17, max_terms=200000):
    tau = 0.5 * float(t) * float(ub)
    # number of terms: I_j(tau) e^-tau decays like exp(-j^2 / (2 tau)) once j < tau
    jmax = int(np.ceil(np.sqrt(2.0 * max(tau, 1.0) * 45.0))) + 30
    jmax = min(jmax, max_terms)
    j = np.arange(jmax + 1)
    c = scipy.special.ive(j, tau) * np.where(j % 2 == 0, 1.0, -1.0) * 2.0
    c[0] *= 0.5
    tail = np.cumsum(np.abs(c[::-1]))[::-1]
    cut = np.nonzero(tail < rtol * abs(c[0]))[0]
    J = int(cut[0]) if len(cut) else len(c)
    return c[:max(J, 2)]


def expm_multiply(A: BSR, X, t, ub=None, shift=0.0):
    """exp(-t (A - shift I)) X for a PSD BSR operator A (X: (rows,) or
    (rows, r)); shift must not exceed A's smallest eigenvalue.  t is a
    scalar, or one diffusion time per column of X: the recurrence T_j X then
    runs once, to the longest series, and each column accumulates its own
    coefficients.

    The series' error is absolute -- about 1e-17 of |X| -- so when exp(-t A)
    X is itself tiny (a large t over a spectrum bounded away from 0, e.g. a
    connection Laplacian on a curved surface) its relative accuracy is lost;
    with shift = lambda_min the result keeps its full relative accuracy and
    exp(-t shift) is the caller's (normalised diffusion divides it out)."""
    if ub is None:
        ub = A.spectral_upper_bound() * 1.0001 + 1e-12
    shift = float(shift)
    if shift > 0.0:
        # the series of exp(-t lambda') on [0, ub - shift] in lambda' = lambda - shift
        return _expm_multiply_shifted(A, X, t, ub, shift)
    X2 = X.reshape(-1, 1) if X.dim() == 1 else X
    ts = np.atleast_1d(np.asarray(t, dtype=np.float64))
    if ts.size > 1:
        if ts.size != X2.shape[1]:
            raise ValueError(f"{ts.size} diffusion times for {X2.shape[1]} columns")
        cs = [cheb_coeffs(tc, ub) for tc in ts]
        C = np.zeros((max(len(c) for c in cs), ts.size))
        for col, c in enumerate(cs):
            C[:len(c), col] = c
        Cd = torch.as_tensor(C, dtype=X2.dtype, device=X2.device)
        out = X2 * Cd[0]
        T1 = spmm(A, X2, alpha=2.0 / ub, beta=-1.0)
        out.addcmul_(T1, Cd[1])
        prev, cur = X2.clone(), T1
        for j in range(2, C.shape[0]):
            spmm(A, cur, Y=prev, alpha=4.0 / ub, beta=-2.0, Z=prev, gamma=-1.0)
            out.addcmul_(prev, Cd[j])
            prev, cur = cur, prev
        return out.reshape(X.shape)
    c = cheb_coeffs(float(ts[0]), ub)
    out = X2 * float(c[0])
    if len(c) > 1:
        # T_1 = (2/ub) A X - X
        T1 = spmm(A, X2, alpha=2.0 / ub, beta=-1.0, acc=out, delta=float(c[1]))
        prev, cur = X2.clone(), T1
        for j in range(2, len(c)):
            # T_j = (4/ub) A T_{j-1} - 2 T_{j-1} - T_{j-2}, written over T_{j-2}
            spmm(A, cur, Y=prev, alpha=4.0 / ub, beta=-2.0, Z=prev, gamma=-1.0, acc=out,
                 delta=float(c[j]))
            prev, cur = cur, prev
    return out.reshape(X.shape)


def _expm_multiply_shifted(A, X, t, ub, shift):
    ubs = ub - shift
    X2 = X.reshape(-1, 1) if X.dim() == 1 else X
    ts = np.atleast_1d(np.asarray(t, dtype=np.float64))
    cs = [cheb_coeffs(tc, ubs) for tc in ts]
    C = np.zeros((max(len(c) for c in cs), X2.shape[1]))
    for col in range(X2.shape[1]):
        c = cs[col if ts.size > 1 else 0]
        C[:len(c), col] = c
    Cd = torch.as_tensor(C, dtype=X2.dtype, device=X2.device)
    # T_1 = (2/ubs)(A - shift) X - X ; T_j = (4/ubs)(A - shift) T_{j-1} - 2 T_{j-1} - T_{j-2}
    out = X2 * Cd[0]
    if C.shape[0] > 1:
        T1 = spmm(A, X2, alpha=2.0 / ubs, beta=-1.0 - 2.0 * shift / ubs)
        out.addcmul_(T1, Cd[1])
        prev, cur = X2.clone(), T1
        for j in range(2, C.shape[0]):
            spmm(A, cur, Y=prev, alpha=4.0 / ubs, beta=-2.0 - 4.0 * shift / ubs, Z=prev,
                 gamma=-1.0)
            out.addcmul_(prev, Cd[j])
            prev, cur = cur, prev
    return out.reshape(X.shape)


def lambda_min_lower(A: BSR, evals=None, residual=None):
    """A value at or below A's smallest eigenvalue, for the heat-kernel
    shift: from a computed smallest eigenvalue and its residual bound
    (|theta - lambda| <= |A x - theta x|), or from a one-pair eigensolve."""
    scale = A.gershgorin_bound()
    if evals is None:
        from .eigsolve import chfsi
        w, _, st = chfsi(A, 1, precision="fp64")
        theta, res = float(w[0].item()), float(st["residual"]) * scale
    else:
        theta, res = float(evals), float(residual) * scale
    return max(0.0, theta - 2.0 * res - 1e-12 * scale)


def vector_diffusion_device(c, t, Lc: BSR, L: BSR, lc_min=0.0):
    """smoothing.py:37-64 (method="matrix_exp", normalise=True) on device.

    c: (n, d) local-frame coefficients, or (F, n, d) for F fields at once;
    t: a diffusion time, or (F,) times, one per field.  lc_min: a lower bound
    of Lc's smallest eigenvalue; the connection diffusion runs shifted by it
    (exp(-t lc_min) is common to every node and cancels in the
    normalisation), which keeps the smoothed directions accurate however
    small exp(-t Lc) c is."""
    batched = c.dim() == 3
    C = c if batched else c.unsqueeze(0)
    F, n, d = C.shape
    ts = np.atleast_1d(np.asarray(t, dtype=np.float64))
    per_field = ts.size > 1
    if per_field and ts.size != F:
        raise ValueError(f"{ts.size} diffusion times for {F} fields")
    # connection diffusion of every field: columns = fields
    Xc = C.reshape(F, n * d).T.contiguous()
    out = expm_multiply(Lc, Xc, ts if per_field else float(ts[0]),
                        shift=lc_min).T.reshape(F, n, d).contiguous()
    # scalar diffusions exp(-tL)|x| and exp(-tL)1 share one multi-column pass
    # (one column of ones per distinct time)
    n1 = F if per_field else 1
    xs = torch.empty((n, F + n1), dtype=torch.float64, device=c.device)
    xs[:, :F] = torch.linalg.vector_norm(C, dim=2).T
    xs[:, F:] = 1.0
    sd = expm_multiply(L, xs, np.concatenate([ts, ts]) if per_field else float(ts[0]))
    for f in range(F):
        one = F + (f if per_field else 0)
        _lib.call("rvgp_diffusion_normalise", n, d, _lib.ptr(out[f]), d, _lib.ptr(sd[:, f:]),
                  F + n1, _lib.ptr(sd[:, one:]), F + n1, _lib.stream())
    return out if batched else out[0]


def spectral_diffusion_device(X, t, evals, evecs):
    """evecs diag(exp(-t evals)) evecs^T X on the device (smoothing.py:17-32:
    the heat kernel restricted to the given eigenpairs, eigenvectors taken as
    they are passed).  X: (rows,) or (rows, r)."""
    from . import dense
    X2 = X.reshape(-1, 1) if X.dim() == 1 else X
    V = evecs.contiguous()
    C = dense.gemm(V, X2.contiguous(), transA=True)                 # (k, r)
    C = C * torch.exp(-float(t) * evals.to(torch.float64)).reshape(-1, 1)
    return dense.gemm(V, C.contiguous()).reshape(X.shape)


def vector_diffusion_spectral_device(c, t, evals_c, evecs_c, evals_l, evecs_l):
    """smoothing.py:37-64 with method="spectral" and normalise=True on the
    device: the connection heat kernel and the scalar ones (for |c| and 1)
    from eigenpairs.  With normalisation any common scale of the eigenvectors
    (the sqrt(rows) of geometry.py:75) cancels."""
    n, d = c.shape
    out = spectral_diffusion_device(c.reshape(-1, 1), t, evals_c, evecs_c).reshape(n, d)
    xs = torch.empty((n, 2), dtype=torch.float64, device=c.device)
    xs[:, 0] = torch.linalg.vector_norm(c, dim=1)
    xs[:, 1] = 1.0
    sd = spectral_diffusion_device(xs, t, evals_l, evecs_l)
    return out * sd[:, :1] / (sd[:, 1:] * torch.linalg.vector_norm(out, dim=1, keepdim=True))


def _host_eigenpairs(par):
    assert isinstance(par, (list, tuple)) and len(par) == 2, \
        "For spectral method, par must be a tuple of eigenvalues, eigenvectors!"
    dev = _lib.device()
    return (torch.as_tensor(np.asarray(par[0], dtype=np.float64), device=dev),
            torch.as_tensor(np.asarray(par[1], dtype=np.float64), device=dev))


def scalar_diffusion(x, t, method="matrix_exp", par=None):
    """smoothing.py:8-34; returns a numpy array.  "matrix_exp": exp(-t par) x
    for a host scipy operator par; "spectral": par = (eigenvalues,
    eigenvectors), evecs diag(exp(-t evals)) evecs^T x."""
    X = torch.as_tensor(np.asarray(x, dtype=np.float64), device=_lib.device())
    if method == "matrix_exp":
        A = BSR.from_scipy(par, 1)
        return expm_multiply(A, X, t).cpu().numpy()
    if method == "spectral":
        evals, evecs = _host_eigenpairs(par)
        return spectral_diffusion_device(X, t, evals, evecs).cpu().numpy()
    raise NotImplementedError


def vector_diffusion(x, t, Lc, L=None, method="spectral", normalise=True):
    """smoothing.py:37-64; returns numpy (n, d).  method "matrix_exp": Lc, L
    host scipy operators; "spectral": Lc = (evals_Lc, evecs_Lc) and
    L = (evals_L, evecs_L) eigenpairs, e.g. a data object's."""
    n, d = x.shape[0], x.shape[1]
    dev = _lib.device()
    xd = torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)
    if method == "spectral":
        assert len(Lc) == 2, "Lc must be a tuple of eigenvalues, eigenvectors!"
        nd = np.asarray(Lc[1]).shape[0]
        assert (n * d % nd) == 0, "Data dimension must be an integer multiple of the dimensions of the connection Laplacian!"
        ec, vc = _host_eigenpairs(Lc)
        if not normalise:
            return spectral_diffusion_device(xd.reshape(nd, -1), t, ec, vc).reshape(x.shape).cpu().numpy()
        assert L is not None, 'Need Laplacian for normalised diffusion!'
        el, vl = _host_eigenpairs(L)
        return vector_diffusion_spectral_device(xd.reshape(n, -1), t, ec, vc, el, vl) \
            .reshape(x.shape).cpu().numpy()
    if method != "matrix_exp":
        raise NotImplementedError
    nd = Lc.shape[0]
    assert (n * d % nd) == 0, "Data dimension must be an integer multiple of the dimensions of the connection Laplacian!"
    Lcb = BSR.from_scipy(Lc, nd // n if nd % n == 0 else 1)
    if not normalise:
        return expm_multiply(Lcb, xd.reshape(nd, -1), t).reshape(x.shape).cpu().numpy()
    assert L is not None, 'Need Laplacian for normalised diffusion!'
    return vector_diffusion_device(xd, t, Lcb, BSR.from_scipy(L, 1),
                                   lc_min=lambda_min_lower(Lcb)).cpu().numpy()
