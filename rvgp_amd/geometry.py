"""Drop-in for RVGP/geometry.py on MI355X.

Every function keeps the reference's name, arguments, defaults and return
types; the work runs in the HIP kernels of librvgp.so.
"""
import numpy as np
import torch

from . import _lib
from .graph import SymGraph, as_device_graph


def _dev_f64(x):
    dev = _lib.device()
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=dev)


# ---------------------------------------------------------------------------
# a1  geometry.py:126-162
# ---------------------------------------------------------------------------
def furthest_point_sampling_device(x, N=None, spacing=0.1, start_idx=0):
    """Device tensors (perm int32, lambdas f64) of furthest_point_sampling."""
    xd = _dev_f64(x)
    n, D = xd.shape
    n_iter = n if N is None else int(N)
    if n_iter < 1:
        raise IndexError("index 0 is out of bounds for axis 0 with size 0")
    if not (0 <= int(start_idx) < n):
        raise IndexError(f"index {start_idx} is out of bounds for axis 0 with size {n}")
    dev = xd.device
    s = _lib.stream()
    diam2 = None
    if N is None:
        diam2 = torch.zeros(1, dtype=torch.float64, device=dev)
        _lib.call("rvgp_fps_diam2", _lib.ptr(xd), n, D, _lib.ptr(diam2), s)
    ds = torch.empty(n, dtype=torch.float64, device=dev)
    perm = torch.empty(n_iter, dtype=torch.int32, device=dev)
    lambdas = torch.empty(n_iter, dtype=torch.float64, device=dev)
    nsel = torch.empty(1, dtype=torch.int32, device=dev)
    _lib.call("rvgp_fps", _lib.ptr(xd), n, D, n_iter, int(start_idx), float(spacing),
              1 if N is None else 0, _lib.ptr(diam2), _lib.ptr(ds), _lib.ptr(perm),
              _lib.ptr(lambdas), _lib.ptr(nsel), s)
    k = int(nsel.item())
    return perm[:k], lambdas[:k]


def furthest_point_sampling(x, N=None, spacing=0.1, start_idx=0):
    """A greedy farthest-point order (geometry.py:126-162).

    Returns (perm int32, lambdas f64) as numpy arrays, or
    (arange(len(x)), None) when spacing == 0.0 (geometry.py:139-140).
    """
    if spacing == 0.0:
        return np.arange(len(x)), None
    perm, lambdas = furthest_point_sampling_device(x, N, spacing, start_idx)
    return perm.cpu().numpy(), lambdas.cpu().numpy()


# ---------------------------------------------------------------------------
# a2  geometry.py:100-123
# ---------------------------------------------------------------------------
def knn_device(xd, k, query_order=None):
    """nbr (n, k) int32 on the device.  query_order (a permutation of the
    points, e.g. a Morton order) only sets which queries share a wave: the
    result does not depend on it."""
    n, D = xd.shape
    if not (0 < k < n):
        raise ValueError(f"Expected n_neighbors < n_samples_fit, but n_neighbors = {k + 1}, "
                         f"n_samples_fit = {n}, n_samples = {n}")
    # sklearn's check_array rejects NaN / inf; beyond 1e150 the squared
    # distances themselves would overflow and no neighbour could be ranked
    amax = float(xd.abs().max().item())
    if not np.isfinite(amax):
        raise ValueError("Input X contains NaN or infinity.")
    if amax >= 1e150:
        raise ValueError("Input X has coordinates too large for squared distances in float64.")
    nbr = torch.empty((n, k), dtype=torch.int32, device=xd.device)
    ws_bytes = _lib.lib().rvgp_knn_ws_bytes(n, int(k))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=xd.device)
    qp = None
    if query_order is not None:
        qp = torch.as_tensor(query_order, device=xd.device).to(torch.int32).contiguous()
        if qp.shape != (n,) or not torch.equal(
                torch.sort(qp).values, torch.arange(n, dtype=torch.int32, device=xd.device)):
            raise ValueError("query_order must be a permutation of the points")
    _lib.call("rvgp_knn", _lib.ptr(xd), n, D, int(k), _lib.ptr(qp), _lib.ptr(nbr), _lib.ptr(ws),
              ws_bytes, _lib.stream())
    return nbr


def symmetrize_device(nbr):
    n, k = nbr.shape
    dev = nbr.device
    ws_bytes = _lib.lib().rvgp_graph_symmetrize_ws_bytes(n, k)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    rowptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
    cols = torch.empty(n * (2 * k + 1), dtype=torch.int32, device=dev)
    _lib.call("rvgp_graph_symmetrize", _lib.ptr(nbr), n, k, _lib.ptr(rowptr), _lib.ptr(cols),
              _lib.ptr(ws), ws_bytes, _lib.stream())
    nnz = int(rowptr[n].item())
    return rowptr, cols[:nnz].clone()


AFFINITY_SIGMA = 0.1  # geometry.py:116


def affinity_graph(X, sigma=AFFINITY_SIGMA):
    """manifold_graph(typ='affinity') (geometry.py:114-118): the complete
    graph weighted by the Gaussian affinity exp(-|x_i - x_j|^2 / (2 sigma^2))
    of all pairs, self loops (weight 1) included, entries that underflow to 0
    left out (nx.from_numpy_array's rule).  The dense affinity is formed on the
    device (rvgp_affinity); the result is a weighted :class:`SymGraph` whose
    networkx view carries the weights and the 'pos' attribute."""
    xd = _dev_f64(X)
    n, D = xd.shape
    A = torch.empty((n, n), dtype=torch.float64, device=xd.device)
    _lib.call("rvgp_affinity", _lib.ptr(xd), n, D, float(sigma), _lib.ptr(A), n, _lib.stream())
    nz = A != 0
    rowptr = torch.zeros(n + 1, dtype=torch.int64, device=xd.device)
    rowptr[1:] = torch.cumsum(nz.sum(1), 0)
    cols = torch.nonzero(nz)[:, 1].to(torch.int32).contiguous()
    g = SymGraph(rowptr.to(torch.int32), cols, n, positions=X if isinstance(X, np.ndarray) else None)
    g._weights = A[nz].contiguous()
    return g


KNN_ORDER_MIN_N = 4096  # below this the kNN scan is too short for a query order to pay


def manifold_graph(X, typ='knn', n_neighbors=5, query_order=None):
    """Fit graph over a pointset X (geometry.py:100-123).

    Returns a :class:`SymGraph` -- the symmetrised CSR of kNN + I kept on the
    device; ``.to_networkx()`` gives the networkx graph the reference returns.
    query_order (default: a Morton order of the leading principal
    coordinates from n = 4096 on) groups manifold neighbours into the kNN's
    waves; the graph does not depend on it.
    """
    if typ == 'affinity':
        return affinity_graph(X)
    if typ != 'knn':
        raise UnboundLocalError("local variable 'G' referenced before assignment")
    xd = _dev_f64(X)
    if query_order is None and xd.shape[0] >= KNN_ORDER_MIN_N:
        from .reorder import morton_perm
        query_order = morton_perm(xd)
    rowptr, cols = symmetrize_device(knn_device(xd, int(n_neighbors), query_order))
    pos = X if isinstance(X, np.ndarray) else None
    return SymGraph(rowptr, cols, xd.shape[0], positions=pos)


# ---------------------------------------------------------------------------
# a6 / a7  geometry.py:14-63  (host scipy views of the device operators)
# ---------------------------------------------------------------------------
def _networkx_degrees(G):
    """Node degrees as networkx's G.degree() reports them: unweighted, a self
    loop counted twice."""
    if not isinstance(G, SymGraph):
        return np.array(list(dict(G.degree()).values()), dtype=np.float64)
    rp, c = G.csr_numpy()
    rows = np.repeat(np.arange(G.n), np.diff(rp))
    loops = np.bincount(rows[c == rows], minlength=G.n)
    return (np.diff(rp) + loops).astype(np.float64)


def compute_laplacian(G, normalization=False):
    """L = D - A of the symmetrised graph, as scipy csr (geometry.py:55-63);
    normalization=True gives networkx's normalized_laplacian_matrix of the
    same graph (the reference's host call, geometry.py:58-59)."""
    import scipy.sparse as sp
    if normalization:
        import networkx as nx
        Gx = G.to_networkx() if isinstance(G, SymGraph) else G
        return sp.csr_matrix(nx.normalized_laplacian_matrix(Gx), dtype=np.float64)
    g = as_device_graph(G)
    A = g.to_scipy()
    deg = np.asarray(A.sum(1)).ravel()
    return sp.csr_matrix(sp.diags(deg) - A)


def compute_connection_laplacian(G, R, normalization=None):
    """kron(L, ones(d,d)) .multiply(R) as scipy csr (geometry.py:14-52)."""
    import scipy.sparse as sp
    n = len(G)
    dim = R.shape[0] // n
    L = compute_laplacian(G)
    Lc = sp.csr_matrix(sp.kron(L, np.ones([dim, dim])).multiply(R))
    if normalization == "rw":
        # D^-1 Lc with networkx's degrees (geometry.py:45-50): the kNN + I
        # graph's self loop counts twice, so deg_i = neighbours + 2
        deg = _networkx_degrees(G)
        deg_inv = np.where(deg > 0, 1.0 / np.where(deg > 0, deg, 1.0), 0.0).repeat(dim)
        Lc = sp.diags(deg_inv, 0, format='csr') @ Lc
    return Lc


# ---------------------------------------------------------------------------
# a4  geometry.py:83-97
# ---------------------------------------------------------------------------
def manifold_dimension(Sigma, frac_explained=0.9):
    """Estimate manifold dimension based on singular values (geometry.py:83-97).
    Like the reference, squares and normalises ``Sigma`` in place."""
    if isinstance(Sigma, torch.Tensor):
        Sigma = Sigma.cpu().numpy()
    if frac_explained == 1.0:
        return Sigma.shape[1]
    Sigma **= 2
    Sigma /= Sigma.sum(1, keepdims=True)
    Sigma = Sigma.cumsum(1)
    var_exp = Sigma.mean(0) - Sigma.std(0)
    dim_man = np.where(var_exp >= frac_explained)[0][0] + 1
    print("Fraction of variance explained: ", var_exp)
    return int(dim_man)


# ---------------------------------------------------------------------------
# a10 / a11 helpers  geometry.py:165-175 (host numpy, as in the reference)
# ---------------------------------------------------------------------------
def project_to_manifold(x, gauges):
    """Project vectors to local coordinates over manifold (geometry.py:165-168)."""
    coeffs = np.einsum("bij,bi->bj", gauges, x)
    return np.einsum("bj,bij->bi", coeffs, gauges)


def express_in_local_frame(x, gauges, reverse=False):
    """Express vectors in local coordinates over manifold (geometry.py:171-175)."""
    if reverse:
        return np.einsum("bji,bi->bj", gauges, x)
    return np.einsum("bij,bi->bj", gauges, x)


def compute_spectrum(laplacian, n_eigenpairs=None, dtype=None):
    """Smallest eigenpairs, eigenvectors scaled by sqrt(rows) (geometry.py:66-80)."""
    from .eigsolve import spectrum_of_scipy
    return spectrum_of_scipy(laplacian, n_eigenpairs)
