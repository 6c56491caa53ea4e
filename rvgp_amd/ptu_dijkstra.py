"""Drop-in for the reference's native module ``ptu_dijkstra``
(RVGP/lib/ptu_dijkstra.pyx), backed by the HIP kernels of librvgp.so.

``tangent_frames`` and ``connections`` keep the reference's signatures,
validation order, exception types and return types (numpy / scipy).  The
``*_device`` variants are what the data object uses: they take and return
device tensors and never touch the host.
"""
import numpy as np
import torch

from . import _lib
from .graph import as_device_graph


def _check_status(st):
    v = int(st.item())
    if v == _lib.ST_RANK_DEFICIENT:
        raise RuntimeError("Local tangent space approximation failed, at least one geodesic "
                           "neighborhood does not span d-dimensional space")
    if v == _lib.ST_SHORT_COMPONENT:
        raise ValueError("a connected component has fewer nodes than the geodesic "
                         "neighborhood size K+1")
    if v == _lib.ST_POOL_OVERFLOW:
        raise MemoryError("Failed to allocate memory in _geodesic_neigborhood_tangents")
    if v != 0:
        raise RuntimeError(f"device status {v}")


def geo_nbh_device(g, K, max_ws_bytes=16 << 30):
    """(n, K+1) int32 neighbourhoods in heap pop order (pyx:361-394).

    One lane per source and every source in flight at once where the
    workspace cap allows (~27 KB per lane in the first pass, whose node pool
    holds what a search on a kNN graph touches; the few searches that
    outgrow it are redone with the worst-case pool 1 + K * max_degree): the
    heap walks are latency-bound, so the number of lanes in flight sets the
    speed."""
    n = g.n
    pool = 1 + K * g.max_row()
    L = _lib.lib()
    per_lane = max(1, L.rvgp_geo_nbh_ws_bytes(n, 2, pool) - L.rvgp_geo_nbh_ws_bytes(n, 1, pool))
    batch = int(max(1, min(n, max_ws_bytes // per_lane, 1 << 17)))
    ws_bytes = L.rvgp_geo_nbh_ws_bytes(n, batch, pool)
    dev = g.cols.device
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    nbh = torch.empty((n, K + 1), dtype=torch.int32, device=dev)
    st = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.call("rvgp_geo_nbh", _lib.ptr(g.rowptr), _lib.ptr(g.cols), _lib.ptr(g.weights), n, K,
              batch, pool, _lib.ptr(nbh), _lib.ptr(ws), ws_bytes, _lib.ptr(st), _lib.stream())
    _check_status(st)
    return nbh


def tangent_frames_device(Xd, g, K, ncols, nsv, n_check):
    """Centred local PCA of every heap neighbourhood (pyx:397-434).
    Returns (tangents (n, D, ncols) f64, sigma (n, nsv) f64) on the device."""
    n, D = Xd.shape
    nbh = geo_nbh_device(g, K)
    tangents = torch.empty((n, D, ncols), dtype=torch.float64, device=Xd.device)
    sigma = torch.empty((n, nsv), dtype=torch.float64, device=Xd.device)
    st = torch.zeros(1, dtype=torch.int32, device=Xd.device)
    _lib.call("rvgp_gauges", _lib.ptr(Xd), n, D, _lib.ptr(nbh), K + 1, ncols, nsv, n_check,
              _lib.ptr(tangents), _lib.ptr(sigma), _lib.ptr(st), _lib.stream())
    _check_status(st)
    return tangents, sigma


def tangent_frames(X, G, d, K):
    """Tangent frames of geodesic neighbourhoods (pyx:33-125).

    Returns (tangents (N, D, d), Sigma (N, d)) as float64 numpy arrays.
    """
    N, D = X.shape[0], X.shape[1]
    if K >= N:
        raise ValueError("Geodesic neighborhood size must be less than the total number of samples")
    if K < d:
        raise ValueError("Geodesic neighborhood size must be larger or equal to the embedding dimension")
    if D < d:
        raise ValueError("Embedding dimension must be less or equal to the ambient dimension of input data")
    K_t = int(K)
    g = as_device_graph(G)
    if np.any(g.weights.cpu().numpy() < 0):
        import warnings
        warnings.warn("Graph has negative weights: negative distances are not allowed.")
    from .geometry import _dev_f64
    Xd = _dev_f64(X)
    mn = min(D, K_t + 1)
    T, S = tangent_frames_device(Xd, g, K_t, min(d, mn), min(d, mn), min(d, mn))
    return T.cpu().numpy(), S.cpu().numpy()


def connections_device(tangents, g, d, want_R=True, lc64=True, lc32=True, scalar=True):
    """R blocks (nnz, d, d) and the Laplacian values in block-CSR order
    (scalar=False skips the scalar Laplacian's values)."""
    n, D, ldt = tangents.shape
    if D < d:
        raise ValueError("Embedding dimension must be less or equal to the ambient dimension of input data")
    dev = tangents.device
    nnz = g.nnz
    out = {
        "R": torch.empty((nnz, d, d), dtype=torch.float64, device=dev) if want_R else None,
        "lc64": torch.empty((nnz, d, d), dtype=torch.float64, device=dev) if lc64 else None,
        "lc32": torch.empty((nnz, d, d), dtype=torch.float32, device=dev) if lc32 else None,
        "l64": torch.empty(nnz, dtype=torch.float64, device=dev) if (lc64 and scalar) else None,
        "l32": torch.empty(nnz, dtype=torch.float32, device=dev) if (lc32 and scalar) else None,
    }
    _lib.call("rvgp_connections", _lib.ptr(tangents.contiguous()), n, D, ldt, d,
              _lib.ptr(g.rowptr), _lib.ptr(g.cols), nnz, _lib.ptr(out["R"]), _lib.ptr(out["lc64"]),
              _lib.ptr(out["lc32"]), _lib.ptr(out["l64"]), _lib.ptr(out["l32"]), _lib.stream())
    return out


def blocks_to_coo(R, rowptr, cols):
    """The (n*d, n*d) COO matrix of pyx:194-214 from CSR-ordered blocks."""
    import scipy.sparse as sp
    n = len(rowptr) - 1
    d = R.shape[-1]
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    p, q = np.meshgrid(np.arange(d), np.arange(d), indexing="ij")
    r = (rows[:, None, None] * d + p[None]).ravel()
    c = (cols[:, None, None] * d + q[None]).ravel()
    return sp.coo_matrix((R.ravel(), (r, c)))


def connections(tangents, G, d):
    """Parallel-transport maps between neighbouring tangent frames
    (pyx:128-199).  Returns the scipy COO matrix of the reference."""
    if tangents.shape[1] < d:
        raise ValueError("Embedding dimension must be less or equal to the ambient dimension of input data")
    g = as_device_graph(G)
    from .geometry import _dev_f64
    T = _dev_f64(tangents)
    out = connections_device(T, g, d, lc64=False, lc32=False)
    rp, c = g.csr_numpy()
    return blocks_to_coo(out["R"].cpu().numpy(), rp, c)
