"""FP64 CPU restatement of the RVGP connection-Laplacian GP path (the oracle).

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.

Each function cites the reference lines it restates.  How each one is pinned
(checked against something other than itself) is recorded here and exercised by
tests/test_oracle.py:

  fps                 geometry.py:126-162   pinned: same loop over sklearn
                      ``pairwise_distances`` (``metric="sklearn"``) vs the
                      direct-formula C loop; brute-force invariants.
  knn / sym_csr       geometry.py:100-112   pinned: sklearn ``kneighbors_graph``
                      + networkx ``adjacency_matrix`` on the same points.
  affinity_graph      geometry.py:114-118   pinned: brute-force Gaussian
                      weights on small inputs; symmetric, unit diagonal.
  geo_nbh             ptu_dijkstra.pyx:361-404  pinned: golden Sigma/tangents
                      written by the reference's own compiled module
                      (tests/golden/heap_*.npz).
  tangent_frames      ptu_dijkstra.pyx:33-125, 406-434  pinned: golden vectors.
  manifold_dimension  geometry.py:83-97     pinned: golden dims + closed form.
  connections         ptu_dijkstra.pyx:128-293  pinned: golden R (COO) and
                      ``scipy.linalg.polar``.
  laplacian           geometry.py:55-63     pinned: ``networkx.laplacian_matrix``.
  connection_laplacian geometry.py:14-52    pinned: restated scipy kron/multiply
                      on the golden R; ring-holonomy closed form.
  compute_spectrum    geometry.py:66-80     pinned: ``numpy.linalg.eigh`` and the
                      ring-graph closed form 2-2cos(2 pi j/n).
  heat                smoothing.py:8-15     pinned: ``scipy.linalg.expm`` vs the
                      eigen form; exp(-tL)1 = 1.
  vector_diffusion    smoothing.py:37-64    pinned: closed form on a trivial
                      connection (L (x) I, one-direction field: the result is
                      e^{-tL}a along v); for any connection |out_i| =
                      (e^{-tL}|x|)_i / (e^{-tL}1)_i with out_i parallel to
                      (e^{-tLc}x)_i from Lc's own eigendecomposition; t -> 0.
                      "spectral" (smoothing.py:17-32): over the full
                      eigenbasis it equals "expm"; the sqrt(rows) scale of
                      stored eigenvectors cancels in the normalisation.
  project_to_manifold geometry.py:165-168   pinned: projector identities
                      (P^2 = P, P T = T, residual orthogonal to T, |Px| <= |x|).
  express_in_local_frame geometry.py:171-175  pinned: T^T (T c) = c,
                      T (T^T x) = P x, c -> O^T c under the gauge T -> T O.
  Data (Phi assembly) dataclass.py:57-59    pinned: Phi^T Phi = U^T U, rows in
                      the tangent planes, B Lc B^T Phi = Phi diag(evals_Lc).
  smooth_vector_field dataclass.py:105-120  composition of the three above.
  spectral_filter     kernels.py:41-53      pinned: closed form + finite diffs.
  gpr_*               GPflow GPR (main.py:98-116)  pinned: textbook GPR
                      (Rasmussen & Williams Alg. 2.1) vs a dense
                      ``scipy.stats.multivariate_normal`` log-density, central
                      finite differences; GPflow itself is absent, so the
                      GPflow-specific optimiser trajectory is parity unpinned.
  train_test_split    main.py:40-45         pinned: sklearn ``train_test_split``.
  fit_lbfgs           main.py:87-95         pinned: the returned point is a
                      stationary minimum (gradient ~0 on the free parameters,
                      no +-1e-3 coordinate step lowers the NLML); the GPflow
                      trajectory itself is parity unpinned.
"""
import ctypes
import os
import subprocess

import numpy as np
import scipy.linalg
import scipy.special
import scipy.sparse as sp
import scipy.sparse.linalg

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compile oracle/oracle.c into oracle/liboracle.so (gcc)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B", "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.oracle_diam2.restype = ctypes.c_double
        L.oracle_diam2.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.oracle_fps.restype = ctypes.c_int
        L.oracle_fps.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, P, P]
        L.oracle_knn.restype = ctypes.c_int
        L.oracle_knn.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
        L.oracle_sym_csr.restype = ctypes.c_int
        L.oracle_sym_csr.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P]
        L.oracle_geo_nbh.restype = ctypes.c_int
        L.oracle_geo_nbh.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_int, P, P]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# a1  furthest point sampling  (geometry.py:126-162)
# --------------------------------------------------------------------------
def fps(x, N=None, spacing=0.1, start_idx=0, metric="direct"):
    """Greedy farthest-point order.  ``metric="direct"`` runs the C loop with
    the fixed direct distance formula; ``metric="sklearn"`` runs the same loop
    over ``sklearn.metrics.pairwise_distances`` as the reference does."""
    if spacing == 0.0:
        return np.arange(len(x)), None
    x = np.ascontiguousarray(x, dtype=np.float64)
    n, D = x.shape
    if metric == "sklearn":
        from sklearn.metrics import pairwise_distances
        Dm = pairwise_distances(x)
        nn = Dm.shape[0] if N is None else N
        diam = Dm.max()
        perm = np.zeros(nn, dtype=np.int32)
        perm[0] = start_idx
        lambdas = np.zeros(nn)
        ds = Dm[start_idx, :]
        for i in range(1, nn):
            idx = np.argmax(ds)
            perm[i] = idx
            lambdas[i] = ds[idx]
            ds = np.minimum(ds, Dm[idx, :])
            if N is None and lambdas[i] / diam < spacing:
                perm, lambdas = perm[:i], lambdas[:i]
                break
        return perm, lambdas
    nn = n if N is None else int(N)
    perm = np.zeros(nn, dtype=np.int32)
    lambdas = np.zeros(nn)
    cnt = lib().oracle_fps(_p(x), n, D, -1 if N is None else nn, float(spacing), int(start_idx),
                           _p(perm), _p(lambdas))
    if cnt < 0:
        raise MemoryError("oracle_fps")
    return perm[:cnt], lambdas[:cnt]


def diameter(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return float(np.sqrt(lib().oracle_diam2(_p(x), x.shape[0], x.shape[1])))


# --------------------------------------------------------------------------
# a2  kNN graph  (geometry.py:100-112)
# --------------------------------------------------------------------------
def knn(x, k):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n, D = x.shape
    nbr = np.empty((n, k), dtype=np.int32)
    rc = lib().oracle_knn(_p(x), n, D, int(k), _p(nbr))
    if rc != 0:
        raise ValueError("oracle_knn: bad k")
    return nbr


def sym_csr(nbr):
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    n, k = nbr.shape
    rowptr = np.empty(n + 1, dtype=np.int32)
    cols = np.empty(n * (2 * k + 1), dtype=np.int32)
    nnz = lib().oracle_sym_csr(_p(nbr), n, k, _p(rowptr), _p(cols))
    if nnz < 0:
        raise MemoryError("oracle_sym_csr")
    return rowptr, cols[:nnz].copy()


def affinity_graph(x, sigma=0.1):
    """manifold_graph(typ='affinity') (geometry.py:114-118): sklearn
    pairwise_distances, the Gaussian affinity, nx.from_numpy_array; returns
    the networkx graph and its weighted adjacency (scipy csr)."""
    import networkx as nx
    from sklearn.metrics import pairwise_distances
    Dm = pairwise_distances(np.asarray(x, dtype=np.float64))
    A = np.exp(-Dm ** 2 / (2 * sigma ** 2))
    G = nx.from_numpy_array(A)
    W = sp.csr_matrix(nx.adjacency_matrix(G, weight="weight"), dtype=np.float64)
    W.sort_indices()
    return G, W


def graph_csr_networkx(x, n_neighbors):
    """The symmetrised CSR exactly as the reference obtains it: sklearn
    kneighbors_graph + I (geometry.py:103-112) -> networkx -> adjacency matrix
    -> max(A, A^T) (ptu_dijkstra.pyx:84-103)."""
    import networkx as nx
    from sklearn.neighbors import kneighbors_graph
    A = kneighbors_graph(x, n_neighbors, mode="connectivity", metric="minkowski", p=2,
                         metric_params=None, include_self=False, n_jobs=None)
    A += sp.eye(A.shape[0])
    G = nx.from_scipy_sparse_array(A)
    C = sp.csr_matrix(nx.adjacency_matrix(G, weight="weight"), dtype=np.float64)
    S = C.maximum(C.T.tocsr()).tocsr()
    S.sort_indices()
    return S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data


# --------------------------------------------------------------------------
# a3  heap neighbourhoods + tangent frames  (ptu_dijkstra.pyx:33-125, 300-437)
# --------------------------------------------------------------------------
def geo_nbh(x, rowptr, cols, K, weights=None, centred=True):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n, D = x.shape
    w = np.ones(len(cols)) if weights is None else np.ascontiguousarray(weights, np.float64)
    nbh = np.empty((n, K + 1), dtype=np.int32)
    cm = np.empty((n, D, K + 1)) if centred else None
    rc = lib().oracle_geo_nbh(_p(x), n, D, _p(np.ascontiguousarray(rowptr, np.int32)),
                              _p(np.ascontiguousarray(cols, np.int32)), _p(w), int(K), _p(nbh),
                              _p(cm) if centred else None)
    if rc == -3:
        raise MemoryError("Failed to allocate memory in geo_nbh")
    if rc <= -1000:
        raise ValueError(f"node {-rc - 1001} reaches fewer than K+1={K + 1} nodes")
    return nbh, cm


def tangent_frames(x, rowptr, cols, d, K, allow_rank_deficient_k=False):
    """ptu_dijkstra.pyx:33-125.  Validation as pyx:68-82.  With
    ``allow_rank_deficient_k`` the K<d check is relaxed to the build's
    documented extension (min(D, K+1) singular values)."""
    N, D = x.shape
    K = int(K)
    if K >= N:
        raise ValueError("Geodesic neighborhood size must be less than the total number of samples")
    if K < d and not allow_rank_deficient_k:
        raise ValueError("Geodesic neighborhood size must be larger or equal to the embedding dimension")
    if D < d:
        raise ValueError("Embedding dimension must be less or equal to the ambient dimension of input data")
    _, cm = geo_nbh(x, rowptr, cols, K)
    U, S, _ = np.linalg.svd(cm, full_matrices=False)
    d_eff = min(d, S.shape[1])
    # pyx:426-428 checks the first d values.  In the extension, K+1 centred
    # points span at most K directions, so only the first min(d, K) are checked.
    n_check = min(d_eff, K) if allow_rank_deficient_k else d_eff
    if np.any(S[:, :n_check] < 1e-10):
        raise RuntimeError("Local tangent space approximation failed, at least one geodesic "
                           "neighborhood does not span d-dimensional space")
    return np.ascontiguousarray(U[:, :, :d_eff]), np.ascontiguousarray(S[:, :d_eff])


def manifold_dimension(Sigma, frac_explained=0.9):
    """geometry.py:83-97 (in place on a copy)."""
    if frac_explained == 1.0:
        return Sigma.shape[1]
    S = np.array(Sigma, dtype=np.float64) ** 2
    S /= S.sum(1, keepdims=True)
    S = S.cumsum(1)
    var_exp = S.mean(0) - S.std(0)
    return int(np.where(var_exp >= frac_explained)[0][0] + 1)


# --------------------------------------------------------------------------
# a5  connections  (ptu_dijkstra.pyx:128-293)
# --------------------------------------------------------------------------
def connections(tangents, rowptr, cols):
    """R_e = U V^T of SVD(T_i^T T_j) for every CSR entry (i, j), self loops
    included, in row-major sorted order (pyx:259-293).  Returns (nnz, d, d)."""
    n = tangents.shape[0]
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    M = np.einsum("eki,ekj->eij", tangents[rows], tangents[cols])
    U, _, Vt = np.linalg.svd(M)
    return U @ Vt


def connections_coo(R, rowptr, cols):
    """The COO matrix that ptu_dijkstra.connections returns (pyx:194-214)."""
    n = len(rowptr) - 1
    d = R.shape[-1]
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    p, q = np.meshgrid(np.arange(d), np.arange(d), indexing="ij")
    r = (rows[:, None, None] * d + p[None]).ravel()
    c = (cols[:, None, None] * d + q[None]).ravel()
    return sp.coo_matrix((R.ravel(), (r, c)), shape=(n * d, n * d))


# --------------------------------------------------------------------------
# a6 / a7  Laplacians  (geometry.py:14-63)
# --------------------------------------------------------------------------
def laplacian(rowptr, cols):
    """L = D - A of the simple symmetrised graph (self loops cancel)."""
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    A = sp.csr_matrix((np.ones(len(cols)), cols, rowptr), shape=(n, n))
    deg = np.asarray(A.sum(1)).ravel()
    return (sp.diags(deg) - A).tocsr()


def connection_laplacian(rowptr, cols, R):
    """kron(L, ones(d,d)) .multiply(R) (geometry.py:37-42) on the block
    pattern; returns csr (nd x nd)."""
    d = R.shape[-1]
    L = laplacian(rowptr, cols)
    Lk = sp.kron(L, np.ones((d, d)))
    return sp.csr_matrix(Lk.multiply(connections_coo(R, rowptr, cols)))


def lc_blocks(rowptr, cols, R):
    """Block values of Lc in CSR order: L_ij R_ij (diag: deg_i R_ii)."""
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    deg = np.diff(rowptr) - 1
    lv = np.where(rows == cols, deg[rows], -1.0)
    return R * lv[:, None, None]


# --------------------------------------------------------------------------
# a8  spectrum  (geometry.py:66-80)
# --------------------------------------------------------------------------
def compute_spectrum(A, n_eigenpairs=None, dense=False):
    """eigsh(k, which='SM'), ascending, eigenvectors scaled by sqrt(rows).
    ``dense=True`` uses numpy.linalg.eigh (the library cross-check)."""
    N = A.shape[0]
    k = N if n_eigenpairs is None else min(int(n_eigenpairs), N)
    if dense:
        w, V = np.linalg.eigh(A.toarray() if sp.issparse(A) else A)
        w, V = w[:k], V[:, :k]
    else:
        if k >= N - 1:
            raise TypeError("Cannot use scipy.linalg.eigh for sparse A with k >= N - 1.")
        w, V = scipy.sparse.linalg.eigsh(A, k=k, which="SM")
        o = np.argsort(w)
        w, V = w[o], V[:, o]
    return w, V * np.sqrt(len(V))


# --------------------------------------------------------------------------
# a9  data object  (dataclass.py:22-89)
# --------------------------------------------------------------------------
class Data:
    def __init__(self, vertices, vectors=None, n_neighbors=10, frac_geodesic_neighbours=1.5,
                 explained_variance=0.8, n_eigenpairs=None, allow_rank_deficient_k=False,
                 graph="direct", dense_eigh=False, spectra=True):
        X = np.ascontiguousarray(vertices, dtype=np.float64)
        if graph == "networkx":
            rowptr, cols, _ = graph_csr_networkx(X, n_neighbors)
        else:
            rowptr, cols = sym_csr(knn(X, n_neighbors))
        D = X.shape[1]
        K = int(n_neighbors * frac_geodesic_neighbours)
        gauges, Sigma = tangent_frames(X, rowptr, cols, D, K, allow_rank_deficient_k)
        dim_man = manifold_dimension(Sigma, explained_variance)
        gauges = np.ascontiguousarray(gauges[:, :, :dim_man])
        R = connections(gauges, rowptr, cols)
        self.vertices, self.n, self.dim_man = X, X.shape[0], dim_man
        self.rowptr, self.cols, self.gauges, self.Sigma, self.Rblocks = rowptr, cols, gauges, Sigma, R
        self.L = laplacian(rowptr, cols)
        self.Lc = connection_laplacian(rowptr, cols, R)
        self.vectors = vectors
        if spectra:
            self.evals_L, self.evecs_L = compute_spectrum(self.L, n_eigenpairs, dense_eigh)
            evals_Lc, U = compute_spectrum(self.Lc, n_eigenpairs, dense_eigh)
            k = U.shape[1]
            self.evals_Lc = evals_Lc
            self.U = U
            self.evecs_Lc = np.einsum("bij,bjk->bik", gauges, U.reshape(-1, dim_man, k)).reshape(-1, k)

    # a10  dataclass.py:91-103
    def random_vector_field(self, seed=0):
        np.random.seed(seed)
        v = np.random.uniform(low=-0.5, high=0.5, size=(len(self.vertices), self.vertices.shape[1]))
        v = project_to_manifold(v, self.gauges[..., :self.dim_man])
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        self.vectors = v

    # a11  dataclass.py:105-120
    def smooth_vector_field(self, t=100, method="expm"):
        c = express_in_local_frame(self.vectors, self.gauges)
        c = vector_diffusion(c, t, self.Lc, self.L, method)
        self.vectors = express_in_local_frame(c, self.gauges, reverse=True)


def project_to_manifold(x, gauges):
    """geometry.py:165-168."""
    coeffs = np.einsum("bij,bi->bj", gauges, x)
    return np.einsum("bj,bij->bi", coeffs, gauges)


def express_in_local_frame(x, gauges, reverse=False):
    """geometry.py:171-175."""
    if reverse:
        return np.einsum("bji,bi->bj", gauges, x)
    return np.einsum("bij,bi->bj", gauges, x)


def heat(A, t, x, method="expm"):
    """exp(-tA) x.  ``expm``: scipy.linalg.expm of the dense matrix, as
    smoothing.py:13-15; ``eigh``: V exp(-t w) V^T x, the same operator."""
    Ad = A.toarray() if sp.issparse(A) else np.asarray(A)
    if method == "expm":
        return scipy.linalg.expm(-t * Ad) @ x
    w, V = np.linalg.eigh(Ad)
    return V @ (np.exp(-t * w)[:, None] * (V.T @ x))


def heat_spectral(evals, evecs, t, x):
    """smoothing.py:17-32 (the "spectral" method, as the reference means it):
    evecs diag(exp(-t evals)) evecs^T x."""
    return evecs @ (np.exp(-t * np.asarray(evals))[:, None] * (evecs.T @ x))


def vector_diffusion(x, t, Lc, L, method="expm"):
    """smoothing.py:37-64 with normalise=True.  method "expm" / "eigh": Lc, L
    operators; "spectral": Lc, L = (eigenvalues, eigenvectors) pairs."""
    n, d = x.shape
    if method == "spectral":
        rows = np.asarray(Lc[1]).shape[0]
        out = heat_spectral(Lc[0], Lc[1], t, x.reshape(rows, -1)).reshape(x.shape)
        x_abs = np.linalg.norm(x, axis=-1, keepdims=True)
        out_abs = heat_spectral(L[0], L[1], t, x_abs)
        ind = heat_spectral(L[0], L[1], t, np.ones([n, 1]))
        return out * out_abs / (ind * np.linalg.norm(out, axis=-1, keepdims=True))
    out = heat(Lc, t, x.reshape(n * d, -1), method).reshape(x.shape)
    x_abs = np.linalg.norm(x, axis=-1, keepdims=True)
    out_abs = heat(L, t, x_abs, method)
    ind = heat(L, t, np.ones([n, 1]), method)
    return out * out_abs / (ind * np.linalg.norm(out, axis=-1, keepdims=True))


# --------------------------------------------------------------------------
# a12  spectral filter  (kernels.py:41-53)
# --------------------------------------------------------------------------
def spectral_filter(evals, nu, kappa, sigma_f, n_rows, typ="matern"):
    """S and dS/d(nu, kappa, sigma_f) including the normalisation
    S = sigma_f * n_rows * r / sum(r)."""
    lam = np.asarray(evals, dtype=np.float64)
    if typ == "matern":
        base = lam + 2.0 * nu / kappa ** 2
        r = base ** (-nu)
        dlog_nu = -np.log(base) - nu * (2.0 / kappa ** 2) / base
        dlog_kappa = 4.0 * nu ** 2 / (kappa ** 3 * base)
    else:
        r = np.exp(-0.5 * lam * kappa ** 2)
        dlog_nu = np.zeros_like(lam)
        dlog_kappa = -lam * kappa
    sr = r.sum()
    S = sigma_f * n_rows * r / sr
    w = r / sr
    dS_nu = S * (dlog_nu - (w * dlog_nu).sum())
    dS_kappa = S * (dlog_kappa - (w * dlog_kappa).sum())
    dS_sigma = S / sigma_f
    return S, np.stack([dS_nu, dS_kappa, dS_sigma])


# --------------------------------------------------------------------------
# a14 / a16  GP regression, textbook (Rasmussen & Williams Alg. 2.1)
# --------------------------------------------------------------------------
def gpr_nlml(Phi, y, S, noise):
    """-log N(y | 0, Phi diag(S) Phi^T + noise I), function space.  A 2-D y
    (M, R) holds R right-hand sides sharing the hyperparameters; their
    NLMLs are summed (GPflow's GPR.log_marginal_likelihood sums the
    log-density over the columns of Y)."""
    K = (Phi * S) @ Phi.T
    M = K.shape[0]
    R = y.shape[1] if np.ndim(y) == 2 else 1
    L = np.linalg.cholesky(K + noise * np.eye(M))
    a = scipy.linalg.solve_triangular(L, y, lower=True)
    nlml = (0.5 * float((a * a).sum()) + R * float(np.log(np.diag(L)).sum())
            + 0.5 * R * M * np.log(2 * np.pi))
    return nlml


def gpr_grad(Phi, y, S, noise):
    """(dNLML/dS (k,), dNLML/dnoise) from 1/2 tr((R Ky^-1 - A A^T) dKy),
    A = Ky^-1 Y over the R right-hand sides."""
    K = (Phi * S) @ Phi.T
    M = K.shape[0]
    Y = y.reshape(M, -1)
    R = Y.shape[1]
    Ky = K + noise * np.eye(M)
    Kinv = np.linalg.inv(Ky)
    A = Kinv @ Y
    PtKP = ((Kinv @ Phi) * Phi).sum(0)   # diag(Phi^T Ky^-1 Phi)
    PtA = Phi.T @ A
    gS = 0.5 * (R * PtKP - (PtA ** 2).sum(1))
    gn = 0.5 * (R * np.trace(Kinv) - (A * A).sum())
    return gS, gn


def gpr_predict(Phi, y, S, noise, Phi_star):
    """GPflow GPR.predict_f(full_cov=False): mean and marginal variance."""
    K = (Phi * S) @ Phi.T
    M = K.shape[0]
    L = np.linalg.cholesky(K + noise * np.eye(M))
    Kmn = (Phi * S) @ Phi_star.T
    A = scipy.linalg.solve_triangular(L, Kmn, lower=True)
    var = np.einsum("ij,j,ij->i", Phi_star, S, Phi_star) - (A * A).sum(0)
    a = scipy.linalg.solve_triangular(L, y, lower=True)
    mean = A.T @ a
    return mean, var[:, None]


# --------------------------------------------------------------------------
# n_inducing_points branch (main.py:54-67): GPflow's SGPR, the collapsed
# (Titsias) bound, with the spectral kernel Phi diag(S) Phi^T
# --------------------------------------------------------------------------
SGPR_JITTER = 1e-6  # gpflow.config.default_jitter()


def sgpr_nlml(Phi, y, S, noise, Z, jitter=SGPR_JITTER):
    """-ELBO of gpflow.models.SGPR, written the way SGPR.elbo forms it:
    A = L^-1 K_uf / sigma, B = I + A A^T, c = L_B^-1 A y / sigma."""
    N = Phi.shape[0]
    y = y.reshape(N, -1)
    R = y.shape[1]
    m = Z.shape[0]
    Kdiag = np.einsum("ij,j,ij->i", Phi, S, Phi)
    kuf = (Z * S) @ Phi.T
    kuu = (Z * S) @ Z.T + jitter * np.eye(m)
    L = np.linalg.cholesky(kuu)
    sigma = np.sqrt(noise)
    A = scipy.linalg.solve_triangular(L, kuf, lower=True) / sigma
    AAT = A @ A.T
    LB = np.linalg.cholesky(AAT + np.eye(m))
    c = scipy.linalg.solve_triangular(LB, A @ y, lower=True) / sigma
    bound = -0.5 * N * R * np.log(2 * np.pi)
    bound -= R * np.log(np.diag(LB)).sum()
    bound -= 0.5 * N * R * np.log(noise)
    bound -= 0.5 * (y * y).sum() / noise
    bound += 0.5 * (c * c).sum()
    bound -= 0.5 * R * Kdiag.sum() / noise
    bound += 0.5 * R * np.trace(AAT)
    return -bound


def sgpr_predict(Phi, y, S, noise, Z, Phi_star, jitter=SGPR_JITTER):
    """gpflow.models.SGPR.predict_f(full_cov=False): mean (m*, R), variance (m*,)."""
    N = Phi.shape[0]
    y = y.reshape(N, -1)
    m = Z.shape[0]
    kuf = (Z * S) @ Phi.T
    kuu = (Z * S) @ Z.T + jitter * np.eye(m)
    kus = (Z * S) @ Phi_star.T
    L = np.linalg.cholesky(kuu)
    sigma = np.sqrt(noise)
    A = scipy.linalg.solve_triangular(L, kuf, lower=True) / sigma
    LB = np.linalg.cholesky(A @ A.T + np.eye(m))
    c = scipy.linalg.solve_triangular(LB, A @ y, lower=True) / sigma
    tmp1 = scipy.linalg.solve_triangular(L, kus, lower=True)
    tmp2 = scipy.linalg.solve_triangular(LB, tmp1, lower=True)
    mean = tmp2.T @ c
    var = (np.einsum("ij,j,ij->i", Phi_star, S, Phi_star) + (tmp2 * tmp2).sum(0)
           - (tmp1 * tmp1).sum(0))
    return mean, var


# --------------------------------------------------------------------------
# kernel='rbf' branch (main.py:33-37): gpflow.kernels.RBF on evecs_L rows,
# the D vector components as D output columns of one GPR
# --------------------------------------------------------------------------
def rbf_kernel(X, X2, variance, lengthscale):
    """variance * exp(-|x - x'|^2 / (2 lengthscale^2)) (squared distances by
    scipy.spatial.distance.cdist)."""
    from scipy.spatial.distance import cdist
    return variance * np.exp(-0.5 * cdist(X, X2, "sqeuclidean") / lengthscale ** 2)


def rbf_gpr_nlml(X, Y, variance, lengthscale, noise):
    """Summed NLML of the R = Y.shape[1] output columns."""
    M, R = Y.shape
    L = np.linalg.cholesky(rbf_kernel(X, X, variance, lengthscale) + noise * np.eye(M))
    a = scipy.linalg.solve_triangular(L, Y, lower=True)
    return (0.5 * float((a * a).sum()) + R * float(np.log(np.diag(L)).sum())
            + 0.5 * R * M * np.log(2 * np.pi))


def rbf_gpr_grad(X, Y, variance, lengthscale, noise):
    """d NLML / d(variance, lengthscale, noise) = 1/2 tr((R K^-1 - A A^T) dK),
    A = K^-1 Y."""
    from scipy.spatial.distance import cdist
    M, R = Y.shape
    D2 = cdist(X, X, "sqeuclidean")
    Kf = variance * np.exp(-0.5 * D2 / lengthscale ** 2)
    Kinv = np.linalg.inv(Kf + noise * np.eye(M))
    A = Kinv @ Y
    W = R * Kinv - A @ A.T
    return np.array([0.5 * (W * Kf).sum() / variance,
                     0.5 * (W * Kf * D2).sum() / lengthscale ** 3,
                     0.5 * np.trace(W)])


def rbf_gpr_predict(X, Y, variance, lengthscale, noise, Xs):
    """Mean (m*, R) and marginal variance (m*,)."""
    M = X.shape[0]
    L = np.linalg.cholesky(rbf_kernel(X, X, variance, lengthscale) + noise * np.eye(M))
    Ks = rbf_kernel(X, Xs, variance, lengthscale)
    V = scipy.linalg.solve_triangular(L, Ks, lower=True)
    a = scipy.linalg.solve_triangular(L, Y, lower=True)
    return V.T @ a, variance - (V * V).sum(0)


# --------------------------------------------------------------------------
# a13 / a15  fit driver  (main.py:11-95)
# --------------------------------------------------------------------------
def train_test_split_indices(n, test_size=0.2, seed=0):
    """sklearn ShuffleSplit as used by train_test_split(random_state=seed):
    test = perm[:ceil(test_size n)], train = perm[n_test:]."""
    n_test = int(np.ceil(test_size * n))
    perm = np.random.RandomState(seed).permutation(n)
    return perm[n_test:], perm[:n_test]


def softplus(u):
    return np.logaddexp(0.0, u)


def inv_softplus(x):
    return x + np.log(-np.expm1(-x))


def fit_lbfgs(Phi, y, evals, n_rows, theta0=(1.5, 5.0, 1.0, 1.0), lower=(0.0, 0.0, 0.0, 1e-6),
              maxiter=1000):
    """L-BFGS-B on softplus-unconstrained (nu, kappa, sigma_f, noise)."""
    import scipy.optimize
    lower = np.asarray(lower, dtype=np.float64)
    u0 = inv_softplus(np.asarray(theta0, dtype=np.float64) - lower)

    def f(u):
        th = softplus(u) + lower
        S, dS = spectral_filter(evals, th[0], th[1], th[2], n_rows)
        val = gpr_nlml(Phi, y, S, th[3])
        gS, gn = gpr_grad(Phi, y, S, th[3])
        g = np.array([dS[0] @ gS, dS[1] @ gS, dS[2] @ gS, gn])
        return val, g * scipy.special.expit(u)   # d softplus / du, overflow-free

    res = scipy.optimize.minimize(f, u0, jac=True, method="L-BFGS-B", options={"maxiter": maxiter})
    return softplus(res.x) + lower, res
