"""Write golden vectors for the reference's native geometry kernel.

The neighbourhoods of ``tangent_frames`` (RVGP/lib/ptu_dijkstra.pyx:361-394)
are chosen by the pop order of a Fibonacci heap among equal hop distances, a
choice that no closed form or library routine fixes.  This script compiles the
reference's own ``ptu_dijkstra.pyx`` with Cython into a temporary directory in
this container, runs ``tangent_frames`` and ``connections`` on small seeded
point clouds, and stores inputs and outputs under ``tests/golden/`` as
``heap_<case>.npz``.  Nothing from the reference is kept: the compiled module
lives in a temporary directory that is deleted on exit.

The graph is built exactly as RVGP/geometry.py:100-112 builds it
(sklearn ``kneighbors_graph`` + identity, networkx undirected graph).

Run from the repository root:  python tests/golden/make_heap_fixtures.py
"""
import os
import subprocess
import sys
import sysconfig
import tempfile

import numpy as np

REF_PYX = "/root/reference/RVGP/lib/ptu_dijkstra.pyx"
HERE = os.path.dirname(os.path.abspath(__file__))


def build_module(tmp):
    c_path = os.path.join(tmp, "ptu_dijkstra.c")
    subprocess.check_call(["cython", "-3", REF_PYX, "-o", c_path])
    so_path = os.path.join(tmp, "ptu_dijkstra" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = ["gcc", "-O2", "-fPIC", "-shared", c_path, "-o", so_path,
           "-I" + sysconfig.get_paths()["include"], "-I" + np.get_include(),
           "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION", "-w"]
    subprocess.check_call(cmd)
    sys.path.insert(0, tmp)
    import ptu_dijkstra  # noqa: E402  (built just above)

    # The reference was written against scipy 1.10 / networkx 3.1, whose
    # adjacency matrices carry int32 indices for graphs this small; the
    # installed scipy 1.15 keeps int64, which the typed memoryviews of
    # ptu_dijkstra.pyx:221-229 / 300-311 reject.  Hand it the int32 form the
    # pinned versions produced; values and pattern are unchanged.
    import networkx as nx

    class _NxInt32:
        def __getattr__(self, name):
            return getattr(nx, name)

        @staticmethod
        def adjacency_matrix(G, weight="weight"):
            A = nx.adjacency_matrix(G, weight=weight).tocsr()
            A.indices = A.indices.astype(np.int32)
            A.indptr = A.indptr.astype(np.int32)
            return A

    ptu_dijkstra.nx = _NxInt32()
    return ptu_dijkstra


def sphere(n, seed):
    x = np.random.default_rng(seed).normal(size=(n, 3))
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def torus(n, seed, R=2.0, r=0.7):
    rng = np.random.default_rng(seed)
    th, ph = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    return np.stack([(R + r * np.cos(ph)) * np.cos(th), (R + r * np.cos(ph)) * np.sin(th),
                     r * np.sin(ph)], 1)


def half_ellipsoid(n, seed):
    x = sphere(4 * n, seed) * np.array([1.0, 0.9, 0.8])
    return x[x[:, 2] > 0][:n]


def flat_torus3(n, seed, D=8):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 2 * np.pi, (n, 3))
    E = np.concatenate([np.cos(a), np.sin(a)], 1)
    Q, Rq = np.linalg.qr(rng.normal(size=(D, D)))
    Q = Q * np.sign(np.diag(Rq))
    return E @ Q[:6] + 0.01 * rng.normal(size=(n, D))


CASES = {
    # name: (points, n_neighbors, dim passed to tangent_frames, dim_man for connections)
    "sphere2000_k10": (lambda: sphere(2000, 0), 10),
    "torus600_k15": (lambda: torus(600, 1), 15),
    "ellipsoid400_k12": (lambda: half_ellipsoid(400, 2), 12),
    "flat3torus500_D8_k10": (lambda: flat_torus3(500, 3), 10),
}


def main():
    import networkx as nx
    from scipy import sparse
    from sklearn.neighbors import kneighbors_graph

    with tempfile.TemporaryDirectory() as tmp:
        ptu = build_module(tmp)
        for name, (gen, k) in CASES.items():
            X = gen()
            A = kneighbors_graph(X, k, mode="connectivity", metric="minkowski", p=2,
                                 metric_params=None, include_self=False, n_jobs=None)
            A += sparse.eye(A.shape[0])
            G = nx.from_scipy_sparse_array(A)
            K = k * 1.5
            D = X.shape[1]
            tangents, Sigma = ptu.tangent_frames(X, G, D, K)
            # dimension as estimated by geometry.manifold_dimension(Sigma, 0.8)
            S = Sigma.copy() ** 2
            S /= S.sum(1, keepdims=True)
            S = S.cumsum(1)
            var_exp = S.mean(0) - S.std(0)
            dim = int(np.where(var_exp >= 0.8)[0][0] + 1)
            gauges = np.ascontiguousarray(tangents[:, :, :dim])
            R = ptu.connections(gauges, G, dim).tocoo()
            knn = A.tocsr()
            np.savez_compressed(
                os.path.join(HERE, f"heap_{name}.npz"),
                X=X, k=k, K=int(K), knn_indptr=knn.indptr, knn_indices=knn.indices,
                Sigma=Sigma, tangents=tangents, dim=dim,
                R_row=R.row.astype(np.int32), R_col=R.col.astype(np.int32), R_data=R.data)
            print(name, X.shape, "dim", dim, "nnz(R)", R.nnz)


if __name__ == "__main__":
    main()
