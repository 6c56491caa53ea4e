"""Device-resident symmetrised kNN graph.

The reference builds a networkx graph (RVGP/geometry.py:100-123) and every
consumer converts it back to a symmetrised CSR matrix
(RVGP/lib/ptu_dijkstra.pyx:84-103).  Here the graph *is* that CSR, kept in
HBM: ``rowptr`` (n+1) and ``cols`` (nnz) int32, columns sorted within each
row, one self loop per node, edge weights 1.0 (kNN) or per-edge values
(the affinity graph).  A networkx view is built on the host only when someone
asks for it (``to_networkx``).
"""
import numpy as np
import torch

from . import _lib


class SymGraph:
    def __init__(self, rowptr, cols, n, positions=None):
        self.rowptr = rowptr
        self.cols = cols
        self.n = int(n)
        self.nnz = int(cols.numel())
        self._positions = positions
        self._weights = None
        self._nx = None

    def __len__(self):
        return self.n

    @property
    def weights(self):
        if self._weights is None:
            self._weights = torch.ones(self.nnz, dtype=torch.float64, device=self.cols.device)
        return self._weights

    def max_row(self):
        return int(torch.diff(self.rowptr).max().item())

    def csr_numpy(self):
        return self.rowptr.cpu().numpy(), self.cols.cpu().numpy()

    def to_scipy(self):
        import scipy.sparse as sp
        rp, c = self.csr_numpy()
        w = np.ones(len(c)) if self._weights is None else self._weights.cpu().numpy()
        return sp.csr_matrix((w, c, rp), shape=(self.n, self.n))

    def to_networkx(self):
        """The networkx graph the reference returns from manifold_graph:
        undirected, edge weights (1.0 for the kNN graph, the affinities for
        typ='affinity'), self loops, node attribute 'pos'."""
        if self._nx is None:
            import networkx as nx
            G = nx.from_scipy_sparse_array(self.to_scipy())
            if self._positions is not None:
                pos = self._positions
                nx.set_node_attributes(G, {i: pos[i] for i in G.nodes}, "pos")
            self._nx = G
        return self._nx


def from_networkx(G, device=None):
    """Symmetrised CSR of a networkx graph exactly as ptu_dijkstra.pyx:84-103
    derives it (adjacency matrix, max(A, A^T)), uploaded to the device."""
    import networkx as nx
    import scipy.sparse as sp
    A = sp.csr_matrix(nx.adjacency_matrix(G, weight="weight"), dtype=np.float64)
    S = A.maximum(A.T.tocsr()).tocsr()
    S.sort_indices()
    dev = device or _lib.device()
    rowptr = torch.as_tensor(S.indptr.astype(np.int32), device=dev)
    cols = torch.as_tensor(S.indices.astype(np.int32), device=dev)
    g = SymGraph(rowptr, cols, S.shape[0])
    g._weights = torch.as_tensor(S.data.astype(np.float64), device=dev)
    return g


def as_device_graph(G):
    if isinstance(G, SymGraph):
        return G
    return from_networkx(G)
