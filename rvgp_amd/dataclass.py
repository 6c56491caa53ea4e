"""Drop-in for RVGP/dataclass.py (``RVGP.create_data_object``).

The constructor runs the whole geometry path on the GPU -- kNN graph,
heap-exact neighbourhoods and gauges, manifold dimension, connections,
block-CSR Laplacians and both eigensolves (dataclass.py:22-89) -- and keeps
every array in HBM.  The reference's attributes (vertices, n, dim_man, G,
gauges, R, L, Lc, evals_L, evecs_L, evals_Lc, evecs_Lc, vectors) are
available with the reference's types (numpy / scipy / networkx); they are
materialised on the host only when read, and assigning to them (as the
reference's examples do, e.g. truncating evecs_Lc) is honoured by fit().
"""
import math
import threading

import numpy as np
import torch

from . import _lib, geometry
from .eigsolve import chfsi
from .ptu_dijkstra import connections_device, tangent_frames_device
from .reorder import bsr_permutation, morton_perm, permute_bsr
from .smoothing import lambda_min_lower, vector_diffusion_device, vector_diffusion_spectral_device
from .sparse import BSR, graph_laplacian


class data:
    def __init__(self, vertices, vectors=None, dim_man=2, n_neighbors=10,
                 frac_geodesic_neighbours=1.5, explained_variance=0.8, n_eigenpairs=None, *,
                 precision="fp64", reference_semantics=True, eig_tol=None, verbose=True,
                 timings=None, eig_verbose=False, reorder=True):
        """precision: "fp64" (all FP64) or "mixed" (FP32 eigensolver body with
        FP64 Rayleigh-Ritz and polish).  eig_tol: eigen-residual tolerance
        relative to the spectral bound; default 1e-9 for "fp64" and 1e-7 for
        "mixed" (predictions then agree with the FP64 form to ~1e-7, far
        inside the 1e-4 the mixed form is held to).  reference_semantics=False
        enables the documented extension for D > K (min(D, K+1) singular
        values)."""
        if eig_tol is None:
            eig_tol = 1e-7 if precision == "mixed" else 1e-9
        t = timings if timings is not None else {}
        main = torch.cuda.current_stream()
        sync = main.synchronize  # this stream only: the scalar spectrum runs beside it
        import time
        t0 = time.perf_counter()
        if verbose:
            print('Fit graph')
        Xd = geometry._dev_f64(vertices)
        N, D = Xd.shape
        # Morton renumbering of the leading principal coordinates: groups
        # manifold neighbours into the kNN's waves and, below, gives the
        # block SpMM its locality
        perm = morton_perm(Xd)
        G = geometry.manifold_graph(Xd, n_neighbors=n_neighbors, query_order=perm)
        if isinstance(vertices, np.ndarray):
            G._positions = vertices
        sync()
        t["graph"] = time.perf_counter() - t0

        # The scalar Laplacian depends on the graph alone: its spectrum starts
        # now, on a second (default-priority) stream from a host thread, and
        # runs beside the tangent spaces, connections and the connection
        # spectrum on this stream -- each fills the GPU while the other waits
        # on latency-bound steps (heap walks, small eigenproblems, host round
        # trips).
        kL = N if n_eigenpairs is None else min(int(n_eigenpairs), N)
        perm = perm if reorder else None
        t0 = time.perf_counter()
        Lb = graph_laplacian(G)
        # the renumbered pattern, shared by both Laplacians of the graph
        plan = bsr_permutation(G.rowptr, G.cols, perm) if reorder else None
        Lb_e = permute_bsr(Lb, perm, plan) if reorder else Lb
        sync()
        t["laplacian_L"] = time.perf_counter() - t0
        resL = {}
        side = _side_stream(main.device)
        side.wait_stream(main)

        def solve_scalar():
            try:
                t1 = time.perf_counter()
                with torch.cuda.stream(side):
                    # unit-norm columns times sqrt(N) (geometry.py:75), rows
                    # mapped back to the original node numbering
                    w, V, st = chfsi(Lb_e, kL, precision=precision, tol=eig_tol,
                                     verbose=eig_verbose, perm=perm, scale=math.sqrt(N))
                side.synchronize()
                resL.update(w=w, V=V, st=st, t=time.perf_counter() - t1)
            except BaseException as e:  # re-raised on the calling thread
                resL["error"] = e

        worker = threading.Thread(target=solve_scalar, name="rvgp-spectrum-L")
        worker.start()
        try:
            if verbose:
                print('Fit tangent spaces')
            t0 = time.perf_counter()
            K = n_neighbors * frac_geodesic_neighbours
            if K >= N:
                raise ValueError("Geodesic neighborhood size must be less than the total number of samples")
            if K < D and reference_semantics:
                raise ValueError("Geodesic neighborhood size must be larger or equal to the embedding dimension")
            K_t = int(K)
            mn = min(D, K_t + 1)
            n_check = mn if reference_semantics else min(D, K_t)
            tangents, Sigma = tangent_frames_device(Xd, G, K_t, mn, mn, n_check)
            sync()
            t["gauges"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            dm = geometry.manifold_dimension(Sigma.cpu().numpy(), frac_explained=explained_variance) \
                if verbose else _quiet_dimension(Sigma, explained_variance)
            gauges = tangents[:, :, :dm].contiguous()
            t["dimension"] = time.perf_counter() - t0
            if verbose:
                print('Predicted manifold dimension is {}'.format(dm))

            t0 = time.perf_counter()
            conn = connections_device(gauges, G, dm, scalar=False)
            sync()
            t["connections"] = time.perf_counter() - t0
            if verbose:
                print('Fit connections')
                print('Compute Laplacians')
            Lcb = BSR(G.rowptr, G.cols, conn["lc64"].reshape(-1), conn["lc32"].reshape(-1), dm)

            if verbose:
                print('Compute eigendecompositions')
            t0 = time.perf_counter()
            kLc = N * dm if n_eigenpairs is None else min(int(n_eigenpairs), N * dm)
            Lcb_e = permute_bsr(Lcb, perm, plan) if reorder else Lcb
            sync()
            t["permute_Lc"] = time.perf_counter() - t0
            t0 = time.perf_counter()
            # intrinsic (n*d, k) eigenvectors, unit-norm columns times
            # sqrt(n*d) (geometry.py:75), in the original node numbering
            wLc, VLc, stLc = chfsi(Lcb_e, kLc, precision=precision, tol=eig_tol,
                                   verbose=eig_verbose, perm=perm, scale=math.sqrt(N * dm))
            sync()
            t["eig_Lc"] = time.perf_counter() - t0
        finally:
            worker.join()
        if "error" in resL:
            raise resL["error"]
        main.wait_stream(side)
        t0 = time.perf_counter()
        wL, VL, stL = resL["w"], resL["V"], resL["st"]
        t["eig_L"] = resL["t"]
        t["eig_L_wait"] = time.perf_counter() - t0
        self.eig_stats = {"L": stL, "Lc": stLc}

        self._X = Xd
        self._vertices_in = vertices
        self.n = N
        self.dim_man = dm
        self._graph = G
        self._gauges = gauges
        self._tangents_full = tangents
        self._Sigma = Sigma
        self._R = conn["R"]
        self._Lb, self._Lcb = Lb, Lcb
        self._evals_L = wL
        self._evecs_L = VL
        self._evals_Lc = wLc
        self._U = VLc                              # intrinsic (n*d, k)
        self.n_eigenpairs = kLc
        self._host = {}
        self._vectors = None
        self.vectors = vectors

    # ------------------------------------------------------------------ attributes
    @property
    def vertices(self):
        return self._vertices_in

    @property
    def G(self):
        return self._graph.to_networkx()

    @property
    def gauges(self):
        if "gauges" not in self._host:
            self._host["gauges"] = self._gauges.cpu().numpy()
        return self._host["gauges"]

    @property
    def R(self):
        if "R" not in self._host:
            from .ptu_dijkstra import blocks_to_coo
            rp, c = self._graph.csr_numpy()
            self._host["R"] = blocks_to_coo(self._R.cpu().numpy(), rp, c)
        return self._host["R"]

    @property
    def L(self):
        if "L" not in self._host:
            self._host["L"] = self._Lb.to_scipy()
        return self._host["L"]

    @property
    def Lc(self):
        if "Lc" not in self._host:
            self._host["Lc"] = self._Lcb.to_scipy()
        return self._host["Lc"]

    def _getset(name, dev_attr):
        def get(self):
            if name not in self._host:
                self._host[name] = getattr(self, dev_attr).cpu().numpy()
            return self._host[name]

        def set_(self, v):
            self._host[name] = np.asarray(v)
            self._overrides = getattr(self, "_overrides", set()) | {name}
        return property(get, set_)

    evals_L = _getset("evals_L", "_evals_L")
    evecs_L = _getset("evecs_L", "_evecs_L")
    evals_Lc = _getset("evals_Lc", "_evals_Lc")

    @property
    def evecs_Lc(self):
        """Ambient eigenvectors Phi = gauges U, shape (n*D, k) (dataclass.py:57-59)."""
        if "evecs_Lc" not in self._host:
            nodes = torch.arange(self.n, dtype=torch.int32, device=self._X.device)
            self._host["evecs_Lc"] = self.phi_rows(nodes).cpu().numpy()
        return self._host["evecs_Lc"]

    @evecs_Lc.setter
    def evecs_Lc(self, v):
        self._host["evecs_Lc"] = np.asarray(v)
        self._overrides = getattr(self, "_overrides", set()) | {"evecs_Lc"}

    del _getset

    def _overridden(self, name):
        return name in getattr(self, "_overrides", set())

    def spectrum_device(self):
        """(evals (k,), device) honouring assignments to evals_Lc."""
        if self._overridden("evals_Lc"):
            return torch.as_tensor(np.asarray(self._host["evals_Lc"], np.float64), device=self._X.device)
        return self._evals_Lc

    def intrinsic_rows(self, nodes):
        """Rows of the intrinsic eigenvectors U for the given nodes,
        ((len(nodes)*d, k) f64), node-major -- or None once evecs_Lc was
        assigned on the host.  Phi = blockdiag(T_i) U and T_i^T T_i = I, so
        Phi^T Phi = U^T U and Phi^T y = U^T (T^T y): the weight-space
        statistics need these d-row blocks instead of the D-row ones."""
        if self._overridden("evecs_Lc"):
            return None
        nodes = torch.as_tensor(nodes, device=self._X.device).long()
        if nodes.numel() and (int(nodes.min()) < 0 or int(nodes.max()) >= self.n):
            raise IndexError(f"node index out of bounds for axis 0 with size {self.n}")
        d = self.dim_man
        rows = (nodes.reshape(-1, 1) * d + torch.arange(d, device=nodes.device)).reshape(-1)
        return self._U.index_select(0, rows).contiguous()

    def phi_rows(self, nodes, dtype=torch.float64):
        """Rows of evecs_Lc for the given nodes, (len(nodes)*D, k) on device."""
        D = self._X.shape[1]
        if self._overridden("evecs_Lc"):
            E = np.asarray(self._host["evecs_Lc"]).reshape(self.n, -1)
            idx = nodes.cpu().numpy() if isinstance(nodes, torch.Tensor) else np.asarray(nodes)
            rows = E[idx].reshape(-1, E.shape[1] // D)
            return torch.as_tensor(np.ascontiguousarray(rows), device=self._X.device).to(dtype)
        nodes = torch.as_tensor(nodes, dtype=torch.int32, device=self._X.device).contiguous()
        m = nodes.numel()
        k = self._U.shape[1]
        out = torch.empty((m * D, k), dtype=dtype, device=self._X.device)
        T = self._gauges
        _lib.call("rvgp_phi_rows", int(dtype == torch.float32), m, self.n, D, self.dim_man, _lib.ptr(T),
                  T.shape[2], k, _lib.ptr(self._U), self._U.stride(0), _lib.ptr(nodes),
                  _lib.ptr(out), k, _lib.stream())
        return out

    @property
    def vectors(self):
        if self._vectors is None:
            return None
        if "vectors" not in self._host:
            self._host["vectors"] = self._vectors.cpu().numpy()
        return self._host["vectors"]

    @vectors.setter
    def vectors(self, v):
        self._host.pop("vectors", None)
        if v is None:
            self._vectors = None
        elif isinstance(v, torch.Tensor):
            self._vectors = v.to(self._X.device, torch.float64).contiguous()
        else:
            self._vectors = torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=self._X.device)

    def vectors_device(self):
        return self._vectors

    # ------------------------------------------------------------------ fields
    def _frame(self, mode, v, out_cols):
        F = v.shape[0] if v.dim() == 3 else 1
        out = torch.empty((F, self.n, out_cols), dtype=torch.float64, device=v.device)
        T = self._gauges
        _lib.call("rvgp_frame_apply", mode, F, self.n, T.shape[1], self.dim_man, _lib.ptr(T),
                  T.shape[2], _lib.ptr(v.contiguous()), _lib.ptr(out), _lib.stream())
        return out if v.dim() == 3 else out[0]

    def random_vector_field(self, seed=0):
        """Generate random vector field over manifold (dataclass.py:91-103).
        Uses numpy's legacy RandomState stream exactly as the reference."""
        np.random.seed(seed)
        v = np.random.uniform(low=-0.5, high=0.5, size=(self.n, self._X.shape[1]))
        vd = torch.as_tensor(v, device=self._X.device)
        vd = self._frame(2, vd, self._X.shape[1])
        vd = vd / torch.linalg.vector_norm(vd, dim=1, keepdim=True)
        self.vectors = vd

    def smooth_vector_field(self, t=100, method="matrix_exp"):
        """Vector diffusion smoothing (dataclass.py:105-120).  vectors may be
        (n, D) or a batch (F, n, D) of fields smoothed in one pass.
        method="matrix_exp" (the reference's) applies exp(-t Lc) and exp(-t L)
        to the operators; method="spectral" (smoothing.py:17-32, an option
        the reference's data object does not expose) uses the data object's
        own eigenpairs -- the heat kernel restricted to the computed spectrum."""
        if self._vectors is None:
            raise TypeError("No vectors to smooth: data.vectors is None")
        c = self._frame(0, self._vectors, self.dim_man)
        if method == "spectral":
            if c.dim() != 2 or np.ndim(t) != 0:
                raise ValueError("method='spectral' smooths one field with one diffusion time")
            c = vector_diffusion_spectral_device(c, float(t), self.spectrum_device(), self._U,
                                                 self._evals_L, self._evecs_L)
        elif method == "matrix_exp":
            st = self.eig_stats["Lc"]
            lc_min = lambda_min_lower(self._Lcb, float(self._evals_Lc[0].item()), st["residual"])
            c = vector_diffusion_device(c, t, self._Lcb, self._Lb, lc_min=lc_min)
        else:
            raise NotImplementedError
        self.vectors = self._frame(1, c, self._X.shape[1])


_SIDE_STREAMS = {}
_SIDE_LOCK = threading.Lock()


def _side_stream(device):
    """The second (default-priority) stream of a device, created once and
    reused by every data object: the caching allocator keeps its blocks per
    stream, so a fresh stream per data object found no cached blocks and sent
    the scalar spectrum's first allocations to hipMalloc, which can wait for
    the other stream's queued work.  Data objects built concurrently from
    several host threads on one device share it (their scalar spectra then
    run one after the other)."""
    key = torch.device(device).index
    with _SIDE_LOCK:
        if key not in _SIDE_STREAMS:
            _SIDE_STREAMS[key] = torch.cuda.Stream(device=device)
        return _SIDE_STREAMS[key]


def _quiet_dimension(Sigma, frac):
    """manifold_dimension (geometry.py:83-97) without its print, reduced on
    the device: only the explained-variance profile (one value per singular
    value) comes to the host for the threshold."""
    if frac == 1.0:
        return Sigma.shape[1]
    S = Sigma * Sigma
    S = torch.cumsum(S / S.sum(1, keepdim=True), 1)
    var_exp = (S.mean(0) - S.std(0, unbiased=False)).cpu().numpy()
    return int(np.where(var_exp >= frac)[0][0] + 1)
