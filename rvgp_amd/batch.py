"""Many independent vector-field fits on one shared manifold.

The reference fits one field per call (RVGP.fit, main.py:11-84, GPflow's
GPR under the scipy optimiser).  Here F fields -- each with its own training
subset, of any size, and its own hyperparameters -- are fitted together:

  * set-up: the weight-space statistics of every fit in two launches
    (rvgp_batch_gram: G_f = Phi_f^T Phi_f over the fit's gathered intrinsic
    eigenvector rows, upper tiles on the matrix cores; rvgp_batch_rhs:
    b_f = Phi_f^T y_f and y_f^T y_f from the targets in the local frames);
  * each optimiser step: ONE launch of rvgp_batch_step -- a persistent
    kernel, one workgroup per fit, that forms A_f, factors it, evaluates the
    NLML and its adjoint and applies Adam -- captured in a HIP graph and
    replayed for every step.

The NLML of each fit is the one of gp.GPREngine's weight-space form (exact for
any subset size, identical to GPflow's function-space value).  Launch shapes
(Gram tile, Cholesky panel width) are chosen by the autotuner.
"""
import numpy as np
import torch

from . import _lib, autotune, dense
from .gp import inv_softplus, softplus
from .kernels import ManifoldKernel

GAUSSIAN_VARIANCE_LOWER_BOUND = 1e-6


class BatchedGPR:
    """F weight-space GP fits sharing the spectrum of one data object.

    fields: (F, n, D) ambient tangent fields (device or host); train_sets: F
    integer arrays of training nodes; theta0: initial (nu, kappa, sigma_f,
    noise) -- one 4-tuple for all fits or an (F, 4) array; gram_precision
    "fp64" (FP64 eigenvector rows) or "mixed" (FP32 rows widened to FP64 on
    the matrix cores)."""

    def __init__(self, data, fields, train_sets, theta0=(1.5, 5.0, 1.0, 1e-3), typ="matern",
                 lower=None, positivity_constraint=1e-2, gram_precision="fp64", gram_tile=None,
                 panel=None, grid=None):
        dev = _lib.device()
        self.dev = dev
        self.data = data
        F = len(train_sets)
        if F < 1:
            raise ValueError("no fits")
        self.F = F
        V = fields if isinstance(fields, torch.Tensor) else torch.as_tensor(
            np.ascontiguousarray(fields, np.float64))
        V = V.to(dev, torch.float64).contiguous()
        if V.dim() != 3 or V.shape[0] != F or V.shape[1] != data.n:
            raise ValueError(f"fields must be (F={F}, n={data.n}, D), got {tuple(V.shape)}")
        self.D = V.shape[2]
        self.d = data.dim_man
        U = data._U
        self.k = U.shape[1]
        sets = [np.asarray(s).astype(np.int64).ravel() for s in train_sets]
        for s in sets:
            if s.size == 0:
                raise ValueError("every fit needs at least one training node")
            if s.min() < 0 or s.max() >= data.n:
                raise IndexError("training node out of range")
        sizes = np.array([s.size for s in sets])
        self.sizes = sizes
        nodes = torch.as_tensor(np.concatenate(sets).astype(np.int32), device=dev)
        off = torch.as_tensor(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32), device=dev)
        fit_of = torch.repeat_interleave(torch.arange(F, device=dev), torch.as_tensor(sizes, device=dev))
        nl = nodes.long()
        C = data._frame(0, V, self.d)                          # (F, n, d) = T_i^T v_i
        c_rows = C[fit_of, nl].contiguous()                    # (sum m_f, d)
        ynorm2 = (V[fit_of, nl] ** 2).sum(1).contiguous()      # (sum m_f,)
        self.nodes, self.off = nodes, off

        f64 = dict(dtype=torch.float64, device=dev)
        k = self.k
        self.G = torch.empty((F, k, k), **f64)
        self.b = torch.empty((F, k), **f64)
        self.yy = torch.empty(F, **f64)
        self.M = torch.as_tensor(sizes * float(self.D), **f64)
        Ug = U if gram_precision == "fp64" else U.float().contiguous()
        self._gram_args = (int(Ug.dtype == torch.float64), F, k, self.d, _lib.ptr(Ug), Ug.stride(0),
                           _lib.ptr(off), _lib.ptr(nodes))
        self._Ug = Ug

        def run_gram(tile):
            _lib.call("rvgp_batch_gram", *self._gram_args, int(tile), _lib.ptr(self.G), _lib.stream())

        if gram_tile is None:
            gram_tile = autotune.choose("batch_gram", f"{gram_precision},F{F},k{k},d{self.d},"
                                        f"rows{int(sizes.sum()) * self.d}", (32, 64, 128), run_gram)
        self.gram_tile = int(gram_tile)
        run_gram(self.gram_tile)
        _lib.call("rvgp_batch_rhs", F, k, self.d, _lib.ptr(U), U.stride(0), _lib.ptr(off),
                  _lib.ptr(nodes), _lib.ptr(c_rows), _lib.ptr(ynorm2), _lib.ptr(self.b),
                  _lib.ptr(self.yy), _lib.stream())

        # hyperparameters and optimiser state, one row per fit
        self.typ = 0 if typ == "matern" else 1
        self.n_rows = float(data.n * self.D)  # rows of evecs_Lc (kernels.py:28)
        self.evals = data.spectrum_device().to(torch.float64).contiguous()
        if lower is None:
            lower = (positivity_constraint,) * 3 + (GAUSSIAN_VARIANCE_LOWER_BOUND,)
        self.lower = torch.as_tensor(np.asarray(lower, np.float64), device=dev)
        th0 = np.broadcast_to(np.asarray(theta0, np.float64), (F, 4))
        self.u = torch.as_tensor(inv_softplus(th0 - np.asarray(lower, np.float64)), **f64).contiguous()
        self.m1 = torch.zeros((F, 4), **f64)
        self.m2 = torch.zeros((F, 4), **f64)
        self.t = torch.zeros(F, **f64)
        self.trainable = torch.ones(4, dtype=torch.int32, device=dev)
        self.out = torch.zeros((F, 9), **f64)
        self.status = torch.zeros(F, dtype=torch.int32, device=dev)
        ws_bytes = _lib.lib().rvgp_batch_step_ws_bytes(F, k)
        self.ws = torch.empty(ws_bytes // 8, **f64)
        self.ws_bytes = ws_bytes
        if grid is None:
            grid = torch.cuda.get_device_properties(dev).multi_processor_count
        self.grid = int(grid)
        if panel is None:
            panel = autotune.choose("batch_step", f"F{F},k{k},grid{self.grid}",
                                    ((16, 512), (16, 1024), (32, 512), (32, 1024)),
                                    lambda c: self._step(0, 0.0, panel=c))
        # panel: the Cholesky block width, or (width, threads per workgroup)
        self.panel, self.threads = (int(panel), 512) if np.isscalar(panel) else map(int, panel)

    # ------------------------------------------------------------------ steps
    def _step(self, do_adam, lr, b1=0.9, b2=0.999, eps=1e-8, panel=None):
        nb, nt = (self.panel, self.threads) if panel is None else (
            (int(panel), 512) if np.isscalar(panel) else panel)
        _lib.call("rvgp_batch_step", self.F, self.k, self.typ, _lib.ptr(self.evals), self.n_rows,
                  _lib.ptr(self.G), _lib.ptr(self.b), _lib.ptr(self.yy), _lib.ptr(self.M),
                  _lib.ptr(self.lower), _lib.ptr(self.u), _lib.ptr(self.trainable),
                  _lib.ptr(self.m1), _lib.ptr(self.m2), _lib.ptr(self.t), int(do_adam), float(lr),
                  float(b1), float(b2), float(eps), int(nb), int(nt), self.grid,
                  _lib.ptr(self.ws), self.ws_bytes, _lib.ptr(self.out), _lib.ptr(self.status),
                  _lib.stream())

    def evaluate(self):
        """NLML and adjoint of every fit at the current hyperparameters;
        out (F, 9) = [nlml, d/d(nu, kappa, sigma_f, s2), d/du(...)]."""
        self._step(0, 0.0)
        return self.out

    def adam_step(self, lr=0.05, b1=0.9, b2=0.999, eps=1e-8):
        """One device-resident Adam step of every fit (one launch)."""
        self._step(1, lr, b1, b2, eps)

    def capture_adam_step(self, lr=0.05, b1=0.9, b2=0.999, eps=1e-8):
        """The step as a HIP graph (torch.cuda.CUDAGraph); it allocates
        nothing, so no warm-up is needed before the capture."""
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._step(1, lr, b1, b2, eps)
        torch.cuda.synchronize()
        return g

    def fit(self, epochs=200, lr=0.05, use_graph=True):
        if use_graph and epochs > 1:
            g = self.capture_adam_step(lr)
            for _ in range(int(epochs)):
                g.replay()
        else:
            for _ in range(int(epochs)):
                self.adam_step(lr)
        torch.cuda.synchronize()
        return self

    def check_status(self):
        bad = np.nonzero(self.status.cpu().numpy())[0]
        if bad.size:
            raise dense.CholeskyError(f"Cholesky decomposition was not successful for fits "
                                      f"{bad.tolist()}: the GP system matrix is not positive definite")

    # ------------------------------------------------------------------ host API
    def theta_host(self):
        """(F, 4) hyperparameters (nu, kappa, sigma_f, noise_variance)."""
        return softplus(self.u.cpu().numpy()) + self.lower.cpu().numpy()

    def set_theta(self, theta):
        th = np.broadcast_to(np.asarray(theta, np.float64), (self.F, 4))
        self.u.copy_(torch.as_tensor(inv_softplus(th - self.lower.cpu().numpy()), device=self.dev))

    def training_loss(self):
        out = self.evaluate().cpu().numpy()
        self.check_status()
        return out[:, 0].copy()

    def _state(self, f):
        k = self.k
        base = f * (2 * k * k + 9 * k)
        W = self.ws[base: base + 2 * k * k].view(k, 2 * k)
        vec = self.ws[base + 2 * k * k: base + 2 * k * k + 9 * k]
        return W[:, k:], vec[7 * k: 8 * k]  # X = L^-1 (ld 2k), w = A^-1 b

    def predict(self, f, Phi_star):
        """Mean (m*,) and marginal variance (m*,) of fit f at ambient rows
        Phi_star (m*, k): Phi* w and s2 |L^-1 phi*|^2 (gp.GPREngine.predict,
        weight space)."""
        self.evaluate()
        self.check_status()
        X, w = self._state(f)
        Ps = Phi_star.to(torch.float64).contiguous()
        mean = dense.gemm(Ps, w.reshape(-1, 1).contiguous()).reshape(-1)
        Z = dense.gemm(X, Ps, transB=True)  # (k, m*); X is lower triangular
        s2 = float(self.theta_host()[f, 3])
        return mean, s2 * dense.colsumsq(Z)

    def transform(self, data, f, test_ind):
        """Predictive mean and variance of fit f at test nodes, (m*, D) each
        (main.py:102-116 for one field)."""
        idx = np.asarray(test_ind).astype(np.int32).ravel()
        nodes = torch.as_tensor(idx, device=self.dev)
        mean, var = self.predict(f, data.phi_rows(nodes))
        return (mean.cpu().numpy().reshape(len(idx), -1), var.cpu().numpy().reshape(len(idx), -1))


def fit_batched(data, train_sets, vectors=None, epochs=200, lr=0.05, kernel=None,
                noise_variance=1e-3, theta0=None, positivity_constraint=1e-2,
                gram_precision="fp64", use_graph=True, timings=None):
    """Fit one GP per field of `vectors` (F, n, D) -- default data.vectors --
    each on its own node subset train_sets[f], all at once (Adam, device
    resident, one graph-captured launch per step).  Hyperparameters start at
    the Matern defaults of RVGP.fit (nu=3/2, kappa=5, sigma_f=1, main.py:30)
    and the given noise variance, or at theta0 ((4,) or (F, 4))."""
    import time
    if kernel not in (None, "matern"):
        raise NotImplementedError("batched fits use the Matern spectral kernel")
    V = data.vectors_device() if vectors is None else vectors
    if V is None:
        raise TypeError("no vector fields to fit")
    if theta0 is None:
        kern = ManifoldKernel(data, nu=3 / 2, kappa=5, typ="matern", sigma_f=1.0)
        theta0 = (float(kern.nu), float(kern.kappa), float(kern.sigma_f), float(noise_variance))
    t0 = time.perf_counter()
    gp = BatchedGPR(data, V, train_sets, theta0=theta0,
                    positivity_constraint=positivity_constraint, gram_precision=gram_precision)
    if timings is not None:
        torch.cuda.synchronize()
        timings["batch_setup"] = time.perf_counter() - t0
        t0 = time.perf_counter()
    gp.fit(epochs=epochs, lr=lr, use_graph=use_graph)
    if timings is not None:
        timings["batch_optimise"] = time.perf_counter() - t0
    return gp
