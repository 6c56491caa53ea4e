"""Drop-in for RVGP/main.py: ``RVGP.fit`` (train_gp) and the fitted model's
``transform``.

With ``reference_semantics=True`` (the default) the reference's behaviour is
kept, quirks included: the 80/20 re-split of ``train_ind`` with
sklearn's train_test_split rule (main.py:40-45), the noise variance starting
at 1.0 whatever ``noise_variance`` says (main.py:99-100 passes None to
GPflow), the kernel parameters created before
``set_default_positive_minimum`` (main.py:30 vs 52), AttributeError for
``kernel_variance`` / ``kernel_lengthscale`` (ManifoldKernel has no
``variance`` / ``lengthscales``), L-BFGS-B with ``maxiter=epochs``, and the
int / float / other dispatch of ``transform`` (main.py:104-109).
``reference_semantics=False`` honours ``noise_variance`` and skips the
re-split.  All numerical work (NLML, adjoint, predictions) runs on the GPU
through gp.GPREngine.
"""
import numpy as np
import torch

from . import _lib
from .gp import GPREngine, weight_space_stats
from .kernels import RBF, ManifoldKernel, set_default_positive_minimum

GAUSSIAN_VARIANCE_LOWER_BOUND = 1e-6  # gpflow.likelihoods.Gaussian default


def checked_nodes(idx, n):
    """Node indices as numpy indexing reads them (main.py:27, 41, 104):
    integer indices in [-n, n), negatives wrapped to idx + n; anything else
    raises IndexError before a kernel sees it (an index outside [0, n) would
    otherwise be an out-of-bounds device read)."""
    idx = np.asarray(idx)
    if idx.dtype == bool:
        if idx.shape != (n,):
            raise IndexError(f"boolean index did not match indexed array along axis 0; size of "
                             f"axis is {n} but size of corresponding boolean axis is {idx.size}")
        return np.nonzero(idx)[0].astype(np.int64)
    if idx.size and not np.issubdtype(idx.dtype, np.integer):
        raise IndexError("arrays used as indices must be of integer (or boolean) type")
    idx = idx.astype(np.int64).reshape(-1)
    bad = (idx < -n) | (idx >= n)
    if bad.any():
        raise IndexError(f"index {int(idx[bad][0])} is out of bounds for axis 0 with size {n}")
    return np.where(idx < 0, idx + n, idx)


def train_test_split_indices(n, test_size=0.2, seed=0):
    """Index rule of sklearn.model_selection.train_test_split(random_state=seed)
    for a float test_size: n_test = ceil(test_size * n), permutation from
    RandomState(seed), test = perm[:n_test], train = perm[n_test:]."""
    n_test = int(np.ceil(test_size * n))
    perm = np.random.RandomState(seed).permutation(n)
    return perm[n_test:], perm[:n_test]


class manifold_GPR:
    """The fitted model (main.py:98-116)."""

    def __init__(self, data, kernel, train_nodes, engine, reference_semantics=True):
        self.data = data
        self.kernel = kernel
        self.train_nodes = train_nodes
        self.engine = engine
        self.reference_semantics = reference_semantics

    @property
    def hyperparameters(self):
        th = self.engine.theta_host()
        return {"nu": th[0], "kappa": th[1], "sigma_f": th[2], "noise_variance": th[3]}

    def training_loss(self):
        u = self.engine.u.cpu().numpy()
        return self.engine.nlml_grad_u(u)[0]

    def predict_f(self, Xnew):
        """Mean and marginal variance at spectral rows Xnew ((m, k) array or
        device tensor), as GPflow's predict_f(full_cov=False)."""
        dev = self.engine.dev
        Xd = Xnew if isinstance(Xnew, torch.Tensor) else torch.as_tensor(
            np.asarray(Xnew, np.float64), device=dev)
        mean, var = self.engine.predict(Xd)
        if self.engine.R > 1:  # (m, R): one column per right-hand side
            return mean, var.reshape(-1, 1)
        return mean.reshape(-1, 1), var.reshape(-1, 1)

    def predict_nodes(self, data, nodes):
        """predict_f at the ambient rows of the given nodes: through the
        intrinsic rows and frames when the weight-space system holds the fit,
        else through data.phi_rows."""
        if isinstance(nodes, torch.Tensor):
            nodes = nodes.cpu().numpy()
        nodes = torch.as_tensor(checked_nodes(nodes, data.n), device=self.engine.dev)
        e = self.engine
        weight = hasattr(e, "predict_nodes") and (e.form == "weight" or e.precision == "mixed")
        U = data.intrinsic_rows(nodes) if weight else None
        if U is None:
            return self.predict_f(data.phi_rows(nodes.int()))
        T = data._gauges.index_select(0, nodes)[:, :, :data.dim_man]
        mean, var = e.predict_nodes(U, T)
        if e.R > 1:
            return mean, var.reshape(-1, 1)
        return mean.reshape(-1, 1), var.reshape(-1, 1)

    def transform(self, data, test_ind):
        """Predictive mean and variance at test nodes (main.py:102-116)."""
        first = test_ind[0]
        ints = (int,) if self.reference_semantics else (int, np.integer)
        floats = (float,) if self.reference_semantics else (float, np.floating)
        if isinstance(first, ints):
            # a Python bool list indexes as a mask in numpy
            mean, var = self.predict_nodes(data, checked_nodes(test_ind, data.n))
        elif isinstance(first, floats):
            # positional encodings: spectral rows (k values each) given
            # directly (main.py:108-109).  The reference then reshapes to
            # len(test_ind) rows -- the number of floats, which cannot hold
            # the predictions -- so here each spectral row gets one row of
            # output: mean (rows, R) and variance (rows, 1).
            test_x = np.asarray(test_ind, dtype=np.float64).reshape(-1, self.engine.k)
            mean, var = self.predict_f(test_x)
            rows = test_x.shape[0]
            return mean.cpu().numpy().reshape(rows, -1), var.cpu().numpy().reshape(rows, -1)
        else:  # np.int64 / np.bool_ match neither branch of main.py:104-109
            raise UnboundLocalError("local variable 'test_x' referenced before assignment")
        if self.engine.R > 1:
            # a batch of R fields fitted with shared hyperparameters:
            # mean (R, len(test_ind), D), variance (len(test_ind), D)
            mean = mean.cpu().numpy().reshape(len(test_ind), -1, self.engine.R)
            return np.ascontiguousarray(mean.transpose(2, 0, 1)), \
                var.cpu().numpy().reshape(len(test_ind), -1)
        mean = mean.cpu().numpy().reshape(len(test_ind), -1)
        var = var.cpu().numpy().reshape(len(test_ind), -1)
        return mean, var


def train_gp(data, train_ind=None, n_inducing_points=None, test_size=0.2, kernel=None,
             noise_variance=0.001, kernel_lengthscale=None, kernel_variance=None, epochs=1000,
             positivity_constraint=1e-2, seed=0, *, reference_semantics=True, optimizer="lbfgs",
             lr=0.05, form="auto", precision="fp64", verbose=True, use_graph=True,
             timings=None):
    """Fit the vector-field GP (main.py:11-84)."""
    if train_ind is None:
        train_ind = np.arange(data.n)
    train_ind = checked_nodes(train_ind, data.n)
    if kernel is None:
        kernel = ManifoldKernel(data, nu=3 / 2, kappa=5, typ='matern', sigma_f=1.)
    elif isinstance(kernel, str) and kernel == 'rbf':
        kernel = RBF()
        if verbose:
            print('Using RBF kernel, treating vectors channel-wise.')
    else:
        raise UnboundLocalError("local variable 'input' referenced before assignment")

    if reference_semantics:
        tr_pos, te_pos = train_test_split_indices(len(train_ind), test_size, seed)
        nodes_tr, nodes_te = train_ind[tr_pos], train_ind[te_pos]
    else:
        nodes_tr, nodes_te = train_ind, train_ind[:0]

    set_default_positive_minimum(positivity_constraint)
    if isinstance(kernel, RBF):
        if n_inducing_points is not None:
            raise NotImplementedError("SGPR inducing points with kernel='rbf' are not built")
        return _train_rbf(data, kernel, nodes_tr, nodes_te, noise_variance, kernel_lengthscale,
                          kernel_variance, epochs, optimizer, reference_semantics, verbose, timings)
    if kernel_variance is not None:
        raise AttributeError("'ManifoldKernel' object has no attribute 'variance'")
    if kernel_lengthscale is not None:
        raise AttributeError("'ManifoldKernel' object has no attribute 'lengthscales'")

    dev = _lib.device()
    clock = _Clock(timings)
    nd_tr = torch.as_tensor(nodes_tr.astype(np.int32), device=dev)
    V = data.vectors_device()
    if V is None:
        raise TypeError("data.vectors is None: nothing to fit")
    idx_tr = torch.as_tensor(nodes_tr, device=dev).long()
    D = V.shape[-1]
    M = len(nodes_tr) * D
    weight = n_inducing_points is None and (form == "weight" or (form == "auto" and M > 8192))
    U = data.intrinsic_rows(idx_tr) if weight else None
    if U is not None:
        # weight form from the d-row intrinsic blocks (Phi^T Phi = U^T U):
        # the D-row ambient rows are never formed
        Phi = None
        c = data._frame(0, V, data.dim_man)  # T_i^T v_i, (n, d) or (R, n, d)
        if V.dim() == 3:
            cy = c[:, idx_tr].permute(1, 2, 0).reshape(-1, V.shape[0]).contiguous()
            yv = V[:, idx_tr]
        else:
            cy = c[idx_tr].reshape(-1, 1).contiguous()
            yv = V[idx_tr]
        yy = (yv * yv).sum().reshape(1)
    else:
        Phi = data.phi_rows(nd_tr)
        if V.dim() == 3:
            # a batch (R, n, D) of fields observed at the same nodes: R
            # right-hand sides of one GP with shared hyperparameters, NLML summed
            y = V[:, idx_tr].permute(1, 2, 0).reshape(-1, V.shape[0]).contiguous()
        else:
            y = V[idx_tr].reshape(-1)
    noise0 = 1.0 if reference_semantics else float(noise_variance)
    lower = (kernel.nu.lower, kernel.kappa.lower, kernel.sigma_f.lower,
             GAUSSIAN_VARIANCE_LOWER_BOUND)
    if n_inducing_points is not None:
        return _train_sgpr(data, kernel, nodes_tr, nodes_te, Phi, y, V, int(n_inducing_points),
                           lower, noise0, epochs, optimizer, reference_semantics, verbose, clock)
    if Phi is None:
        G, b = weight_space_stats(U, cy, precision)
        eng = GPREngine(None, None, data.spectrum_device(), kernel.num_verticies, kernel.typ,
                        form=form, precision=precision, lower=lower, stats=(G, b, yy, M))
    else:
        eng = GPREngine(Phi, y, data.spectrum_device(), kernel.num_verticies, kernel.typ, form=form,
                        precision=precision, lower=lower)
    eng.set_theta([float(kernel.nu), float(kernel.kappa), float(kernel.sigma_f), noise0])
    clock.tick("gp_setup")
    if optimizer == "lbfgs":
        eng.fit_lbfgs(maxiter=epochs, disp=verbose)
    elif optimizer == "adam":
        eng.adam_setup()
        # a step that is already one launch gains nothing from a graph (the
        # capture's warm-up step and instantiation cost more than they save)
        if use_graph and epochs > 1 and not eng.fused():
            g = eng.capture_adam_step(lr=lr)
            for _ in range(int(epochs)):
                g.replay()
        else:
            for _ in range(int(epochs)):
                eng.adam_step(lr=lr)
        torch.cuda.synchronize()
    else:
        raise ValueError("optimizer must be 'lbfgs' or 'adam'")
    clock.tick("gp_optimise")
    th = eng.theta_host()
    kernel.nu.assign(th[0])
    kernel.kappa.assign(th[1])
    kernel.sigma_f.assign(th[2])
    model = manifold_GPR(data, kernel, nodes_tr, eng, reference_semantics)

    if len(nodes_te):
        idx_te = torch.as_tensor(nodes_te, device=dev).long()
        mean, _ = model.predict_nodes(data, idx_te)
        out_test = (V[:, idx_te].permute(1, 2, 0).reshape(-1, V.shape[0]) if V.dim() == 3
                    else V[idx_te].reshape(-1, 1))
        # np.linalg.norm(out_test - out_pred, axis=1).mean() on (rows, 1) arrays
        err = (out_test - mean).abs().mean().item()
        model.heldout_error = err
        if verbose:
            print("Relative l2 error is {}".format(err))
    clock.tick("gp_heldout_predict")
    return model


class manifold_SGPR(manifold_GPR):
    """The fitted n_inducing_points model (main.py:119-137): GPflow's SGPR
    with the manifold kernel; ``inducing_nodes`` are the FPS-selected rows."""

    def training_loss(self):
        return self.engine.nlml_grad_u(self.engine.u)[0]


def _train_sgpr(data, kernel, nodes_tr, nodes_te, Phi, y, V, n_inducing, lower, noise0, epochs,
                optimizer, reference_semantics, verbose, clock):
    """main.py:58-67: inducing rows = furthest_point_sampling(in_train,
    N=n_inducing_points) of the training rows of evecs_Lc, then SGPR;
    L-BFGS-B with maxiter=epochs (main.py:87-95)."""
    from .geometry import furthest_point_sampling_device
    from .sgpr import SGPREngine
    if optimizer != "lbfgs":
        raise NotImplementedError("SGPR is fitted with L-BFGS-B, as in the reference")
    if y.dim() == 2 and y.shape[1] > 1:
        raise ValueError("SGPR fits one field (vectors of shape (n, D))")
    ind, _ = furthest_point_sampling_device(Phi, N=n_inducing)
    Z = Phi[ind.long()].contiguous()
    eng = SGPREngine(Phi, y, Z, data.spectrum_device(), kernel.num_verticies, kernel.typ,
                     lower=lower)
    eng.inducing_rows = ind.cpu().numpy()
    eng.set_theta([float(kernel.nu), float(kernel.kappa), float(kernel.sigma_f), noise0])
    eng.R = 1
    clock.tick("gp_setup")
    eng.fit_lbfgs(maxiter=epochs, disp=verbose)
    clock.tick("gp_optimise")
    th = eng.theta_host()
    kernel.nu.assign(th[0])
    kernel.kappa.assign(th[1])
    kernel.sigma_f.assign(th[2])
    model = manifold_SGPR(data, kernel, nodes_tr, eng, reference_semantics)
    if len(nodes_te):
        dev = eng.dev
        mean, _ = model.predict_f(data.phi_rows(torch.as_tensor(nodes_te.astype(np.int32),
                                                                device=dev)))
        out_test = V[torch.as_tensor(nodes_te, device=dev).long()].reshape(-1, 1)
        model.heldout_error = (out_test - mean).abs().mean().item()
        if verbose:
            print("Relative l2 error is {}".format(model.heldout_error))
    clock.tick("gp_heldout_predict")
    return model


class rbf_GPR:
    """The fitted kernel='rbf' model: one GPR over the D vector components
    as output columns, inputs = rows of evecs_L (main.py:33-37, 98-116)."""

    def __init__(self, data, kernel, train_nodes, engine, reference_semantics=True):
        self.data = data
        self.kernel = kernel
        self.train_nodes = train_nodes
        self.engine = engine
        self.reference_semantics = reference_semantics

    @property
    def hyperparameters(self):
        th = self.engine.theta_host()
        return {"variance": th[0], "lengthscales": th[1], "noise_variance": th[2]}

    def training_loss(self):
        return self.engine.nlml_grad_u(self.engine.u.cpu().numpy())[0]

    def predict_f(self, Xnew):
        dev = self.engine.dev
        Xd = Xnew if isinstance(Xnew, torch.Tensor) else torch.as_tensor(
            np.asarray(Xnew, np.float64), device=dev)
        if Xd.shape[1] != self.engine.k:
            raise ValueError(f"RBF inputs have {Xd.shape[1]} columns; the model was trained on "
                             f"{self.engine.k} (rows of evecs_L)")
        mean, var = self.engine.predict(Xd)
        return mean, var.reshape(-1, 1).expand(-1, self.engine.R)

    def transform(self, data, test_ind):
        """main.py:102-116.  With reference_semantics the reference's int
        branch is kept as it is: it reads rows of evecs_Lc (not evecs_L)
        reshaped to (len * D, k), which only runs when both spectra have the
        same number of eigenpairs, and returns (len, D * D) arrays.  Without,
        the test nodes' evecs_L rows are used: (len, D) mean and variance."""
        first = test_ind[0]
        ints = (int,) if self.reference_semantics else (int, np.integer)
        floats = (float,) if self.reference_semantics else (float, np.floating)
        if isinstance(first, ints):
            idx = torch.as_tensor(checked_nodes(test_ind, data.n), device=self.engine.dev)
            if self.reference_semantics:
                E = torch.as_tensor(np.asarray(data.evecs_Lc), device=self.engine.dev)
                test_x = E.reshape(data.n, -1)[idx].reshape(-1, E.shape[1])
            else:
                test_x = data._evecs_L[idx]
        elif isinstance(first, floats):
            test_x = np.asarray(test_ind, dtype=np.float64).reshape(-1, self.engine.k)
        else:
            raise UnboundLocalError("local variable 'test_x' referenced before assignment")
        mean, var = self.predict_f(test_x)
        return (mean.cpu().numpy().reshape(len(test_ind), -1),
                var.cpu().numpy().reshape(len(test_ind), -1))


def _train_rbf(data, kernel, nodes_tr, nodes_te, noise_variance, kernel_lengthscale,
               kernel_variance, epochs, optimizer, reference_semantics, verbose, timings):
    """kernel='rbf' (main.py:33-37): inputs = evecs_L rows of the training
    nodes, outputs = their D vector components as D columns; GPflow's GPR
    with the Gaussian noise starting at 1.0 under reference semantics;
    kernel_variance / kernel_lengthscale fix those parameters
    (main.py:69-75); L-BFGS-B with maxiter=epochs (main.py:87-95)."""
    from .rbf import RBFEngine
    if optimizer != "lbfgs":
        raise NotImplementedError("kernel='rbf' is fitted with L-BFGS-B, as in the reference")
    dev = _lib.device()
    clock = _Clock(timings)
    V = data.vectors_device()
    if V is None:
        raise TypeError("data.vectors is None: nothing to fit")
    if V.dim() != 2:
        raise ValueError("kernel='rbf' fits one field (vectors of shape (n, D))")
    idx = torch.as_tensor(nodes_tr, device=dev).long()
    X = data._evecs_L[idx].contiguous()
    Y = V[idx].contiguous()
    noise0 = 1.0 if reference_semantics else float(noise_variance)
    trainable = [True, True, True]
    if kernel_variance is not None:
        kernel.variance.assign(kernel_variance)
        trainable[0] = False
    if kernel_lengthscale is not None:
        kernel.lengthscales.assign(kernel_lengthscale)
        trainable[1] = False
    eng = RBFEngine(X, Y, lower=(kernel.variance.lower, kernel.lengthscales.lower,
                                 GAUSSIAN_VARIANCE_LOWER_BOUND), trainable=trainable)
    eng.set_theta([float(kernel.variance), float(kernel.lengthscales), noise0])
    clock.tick("gp_setup")
    eng.fit_lbfgs(maxiter=epochs, disp=verbose)
    clock.tick("gp_optimise")
    th = eng.theta_host()
    kernel.variance.assign(th[0])
    kernel.lengthscales.assign(th[1])
    model = rbf_GPR(data, kernel, nodes_tr, eng, reference_semantics)
    if len(nodes_te):
        mean, _ = model.predict_f(data._evecs_L[torch.as_tensor(nodes_te, device=dev).long()])
        out_test = V[torch.as_tensor(nodes_te, device=dev).long()]
        model.heldout_error = torch.linalg.vector_norm(out_test - mean, dim=1).mean().item()
        if verbose:
            print("Relative l2 error is {}".format(model.heldout_error))
    clock.tick("gp_heldout_predict")
    return model


class _Clock:
    """Optional stage timer (synchronises the device at each tick)."""

    def __init__(self, timings):
        import time
        self.t, self.time = timings, time
        self.last = time.perf_counter()

    def tick(self, name):
        if self.t is None:
            return
        torch.cuda.synchronize()
        now = self.time.perf_counter()
        self.t[name] = now - self.last
        self.last = now
