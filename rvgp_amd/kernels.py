"""Drop-in for RVGP/kernels.py: the spectral Matern / SE kernel on the
connection-Laplacian eigenbasis, evaluated on the GPU.

GPflow is not used: ``Parameter`` reproduces the parts of gpflow.Parameter
the RVGP path relies on -- a constrained value theta = softplus(u) + lower
with ``.numpy()``, ``.assign()`` and a trainable flag.  The lower bound is
taken from the module-level default positive minimum at construction time,
as gpflow.utilities.positive() does (RVGP/main.py:52 sets that default
after the kernel's parameters already exist).
"""
import numpy as np
import torch

from . import _lib, dense

_DEFAULT_POSITIVE_MINIMUM = 0.0


def set_default_positive_minimum(v):
    """gpflow.config.set_default_positive_minimum (global state)."""
    global _DEFAULT_POSITIVE_MINIMUM
    _DEFAULT_POSITIVE_MINIMUM = float(v)


def default_positive_minimum():
    return _DEFAULT_POSITIVE_MINIMUM


class Parameter:
    def __init__(self, value, lower=None, name=None, trainable=True):
        self.lower = default_positive_minimum() if lower is None else float(lower)
        self.name = name
        self.trainable = trainable
        self._value = float(value)

    def numpy(self):
        return np.float64(self._value)

    def assign(self, v):
        self._value = float(v)

    def __float__(self):
        return self._value

    def __repr__(self):
        return f"Parameter({self.name}={self._value:.6g}, lower={self.lower})"


class ManifoldKernel:
    """Matern kernel on manifold (kernels.py:7-67).

    K(X, X2) = (X * S) X2^T with S the spectral filter of the connection
    Laplacian eigenvalues, normalised to sum to the number of rows of the
    ambient eigenvector matrix and scaled by sigma_f (kernels.py:41-53).
    """

    def __init__(self, data, nu=3, kappa=4, sigma_f=1, typ='matern', dtype=None):
        self.data = data
        self.eigenvalues = data.evals_Lc
        self.num_verticies = float(data.n * data.vertices.shape[1])
        self.typ = typ
        if typ == 'matern':
            self.nu = Parameter(nu, name='nu')
        else:
            self.nu = Parameter(1.0, name='nu', trainable=False)
        self.kappa = Parameter(kappa, name='kappa')
        self.sigma_f = Parameter(sigma_f, name='sigma_f')

    def _evals_device(self):
        return self.data.spectrum_device()

    def eval_S(self, typ=None, with_grad=False):
        """Spectral filter S (k,) on the device (and dS/d(nu, kappa, sigma_f))."""
        typ = self.typ if typ is None else typ
        lam = self._evals_device()
        k = lam.numel()
        th = torch.tensor([float(self.nu), float(self.kappa), float(self.sigma_f)],
                          dtype=torch.float64, device=lam.device)
        S = torch.empty(k, dtype=torch.float64, device=lam.device)
        dS = torch.empty(3 * k, dtype=torch.float64, device=lam.device) if with_grad else None
        _lib.call("rvgp_spectral_filter", 0 if typ == 'matern' else 1, k, _lib.ptr(lam),
                  _lib.ptr(th), self.num_verticies, _lib.ptr(S), _lib.ptr(dS), _lib.stream())
        return (S, dS.reshape(3, k)) if with_grad else S

    def K(self, X, X2=None):
        """(X * S) @ X2^T (kernels.py:55-61), X rows of evecs_Lc."""
        dev = _lib.device()
        Xd = torch.as_tensor(np.asarray(X, np.float64), device=dev) if not isinstance(X, torch.Tensor) else X
        X2d = Xd if X2 is None else (torch.as_tensor(np.asarray(X2, np.float64), device=dev)
                                     if not isinstance(X2, torch.Tensor) else X2)
        S = self.eval_S()
        out = dense.gemm(Xd.double().contiguous(), X2d.double().contiguous(), transB=True, kscale=S)
        return out if isinstance(X, torch.Tensor) else out.cpu().numpy()

    def K_diag(self, X):
        """diag((X * S) X^T) (kernels.py:63-67)."""
        dev = _lib.device()
        Xd = torch.as_tensor(np.asarray(X, np.float64), device=dev) if not isinstance(X, torch.Tensor) else X
        S = self.eval_S()
        d = ((Xd.double() ** 2) * S).sum(1)
        return d if isinstance(X, torch.Tensor) else d.cpu().numpy()

    def __call__(self, X, X2=None):
        return self.K(X, X2)


class RBF:
    """gpflow.kernels.RBF as the reference creates it for kernel='rbf'
    (main.py:33-37): variance and lengthscales start at 1.0, positive with
    the default minimum in force when the kernel is created."""

    def __init__(self, variance=1.0, lengthscales=1.0):
        self.variance = Parameter(variance, name='variance')
        self.lengthscales = Parameter(lengthscales, name='lengthscales')
