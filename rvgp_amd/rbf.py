"""The kernel='rbf' branch of RVGP.fit (RVGP/main.py:33-37) on the GPU.

The reference swaps its manifold kernel for gpflow.kernels.RBF on the rows of
the scalar Laplacian's eigenvectors (evecs_L) and fits the D ambient vector
components as D output columns of one GPflow GPR ("treating vectors
channel-wise"): hyperparameters variance, lengthscale (both 1.0 at start)
and the Gaussian noise variance, the NLML summed over the D columns.

Here: the Gram X X^T on the matrix cores (rvgp_gemm), the kernel matrix
from it (rvgp_rbf_kernel), the FP64 Cholesky and solves of the library, and
the hand-written adjoint rvgp_rbf_grad; L-BFGS-B on the host over the
device loss and gradient, or device-resident Adam.
"""
import numpy as np
import torch

from . import _lib, dense
from .gp import inv_softplus, softplus


class RBFEngine:
    """theta = (variance, lengthscale, noise) = softplus(u) + lower."""

    def __init__(self, X, Y, lower=(0.0, 0.0, 1e-6), trainable=(1, 1, 1)):
        dev = X.device
        self.dev = dev
        self.X = X.to(torch.float64).contiguous()
        Y = Y.to(torch.float64)
        self.y = (Y.reshape(-1, 1) if Y.dim() == 1 else Y).contiguous()
        self.M, self.k = self.X.shape
        self.R = self.y.shape[1]
        f64 = dict(dtype=torch.float64, device=dev)
        self.lower = torch.as_tensor(np.asarray(lower, np.float64), device=dev)
        self.trainable = np.asarray(trainable, bool)
        self.u = torch.zeros(3, **f64)
        self.theta = torch.zeros(3, **f64)
        self.G = dense.gemm(self.X, self.X, transB=True)        # X X^T (matrix cores)
        self.sqn = torch.diagonal(self.G).clone()
        self.K = torch.empty((self.M, self.M), **f64)
        self.Kinv = torch.empty((self.M, self.M), **f64)
        self.AAt = torch.empty((self.M, self.M), **f64)
        self.a = torch.empty_like(self.y)
        self.ya = torch.zeros(1, **f64)
        self.out = torch.zeros(7, **f64)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.eye = torch.eye(self.M, **f64)
        self._ws = torch.empty(_lib.lib().rvgp_rbf_grad_ws_bytes(self.M) // 8 + 1, **f64)
        self._dot_ws = torch.empty(_lib.lib().rvgp_dot_ws_bytes(self.M * self.R) // 8 + 1, **f64)

    def set_theta(self, theta):
        th = np.asarray(theta, np.float64) - self.lower.cpu().numpy()
        self.u.copy_(torch.as_tensor(inv_softplus(th), device=self.dev))

    def theta_host(self):
        return softplus(self.u.cpu().numpy()) + self.lower.cpu().numpy()

    def evaluate(self):
        """Enqueue NLML + adjoint: out = [nlml, d/d(variance, lengthscale,
        noise), d/du(...)] on the device."""
        s = _lib.stream()
        _lib.call("rvgp_rbf_params", _lib.ptr(self.u), _lib.ptr(self.lower), _lib.ptr(self.theta), s)
        _lib.call("rvgp_rbf_kernel", self.M, self.M, _lib.ptr(self.G), self.M, _lib.ptr(self.sqn),
                  _lib.ptr(self.sqn), _lib.ptr(self.theta), 1, 1, _lib.ptr(self.K), self.M, s)
        self.status.zero_()
        _lib.call("rvgp_potrf", 1, self.M, _lib.ptr(self.K), self.M, _lib.ptr(self.status), s)
        ldh = dense.chol_logdet_half(self.K)
        # a = K^-1 Y: L^-1 Y first (y^T K^-1 y = |L^-1 Y|^2), then L^-T
        self.a.copy_(self.y)
        dense.trsm(self.K, self.a)
        _lib.call("rvgp_dot", self.a.numel(), _lib.ptr(self.a), _lib.ptr(self.a), _lib.ptr(self.ya),
                  _lib.ptr(self._dot_ws), self._dot_ws.numel() * 8, s)
        dense.trsm(self.K, self.a, trans=True)
        # K^-1 = L^-T L^-1 (needed in full: dK/dl is not low rank)
        Linv = self.eye.clone()
        dense.trsm(self.K, Linv)
        dense.gemm(Linv, Linv, self.Kinv, transA=True)
        dense.gemm(self.a, self.a, self.AAt, transB=True)
        _lib.call("rvgp_rbf_grad", self.M, self.R, _lib.ptr(self.Kinv), self.M, _lib.ptr(self.AAt),
                  self.M, _lib.ptr(self.G), self.M, _lib.ptr(self.sqn), _lib.ptr(self.theta),
                  _lib.ptr(self.ya), _lib.ptr(ldh), _lib.ptr(self.u), _lib.ptr(self._ws),
                  self._ws.numel() * 8, _lib.ptr(self.out), s)
        return self.out

    def nlml_grad_u(self, u):
        self.u.copy_(torch.as_tensor(np.asarray(u, np.float64), device=self.dev))
        out = self.evaluate().cpu().numpy()
        if int(self.status.item()) != 0:
            raise dense.CholeskyError("Cholesky decomposition was not successful: the RBF kernel "
                                      "matrix is not positive definite")
        return float(out[0]), out[4:7].copy(), out

    def fit_lbfgs(self, maxiter=1000, disp=False):
        """L-BFGS-B over the trainable unconstrained parameters (a parameter
        fixed by kernel_variance / kernel_lengthscale keeps its value, as
        gpflow.set_trainable(..., False) does, main.py:69-75)."""
        import scipy.optimize
        u0 = self.u.cpu().numpy().copy()
        tr = self.trainable

        def f(x):
            u = u0.copy()
            u[tr] = x
            val, g, _ = self.nlml_grad_u(u)
            return val, g[tr]

        res = scipy.optimize.minimize(f, u0[tr], jac=True, method="L-BFGS-B",
                                      options={"disp": disp, "maxiter": maxiter})
        u = u0.copy()
        u[tr] = res.x
        self.u.copy_(torch.as_tensor(u, device=self.dev))
        self.last_nfev = int(res.nfev)
        return res

    def predict(self, Xs):
        """Mean (m*, R) and marginal variance (m*,) at inputs Xs (m*, k)."""
        self.evaluate()
        Xs = Xs.to(torch.float64).contiguous()
        m = Xs.shape[0]
        Gx = dense.gemm(self.X, Xs, transB=True)  # (M, m*)
        cn = (Xs * Xs).sum(1).contiguous()
        Ks = torch.empty((self.M, m), dtype=torch.float64, device=self.dev)
        _lib.call("rvgp_rbf_kernel", self.M, m, _lib.ptr(Gx), m, _lib.ptr(self.sqn), _lib.ptr(cn),
                  _lib.ptr(self.theta), 0, 0, _lib.ptr(Ks), m, _lib.stream())
        mean = dense.gemm(Ks, self.a, transA=True)  # (m*, R)
        dense.trsm(self.K, Ks)
        var = self.theta[0] - dense.colsumsq(Ks)
        return mean, var
