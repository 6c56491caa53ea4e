"""The n_inducing_points branch of RVGP.fit (RVGP/main.py:54-67) on the GPU:
GPflow's SGPR with the spectral connection-Laplacian kernel.

The reference picks the inducing points by furthest-point sampling of the
training rows of evecs_Lc (main.py:59-60) and fits gpflow.models.SGPR, whose
objective is Titsias' collapsed bound.  With the kernel x diag(S) x'^T the
bound only needs the weight-space statistics G = X^T X, b = X^T y, y^T y of
the N training rows and m x k / m x m products with the inducing rows Z
(csrc/sgpr.hip states the bound in GPflow's factors and its adjoint): the matrix-core
GEMMs of rvgp_gemm, the FP64 Cholesky and solves of the library, and
rvgp_sgpr_finish for the bound and its gradient.  L-BFGS-B over the softplus
parameters on the host, as gpflow.optimizers.Scipy does.
"""
import numpy as np
import torch

from . import _lib, dense
from .gp import inv_softplus, softplus

SGPR_JITTER = 1e-6  # gpflow.config.default_jitter()


class SGPREngine:
    """theta = (nu, kappa, sigma_f, s2) = softplus(u) + lower."""

    def __init__(self, X, y, Z, evals, n_rows, typ="matern", lower=(0.0, 0.0, 0.0, 1e-6),
                 jitter=SGPR_JITTER):
        dev = X.device
        self.dev = dev
        X = X.to(torch.float64).contiguous()
        self.N, self.k = X.shape
        self.Z = Z.to(torch.float64).contiguous()
        self.m = self.Z.shape[0]
        self.typ = 0 if typ == "matern" else 1
        self.n_rows = float(n_rows)
        self.jitter = float(jitter)
        f64 = dict(dtype=torch.float64, device=dev)
        y = y.to(torch.float64).reshape(-1, 1).contiguous()
        self.G = dense.gemm(X, X, transA=True)
        self.G = (0.5 * (self.G + self.G.T)).contiguous()
        self.b = dense.gemm(X, y, transA=True)
        self.yy = torch.zeros(1, **f64)
        dws = torch.empty(_lib.lib().rvgp_dot_ws_bytes(self.N) // 8 + 1, **f64)
        _lib.call("rvgp_dot", self.N, _lib.ptr(y), _lib.ptr(y), _lib.ptr(self.yy), _lib.ptr(dws),
                  dws.numel() * 8, _lib.stream())
        self.evals = evals.to(torch.float64).contiguous()
        self.lower = np.asarray(lower, np.float64)
        self.u = np.zeros(4)
        self.theta_d = torch.zeros(3, **f64)
        self.u_d = torch.zeros(4, **f64)
        self.S = torch.empty(self.k, **f64)
        self.dS = torch.empty(3 * self.k, **f64)
        self.gS = torch.empty(self.k, **f64)
        self.out = torch.zeros(9, **f64)
        self.scal = torch.zeros(3, **f64)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.eye_m = torch.eye(self.m, **f64)
        self._dws = torch.empty(_lib.lib().rvgp_dot_ws_bytes(self.m) // 8 + 1, **f64)

    def set_theta(self, theta):
        self.u = inv_softplus(np.asarray(theta, np.float64) - self.lower)

    def theta_host(self):
        return softplus(self.u) + self.lower

    def _factor(self, L):
        self.status.zero_()
        _lib.call("rvgp_potrf", 1, self.m, _lib.ptr(L), self.m, _lib.ptr(self.status),
                  _lib.stream())
        return L

    def _chol_solve(self, L, B):
        dense.trsm(L, B)
        dense.trsm(L, B, trans=True)
        return B

    def evaluate(self, u=None):
        """NLML = -ELBO and its gradient; out (9) on the device."""
        if u is not None:
            self.u = np.asarray(u, np.float64).copy()
        th = self.theta_host()
        s2 = float(th[3])
        self.theta_d.copy_(torch.as_tensor(th[:3], device=self.dev))
        self.u_d.copy_(torch.as_tensor(self.u, device=self.dev))
        s = _lib.stream()
        _lib.call("rvgp_spectral_filter", self.typ, self.k, _lib.ptr(self.evals),
                  _lib.ptr(self.theta_d), self.n_rows, _lib.ptr(self.S), _lib.ptr(self.dS), s)
        Z, S = self.Z, self.S
        # GPflow's factors (SGPR.elbo): Kuu = L L^T, A A^T = L^-1 K_uf K_fu
        # L^-T / s2 = R1 S Q^T / s2, B = I + A A^T = L_B L_B^T -- B has
        # eigenvalues >= 1, so its factor stays accurate however low the rank
        # of the kernel is
        Kuu = dense.gemm(Z, Z, transB=True, kscale=S)                     # Z S Z^T
        torch.diagonal(Kuu).add_(self.jitter)
        L = self._factor(Kuu)
        st = int(self.status.item())
        Q = Z.clone()
        dense.trsm(L, Q)                                                  # L^-1 Z
        R1 = dense.gemm(Q, self.G, kscale=S)                              # Q S G
        B = self.eye_m.clone()
        dense.gemm(R1, Q, B, transB=True, kscale=S, alpha=1.0 / s2, beta=1.0)
        LB = self._factor(B)
        if st != 0 or int(self.status.item()) != 0:
            raise dense.CholeskyError("Cholesky decomposition was not successful: the SGPR "
                                      "system matrix is not positive definite")
        Qb = Q.clone()
        dense.trsm(LB, Qb)                                                # L_B^-1 Q
        R1b = R1.clone()
        dense.trsm(LB, R1b)                                               # L_B^-1 R1
        T = dense.gemm(Q, Q, transA=True)                                 # Q^T Q
        W1 = dense.gemm(R1, T, kscale=S)                                  # R1 S Q^T Q
        cc = dense.gemm(Q, self.b, kscale=S)                              # Q S b
        dense.trsm(LB, cc)                                                # L_B^-1 Q S b
        zw = dense.gemm(Qb, cc, transA=True)
        hw = dense.gemm(R1b, cc, transA=True)
        self.scal[0:1].copy_(dense.chol_logdet_half(LB))
        _lib.call("rvgp_dot", self.m, _lib.ptr(cc), _lib.ptr(cc), _lib.ptr(self.scal[2:]),
                  _lib.ptr(self._dws), self._dws.numel() * 8, s)
        _lib.call("rvgp_sgpr_finish", self.m, self.k, float(self.N), _lib.ptr(Q), _lib.ptr(Qb),
                  _lib.ptr(R1), _lib.ptr(R1b), _lib.ptr(W1), _lib.ptr(self.G), _lib.ptr(self.b),
                  _lib.ptr(zw), _lib.ptr(hw), _lib.ptr(S), _lib.ptr(self.dS), _lib.ptr(self.scal),
                  _lib.ptr(self.yy), s2, _lib.ptr(self.u_d), _lib.ptr(self.gS), _lib.ptr(self.out),
                  s)
        self._cache = (L, LB, cc, s2)
        return self.out

    def nlml_grad_u(self, u):
        out = self.evaluate(u).cpu().numpy()
        return float(out[0]), out[5:9].copy(), out

    def fit_lbfgs(self, maxiter=1000, disp=False):
        import scipy.optimize

        def f(u):
            val, g, _ = self.nlml_grad_u(u)
            return val, g

        res = scipy.optimize.minimize(f, self.u.copy(), jac=True, method="L-BFGS-B",
                                      options={"disp": disp, "maxiter": maxiter})
        self.u = res.x.copy()
        self.last_nfev = int(res.nfev)
        return res

    def predict(self, Phi_star):
        """SGPR.predict_f(full_cov=False): with Kus = Z S X*^T, tmp1 = L^-1 Kus,
        tmp2 = L_B^-1 tmp1: mean tmp2^T cc / s2, variance
        diag(K**) + |tmp2|^2 - |tmp1|^2 per column."""
        self.evaluate()
        L, LB, cc, s2 = self._cache
        Ps = Phi_star.to(torch.float64).contiguous()
        t1 = dense.gemm(self.Z, Ps, transB=True, kscale=self.S)           # Kus (m, m*)
        dense.trsm(L, t1)
        t2 = t1.clone()
        dense.trsm(LB, t2)
        mean = dense.gemm(t2, cc, transA=True).reshape(-1) / s2
        kss = dense.colsumsq((Ps * torch.sqrt(self.S)).T.contiguous())
        var = kss + dense.colsumsq(t2) - dense.colsumsq(t1)
        return mean, var
