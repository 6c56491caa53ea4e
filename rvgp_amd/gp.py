"""Device GP regression with the spectral connection-Laplacian kernel.

Replaces GPflow's GPR (RVGP/main.py:98-116 subclasses gpflow.models.GPR):
log marginal likelihood, its gradient (TF autodiff in the reference, a
hand-written adjoint here) and predict_f.

Two exact forms of the same model y ~ N(0, Phi diag(S) Phi^T + s2 I):
  * function space: M x M Gram on the matrix cores (rvgp_gemm with the
    spectral filter as k-scale), blocked Cholesky, triangular solves;
    in mixed precision alpha comes from the FP64 k x k system below
    (Woodbury), refined against the matrix-free FP64 K_y, and no M x M
    matrix is formed.
  * weight space (Woodbury / determinant lemma): k x k system
    A = s2 S^-1 + Phi^T Phi; G = Phi^T Phi and b = Phi^T y are formed once per
    training set, so an optimiser step costs O(k^3) and never touches an
    M x M matrix -- the only feasible form at M = 768k.
The adjoint forms dK = Phi diag(dS) Phi^T implicitly: only k-vectors
(diag(Phi^T K^-1 Phi), Phi^T alpha) are needed, never an M x M dK.
"""
import math

import numpy as np
import torch

from . import _lib, dense

TWO_PI = 2.0 * math.pi


def softplus(x):
    return np.logaddexp(0.0, x)


def inv_softplus(y):
    y = np.asarray(y, dtype=np.float64)
    return y + np.log(-np.expm1(-y))


class NLMLFunction(torch.autograd.Function):
    """NLML as a torch.autograd.Function whose backward is the hand-written
    adjoint (replacing TF autodiff through GPflow's training_loss,
    main.py:89-91).  apply(x, engine, unconstrained): x = (nu, kappa,
    sigma_f, noise_variance) -- or, with unconstrained=True, the
    softplus-unconstrained u -- as a (4,) FP64 device tensor; returns the
    scalar NLML.  forward runs the engine's evaluate() (NLML and gradient in
    one pass); backward only scales the cached gradient."""

    @staticmethod
    def forward(ctx, x, engine, unconstrained=False):
        xd = x.detach().to(engine.dev, torch.float64)
        if unconstrained:
            engine.u.copy_(xd)
        else:
            y = xd - engine.lower
            engine.u.copy_(y + torch.log(-torch.expm1(-y)))  # inverse softplus
        out = engine.evaluate()
        ctx.save_for_backward((out[5:9] if unconstrained else out[1:5]).clone())
        return out[0].clone()

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None, None


def weight_space_stats(U, c, precision="fp64"):
    """Phi^T Phi, Phi^T y and y^T y of a GP whose rows Phi = blockdiag(T_i) U
    have orthonormal frames T_i, from the intrinsic rows U (m*d, k) and the
    frame coordinates c = T^T y (m*d, R): Phi^T Phi = U^T U, Phi^T y = U^T c.
    y^T y is not recoverable from c (the normal components of y count) and is
    taken by the caller.  precision "mixed": FP32 U tiles with the products
    accumulated in FP64 on the matrix cores, as for ambient rows."""
    if precision == "fp64":
        G = dense.gemm(U, U, transA=True, mode=dense.GEMM_F64)
    else:
        Uf = U.float()
        G = dense.gemm(Uf, Uf, transA=True, mode=dense.GEMM_F32_ACC64, out_dtype=torch.float64)
    b = dense.gemm(U, c, transA=True)
    return G, b


FUSED_STEP_MAX_K = 2048  # RVGP_WS_STEP_MAX_K


class GPREngine:
    """Hyperparameters theta = (nu, kappa, sigma_f, s2) = softplus(u) + lower."""

    def __init__(self, Phi, y, evals, n_rows, typ="matern", form="auto", precision="fp64",
                 lower=(0.0, 0.0, 0.0, 1e-6), fs_max=8192, gram_mode=None, stats=None):
        """Phi (M, k), y (M,) or (M, R); or Phi = y = None and stats = (G, b,
        yy, M) -- the weight-space statistics Phi^T Phi (k, k), Phi^T y
        (k, R), y^T y (1,) and the row count M -- for the weight form."""
        if stats is not None:
            self._init_from_stats(stats, evals, n_rows, typ, form, precision, lower)
            return
        dev = Phi.device
        self.dev = dev
        self.M, self.k = Phi.shape
        self.typ = 0 if typ == "matern" else 1
        self.n_rows = float(n_rows)
        self.precision = precision
        self.form = ("function" if self.M <= fs_max else "weight") if form == "auto" else form
        self.Phi = Phi.to(torch.float64).contiguous()
        # y: (M,) or (M, R) -- R right-hand sides sharing the hyperparameters,
        # NLML summed over them
        y = y.to(torch.float64)
        self.y = (y.reshape(-1, 1) if y.dim() == 1 else y).contiguous()
        assert self.y.shape[0] == self.M, (self.y.shape, self.M)
        self.R = self.y.shape[1]
        self.evals = evals.to(torch.float64).contiguous()
        f64 = dict(dtype=torch.float64, device=dev)
        self.u = torch.zeros(4, **f64)
        self.lower = torch.as_tensor(np.asarray(lower, np.float64), device=dev)
        self.theta = torch.zeros(3, **f64)
        self.s2 = torch.zeros(1, **f64)
        self.S = torch.empty(self.k, **f64)
        self.dS = torch.empty(3 * self.k, **f64)
        self.gS = torch.empty(self.k, **f64)
        self.out = torch.zeros(9, **f64)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.fused_step = True
        self.ya = torch.zeros(1, **f64)
        self.aa = torch.zeros(1, **f64)
        self._dot_ws = torch.empty(
            _lib.lib().rvgp_dot_ws_bytes(self.M * self.R) // 8 + 1, **f64)
        self._cs_ws = torch.empty(
            _lib.lib().rvgp_colsumsq_ws_bytes(max(self.M, self.k), max(self.k, 1)) // 8 + 1, **f64)
        if self.form == "weight" or precision == "mixed":
            # k x k normal-equation system (weight space; in the mixed
            # function-space form it carries log det, diag(Phi^T K^-1 Phi) and
            # the Woodbury solve for alpha, all in FP64)
            if gram_mode is None:
                # mixed weight form: FP32 Phi tiles, products accumulated in
                # FP64 on the matrix cores (an FP32-accumulated G over M = 768k
                # rows costs ~1e-4 in the NLML).  The function form's Woodbury
                # refinement divides by s2, which amplifies the ~1e-7 error of
                # FP32-rounded rows until it stops converging at s2 ~ 1e-5; its
                # G (M <= fs_max rows) costs next to nothing in FP64.
                mixed_rows = precision == "mixed" and self.form == "weight"
                gram_mode = dense.GEMM_F32_ACC64 if mixed_rows else dense.GEMM_F64
            self.gram_mode = gram_mode
            P = self.Phi if gram_mode == dense.GEMM_F64 else self.Phi.float()
            acc64 = gram_mode in (dense.GEMM_F64, dense.GEMM_F32_ACC64)
            G = dense.gemm(P, P, transA=True, mode=gram_mode,
                           out_dtype=torch.float64 if acc64 else torch.float32)
            self.G = G.double().contiguous()
            self.G = 0.5 * (self.G + self.G.T)
            self.b = dense.gemm(self.Phi, self.y, transA=True)
            self.yy = torch.empty(1, **f64)
            self._dot(self.y, self.y, self.yy)
            self.A = torch.empty((self.k, self.k), **f64)
            self.w = torch.empty((self.k, self.R), **f64)
            self.Linv = torch.empty((self.k, self.k), **f64)
            self.eye = torch.eye(self.k, **f64)
            self.ainv = torch.empty(self.k, **f64)
        if self.form == "function":
            if precision != "mixed":
                self.K = torch.empty((self.M, self.M), **f64)
            self.a = torch.empty((self.M, self.R), **f64)
            self.Z = torch.empty_like(self.Phi)
            self.q = torch.empty(self.k, **f64)
            self.logdet = torch.empty(1, **f64)

    def _init_from_stats(self, stats, evals, n_rows, typ, form, precision, lower):
        G, b, yy, M = stats
        if form not in ("auto", "weight"):
            raise ValueError("weight-space statistics only serve the weight form")
        dev = G.device
        self.dev = dev
        self.k = G.shape[0]
        self.M = int(M)
        self.typ = 0 if typ == "matern" else 1
        self.n_rows = float(n_rows)
        self.precision = precision
        self.form = "weight"
        self.gram_mode = "stats"
        self.Phi = self.y = None
        f64 = dict(dtype=torch.float64, device=dev)
        self.G = G.to(torch.float64)
        self.G = (0.5 * (self.G + self.G.T)).contiguous()
        self.b = (b.reshape(-1, 1) if b.dim() == 1 else b).to(torch.float64).contiguous()
        self.R = self.b.shape[1]
        self.yy = yy.to(torch.float64).reshape(1).contiguous()
        self.evals = evals.to(torch.float64).contiguous()
        self.u = torch.zeros(4, **f64)
        self.lower = torch.as_tensor(np.asarray(lower, np.float64), device=dev)
        self.theta = torch.zeros(3, **f64)
        self.s2 = torch.zeros(1, **f64)
        self.S = torch.empty(self.k, **f64)
        self.dS = torch.empty(3 * self.k, **f64)
        self.gS = torch.empty(self.k, **f64)
        self.out = torch.zeros(9, **f64)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.fused_step = True
        self._cs_ws = torch.empty(
            _lib.lib().rvgp_colsumsq_ws_bytes(self.k, self.k) // 8 + 1, **f64)
        self.A = torch.empty((self.k, self.k), **f64)
        self.w = torch.empty((self.k, self.R), **f64)
        self.Linv = torch.empty((self.k, self.k), **f64)
        self.eye = torch.eye(self.k, **f64)
        self.ainv = torch.empty(self.k, **f64)

    # ---------------------------------------------------------------- params
    def set_theta(self, theta):
        th = np.asarray(theta, np.float64) - self.lower.cpu().numpy()
        self.u.copy_(torch.as_tensor(inv_softplus(th), device=self.dev))

    def theta_host(self):
        return softplus(self.u.cpu().numpy()) + self.lower.cpu().numpy()

    def _prologue(self):
        s = _lib.stream()
        _lib.call("rvgp_gp_params", _lib.ptr(self.u), _lib.ptr(self.lower), _lib.ptr(self.theta),
                  _lib.ptr(self.s2), s)
        _lib.call("rvgp_spectral_filter", self.typ, self.k, _lib.ptr(self.evals),
                  _lib.ptr(self.theta), self.n_rows, _lib.ptr(self.S), _lib.ptr(self.dS), s)

    def _dot(self, x, y, out):
        """out[0] = sum(x * y) over all entries (fixed order, FP64)."""
        assert x.is_contiguous() and y.is_contiguous() and x.numel() == y.numel()
        _lib.call("rvgp_dot", x.numel(), _lib.ptr(x), _lib.ptr(y), _lib.ptr(out),
                  _lib.ptr(self._dot_ws), self._dot_ws.numel() * 8, _lib.stream())
        return out

    def _colsumsq(self, Z, out):
        _lib.call("rvgp_colsumsq", int(Z.dtype == torch.float64), Z.shape[0], Z.shape[1],
                  _lib.ptr(Z), Z.stride(0), _lib.ptr(out), _lib.ptr(self._cs_ws),
                  self._cs_ws.numel() * 8, _lib.stream())

    # ---------------------------------------------------------------- evaluate
    def evaluate(self):
        """Enqueue NLML + gradient; results in self.out (device):
        [nlml, d/d(nu, kappa, sigma_f, s2), d/du(...)]."""
        self._prologue()
        s = _lib.stream()
        self.status.zero_()
        if self.form == "weight" or self.precision == "mixed":
            _lib.call("rvgp_ws_form", self.k, _lib.ptr(self.G), self.k, _lib.ptr(self.S),
                      _lib.ptr(self.s2), _lib.ptr(self.A), self.k, s)
            _lib.call("rvgp_potrf", 1, self.k, _lib.ptr(self.A), self.k, _lib.ptr(self.status), s)
            # L^-1 once (its column norms are diag(A^-1)); w = L^-T (L^-1 b) by two
            # GEMVs instead of two more blocked triangular solves
            self.Linv.copy_(self.eye)
            dense.trsm(self.A, self.Linv)
            self._colsumsq(self.Linv, self.ainv)
            t = dense.gemm(self.Linv, self.b)
            dense.gemm(self.Linv, t, self.w, transA=True)
        if self.form == "weight":
            Gw = dense.gemm(self.G, self.w)
            _lib.call("rvgp_ws_finish", self.k, float(self.M), self.R, _lib.ptr(self.A), self.k,
                      _lib.ptr(self.ainv), _lib.ptr(self.b), _lib.ptr(self.w), _lib.ptr(Gw),
                      _lib.ptr(self.yy), _lib.ptr(self.S), _lib.ptr(self.dS), _lib.ptr(self.s2),
                      _lib.ptr(self.u), _lib.ptr(self.gS), _lib.ptr(self.out), s)
        else:
            if self.precision == "mixed":
                # alpha = K_y^-1 y = (y - Phi w) / s2 (Woodbury; w = A^-1 Phi^T y
                # above), then two refinement sweeps against the FP64 K_y
                torch.div(self.y - dense.gemm(self.Phi, self.w), self.s2, out=self.a)
                for _ in range(2):
                    self.a += self._woodbury(self.y - self._apply_Ky(self.a))
                self._dot(self.y, self.a, self.ya)
            else:
                self._factor()
                # a = K_y^-1 y: L^-1 y first (ya = |L^-1 y|^2), then L^-T
                self.a.copy_(self.y)
                dense.trsm(self.K, self.a)
                self._dot(self.a, self.a, self.ya)
                dense.trsm(self.K, self.a, trans=True)
            self._dot(self.a, self.a, self.aa)
            if self.precision == "mixed":
                _lib.call("rvgp_lemma_terms", self.k, float(self.M), _lib.ptr(self.S),
                          _lib.ptr(self.s2), _lib.ptr(self.ainv), _lib.ptr(self.A), self.k,
                          _lib.ptr(self.q), _lib.ptr(self.logdet), s)
            else:
                self.Z.copy_(self.Phi)
                dense.trsm(self.K, self.Z)
                self._colsumsq(self.Z, self.q)
            p = dense.gemm(self.Phi, self.a, transA=True)  # (k, R)
            _lib.call("rvgp_fs_finish", self.k, float(self.M), self.R, _lib.ptr(self.q), _lib.ptr(p),
                      _lib.ptr(self.ya), _lib.ptr(self.aa), _lib.ptr(self.logdet), _lib.ptr(self.S),
                      _lib.ptr(self.dS), _lib.ptr(self.s2), _lib.ptr(self.u), _lib.ptr(self.gS),
                      _lib.ptr(self.out), s)
        return self.out

    # function-space helpers -------------------------------------------------
    def _factor(self):
        """K_y = Phi diag(S) Phi^T + s2 I and its FP64 Cholesky factor."""
        self.K.zero_()
        torch.diagonal(self.K).copy_(self.s2.expand(self.M))
        dense.gemm(self.Phi, self.Phi, self.K, transB=True, alpha=1.0, beta=1.0, kscale=self.S)
        _lib.call("rvgp_potrf", 1, self.M, _lib.ptr(self.K), self.M, _lib.ptr(self.status),
                  _lib.stream())
        self.logdet.copy_(dense.chol_logdet_half(self.K))

    def _apply_Ky(self, x):
        """K_y x = Phi (S * (Phi^T x)) + s2 x, matrix-free in FP64."""
        t = dense.gemm(self.Phi, x, transA=True)
        t = t * self.S.reshape(-1, 1)
        out = dense.gemm(self.Phi, t.contiguous())
        return out + self.s2 * x

    def _woodbury(self, r):
        """K_y^-1 r = (r - Phi A^-1 Phi^T r) / s2, A^-1 = Linv^T Linv: two
        skinny products with Phi and two k x k products, never an M x M
        matrix."""
        t = dense.gemm(self.Linv, dense.gemm(self.Phi, r, transA=True))
        u = dense.gemm(self.Linv, t, transA=True)
        return (r - dense.gemm(self.Phi, u)) / self.s2

    def nlml(self, x, unconstrained=False):
        """Differentiable NLML at x (see NLMLFunction)."""
        return NLMLFunction.apply(x, self, unconstrained)

    # ---------------------------------------------------------------- host API
    def nlml_grad_u(self, u):
        self.u.copy_(torch.as_tensor(np.asarray(u, np.float64), device=self.dev))
        out = self.evaluate().cpu().numpy()
        if int(self.status.item()) != 0:
            raise dense.CholeskyError("Cholesky decomposition was not successful: the GP system "
                                      "matrix is not positive definite")
        return float(out[0]), out[5:9].copy(), out

    def predict(self, Phi_star):
        """Mean (m*,) -- (m*, R) with R right-hand sides -- and marginal
        variance (m*,) of f at rows Phi_star, as GPflow's
        GPR.predict_f(full_cov=False) (the variance does not depend on y)."""
        self.evaluate()
        Ps = Phi_star.to(torch.float64).contiguous()
        m = Ps.shape[0]
        if self.form == "weight" or self.precision == "mixed":
            # mean = Phi* w, var = s2 |L_A^-1 phi*|^2 from the FP64 k x k system
            mean = dense.gemm(Ps, self.w)
            Zt = Ps.T.contiguous()
            dense.trsm(self.A, Zt)
            v = torch.empty(m, dtype=torch.float64, device=self.dev)
            cs_ws = torch.empty(_lib.lib().rvgp_colsumsq_ws_bytes(self.k, m) // 8 + 1,
                                dtype=torch.float64, device=self.dev)
            _lib.call("rvgp_colsumsq", 1, self.k, m, _lib.ptr(Zt), m, _lib.ptr(v), _lib.ptr(cs_ws),
                      cs_ws.numel() * 8, _lib.stream())
            var = self.s2 * v
        else:
            p = dense.gemm(self.Phi, self.a, transA=True)
            mean = dense.gemm(Ps, (p * self.S.reshape(-1, 1)).contiguous())
            Kms = dense.gemm(self.Phi, Ps, transB=True, kscale=self.S)  # (M, m*)
            dense.trsm(self.K, Kms)
            cs_ws = torch.empty(_lib.lib().rvgp_colsumsq_ws_bytes(max(self.M, self.k), m) // 8 + 1,
                                dtype=torch.float64, device=self.dev)
            v = torch.empty(m, dtype=torch.float64, device=self.dev)
            _lib.call("rvgp_colsumsq", 1, self.M, m, _lib.ptr(Kms), m, _lib.ptr(v), _lib.ptr(cs_ws),
                      cs_ws.numel() * 8, _lib.stream())
            Zs = (Ps.T * torch.sqrt(self.S).reshape(-1, 1)).contiguous()
            d = torch.empty(m, dtype=torch.float64, device=self.dev)
            _lib.call("rvgp_colsumsq", 1, self.k, m, _lib.ptr(Zs), m, _lib.ptr(d), _lib.ptr(cs_ws),
                      cs_ws.numel() * 8, _lib.stream())
            var = d - v
        return (mean.reshape(-1) if self.R == 1 else mean), var

    def predict_nodes(self, U_star, T_star):
        """predict() at the ambient rows phi_ip = T_i[p] U_i of m nodes, from
        their intrinsic rows U_star (m*d, k) and frames T_star (m, D, d):
        mean_i = T_i (U_i w) and var_ip = s2 T_i[p] B_i T_i[p]^T with
        B_i = (L_A^-1 U_i^T)^T (L_A^-1 U_i^T) (d x d), so the triangular solve
        runs on m*d columns instead of m*D.  Weight-space system only."""
        if not (self.form == "weight" or self.precision == "mixed"):
            raise ValueError("predict_nodes needs the k x k weight-space system")
        self.evaluate()
        m, D, d = T_star.shape
        T = T_star.to(torch.float64).contiguous()
        mu = dense.gemm(U_star, self.w)                                  # (m*d, R)
        # mean_i = T_i mu_i for every right-hand side (frames applied per node)
        c = mu.reshape(m, d, self.R).permute(2, 0, 1).contiguous()      # (R, m, d)
        out = torch.empty((self.R, m, D), dtype=torch.float64, device=T.device)
        _lib.call("rvgp_frame_apply", 1, self.R, m, D, d, _lib.ptr(T), d, _lib.ptr(c),
                  _lib.ptr(out), _lib.stream())
        mean = out.permute(1, 2, 0)                                      # (m, D, R)
        Zt = U_star.T.contiguous()
        dense.trsm(self.A, Zt)                                           # L_A^-1 U^T
        # B_i = Z_i Z_i^T (d x d) from the columns of Zt that belong to node i
        Z3 = Zt.reshape(self.k, m, d)
        B = torch.stack([torch.stack([(Z3[:, :, a] * Z3[:, :, b]).sum(0) for b in range(d)], -1)
                         for a in range(d)], -2)                          # (m, d, d)
        var = self.s2 * (T.unsqueeze(-1) * B.unsqueeze(1) * T.unsqueeze(-2)).sum((-1, -2))
        var = var.reshape(-1)
        mean = mean.reshape(m * D, self.R)
        return (mean.reshape(-1) if self.R == 1 else mean), var

    # ---------------------------------------------------------------- optimisers
    def fit_lbfgs(self, maxiter=1000, disp=False):
        import scipy.optimize
        u0 = self.u.cpu().numpy()

        def f(u):
            val, g, _ = self.nlml_grad_u(u)
            return val, g

        res = scipy.optimize.minimize(f, u0, jac=True, method="L-BFGS-B",
                                      options={"disp": disp, "maxiter": maxiter})
        self.u.copy_(torch.as_tensor(res.x, device=self.dev))
        self.last_nfev = int(res.nfev)
        return res

    def fused(self):
        """Whether a step runs as the single-launch tile dataflow
        (rvgp_ws_step): weight form, one right-hand side, k <= 2048."""
        return (self.fused_step and self.form == "weight" and self.R == 1
                and self.k <= FUSED_STEP_MAX_K)

    def _ws_step(self, do_adam, lr=0.0, b1=0.9, b2=0.999, eps=1e-8):
        if getattr(self, "_fws", None) is None:
            L = _lib.lib()
            self._fws = torch.empty(L.rvgp_ws_step_ws_bytes(self.k) // 8, dtype=torch.float64,
                                    device=self.dev)
            # flags of the dataflow: zeroed once, then advanced by every call
            self._fsync = torch.zeros(L.rvgp_ws_step_sync_ints(self.k), dtype=torch.int32,
                                      device=self.dev)
        args = (_lib.ptr(self.m1), _lib.ptr(self.m2), _lib.ptr(self.t),
                _lib.ptr(self.trainable)) if do_adam else (None, None, None, None)
        m1, m2, t, tr = args
        _lib.call("rvgp_ws_step", self.k, self.typ, _lib.ptr(self.evals), self.n_rows,
                  _lib.ptr(self.G), self.G.stride(0), _lib.ptr(self.b), _lib.ptr(self.yy),
                  float(self.M), _lib.ptr(self.lower), _lib.ptr(self.u), tr, m1, m2, t,
                  int(do_adam), float(lr), float(b1), float(b2), float(eps), _lib.ptr(self._fws),
                  self._fws.numel() * 8, _lib.ptr(self._fsync), self._fsync.numel(),
                  _lib.ptr(self.out), _lib.ptr(self.status), _lib.stream())
        return self.out

    def adam_setup(self, trainable=(1, 1, 1, 1)):
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.m1 = torch.zeros(4, **f64)
        self.m2 = torch.zeros(4, **f64)
        self.t = torch.zeros(1, **f64)
        self.trainable = torch.as_tensor(np.asarray(trainable, np.int32), device=self.dev)
        if self.fused():
            self._ws_step(0)  # allocates the dataflow's workspace before any capture

    def adam_step(self, lr=0.05, b1=0.9, b2=0.999, eps=1e-8):
        """One device-resident optimiser step: NLML, adjoint, Adam update.
        No host synchronisation, so it can be captured in a HIP graph.  In the
        weight form with one right-hand side the whole step is one launch."""
        if self.fused():
            self._ws_step(1, lr, b1, b2, eps)
            return
        self.evaluate()
        _lib.call("rvgp_adam_step", _lib.ptr(self.u), _lib.ptr(self.out[5:]), _lib.ptr(self.m1),
                  _lib.ptr(self.m2), _lib.ptr(self.t), _lib.ptr(self.trainable), float(lr),
                  float(b1), float(b2), float(eps), _lib.stream())

    def capture_adam_step(self, lr=0.05, b1=0.9, b2=0.999, eps=1e-8):
        """Capture one adam_step into a HIP graph (torch.cuda.CUDAGraph).
        Every buffer the step touches is allocated beforehand by an eager
        warm-up step, whose effect on the optimiser state is undone."""
        from . import instrument
        saved = [x.clone() for x in (self.u, self.m1, self.m2, self.t)]
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.adam_step(lr, b1, b2, eps)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        for x, s in zip((self.u, self.m1, self.m2, self.t), saved):
            x.copy_(s)
        g = torch.cuda.CUDAGraph()
        prev = instrument._ACTIVE
        instrument._ACTIVE = None  # no timing events inside the capture
        try:
            with torch.cuda.graph(g):
                self.adam_step(lr, b1, b2, eps)
        finally:
            instrument._ACTIVE = prev
        torch.cuda.synchronize()
        return g
