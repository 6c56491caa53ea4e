#!/usr/bin/env python3
"""Headline benchmark: full RVGP fit on the 3-torus in R^32 (n = 100,000,
k = 20, dim_man = 3, 500 eigenpairs, 30% of nodes train), mixed precision +
adjoint vs the plain FP64 GPU form, on one MI355X.

A "step" is one full fit: kNN graph, heap neighbourhoods + gauges,
connections, block-CSR Laplacians, both eigensolves (create_data_object),
then RVGP.fit with 200 Adam steps on the NLML through the hand-written
adjoint.  Points and the target tangent field are synthetic (seeded) and
resident in HBM before the timed region.

Prints ONE JSON line:
  value          wall time of the mixed-precision fit (s, lower is better)
  fp64_s         wall time of the FP64 GPU fit on the same input
  rel_err_pred   |f_mixed - f_fp64| / |f_fp64| over test-node predictions
  roofline       dominant kernel (the block SpMM of the Lc eigensolve -- the
                 FP64, FP32 or binary16 form with the largest summed time
                 in the timed fit): algorithmic bytes / HIP-event-timed
                 duration vs 8 TB/s
  cpu_baseline   the FP64 CPU oracle (oracle/) on a bounded sample of the
                 same workload, with the GPU time on that sample beside it
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fit wall-time @n=1e5,500 eigs: mixed-prec+adjoint vs FP64 GPU at <=1e-4 rel err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------- workload
def flat_torus3(n, D=32, seed=0, noise=0.01):
    """Angles a ~ U[0, 2pi)^3, E = [cos a, sin a] in R^6, mapped into R^D by the
    first 6 rows of a seeded orthogonal Q, plus isotropic N(0, noise^2)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 2 * np.pi, (n, 3))
    E = np.concatenate([np.cos(a), np.sin(a)], 1)
    Qm, Rq = np.linalg.qr(np.random.default_rng(seed + 1000).normal(size=(D, D)))
    Qm = Qm * np.sign(np.diag(Rq))
    X = E @ Qm[:6] + noise * rng.normal(size=(n, D))
    return X, a, Qm[:6]


def torus_field(a, Q6):
    """A smooth tangent field sum_i f_i(a) dE/da_i pushed into R^D."""
    import numpy as np
    f = np.stack([np.sin(a[:, 1]) + 0.5 * np.cos(a[:, 2]), np.cos(a[:, 0] + a[:, 2]),
                  np.sin(a[:, 0]) * np.cos(a[:, 1])], 1)
    v6 = np.zeros((len(a), 6))
    for i in range(3):
        v6[:, i] += -np.sin(a[:, i]) * f[:, i]
        v6[:, 3 + i] += np.cos(a[:, i]) * f[:, i]
    return v6 @ Q6


def run_fit(X, V, train, precision, args, timings):
    import torch
    import rvgp_amd as RVGP
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d = RVGP.create_data_object(X, vectors=V, n_neighbors=args.k_nn, n_eigenpairs=args.k_eig,
                                precision=precision, reference_semantics=False, verbose=False,
                                timings=timings, eig_verbose=args.verbose,
                                eig_tol=args.eig_tol_mixed if precision == "mixed" else None)
    t1 = time.perf_counter()
    gp = RVGP.fit(d, train_ind=train, epochs=args.adam_steps, optimizer="adam", lr=0.05,
                  precision=precision, verbose=False, timings=timings)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    timings["data_object"] = t1 - t0
    timings["gp_fit"] = t2 - t1
    return d, gp, t2 - t0


def kernel_stats(rec):
    out = {}
    for (kind, key), s in rec.summary().items():
        name = f"{kind}[{key[0]},b{key[1]}]" if isinstance(key, tuple) else f"{kind}[{key}]"
        out[name] = {"count": s["count"], "avg_us": 1e3 * s["ms"] / max(1, s["count"]),
                     "total_ms": s["ms"], "GBps": s["bytes"] / (s["ms"] * 1e6) if s["ms"] else 0.0,
                     "TFLOPs": s["flops"] / (s["ms"] * 1e9) if s["ms"] else 0.0}
    return out


SPMM_SHORT = {"float64": "f64", "float32": "f32", "float16": "f16"}


def traffic_from_profile(dtype="float32", b=3):
    """HBM bytes per launch of the block SpMM of this dtype from the
    committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by a
    plain copy kernel), or None when no such summary is present."""
    path = os.path.join(ROOT, "profiles", f"spmm_{SPMM_SHORT[dtype]}_b{b}_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


# ---------------------------------------------------------------- CPU baseline
STAGE_CAP_S = 1800.0  # per-stage wall-time cap of the CPU baseline


def _full_size_markers(n, k_nn, D=32, dim_man=3):
    """Stages of the reference's CPU path that cannot run at the full
    workload, with the reason (sizes in bytes of FP64)."""
    K = int(k_nn * 1.5)
    N = n * dim_man
    M = int(0.3 * n * 0.8) * D
    return {
        "gauges": (f"reference raises ValueError at K={K} < D={D} "
                   f"(RVGP/lib/ptu_dijkstra.pyx:73-77); tangent_frames and connections "
                   f"each allocate an unused {n}x{n} distance buffer "
                   f"({8 * n * n / 1e9:.0f} GB each, pyx:93 and pyx:171); the oracle runs "
                   f"the min(D, K+1) extension"),
        "smoothing": (f"dense scipy.linalg.expm of the {N}x{N} connection Laplacian "
                      f"(RVGP/smoothing.py:13-15) needs {8 * N * N / 1e9:.0f} GB per matrix"),
        "nlml_grad_step": (f"dense GPR Gram of the {M} training rows "
                           f"(RVGP/kernels.py:55-61) needs {8 * M * M / 1e12:.1f} TB"),
    }


def cpu_baseline(args):
    """The reference's CPU path restated by the FP64 oracle (oracle/cpu.py: C
    kNN + Fibonacci-heap neighbourhoods, LAPACK SVD gauges and Procrustes
    connections, scipy Laplacians, eigsh(which='SM'), dense expm, numpy
    Cholesky GPR), timed stage by stage on the same generator at
    n = args.cpu_n.  Stages the full workload cannot run on a CPU are marked
    with the reason.  The fit total counts the stages the GPU fit runs: graph,
    gauges, connections, Laplacians, both eigensolves, and the NLML + gradient
    of one Adam step times the number of steps."""
    import numpy as np
    from threadpoolctl import threadpool_info, threadpool_limits
    from oracle import cpu
    # BLAS pools sized to the CPU share this job runs with, not to the host's
    # core count (an oversubscribed pool is many times slower)
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threadpool_limits(limits=min(16, share))
    cores = max([t.get("num_threads", 1) for t in threadpool_info()] + [1])
    n, k = args.cpu_n, args.cpu_k_eig
    X, a, Q6 = flat_torus3(n, seed=0)
    V = torus_field(a, Q6)
    train = np.random.default_rng(3).permutation(n)[: int(0.3 * n)]
    stages = {}

    def stage(name, fn):
        t0 = time.perf_counter()
        out = fn()
        stages[name] = round(time.perf_counter() - t0, 4)
        log(f"[bench] CPU oracle {name}: {stages[name]:.2f}s")
        return out

    D = X.shape[1]
    K = int(args.k_nn * 1.5)
    rowptr, cols = stage("knn", lambda: cpu.sym_csr(cpu.knn(X, args.k_nn)))
    gauges, Sigma = stage("gauges", lambda: cpu.tangent_frames(X, rowptr, cols, D, K,
                                                               allow_rank_deficient_k=True))
    dim_man = cpu.manifold_dimension(Sigma.copy(), 0.8)
    gauges = np.ascontiguousarray(gauges[:, :, :dim_man])
    R = stage("connections", lambda: cpu.connections(gauges, rowptr, cols))
    L, Lc = stage("laplacians", lambda: (cpu.laplacian(rowptr, cols),
                                         cpu.connection_laplacian(rowptr, cols, R)))

    def spectra():
        wl, _ = cpu.compute_spectrum(L, k)
        w, U = cpu.compute_spectrum(Lc, k)
        return w, np.einsum("bij,bjk->bik", gauges, U.reshape(-1, dim_man, k)).reshape(-1, k)
    evals, Phi_all = stage("eigensolve", spectra)
    c = cpu.express_in_local_frame(V, gauges)
    stage("smoothing", lambda: cpu.vector_diffusion(c, 100.0, Lc, L))
    tr, te = cpu.train_test_split_indices(len(train), 0.2, 0)
    nodes = train[tr]
    Phi = Phi_all.reshape(n, -1)[nodes].reshape(-1, k)
    y = V[nodes].reshape(-1, 1)
    evals_n = 1

    def step():
        for _ in range(evals_n):
            S, _ = cpu.spectral_filter(evals, 1.5, 5.0, 1.0, n * D)
            cpu.gpr_nlml(Phi, y, S, 1.0)
            cpu.gpr_grad(Phi, y, S, 1.0)
    stage("nlml_grad_step", step)
    stages["nlml_grad_step"] = round(stages["nlml_grad_step"] / evals_n, 4)
    test = np.setdiff1d(np.arange(n), train)
    Ps = Phi_all.reshape(n, -1)[test].reshape(-1, k)
    S, _ = cpu.spectral_filter(evals, 1.5, 5.0, 1.0, n * D)
    stage("transform", lambda: cpu.gpr_predict(Phi, y, S, 1.0, Ps))
    fit_stages = ("knn", "gauges", "connections", "laplacians", "eigensolve")
    total = sum(stages[s] for s in fit_stages) + args.adam_steps * stages["nlml_grad_step"]
    return {"value": round(total, 3), "unit": "s per fit", "cores": int(cores), "kind": "port",
            "sample": (f"oracle on the same generator at n={n} (k={args.k_nn}, {k} eigenpairs, "
                       f"M={Phi.shape[0]} GP rows); fit = graph + gauges + connections + "
                       f"Laplacians + eigensolves + {args.adam_steps} x one NLML+gradient step "
                       f"(timed once); smoothing and transform timed beside it; "
                       "FPS is not part of this workload (its points are used as given)"),
            "stages_s": stages,
            "stage_cap_s": STAGE_CAP_S,
            "infeasible_at_full_size": _full_size_markers(args.n, args.k_nn)}


def gp_step_utilisation(engine, reps=20):
    """The optimiser step's k x k Cholesky + triangular inverse (2k^3/3 flops,
    the dominant arithmetic of a step) against the FP64 matrix-core peak:
    `reps` evaluations of the single-launch step timed with HIP events on
    the launch stream (the hyperparameters do not move)."""
    import torch
    if not engine.fused():
        return None
    engine.adam_setup()
    engine._ws_step(0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        engine._ws_step(0)
    e1.record()
    e1.synchronize()
    us = 1e3 * e0.elapsed_time(e1) / reps
    k = engine.k
    tf = 2.0 * k ** 3 / 3.0 / (us * 1e-6) / 1e12
    return {"kernel": "ws_step_kernel", "k": k, "avg_us": round(us, 1), "achieved": round(tf, 3),
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / FP64_PEAK_TFLOPS, 5),
            "bound": "latency: a chain of ceil(k/32) diagonal 32x32 tile factorisations"}


def gpu_same_sample(args):
    """The mixed GPU fit on the CPU baseline's sample (warm), with its stages."""
    import numpy as np
    n = args.cpu_n
    X, a, Q6 = flat_torus3(n, seed=0)
    V = torus_field(a, Q6)
    train = np.random.default_rng(3).permutation(n)[: int(0.3 * n)]
    a2 = argparse.Namespace(**vars(args))
    a2.k_eig = args.cpu_k_eig
    run_fit(X, V, train, "mixed", a2, {})  # warm
    st = {}
    _, _, t = run_fit(X, V, train, "mixed", a2, st)
    return t, {k: round(v, 4) for k, v in st.items()}


def roofline_only(args):
    """Launch the dominant kernel alone at the headline shape: the block SpMM
    (b = 3, buffer form; FP64, FP32 or binary16 storage by
    --roofline-dtype) of the Lc eigensolve on the Morton-renumbered operator,
    with the Chebyshev-step epilogue (Z read) at the starting block width
    r = k + guard rounded up to a multiple of 8, then a plain 1 GiB device
    copy (the reference for scaling PMC byte counters).  Used under
    rocprofv3 --pmc; prints one JSON line."""
    import torch
    from rvgp_amd import _lib, geometry
    from rvgp_amd.ptu_dijkstra import connections_device, tangent_frames_device
    from rvgp_amd.reorder import morton_perm, permute_bsr
    from rvgp_amd.sparse import BSR, spmm, spmm_bytes
    _lib.lib()
    X, _, _ = flat_torus3(args.n, seed=0)
    Xd = torch.as_tensor(X, device="cuda")
    G = geometry.manifold_graph(Xd, n_neighbors=args.k_nn)
    K = int(args.k_nn * 1.5)
    D = X.shape[1]
    T, _ = tangent_frames_device(Xd, G, K, min(D, K + 1), min(D, K + 1), min(D, K))
    c = connections_device(T[:, :, :3].contiguous(), G, 3)
    A = permute_bsr(BSR(G.rowptr, G.cols, c["lc64"].reshape(-1), c["lc32"].reshape(-1), 3),
                    morton_perm(Xd))
    r = -(-(args.k_eig + max(8, int(math.ceil(0.15 * args.k_eig)))) // 8) * 8
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    dt = getattr(torch, args.roofline_dtype)
    Xm = torch.randn((A.shape[0], r), generator=gen, device="cuda",
                     dtype=torch.float32).to(dt)
    Z = torch.randn_like(Xm)
    for _ in range(2):
        spmm(A, Xm, Y=Z, alpha=0.5, beta=0.1, Z=Z, gamma=-0.3)
    torch.cuda.synchronize()
    reps = 10
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        spmm(A, Xm, Y=Z, alpha=0.5, beta=0.1, Z=Z, gamma=-0.3)
    e1.record()
    torch.cuda.synchronize()
    us = 1e3 * e0.elapsed_time(e1) / reps
    s = Xm.element_size()
    # binary16 blocks with f32 block values: the values count 4 bytes
    nbytes = spmm_bytes(A, r, s, with_z=True) + (A.nnzb * A.b * A.b * 2 if s == 2 else 0)
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel: reads 1 GiB, writes 1 GiB
    torch.cuda.synchronize()
    kname = {8: "spmm_buf<double,3,2>", 4: "spmm_buf<float,3,4>", 2: "spmm_buf_h<3>"}[s]
    print(json.dumps({"kernel": kname, "r": r, "nb": A.nb, "nnzb": A.nnzb,
                      "launches": reps + 2, "avg_us": us, "alg_bytes": nbytes,
                      "GBps": nbytes / (us * 1e3), "copy_bytes_each_way": 1 << 30}))


# ---------------------------------------------------------------- other configs
def _timed(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0


def bench_batched(args):
    """Batched fits on one shared 3-torus manifold: n = 50,000 (the 3-torus
    in R^32 generator), dim_man = 3, 300 eigenpairs; 256 Gaussian tangent
    fields smoothed with diffusion times log-uniform in [10, 500]; each fit
    has its own training subset of uniform size in [100, 5000] nodes and its
    own hyperparameters; 200 Adam steps, every step of all 256 fits one
    graph-replayed launch.  Timed: the batched fit (statistics + 200 steps)
    on the resident data object and fields; beside it, the same fits run one
    after another through the single-fit engine (a sample of 8 fits, scaled
    to 256)."""
    import numpy as np
    import torch
    import rvgp_amd as RVGP
    from rvgp_amd import _lib
    from rvgp_amd.batch import BatchedGPR
    from rvgp_amd.gp import GPREngine
    _lib.lib()
    dev = _lib.device()
    n, F, k_eig, steps = args.batch_n, args.batch_fields, 300, args.adam_steps
    X, _, _ = flat_torus3(n, seed=0)
    d = RVGP.create_data_object(torch.as_tensor(X, device=dev), n_neighbors=args.k_nn,
                                n_eigenpairs=k_eig, precision="mixed", reference_semantics=False,
                                verbose=False)
    rng = np.random.default_rng(7)
    ts = np.exp(rng.uniform(np.log(10.0), np.log(500.0), F))
    V = torch.as_tensor(rng.normal(size=(F, n, X.shape[1])), device=dev)
    d.vectors = d._frame(2, V, X.shape[1])
    del V
    (_, t_smooth) = _timed(lambda: d.smooth_vector_field(ts))
    log(f"[bench batched] n={n} dim_man={d.dim_man} fields={F}: smoothing {t_smooth:.2f}s")
    sizes = rng.integers(100, 5001, F)
    sets = [rng.choice(n, size=int(m), replace=False) for m in sizes]
    th0 = np.stack([rng.uniform(1.0, 2.5, F), rng.uniform(3.0, 8.0, F), np.ones(F),
                    np.full(F, 0.1)], 1)
    Vd = d.vectors_device()
    RVGP.fit_batched(d, sets, vectors=Vd, epochs=2, theta0=th0)  # warm-up (and autotune)
    gp, t_fit = _timed(lambda: RVGP.fit_batched(d, sets, vectors=Vd, epochs=steps, theta0=th0))
    gp.check_status()
    # per-step launch time over graph replays (HIP events on the launch stream)
    g = gp.capture_adam_step(0.05)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(50):
        g.replay()
    e1.record()
    e1.synchronize()
    step_ms = e0.elapsed_time(e1) / 50
    e0.record()
    gp2 = BatchedGPR(d, Vd, sets, theta0=th0, gram_tile=gp.gram_tile,
                     panel=(gp.panel, gp.threads))
    e1.record()
    e1.synchronize()
    setup_ms = e0.elapsed_time(e1)
    del gp2
    k = gp.k
    rows = int(sizes.sum()) * d.dim_man
    step_flops = F * 2.0 * k ** 3 / 3.0  # Cholesky + triangular inverse per fit
    gram_flops = rows * float(k) * k     # symmetric Gram over the gathered intrinsic rows
    # sequential single fits through the single-fit weight-space engine
    samp = list(range(0, F, F // 8))[:8]
    lower = gp.lower.cpu().numpy()

    def seq():
        for f in samp:
            nodes = torch.as_tensor(sets[f].astype(np.int32), device=dev)
            Phi = d.phi_rows(nodes)
            y = Vd[f, nodes.long()].reshape(-1)
            e = GPREngine(Phi, y, d.spectrum_device(), d.n * X.shape[1], form="weight",
                          precision="fp64", lower=lower)
            e.set_theta(th0[f])
            e.adam_setup()
            gg = e.capture_adam_step(lr=0.05)
            for _ in range(steps):
                gg.replay()
    _, t_seq = _timed(seq)
    t_seq_all = t_seq * F / len(samp)
    out = {
        "metric": "batched fit wall-time: 256 fields x 200 adjoint Adam steps on one manifold",
        "value": round(t_fit, 4), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic (seeded flat 3-torus in R^32, smoothed Gaussian fields)",
        "config": {"workload": (f"3-torus in R^32, n={n}, k={args.k_nn}, dim_man={d.dim_man}, "
                                f"{k_eig} eigenpairs, {F} fields, subsets 100-5000 nodes, "
                                f"{steps} Adam steps"), "fields": F, "n": n},
        "fits_steps_per_s": round(F * steps / t_fit, 1),
        "step_ms": round(step_ms, 4), "setup_ms": round(setup_ms, 3),
        "step_fp64_tflops": round(step_flops / (step_ms * 1e9), 3),
        "gram_rows": rows, "gram_fp64_tflops_incl_rhs": round(gram_flops / (setup_ms * 1e9), 3),
        "autotune": {"gram_tile": gp.gram_tile, "panel": gp.panel, "threads": gp.threads},
        "sequential_single_fit_s": round(t_seq_all, 3),
        "speedup_vs_sequential": round(t_seq_all / t_fit, 2),
        "smoothing_s": round(t_smooth, 3),
    }
    print(json.dumps(out))


def bench_multirhs(args):
    """Scalp-like multi-RHS: n = 5,000 points by FPS from 50,000 on the upper
    half of the ellipsoid (1, 0.9, 0.8), k = 12, dim_man = 2, 150
    eigenpairs; 2,000 fields = 3 smoothed Gaussian tangent modes (t = 20)
    with AR(1) time courses (rho = 0.95, unit stationary variance), observed
    at a fixed set of 64 nodes: one GP with 2,000 right-hand sides and shared
    hyperparameters (summed NLML), 200 Adam steps, then prediction at every
    node.  Timed: fit + prediction on the resident data object."""
    import numpy as np
    import torch
    import rvgp_amd as RVGP
    from rvgp_amd import _lib, geometry
    _lib.lib()
    dev = _lib.device()
    rng = np.random.default_rng(0)
    P = rng.normal(size=(50000, 3))
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    P[:, 2] = np.abs(P[:, 2])
    P *= np.array([1.0, 0.9, 0.8])
    perm, _ = geometry.furthest_point_sampling(P, N=5000)
    X = P[perm]
    d = RVGP.create_data_object(X, n_neighbors=12, dim_man=2, n_eigenpairs=150,
                                precision=args.precision_multirhs, verbose=False)
    n, D = X.shape
    modes = d._frame(2, torch.as_tensor(rng.normal(size=(3, n, D)), device=dev), D)
    d.vectors = modes
    d.smooth_vector_field(20.0)
    modes = d.vectors_device()
    T, rho = args.multirhs_T, 0.95
    a = np.zeros((T, 3))
    a[0] = rng.normal(size=3)
    for t in range(1, T):
        a[t] = rho * a[t - 1] + np.sqrt(1 - rho ** 2) * rng.normal(size=3)
    fields = torch.einsum("tm,mnd->tnd", torch.as_tensor(a, device=dev), modes).contiguous()
    d.vectors = fields
    obs = rng.choice(n, size=64, replace=False)
    test = [int(i) for i in range(n)]

    def run():
        gp = RVGP.fit(d, train_ind=obs, epochs=args.adam_steps, optimizer="adam", lr=0.05,
                      reference_semantics=False, precision=args.precision_multirhs, verbose=False)
        mean, var = gp.transform(d, test)
        return gp, mean, var
    run()  # warm-up
    (gp, mean, var), t = _timed(run)
    err = float(np.linalg.norm(mean[:, obs] - fields[:, obs].cpu().numpy())
                / np.linalg.norm(fields[:, obs].cpu().numpy()))
    out = {
        "metric": "multi-RHS fit + predict wall-time: 2000 fields at 64 observed nodes",
        "value": round(t, 4), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64" if args.precision_multirhs == "fp64" else "f32+f64",
        "data": "synthetic (FPS half-ellipsoid, 3 smoothed modes x AR(1) time courses)",
        "config": {"workload": (f"half-ellipsoid n={n}, k=12, dim_man={d.dim_man}, 150 eigenpairs, "
                                f"{T} right-hand sides at 64 nodes, {args.adam_steps} Adam steps, "
                                f"prediction at all {n} nodes")},
        "fit_form": gp.engine.form, "rhs": gp.engine.R, "observations": gp.engine.M,
        "hyperparameters": {k: round(float(v), 5) for k, v in gp.hyperparameters.items()},
        "rel_err_at_observed_nodes": err,
    }
    print(json.dumps(out))


def bench_sphere(args):
    """Plain method on S^2, FP64: n = 2,000 points uniform on the unit sphere
    (seed 0), k = 10, dim_man = 2, 50 eigenpairs; standard-normal ambient
    vectors projected onto the tangent planes and smoothed with t = 100;
    train on a random 50 % (the reference's 80/20 re-split applies), noise
    variance 1e-3, Matern defaults, L-BFGS-B; predict the other 50 %.  Timed:
    the whole pipeline on the GPU (data object, field, smoothing, fit,
    transform), warm.  Beside it the oracle: data object and dense-expm
    smoothing timed in full, NLML + gradient timed over 3 evaluations and
    scaled to the GPU fit's L-BFGS-B evaluation count."""
    import numpy as np
    import torch
    import rvgp_amd as RVGP
    from rvgp_amd import _lib
    from oracle import cpu
    _lib.lib()
    dev = _lib.device()
    rng = np.random.default_rng(0)
    X = rng.normal(size=(2000, 3))
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    v0 = rng.normal(size=(2000, 3))
    train = rng.permutation(2000)[:1000]
    test = [int(i) for i in np.setdiff1d(np.arange(2000), train)]

    def run():
        d = RVGP.create_data_object(X, n_neighbors=10, dim_man=2, n_eigenpairs=50, verbose=False)
        d.vectors = d._frame(2, torch.as_tensor(v0, device=dev), 3)
        d.smooth_vector_field(t=100)
        gp = RVGP.fit(d, train_ind=train, noise_variance=1e-3, verbose=False)
        mean, var = gp.transform(d, test)
        return d, gp, mean, var
    run()
    (d, gp, mean, var), t_gpu = _timed(run)
    nfev = int(getattr(gp.engine, "last_nfev", 0)) or None
    # oracle on the same input; prediction parity at the GPU's hyperparameters
    t0 = time.perf_counter()
    o = cpu.Data(X, n_neighbors=10, n_eigenpairs=50)
    o.vectors = cpu.project_to_manifold(v0, o.gauges)
    o.smooth_vector_field(t=100)
    t_data = time.perf_counter() - t0
    tr, _ = cpu.train_test_split_indices(len(train), 0.2, 0)
    nodes = train[tr]
    k = o.evecs_Lc.shape[1]
    Phi = o.evecs_Lc.reshape(2000, -1)[nodes].reshape(-1, k)
    y = o.vectors[nodes].reshape(-1, 1)
    th = gp.engine.theta_host()
    S, dS = cpu.spectral_filter(o.evals_Lc, th[0], th[1], th[2], 2000 * 3)
    t0 = time.perf_counter()
    for _ in range(3):
        cpu.gpr_nlml(Phi, y, S, th[3])
        cpu.gpr_grad(Phi, y, S, th[3])
    t_eval = (time.perf_counter() - t0) / 3
    Ps = o.evecs_Lc.reshape(2000, -1)[np.asarray(test)].reshape(-1, k)
    mref, _ = cpu.gpr_predict(Phi, y, S, th[3], Ps)
    rel = float(np.linalg.norm(mean.ravel() - mref.ravel()) / np.linalg.norm(mref))
    evals_rel = float(np.max(np.abs(d.evals_Lc - o.evals_Lc) / np.maximum(np.abs(o.evals_Lc), 1e-9)))
    out = {
        "metric": "full pipeline wall-time, S^2 n=2000, 50 eigenpairs, FP64 (drop-in defaults)",
        "value": round(t_gpu, 4), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic (seeded uniform S^2, projected normal field)",
        "config": {"workload": "S^2 n=2000, k=10, dim_man=2, 50 eigenpairs, t=100, 50% train, "
                               "L-BFGS-B (maxiter 1000), predict 50%"},
        "lbfgs_evaluations": nfev,
        "pred_rel_err_vs_oracle": rel, "evals_Lc_rel_err_vs_oracle": evals_rel,
        "cpu_baseline": {"value": round(t_data + (nfev or 1) * t_eval, 3), "unit": "s",
                         "kind": "port", "cores": None,
                         "sample": (f"oracle data object + dense-expm smoothing {t_data:.2f}s in full; "
                                    f"NLML+gradient {t_eval:.3f}s per evaluation x "
                                    f"{nfev} L-BFGS-B evaluations")},
    }
    print(json.dumps(out))


def bench_torus(args):
    """Torus of radii R = 2, r = 0.7: 20,000 points from uniform (theta, phi),
    FPS with stop_crit 0.015 (~3,956 points), k = 15, dim_man = 2, 200
    eigenpairs; a Gaussian tangent field smoothed with t = 50 plus
    N(0, 0.05^2) ambient noise; 50 % of the nodes train; 200 Adam steps on
    the NLML through the adjoint.  Timed (warm): FPS, data object, field
    smoothing and fit, in the mixed form and in the FP64 form."""
    import numpy as np
    import torch
    import rvgp_amd as RVGP
    from rvgp_amd import _lib, geometry
    _lib.lib()
    dev = _lib.device()
    rng = np.random.default_rng(0)
    th, ph = rng.uniform(0, 2 * np.pi, (2, 20000))
    P = np.stack([(2 + 0.7 * np.cos(ph)) * np.cos(th), (2 + 0.7 * np.cos(ph)) * np.sin(th),
                  0.7 * np.sin(ph)], 1)
    g0 = rng.normal(size=(20000, 3))
    noise = 0.05 * rng.normal(size=(20000, 3))
    res = {}

    def run(prec):
        perm, _ = geometry.furthest_point_sampling(P, spacing=0.015)
        X = P[perm]
        d = RVGP.create_data_object(X, n_neighbors=15, dim_man=2, n_eigenpairs=200,
                                    precision=prec, verbose=False)
        n = len(perm)
        d.vectors = d._frame(2, torch.as_tensor(g0[:n], device=dev), 3)
        d.smooth_vector_field(t=50)
        d.vectors = d.vectors_device() + torch.as_tensor(noise[:n], device=dev)
        train = np.random.default_rng(3).permutation(n)[: n // 2]
        gp = RVGP.fit(d, train_ind=train, epochs=200, optimizer="adam", lr=0.05, precision=prec,
                      reference_semantics=False, verbose=False)
        test = [int(i) for i in np.setdiff1d(np.arange(n), train)]
        mean, _ = gp.transform(d, test)
        return n, mean
    for prec in ("mixed", "fp64"):
        run(prec)
        (n, mean), t = _timed(lambda: run(prec))
        res[prec] = (t, mean)
    rel = float(np.linalg.norm(res["mixed"][1] - res["fp64"][1]) / np.linalg.norm(res["fp64"][1]))
    out = {
        "metric": "FPS + full fit wall-time, torus n=20k -> FPS(0.015), 200 eigenpairs, 200 Adam steps",
        "value": round(res["mixed"][0], 4), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f32+f64", "data": "synthetic (seeded torus R=2 r=0.7, smoothed Gaussian field + noise)",
        "config": {"workload": (f"torus R=2 r=0.7, 20000 points -> FPS(0.015) = {n} points, k=15, "
                                f"dim_man=2, 200 eigenpairs, t=50, 50% train, 200 Adam steps")},
        "fp64_s": round(res["fp64"][0], 4),
        "speedup_vs_fp64": round(res["fp64"][0] / res["mixed"][0], 3),
        "rel_err_pred_mixed_vs_fp64": rel,
    }
    print(json.dumps(out))


# ---------------------------------------------------------------- main
def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=100000)
    p.add_argument("--k-eig", dest="k_eig", type=int, default=500)
    p.add_argument("--k-nn", dest="k_nn", type=int, default=20)
    p.add_argument("--adam-steps", dest="adam_steps", type=int, default=200)
    p.add_argument("--skip-fp64", action="store_true")
    p.add_argument("--skip-cpu", action="store_true")
    p.add_argument("--cpu-n", dest="cpu_n", type=int, default=1000)
    p.add_argument("--eig-tol-mixed", dest="eig_tol_mixed", type=float, default=None,
                   help="eigensolver residual tolerance of the mixed fit (default: the data "
                        "object's, 1e-7; the FP64 fit uses 1e-9)")
    p.add_argument("--cpu-k-eig", dest="cpu_k_eig", type=int, default=50)
    p.add_argument("--verbose", action="store_true")
    p.add_argument("--roofline-only", dest="roofline_only", action="store_true")
    p.add_argument("--roofline-dtype", dest="roofline_dtype", default="float64",
                   choices=sorted(SPMM_SHORT))
    p.add_argument("--config", default="fit",
                   choices=["fit", "batched", "multirhs", "sphere", "torus"],
                   help="fit: the headline full fit (default); batched: 256 fits on one "
                        "manifold; multirhs: 2000 right-hand sides at 64 nodes; sphere: the "
                        "plain FP64 method on S^2 n=2000; torus: FPS torus, 200 eigenpairs")
    p.add_argument("--batch-n", dest="batch_n", type=int, default=50000)
    p.add_argument("--batch-fields", dest="batch_fields", type=int, default=256)
    p.add_argument("--multirhs-T", dest="multirhs_T", type=int, default=2000)
    p.add_argument("--precision-multirhs", dest="precision_multirhs", default="fp64",
                   choices=["fp64", "mixed"])
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is a design only; the figures are for one GPU.", file=sys.stderr)
        sys.exit(2)
    if args.roofline_only:
        roofline_only(args)
        return
    if args.config == "batched":
        bench_batched(args)
        return
    if args.config == "multirhs":
        bench_multirhs(args)
        return
    if args.config == "sphere":
        bench_sphere(args)
        return
    if args.config == "torus":
        bench_torus(args)
        return

    import numpy as np
    import torch
    from rvgp_amd import _lib, instrument
    _lib.lib()
    dev = _lib.device()

    X, a, Q6 = flat_torus3(args.n, seed=0)
    V = torus_field(a, Q6)
    train = np.random.default_rng(3).permutation(args.n)[: int(0.3 * args.n)]
    Xd = torch.as_tensor(X, device=dev)
    Vd = torch.as_tensor(V, device=dev)
    log(f"[bench] n={args.n} D={X.shape[1]} k_eig={args.k_eig} train={len(train)}")

    for w in range(args.warmup):
        log(f"[bench] warmup {w + 1}/{args.warmup} (mixed)")
        run_fit(Xd, Vd, train, "mixed", args, {})

    times, stages = [], {}
    rec = instrument.Recorder()
    for s in range(args.steps):
        log(f"[bench] timed step {s + 1}/{args.steps} (mixed)")
        torch.cuda.synchronize()
        with rec if s == args.steps - 1 else _Null():
            d_mx, gp_mx, t = run_fit(Xd, Vd, train, "mixed", args, stages)
        torch.cuda.synchronize()
        times.append(t)
        log(f"[bench] mixed fit {t:.3f}s stages={ {k: round(v, 3) for k, v in stages.items()} }")
    t_mixed = sum(times) / len(times)
    kstats = kernel_stats(rec)

    rng = np.random.default_rng(5)
    test_nodes = [int(i) for i in rng.choice(np.setdiff1d(np.arange(args.n), train),
                                             size=min(10000, args.n - len(train)), replace=False)]
    f_mx, _ = gp_mx.transform(d_mx, test_nodes)
    result_extra = {}
    t_fp64 = None
    if not args.skip_fp64:
        log("[bench] FP64 fit")
        st64 = {}
        d64, gp64, t_fp64 = run_fit(Xd, Vd, train, "fp64", args, st64)
        log(f"[bench] fp64 fit {t_fp64:.3f}s stages={ {k: round(v, 3) for k, v in st64.items()} }")
        f64, _ = gp64.transform(d64, test_nodes)
        w64 = d64._evals_Lc.cpu().numpy()
        wmx = d_mx._evals_Lc.cpu().numpy()
        ub = d64._Lcb.gershgorin_bound()
        # eigenspace principal angles between the mixed and FP64 runs over
        # clusters of the FP64 eigenvalues delimited by relative gaps > 1e-5
        # (the basis inside a degenerate cluster is free); the last cluster
        # is skipped when it may continue past the k-th eigenvalue
        import torch
        from rvgp_amd import dense
        rows = d64._U.shape[0]
        Mx = dense.gemm(d64._U, d_mx._U, transA=True).cpu().numpy() / rows
        gaps = np.diff(w64) > 1e-5 * np.maximum(np.abs(w64[1:]), 1e-9 * ub)
        cuts = [0] + list(np.nonzero(gaps)[0] + 1) + [len(w64)]
        ang = 0.0
        for a0, a1 in zip(cuts[:-2], cuts[1:-1]):
            sv = np.linalg.svd(Mx[a0:a1, a0:a1], compute_uv=False)
            ang = max(ang, float(np.arcsin(np.sqrt(max(0.0, 1.0 - float(sv.min()) ** 2)))))
        result_extra = {
            "fp64_s": t_fp64,
            "eig_angle_max_rad": ang,
            "eig_clusters": len(cuts) - 2,
            "speedup_vs_fp64": t_fp64 / t_mixed,
            "rel_err_pred": float(np.linalg.norm(f_mx - f64) / np.linalg.norm(f64)),
            "eig_rel_err_max": float(np.max(np.abs(wmx - w64) / np.maximum(np.abs(w64), 1e-9 * ub))),
            "nlml_rel_err": float(abs(gp_mx.training_loss() - gp64.training_loss())
                                  / abs(gp64.training_loss())),
            "stages_fp64_s": {k: round(v, 4) for k, v in st64.items()},
        }
        del d64, gp64

    # roofline of the dominant kernel: the block SpMM (b = dim_man) of the Lc
    # eigensolve -- of its FP64 (polish), FP32 and binary16 (opening passes)
    # forms the one with the largest summed time in the timed fit -- measured
    # over every launch of that kernel
    dm = d_mx.dim_man
    cands = [f"{kind}[{dt},b{dm}]" for dt in SPMM_SHORT
             for kind in (("spmm_buf_h",) if dt == "float16" else
                          ("spmm_buf", "spmm_vec", "spmm_wide"))]
    key = max(cands, key=lambda c: kstats.get(c, {}).get("total_ms", 0.0))
    ks = kstats.get(key, {"count": 0, "avg_us": 0.0, "GBps": 0.0, "total_ms": 0.0})
    kdt = key.split("[")[1].split(",")[0]
    roof = {"kernel": key, "bound": "hbm", "achieved": round(ks["GBps"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ks["GBps"] / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile(kdt, dm), "launches": ks["count"],
            "avg_us": round(ks["avg_us"], 2), "total_ms": round(ks["total_ms"], 2)}
    gram = kstats.get("gemm[f32acc64]", None)
    step_util = gp_step_utilisation(gp_mx.engine)

    cpu_b = None
    if not args.skip_cpu:
        log("[bench] CPU baseline (oracle)")
        cpu_b = cpu_baseline(args)
        cpu_b["gpu_same_sample_s"], cpu_b["gpu_same_sample_stages_s"] = gpu_same_sample(args)

    out = {
        "metric": METRIC, "value": round(t_mixed, 4), "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * t_mixed, 1),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded flat 3-torus in R^32 + analytic tangent field)",
        "config": {"workload": (f"3-torus in R^32, n={args.n}, k={args.k_nn}, dim_man={dm}, "
                                f"{args.k_eig} eigenpairs, 30% train (80/20 re-split), "
                                f"{args.adam_steps} adjoint Adam steps"),
                   "n": args.n, "k_nn": args.k_nn, "n_eigenpairs": args.k_eig,
                   "adam_steps": args.adam_steps, "parallelism": "single GPU"},
        **result_extra,
        "stages_mixed_s": {k: round(v, 4) for k, v in stages.items()},
        "gauges_connections_points_per_s": round(
            args.n / max(1e-9, stages.get("gauges", 0.0) + stages.get("connections", 0.0)), 1),
        "eig_stats_mixed": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                                for kk, vv in v.items()} for k, v in d_mx.eig_stats.items()},
        "roofline": roof,
        "gram_fp64_mfma": None if gram is None else {
            "achieved": round(gram["TFLOPs"], 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(gram["TFLOPs"] / FP64_PEAK_TFLOPS, 4), "launches": gram["count"]},
        "cholesky_step_fp64_mfma": step_util,
        "kernels": {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in kstats.items()},
        "cpu_baseline": cpu_b,
    }
    print(json.dumps(out))


class _Null:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


if __name__ == "__main__":
    main()
