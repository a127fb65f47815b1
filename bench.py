"""Benchmark of the f(A)b engine on one MI355X.

Workloads (``--config``):

* ``fem1m`` (default) -- the plain method at full scale, FP64: P1 FEM
  stiffness K and lumped mass M on a seeded Delaunay triangulation of the unit
  disc with 1,000,000 nodes (uniform random interior points plus 4096
  equispaced boundary points), A = M^{-1/2} K M^{-1/2} in CSR (reverse Cuthill-McKee
  order from the library's gm_rcm_order, in place of symrcm),
  f1 = exp(-tau kappa A^{alpha/2}), f2 = phi(...), alpha = 1.7, tau = 1e-3,
  kappa = 1, degree-8 Jacobi polynomial preconditioner, Lanczos m = 80, 64
  contour shifts, 20 deflated eigenvectors (from the library's sparse
  Cholesky + shift-invert block Lanczos, in place of eigs(K, M, k, 'sm')).  Its line also carries the other
  two parts of the metric: ``batched_form`` (the ``uneven64`` batch against the
  plain path on the same problems) and ``adjoint_form`` (``adjoint256``: the
  value-and-gradient cost against one forward solve), each with its parity.
* ``dirichlet128`` -- 2D 5-point Dirichlet Laplacian 128 x 128 scaled by 1/h^2,
  b ~ N(0,1) seed 0, alpha = 1.5, m = 50, 32 shifts, no preconditioner, no deflation.
* ``batch3d`` -- batched form: 3D 7-point Dirichlet Laplacian 128^3 (scaled by
  1/h^2), 256 problems with u_p ~ N(0,1) seeds 0..255 (512 columns), 64
  shifts, m = 50, alpha = 1.5, tau = 1e-3, no preconditioner; the plain path
  on the same matrix is timed beside it.

* ``eigs512`` -- the deflation set-up of ``ms512``: the 21 smallest
  eigenpairs of the 512^2 anisotropic 9-point operator by the GPU eigensolver
  (Chebyshev-filtered subspace iteration, block 32, degree 200, relative
  residual 1e-10), timed beside scipy's ARPACK shift-invert on the host.

* ``bfab512`` -- the multi-column Krylov engine alone: forward passes of 32
  columns on the ``ms512`` operator (m = 40, no deflation).

* ``msadj512`` -- the gradient of the ``ms512`` run's activation-map mismatch
  with respect to the conductivity ratio, through 1000 deflated time steps of
  its 16 stimuli (the adjoint of the monodomain time stepping, deflation reused).

(``uneven64``, ``adjoint256``, ``ms512`` and ``msadj512`` are described at their
builders.)
One step is one time step of the exponential integrator = 2 f(A)b solves per
problem; inputs are resident in HBM when the timed region starts.  Prints one
JSON line.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
EIGS_INFO = {}   # fem1m: which eigensolver produced the deflation pairs, and its figures


def kernel_symbol(info, mode):
    """Demangled name of the SpMM kernel the solver runs in `mode` (1 = kMid, 2 = kLast)."""
    if info["spmm_kind"] == 0:
        return f"void gm::k_spmm<{info['spmm_param']}, {mode}>(gm::SpmmArgs)"
    c16 = "true" if info["index_bytes"] == 2 else "false"
    pre = "true" if info["spmm_kind"] == 2 else "false"
    return (f"void gm::k_spmm_stream<{info['spmm_param']}, {mode}, false, {c16}, {pre}>"
            "(gm::SpmmArgs)")


def spmm_bytes(info, mode):
    """Algorithmic bytes of one SpMM launch: values and stored column indices,
    row pointers, and 16 B per row for each of x (gathered once), the row's
    own operand(s) and y (kLast adds v_{k-1})."""
    n, nnz = info["n"], info["nnz"]
    return (8 + info["index_bytes"]) * nnz + 4 * (n + 1) + (64 if mode == 2 else 48) * n


def fab_step_bytes(n, nnz, m, d, K, columns, R):
    """Algorithmic bytes of one time step of `columns` f(A)b columns whose SpMMs
    are shared by R columns at a time: per column m(d+1)(S_A/R + 16n) (the
    Lanczos SpMMs: matrix share, x gathered and y written) + 32nm (recurrence:
    v_(k-1), v_k read, v_(k+1) written; one read of V for V*C) + d(S_A/R + 16n)
    (the collapsed Horner apply) + 24n (output) + 16nK (deflation), with
    S_A = 12 nnz + 4(n+1)."""
    SA = 12 * nnz + 4 * (n + 1)
    per_col = (m * (d + 1) * (SA / R + 16 * n) + 32 * n * m + d * (SA / R + 16 * n)
               + 24 * n + 16 * n * K)
    return columns * per_col


def step_roofline(bytes_per_step, seconds_per_step, what):
    """Whole-step roofline: algorithmic bytes of one step over the measured step
    time, against the HBM peak (for forms without a per-kernel timing entry)."""
    gbs = bytes_per_step / seconds_per_step / 1e9
    return {"bound": "hbm", "scope": "whole step: " + what, "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": None, "algorithmic_bytes_per_step": int(bytes_per_step)}


def kernel_key(info, mode):
    """Schedule/mode key of the committed counter summaries."""
    return f"{info['spmm_schedule']}/mode{mode}"


def measured_traffic(workload, key):
    """HBM bytes per launch of the SpMM `key` on `workload` from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, each scaled by what
    the counter reads for a copy kernel of known size), or None when no
    summary covers this kernel."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh)
    k = rec.get("kernels", {}).get(key)
    return None if k is None else int(k["traffic_bytes"])


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_fem1m(args):
    from gpumatfun_amd import problems, quadrature as q
    t0 = time.time()
    Kdefl = 20
    cache = (os.path.join(args.setup_cache, f"fem1m_{args.nodes}_{args.rcm}_{args.eigs}.npz")
             if args.setup_cache else None)
    if cache and os.path.exists(cache):
        import scipy.sparse as sp
        z = np.load(cache)
        A = sp.csr_matrix((z["data"], z["indices"], z["indptr"]), shape=tuple(z["shape"]))
        mass, lam, Q, lN, u0, g0 = z["mass"], z["lam"], z["Q"], float(z["lN"]), z["u0"], z["g0"]
        log(f"set-up read from {cache} ({time.time() - t0:.1f}s)")
    else:
        A, mass, lam, Q, lN, u0, g0 = _fem1m_setup(args, t0, Kdefl)
        if cache:
            os.makedirs(args.setup_cache, exist_ok=True)
            np.savez(cache, data=A.data, indices=A.indices, indptr=A.indptr,
                     shape=np.array(A.shape), mass=mass, lam=lam, Q=Q, lN=lN, u0=u0, g0=g0)
    inp = q.make_inputs(A, u0, g0, m=80, n=64, eta=1e-3, alpha=1.7, dt=1e-3,
                        l1=float(lam[Kdefl]), lN=lN, degree=8, prob_dim=lN / 4.0,
                        Q=Q[:, :Kdefl], lam=lam[:Kdefl], mass=mass)
    cfg = {"workload": "fem_disc_1m_plain", "n": int(A.shape[0]), "nnz": int(A.nnz),
           "m": 80, "poly_degree": 8, "shifts": 64, "deflation": Kdefl, "alpha": 1.7,
           "tau": 1e-3, "kappa": 1.0, "columns_per_step": 2, "mesh_seed": 0,
           "lambda_max": lN, "l1": float(lam[Kdefl]), "rcm": args.rcm}
    if EIGS_INFO:
        cfg["deflation_setup"] = dict(EIGS_INFO)
    return inp, cfg, time.time() - t0


def _fem1m_setup(args, t0, Kdefl):
    """Seeded Delaunay disc, P1 stiffness and lumped mass, symrcm order, the
    Kdefl + 1 smallest eigenpairs and the largest eigenvalue (host)."""
    from gpumatfun_amd import problems
    nodes, tri = problems.disc_mesh(args.nodes, 4096, seed=0)
    log(f"mesh {nodes.shape[0]} nodes, {tri.shape[0]} triangles ({time.time() - t0:.1f}s)")
    K, mdiag = problems.p1_stiffness_lumped_mass(nodes, tri)
    A, mass, perm = problems.fem_operator(K, mdiag,
                                          reorder="native" if args.rcm == "native" else True)
    nodes = nodes[perm]
    log(f"A: n={A.shape[0]} nnz={A.nnz} ({time.time() - t0:.1f}s)")
    te = time.time()
    if args.eigs == "arpack":
        lam, Q = problems.smallest_eigenpairs(A, Kdefl + 1)
        lN = problems.largest_eigenvalue(A)
        EIGS_INFO.update(eigensolver="scipy eigsh (SuperLU + ARPACK shift-invert, 'LA')")
    else:
        from gpumatfun_amd import eigs as ge
        sigma = ge.default_shift(A)
        lam, Q, inf = ge.eigs_shift_invert(A, Kdefl + 1, sigma=sigma, info=True)
        lN, linf = ge.lambda_max(A, info=True)
        EIGS_INFO.update(eigensolver="gm_eigs_shift_invert + gm_lambda_max (host, library)",
                         sigma=sigma, block_solves=inf["blocks"],
                         shift_invert_residual=inf["max_residual"],
                         lambda_max_iterations=linf["iterations"],
                         max_residual_over_lambda_max=float(np.max(np.linalg.norm(
                             A @ Q - Q * lam, axis=0)) / lN))
    EIGS_INFO["seconds"] = round(time.time() - te, 2)
    log(f"eigenpairs: l1={lam[Kdefl]:.6g} lN={lN:.6g} ({time.time() - t0:.1f}s; "
        f"{EIGS_INFO})")
    # initial condition of problem_startup_nonuniform.m:38-52
    x, y = nodes[:, 0], nodes[:, 1]
    r = np.sqrt((x - 0.1) ** 2 + (y + 0.1) ** 2)
    xh = np.zeros(A.shape[0])
    xh[r < 1 / np.sqrt(2900)] = 0.01
    if xh.max() == 0:
        xh[np.argmin(r)] = 0.01
    u0 = mass * xh
    g0 = mass * xh * (1 - xh)
    return A, mass, lam, Q, lN, u0, g0


def build_dirichlet128(args):
    from gpumatfun_amd import problems, quadrature as q
    t0 = time.time()
    A = problems.laplacian_2d_dirichlet(128)
    lam = problems.dirichlet_eigs_2d(128)
    u0 = np.random.default_rng(0).standard_normal(A.shape[0])
    inp = q.make_inputs(A, u0, u0 * (1 - u0), m=50, n=32, eta=1e-3, alpha=1.5, dt=1e-3,
                        l1=float(lam.min()), lN=float(lam.max()), degree=0)
    cfg = {"workload": "dirichlet2d_128_plain", "n": int(A.shape[0]), "nnz": int(A.nnz),
           "m": 50, "poly_degree": 0, "shifts": 32, "deflation": 0, "alpha": 1.5,
           "tau": 1e-3, "kappa": 1.0, "columns_per_step": 2, "b_seed": 0}
    return inp, cfg, time.time() - t0


def build_batch3d(args):
    from gpumatfun_amd import problems, quadrature as q
    t0 = time.time()
    side = args.side3d
    A = problems.laplacian_3d_dirichlet(side)
    order = None
    if args.node_order in ("pencil", "locality"):
        # the same operator with the nodes in 16 x 16 pencils, or in the library's
        # graph-only locality order (P A P^T, and P u for every right-hand side):
        # f(P A P^T) P u = P f(A) u
        import scipy.sparse as sp
        if args.node_order == "pencil":
            order = problems.pencil_order(side, 16)
        else:
            order = problems.locality_order(A, 512)
        A = sp.csr_matrix(A[order][:, order])
        A.sort_indices()
        A.indptr = A.indptr.astype(np.int32)
        A.indices = A.indices.astype(np.int32)
    h = 1.0 / (side + 1)
    l1d = 4.0 * np.sin(np.arange(1, side + 1) * np.pi * h / 2) ** 2 / h ** 2
    lmin, lmax = 3 * l1d.min(), 3 * l1d.max()
    N = A.shape[0]
    P = args.problems
    U = np.empty((N, 2 * P))
    for p in range(P):
        u = np.random.default_rng(p).standard_normal(N)
        if order is not None:
            u = u[order]
        U[:, 2 * p] = u
        U[:, 2 * p + 1] = u * (1 - u)
    inp = q.make_inputs(A, U[:, 0], U[:, 1], m=50, n=64, eta=1e-3, alpha=1.5, dt=1e-3,
                        l1=float(lmin), lN=float(lmax), degree=0)
    cfg = {"workload": f"batch_laplace3d_{side}_{P}rhs", "n": int(N), "nnz": int(A.nnz),
           "problems": P, "m": 50, "poly_degree": 0, "shifts": 64, "deflation": 0,
           "alpha": 1.5, "tau": 1e-3, "kappa": 1.0, "columns_per_step": 2 * P,
           "rhs_seeds": f"0..{P - 1}", "node_order": args.node_order}
    return inp, U, cfg, time.time() - t0


def build_uneven64(args):
    """64 independent 2D 5-point Dirichlet Laplacians, sides log-uniform in
    [32, 448] (seed 0), m = 40, 32 contour nodes each on its own spectrum."""
    from gpumatfun_amd import problems, quadrature as q
    t0 = time.time()
    rng = np.random.default_rng(0)
    sides = np.rint(np.exp(rng.uniform(np.log(32), np.log(448), args.problems))).astype(int)
    inps = []
    for p, side in enumerate(sides):
        A = problems.laplacian_2d_dirichlet(int(side))
        lam = problems.dirichlet_eigs_2d(int(side))
        u0 = np.random.default_rng(1000 + p).standard_normal(A.shape[0])
        inps.append(q.make_inputs(A, u0, u0 * (1 - u0), m=40, n=32, eta=1e-3, alpha=1.5,
                                  dt=1e-3, l1=float(lam.min()), lN=float(lam.max()), degree=0))
    rows = int(sum(i.A.shape[0] for i in inps))
    cfg = {"workload": f"uneven_laplace2d_{len(inps)}", "problems": len(inps),
           "sides": "log-uniform [32, 448], seed 0", "side_min": int(sides.min()),
           "side_max": int(sides.max()), "total_rows": rows,
           "total_nnz": int(sum(i.A.nnz for i in inps)), "m": 40, "poly_degree": 0,
           "shifts": 32, "deflation": 0, "alpha": 1.5, "tau": 1e-3,
           "columns_per_step": 2 * len(inps), "rhs_seeds": "1000.."}
    return inps, cfg, time.time() - t0


def run_uneven(args, steps, warm, use_graph):
    from gpumatfun_amd.engine import LanczosEngine
    from gpumatfun_amd.ubatch import UnevenBatch
    inps, cfg, setup_s = build_uneven64(args)
    P = len(inps)
    log(f"setup {setup_s:.1f}s, {P} problems, {cfg['total_rows']} rows")
    ub = UnevenBatch.from_inputs(inps, clamp=False, autotune=True)
    ub.step(1, use_graph=False)
    ub.synchronize()
    U1 = [ub.get_U(p) for p in range(P)]
    elapsed = timed(lambda k: ub.step(k, use_graph=use_graph), ub.synchronize, steps, warm)
    value = 2.0 * P * steps / elapsed
    log(f"uneven batch: {steps} steps in {elapsed:.4f}s -> {value:.1f} solves/s")
    info = ub.info()
    ub.delete()
    # the plain path: one engine per problem, each stepped in turn (graph replays)
    engs = [LanczosEngine.from_inputs(inp, clamp=False, schedule="blocks") for inp in inps]

    def plain_step(k):
        for _ in range(k):
            for e in engs:
                e.step(1, use_graph=use_graph)

    def plain_sync():
        for e in engs:
            e.synchronize()

    pel = timed(plain_step, plain_sync, steps, warm)
    plain = 2.0 * P * steps / pel
    for e in engs:
        e.delete()
    log(f"plain, problem by problem: {plain:.1f} solves/s; ratio {value / plain:.2f}")
    result = {
        "metric": "f(A)b solves/sec (FP64, uneven batch)",
        "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": round(1000.0 * elapsed / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded right-hand sides)", "config": cfg, "graph": use_graph,
        "row_blocks": info["row_blocks"],
        "spmm_schedule": {k: info[k] for k in ("spmm_tile", "index_bytes", "early_loads")},
        "plain_same_problems": {"value": round(plain, 3), "unit": "solves/s",
                                "how": "one plain engine per problem, stepped in turn"},
        "batched_vs_plain": round(value / plain, 3),
        "roofline": step_roofline(
            sum(fab_step_bytes(i.A.shape[0], i.A.nnz, i.m, 0, 0, 2, 2) for i in inps),
            elapsed / steps, "every kernel of a batched step over all problems"),
        "cpu_baseline": None,
    }
    if not args.no_cpu and not args.profile_only:
        threads = min(16, os.cpu_count() or 1)
        worst, cpu_s = 0.0, 0.0
        for p, inp in enumerate(inps):
            u, gu, t, used = oracle_one_step(inp, inp.U[:, 0], inp.U[:, 1], threads)
            cpu_s += t
            worst = max(worst, rel(U1[p][:, 0], u), rel(U1[p][:, 1], inp.dt * gu))
        result["rel_err_vs_oracle"] = {"max_over_problems": worst, "steps": 1,
                                       "tolerance": 1e-10}
        result["cpu_baseline"] = {"value": round(2.0 * P / cpu_s, 6), "unit": "solves/s",
                                  "cores": used, "kind": "port",
                                  "sample": f"one time step of each of the {P} problems by the "
                                            "oracle's C/OpenMP restatement (oracle/c/fab_oracle.c)",
                                  "seconds_per_step": round(cpu_s, 4),
                                  "host_cpus": os.cpu_count()}
        log(f"oracle: {cpu_s:.2f}s per step of all problems, worst rel err {worst:.3e}")
    return result


def run_adjoint(args, steps, warm):
    """Forward u = f(A)b and the full backward (dL/db, dL/dkappa, dL/dalpha) of
    L = ||u - y||^2 on the 256^2 variable-coefficient Laplacian,
    kappa = exp(0.5 GRF) (correlation length 0.1, seed 1), tau = 1e-3,
    alpha = 1.5, m = 600, 32 contour nodes."""
    import scipy.sparse.linalg as sla
    from gpumatfun_amd import problems, quadrature as q
    from gpumatfun_amd.adjoint import FabEngine, contour_weights
    t0 = time.time()
    n, m, ns, tau, alpha = 256, 600, 32, 1e-3, 1.5
    kappa = np.exp(0.5 * problems.gaussian_random_field(n, 0.1, seed=1))
    A = problems.varcoef_laplacian_2d(kappa, n)
    lo = sla.eigsh(A, k=1, sigma=0, which="LM", return_eigenvectors=False)[0]
    hi = sla.eigsh(A, k=1, which="LA", return_eigenvectors=False)[0]
    c = q.contour_nodes(0.95 * lo, 1.05 * hi, ns, tau, 1.0)
    cw, dc = contour_weights(c["xi"], c["dxidt"], c["s"], tau, alpha)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(n * n)
    y = rng.standard_normal(n * n)
    setup_s = time.time() - t0
    cfg = {"workload": "adjoint_grf256", "n": n * n, "nnz": int(A.nnz), "m": m, "shifts": ns,
           "alpha": alpha, "tau": tau, "kappa": "exp(0.5 GRF), l = 0.1, seed 1",
           "loss": "||f(A)b - y||^2", "gradients": "b, kappa (65536), alpha"}
    out = {}
    for recompute in (False, True):
        eng = FabEngine(A, m=m, xi=c["xi"], wts=cw, dwts=dc, recompute=recompute)
        for _ in range(warm):
            u = eng.forward(b)
            eng.backward(2 * (u - y), kappa)
        t1 = time.perf_counter()
        for _ in range(steps):
            u = eng.forward(b)
        tf = (time.perf_counter() - t1) / steps
        t1 = time.perf_counter()
        for _ in range(steps):
            u = eng.forward(b)
            grads = eng.backward(2 * (u - y), kappa)
        tfb = (time.perf_counter() - t1) / steps
        eng.delete()
        out[recompute] = (tf, tfb, u, grads)
        log(f"recompute={recompute}: forward {tf * 1e3:.2f} ms, forward+backward {tfb * 1e3:.2f} ms")
    tf, tfb, u, (gb, gk, ga) = out[False]
    result = {
        "metric": "f(A)b value-and-gradient evaluations/sec (FP64, adjoint)",
        "value": round(1.0 / tfb, 4), "unit": "evaluations/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": round(1e3 * tfb, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded GRF conductivity, b and y)", "config": cfg,
        "setup_s": round(setup_s, 1),
        "forward_ms": round(1e3 * tf, 4),
        "adjoint_cost_vs_forward": round(tfb / tf, 3),
        "recompute_basis": {"forward_ms": round(1e3 * out[True][0], 4),
                            "forward_backward_ms": round(1e3 * out[True][1], 4),
                            "adjoint_cost_vs_forward": round(out[True][1] / out[True][0], 3)},
        "roofline": None, "cpu_baseline": None,
    }
    if not args.no_cpu and not args.profile_only:
        from oracle import adjoint as oadj
        t1 = time.perf_counter()
        rb, rk, ra = oadj.krylov_gradients(A, kappa, b, 2 * (u - y), c["xi"], cw, dc, m)
        cpu_s = time.perf_counter() - t1
        result["rel_err_vs_oracle"] = {
            "what": "same Krylov approximation (m = 600) on the CPU, NumPy",
            "grad_b": rel(gb, rb), "grad_kappa": rel(gk, rk),
            "grad_alpha": float(abs(ga - ra) / abs(ra))}
        result["cpu_baseline"] = {"value": round(1.0 / cpu_s, 6), "unit": "evaluations/s",
                                  "cores": 1, "kind": "port",
                                  "sample": "one backward pass (both Krylov runs, resolvent "
                                            "blocks, kappa and alpha reductions) of the oracle's "
                                            "NumPy restatement (oracle/adjoint.py)",
                                  "seconds": round(cpu_s, 3), "host_cpus": os.cpu_count()}
        log(f"oracle backward {cpu_s:.1f}s; rel err {result['rel_err_vs_oracle']}")
    return result


def form_legs(args, use_graph):
    """The other two parts of the metric, measured beside the plain path in the default
    run: the uneven batch against the plain path on the same 64 problems, and the
    adjoint's cost against one forward solve, each with its parity number."""
    import copy
    legs = {}
    a = copy.copy(args)
    a.problems = 64
    r = run_uneven(a, 20, 3, use_graph)
    legs["batched_form"] = {
        "workload": r["config"]["workload"], "value": r["value"], "unit": r["unit"],
        "plain_same_problems": r["plain_same_problems"]["value"],
        "batched_vs_plain": r["batched_vs_plain"], "steps": r["steps"],
        "ms_per_step": r["ms_per_step"], "total_rows": r["config"]["total_rows"],
        "rel_err_vs_oracle": r.get("rel_err_vs_oracle")}
    r = run_adjoint(a, 10, 2)
    legs["adjoint_form"] = {
        "workload": r["config"]["workload"], "value": r["value"], "unit": r["unit"],
        "forward_ms": r["forward_ms"], "forward_backward_ms": r["ms_per_step"],
        "adjoint_cost_vs_forward": r["adjoint_cost_vs_forward"],
        "recompute_basis_cost_vs_forward": r["recompute_basis"]["adjoint_cost_vs_forward"],
        "rel_err_vs_oracle": r.get("rel_err_vs_oracle")}
    return legs


def build_ms512(args):
    """Monodomain Mitchell-Schaeffer on a 512^2 anisotropic 9-point operator
    (fibre angle 0 -> pi/3 across x, conductivity ratio 9:1), alpha = 1.7,
    dt = 0.05, 20 deflation vectors reused every step, 16 stimulus patches of
    20 x 20 nodes (one problem each)."""
    import scipy.sparse.linalg as sla
    from gpumatfun_amd import problems, quadrature as q
    t0 = time.time()
    n, K, P, dt = 512, 20, 16, 0.05
    A = problems.anisotropic_9pt(n, ratio=9.0)
    lam, Q = sla.eigsh(A, k=K + 1, sigma=0, which="LM")
    order = np.argsort(lam)
    lam, Q = lam[order], Q[:, order]
    lN = float(sla.eigsh(A, k=1, which="LA", return_eigenvectors=False)[0])
    N = n * n
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    stims = []
    for k in range(P):
        cx, cy = 40 + 120 * (k % 4), 40 + 120 * (k // 4)
        stims.append(((X >= cx) & (X < cx + 20) & (Y >= cy) & (Y < cy + 20))
                     .astype(float).reshape(-1, order="F"))
    from gpumatfun_amd._lib import MS_DEFAULTS as prm
    U = np.empty((N, 2 * P))
    for k, u0 in enumerate(stims):
        U[:, 2 * k] = u0
        U[:, 2 * k + 1] = (u0 * u0 * (1 - u0)) / prm["tau_in"] - u0 / prm["tau_out"]
    inp = q.make_inputs(A, U[:, 0], U[:, 1], m=40, n=32, eta=dt * 1e-3, alpha=1.7, dt=dt,
                        l1=float(lam[K]), lN=lN, degree=0, ode=np.ones(N),
                        Q=Q[:, :K], lam=lam[:K])
    cfg = {"workload": "ms_aniso512_16stim", "n": N, "nnz": int(A.nnz), "problems": P,
           "m": 40, "poly_degree": 0, "shifts": 32, "deflation": K, "alpha": 1.7, "dt": dt,
           "eta": dt * 1e-3, "reaction": "Mitchell-Schaeffer", "fibre": "0 -> pi/3, ratio 9:1",
           "stimuli": "16 patches of 20 x 20", "columns_per_step": 2 * P}
    return inp, U, stims, dict(prm), cfg, time.time() - t0


def run_ms512(args, steps, warm, use_graph):
    from gpumatfun_amd.batch import BatchEngine
    from gpumatfun_amd.engine import LanczosEngine
    inp, U, stims, prm, cfg, setup_s = build_ms512(args)
    N, P = inp.A.shape[0], len(stims)
    log(f"setup {setup_s:.1f}s")
    b = BatchEngine.from_inputs(inp, U, clamp=False, reaction="ms", ode=np.ones((N, P)), ms=prm,
                                autotune=not args.column_slabs, column_slabs=args.column_slabs)
    b.step(1, use_graph=False)
    b.synchronize()
    U1 = b.get_U()
    elapsed = timed(lambda k: b.step(k, use_graph=use_graph), b.synchronize, steps, warm)
    value = 2.0 * P * steps / elapsed
    log(f"batched: {steps} steps in {elapsed:.3f}s -> {value:.1f} solves/s")
    info = b.info()
    b.delete()
    # the plain path: one engine per stimulus, stepped in turn (MS, same deflation)
    engs = []
    for k in range(P):
        inp.U = U[:, 2 * k:2 * k + 2].copy()
        e = LanczosEngine.from_inputs(inp, clamp=False, autotune=(k == 0))
        e.set_ms_params(**prm)
        engs.append(e)
    e0 = engs[0]
    e0.step(1, monodomain=2, use_graph=False)
    e0.synchronize()
    Up1 = e0.get_U()
    psteps = max(5, min(steps, 100))

    def plain_step(k):
        for _ in range(k):
            for e in engs:
                e.step(1, monodomain=2, use_graph=use_graph)

    def plain_sync():
        for e in engs:
            e.synchronize()

    pel = timed(plain_step, plain_sync, psteps, 2)
    plain = 2.0 * P * psteps / pel
    for e in engs:
        e.delete()
    log(f"plain, stimulus by stimulus: {plain:.1f} solves/s; ratio {value / plain:.2f}")
    result = {
        "metric": "f(A)b solves/sec (FP64, batched monodomain)",
        "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": round(1000.0 * elapsed / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded-free stimulus patches)", "config": cfg, "graph": use_graph,
        "chunk_columns": info["chunk_columns"], "column_slabs": info["column_slabs"],
        "setup_s": round(setup_s, 1),
        "plain_same_problems": {"value": round(plain, 3), "unit": "solves/s",
                                "steps": psteps,
                                "how": "one plain engine per stimulus, stepped in turn"},
        "batched_vs_plain": round(value / plain, 3),
        "batched_vs_plain_rel_err_step1": rel(U1[:, 0:2], Up1),
        "roofline": step_roofline(
            fab_step_bytes(N, int(inp.A.nnz), inp.m, 0, cfg["deflation"], 2 * P,
                           info["chunk_columns"]),
            elapsed / steps, "every kernel of a batched step, all stimuli"),
        "cpu_baseline": None,
    }
    if not args.no_cpu and not args.profile_only:
        import oracle
        S = oracle_setup_dict(inp)
        t1 = time.perf_counter()
        u, gu, h = oracle.cpu_steps_ms(S, stims[0], np.ones(N), 1, prm)
        cpu_s = time.perf_counter() - t1
        result["rel_err_vs_oracle"] = {"u": rel(U1[:, 0], u), "dt_g": rel(U1[:, 1], inp.dt * gu),
                                       "steps": 1, "problem": 0, "tolerance": 1e-10}
        result["cpu_baseline"] = {"value": round(2.0 / cpu_s, 6), "unit": "solves/s",
                                  "cores": 1, "kind": "port",
                                  "sample": "one time step of stimulus 0 by the oracle's NumPy "
                                            "restatement (oracle/stepping.py:cpu_steps_ms)",
                                  "seconds_per_step": round(cpu_s, 3),
                                  "host_cpus": os.cpu_count()}
        log(f"oracle step {cpu_s:.2f}s, rel err {result['rel_err_vs_oracle']}")
    print(json.dumps(result), flush=True)


def run_eigs512(args):
    """Seconds to the K + 1 = 21 smallest eigenpairs of the ms512 operator
    (what build_ms512 asks ARPACK for), GPU against host shift-invert ARPACK,
    with the agreement of the two: eigenvalues, and the distance between the
    spanned subspaces up to the first spectral gap."""
    import scipy.sparse.linalg as sla
    from gpumatfun_amd import problems
    from gpumatfun_amd.eigs import eigs_smallest
    n, k, block, degree, tol = 512, 21, 32, 200, 1e-10
    A = problems.anisotropic_9pt(n, ratio=9.0).tocsr()
    eigs_smallest(problems.laplacian_2d_dirichlet(32), 4, block_size=16, degree=10, tol=1e-8)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        lam, Q, info = eigs_smallest(A, k, block_size=block, degree=degree, tol=tol, info=True)
        times.append(time.perf_counter() - t0)
    gpu_s = min(times)
    log(f"GPU eigensolver: {gpu_s:.3f}s (runs {['%.3f' % t for t in times]}), "
        f"{info['passes']} passes, residual {info['max_residual']:.2e}")
    result = {
        "metric": "seconds to the 21 smallest eigenpairs (FP64, deflation set-up)",
        "value": round(gpu_s, 4), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(1000.0 * gpu_s, 2), "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic operator (512^2 anisotropic 9-point, fibre 0 -> pi/3, ratio 9:1)",
        "config": {"workload": "eigs_aniso512_k21", "n": int(A.shape[0]), "nnz": int(A.nnz),
                   "k": k, "block": block, "degree": degree, "tol": tol,
                   "method": "Chebyshev-filtered subspace iteration"},
        "passes": info["passes"], "max_relative_residual": info["max_residual"],
        "roofline": None, "cpu_baseline": None,
    }
    if not args.no_cpu:
        t0 = time.perf_counter()
        lam_a, Q_a = sla.eigsh(A, k=k, sigma=0, which="LM")
        cpu_s = time.perf_counter() - t0
        order = np.argsort(lam_a)
        lam_a, Q_a = lam_a[order], Q_a[:, order]
        gaps = [j for j in range(1, k) if lam_a[j] - lam_a[j - 1] > 1e-6 * lam_a[j]]
        kk = max(j for j in gaps + [k] if j <= k)
        # subspace distance up to the last gap below k: ||Q_a - Q (Q^T Q_a)||
        sub = float(np.abs(Q_a[:, :kk] - Q[:, :kk] @ (Q[:, :kk].T @ Q_a[:, :kk])).max())
        result["agreement_vs_arpack"] = {
            "max_rel_eig_diff": float(np.max(np.abs(lam - lam_a) / lam_a)),
            "subspace_max_abs_diff": sub, "subspace_columns": kk}
        try:
            from threadpoolctl import threadpool_info
            blas_threads = max([t["num_threads"] for t in threadpool_info()] + [1])
        except ImportError:
            blas_threads = None
        result["cpu_baseline"] = {"value": round(cpu_s, 3), "unit": "s", "cores": blas_threads,
                                  "kind": "port",
                                  "sample": "scipy.sparse.linalg.eigsh(A, k=21, sigma=0) "
                                            "(SuperLU factorisation + ARPACK), the host "
                                            "stand-in for eigs(K, M, k, 'sm') in "
                                            "process_nonuniform.m:106; cores = BLAS "
                                            "threads (the factorisation is serial)",
                                  "host_cpus": os.cpu_count()}
        result["gpu_vs_cpu"] = round(cpu_s / gpu_s, 2)
        log(f"ARPACK shift-invert {cpu_s:.2f}s; agreement {result['agreement_vs_arpack']}")
    print(json.dumps(result), flush=True)


def run_msadj512(args, steps):
    """dL/dratio for the activation-map mismatch L = sum_p ||t_p - t_obs,p||^2 on the
    512^2 Mitchell-Schaeffer monodomain (alpha = 1.7, dt = 0.05, m = 40, 32 nodes,
    K = 20 deflation pairs reused every step, eigenpairs from the GPU eigensolver),
    through `steps` steps of P stimuli (the 16 patches of ms512, 20 x 20 each; P = 1
    is a corner patch).  t_p is the smooth activation-time map dt sum_n s(u_n)
    (threshold 0.5, width 0.05); t_obs,p the map of a run at ratio 8 (a synthetic
    observation); the gradient is taken at ratio 9.  States and trajectories stay
    in HBM; all stimuli share one multi-column engine (or, with --adjoint-engine
    streams, one engine each on 4 concurrent host threads).  A run of the same code
    over min(steps, 100) steps is checked against central differences (step 1e-2)
    of the HIP forward with the eigenpairs recomputed at each perturbed ratio."""
    import scipy.sparse.linalg as sla
    from gpumatfun_amd import problems, quadrature as q
    from gpumatfun_amd.eigs import eigs_smallest
    from gpumatfun_amd.monodomain import (MonodomainBatchAdjoint, StimulusSetAdjoint,
                                          activation_times)
    from gpumatfun_amd._lib import MS_DEFAULTS
    import torch
    t0 = time.time()
    n, K, dt, m, ns, r0, r_obs = 512, 20, 0.05, args.adj_m, 32, 9.0, 8.0
    P, workers = args.stimuli, 4
    eta, alpha = dt * 1e-3, 1.7
    E = problems.anisotropic_9pt_ratio_derivative(n)
    N = n * n
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    if P == 1:
        u0s = [((X < 20) & (Y < 20)).astype(float).reshape(-1, order="F")]
    else:
        u0s = []
        for k in range(P):
            cx, cy = 40 + 120 * (k % 4), 40 + 120 * (k // 4)
            u0s.append(((X >= cx) & (X < cx + 20) & (Y >= cy) & (Y < cy + 20))
                       .astype(float).reshape(-1, order="F"))
    h0 = np.ones(N)
    prm = dict(MS_DEFAULTS)
    nfd = min(steps, 100)

    def engine(r, c=None):
        A = problems.anisotropic_9pt(n, ratio=r)
        lam, Q = eigs_smallest(A, K + 1, block_size=32, degree=200, tol=1e-10)
        if c is None:
            lN = float(sla.eigsh(A, k=1, which="LA", return_eigenvectors=False)[0])
            c = q.contour_nodes(0.98 * lam[K], 1.02 * lN, ns, eta, alpha)
        kw = dict(m=m, xi=c["xi"], wts1=c["wts1"], wts2=c["wts2"], dt=dt, ms=prm,
                  Q=Q[:, :K], lam=lam[:K], eta=eta, alpha=alpha)
        if args.adjoint_engine == "batch":
            st = MonodomainBatchAdjoint(A, P, **kw)
        else:
            st = StimulusSetAdjoint(A, P, workers=workers, **kw)
        return st, c

    def tmaps(st, nsteps):
        st.forward(u0s, h0, nsteps)
        return [activation_times(st.states(p), dt) for p in range(P)]

    def loss(ts, obs):
        return float(sum(((t - o) ** 2).sum() for (t, _), o in zip(ts, obs)))

    def grad(st, ts, obs):
        srcs = [so(2 * (t - o)) for (t, so), o in zip(ts, obs)]
        if args.adjoint_engine == "batch":
            ubar = torch.stack([s_[len(s_) - 1] for s_ in srcs], dim=1)
        else:
            ubar = [s_[len(s_) - 1] for s_ in srcs]
        return st.backward(ubar, [E], sources=[s_[:-1] for s_ in srcs])[2][0]

    st_obs, _ = engine(r_obs)
    t_obs = [t for t, _ in tmaps(st_obs, steps)]
    t_obs_fd = [t for t, _ in tmaps(st_obs, nfd)]
    st_obs.delete()
    del st_obs
    st, c = engine(r0)
    setup_s = time.time() - t0
    log(f"set-up {setup_s:.1f}s (eigenpairs on the GPU, observation at ratio {r_obs}, "
        f"{P} stimuli)")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st.forward(u0s, h0, steps)
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - t1
    log(f"forward {fwd_s:.2f}s")
    ts = [activation_times(st.states(p), dt) for p in range(P)]
    L = loss(ts, t_obs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gr = grad(st, ts, t_obs)
    torch.cuda.synchronize()
    bwd_s = time.perf_counter() - t1
    activated = float(sum(float((t < (steps * dt - 1e-3)).sum()) for t, _ in ts)) / (N * P)
    log(f"{steps} steps x {P} stimuli: forward {fwd_s:.2f}s, backward {bwd_s:.2f}s; "
        f"L = {L:.6e}, dL/dratio = {gr:.6e}; {100 * activated:.1f}% of node-stimuli activated")
    # the same code over nfd steps against central differences of the HIP forward
    gfd = grad(st, tmaps(st, nfd), t_obs_fd)
    st.delete()
    hh = 1e-2
    Ls = []
    for r in (r0 + hh, r0 - hh):
        str_, _ = engine(r, c)
        Ls.append(loss(tmaps(str_, nfd), t_obs_fd))
        str_.delete()
    fd = (Ls[0] - Ls[1]) / (2 * hh)
    log(f"{nfd} steps: dL/dratio {gfd:.10e}, central difference {fd:.10e}, "
        f"rel {abs(gfd - fd) / abs(fd):.2e}")
    cfg = {"workload": f"msadj_aniso512_{P}stim", "n": N, "nnz": int(E.nnz), "steps": steps,
           "stimuli": P, "m": m, "shifts": ns, "deflation": K, "alpha": alpha, "dt": dt,
           "eta": eta, "reaction": "Mitchell-Schaeffer",
           "stimulus_patches": "20 x 20" + (" corner" if P == 1 else ", the ms512 layout"),
           "ratio": r0, "observation_ratio": r_obs, "engine": args.adjoint_engine,
           "loss": "sum_p ||t_p - t_obs,p||^2, t = dt sum_n (1 - sigmoid((u_n - 0.5) / 0.05))"}
    result = {
        "metric": "adjoint cost vs forward (FP64, monodomain conductivity-ratio gradient)",
        "value": round(bwd_s / fwd_s, 3), "unit": "x forward", "n_gpus": 1, "steps": steps,
        "warmup": 0, "ms_per_step": round(1e3 * (fwd_s + bwd_s) / steps, 3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (stimulus patches; observation simulated at ratio 8)", "config": cfg,
        "setup_s": round(setup_s, 1), "forward_s": round(fwd_s, 3), "backward_s": round(bwd_s, 3),
        "loss": L, "dL_dratio": gr, "activated_fraction": activated,
        "fd_check": {"steps": nfd, "adjoint": gfd, "central_difference": fd, "step": hh,
                     "rel_diff": abs(gfd - fd) / abs(fd)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def run_bfab512(args, passes):
    """The multi-column Krylov engine alone (gm_bfab_forward): u = f(A) b for R = 32
    columns on the 512^2 anisotropic 9-point operator (ratio 9, fibre 0 -> pi/3),
    m = 40, 32 contour nodes on the operator's spectrum, no deflation; `passes`
    forward passes on one seeded 262,144 x 32 block.  The workload of the
    32-column SpMM's counter runs."""
    import scipy.sparse.linalg as sla
    import torch
    from gpumatfun_amd import problems, quadrature as q
    from gpumatfun_amd.adjoint import BatchFabEngine
    t0 = time.time()
    n, R, m, ns, tau, alpha = 512, 32, 40, 32, 0.05e-3, 1.7
    A = problems.anisotropic_9pt(n, ratio=9.0)
    lo = float(sla.eigsh(A, k=1, sigma=0, which="LM", return_eigenvectors=False)[0])
    hi = float(sla.eigsh(A, k=1, which="LA", return_eigenvectors=False)[0])
    c = q.contour_nodes(0.98 * lo, 1.02 * hi, ns, tau, alpha)
    eng = BatchFabEngine(A, R=R, m=m, xi=c["xi"], wts1=c["wts1"], wts2=c["wts2"])
    B = torch.as_tensor(np.random.default_rng(0).standard_normal((n * n, R)), device="cuda")
    setup_s = time.time() - t0
    eng.forward(B)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(passes):
        U = eng.forward(B)
    torch.cuda.synchronize()
    per = (time.perf_counter() - t1) / passes
    eng.delete()
    result = {
        "metric": "multi-column f(A)b forward passes/sec (FP64, 32 columns)",
        "value": round(1.0 / per, 3), "unit": "passes/s", "n_gpus": 1, "steps": passes,
        "warmup": 1, "ms_per_step": round(1e3 * per, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded block of right-hand sides)",
        "config": {"workload": "bfab_aniso512_32col", "n": n * n, "nnz": int(A.nnz), "columns": R,
                   "m": m, "shifts": ns, "deflation": 0},
        "setup_s": round(setup_s, 1), "roofline": None, "cpu_baseline": None,
        "output_norm": float(torch.linalg.norm(U))}
    print(json.dumps(result), flush=True)


def run_barrier(args):
    """The cost of one cooperative grid barrier against a kernel boundary, on 64, 128
    and 256 resident workgroups (gm_profile_barrier): what a persistent Lanczos
    kernel would pay per phase in place of the launches of a captured step."""
    import ctypes as C
    import torch
    from gpumatfun_amd import _lib
    torch.cuda.init()
    lib = _lib.load()
    rows = []
    for blocks in (64, 128, 256):
        g, k = C.c_double(), C.c_double()
        _lib.check(lib.gm_profile_barrier(0, blocks, 1000, C.byref(g), C.byref(k)))
        rows.append({"blocks": blocks, "grid_sync_us": round(g.value, 3),
                     "kernel_boundary_us": round(k.value, 3)})
        log(f"{blocks} workgroups: grid barrier {g.value:.2f} us, kernel boundary {k.value:.2f} us")
    print(json.dumps({"metric": "grid barrier vs kernel boundary (us)", "value": rows[0]["grid_sync_us"],
                      "unit": "us", "n_gpus": 1, "higher_is_better": False,
                      "config": {"workload": "cooperative_grid_barrier", "iterations": 1000},
                      "rows": rows}), flush=True)


def oracle_setup_dict(inp, A=None, perm=None):
    S = {"A": inp.A if A is None else A, "Q": inp.Q if inp.Q.shape[1] else None,
         "exp_lambda": inp.exp_lambda, "phi_lambda": inp.phi_lambda, "dt": inp.dt,
         "n": inp.extra["n"], "m": inp.m, "s": inp.extra["s"], "gamma": inp.gamma,
         "gamma_s_real": inp.gamma_s_real, "gamma_s_imag": inp.gamma_s_imag,
         "etaxi": inp.extra["etaxi"], "wts1_cpu": inp.extra["wts1_cpu"],
         "wts2_cpu": inp.extra["wts2_cpu"], "mass": inp.mass}
    if perm is not None:
        S["Q"] = None if S["Q"] is None else inp.Q[perm]
        S["mass"] = inp.mass[perm]
    return S


def oracle_one_step(inp, u0, g0, threads, perm_seed=None, variant=0, extended=False, m=None):
    """One time step of the oracle's C restatement (oracle/c/fab_oracle.c) on `threads`
    host threads: the step loop of problem_startup_CPU_nonuniform.m (mass = 1 is
    problem_startup_CPU.m's loop).  With perm_seed, the same problem with rows and
    columns permuted (only the floating-point summation order changes); `variant`
    switches arithmetic choices (oracle.cfab.V_*), `extended` runs it in long double,
    `m` overrides the Lanczos steps (with a matching shift vector)."""
    import scipy.sparse as sp
    from oracle import cfab
    cfab.load(extended)
    if perm_seed is None:
        S = oracle_setup_dict(inp)
        if m is not None:
            S["m"] = int(m)
        t0 = time.perf_counter()
        u, gu, used = cfab.cpu_steps_c(S, u0, g0, 1, threads=threads, variant=variant,
                                       extended=extended)
        return u, gu, time.perf_counter() - t0, used
    N = inp.A.shape[0]
    p = np.random.default_rng(perm_seed).permutation(N)
    ip = np.argsort(p)
    S = oracle_setup_dict(inp, A=sp.csr_matrix(inp.A[p][:, p]), perm=p)
    t0 = time.perf_counter()
    u, gu, used = cfab.cpu_steps_c(S, u0[p], g0[p], 1, threads=threads)
    return u[ip], gu[ip], time.perf_counter() - t0, used


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def parity_study(inp, u0, g0, U1, threads, runs, extd, result):
    """Where a result depends on the FP64 evaluation order: each arithmetic choice of a
    fused GPU evaluation switched on alone in the oracle (oracle.cfab.V_*), which
    stage's FP64 rounding moves the long-double result how far, and the Krylov
    truncation error of m steps against reorthogonalised Lanczos at twice m."""
    from oracle import cfab
    ue, ge = extd
    d = lambda a, b: max(rel(a[0], b[0]), rel(a[1], b[1]))
    gpu = (U1[:, 0], U1[:, 1] / inp.dt)
    base = runs[0]
    out = {"variants": {}, "extended_with_fp64_stage": {}}
    for name, v in (("scale_order", cfab.V_SCALE), ("fma", cfab.V_FMA), ("tree_sums", cfab.V_TREE),
                    ("deflation_fma", cfab.V_DEFL),
                    ("all_four", cfab.V_SCALE | cfab.V_FMA | cfab.V_TREE | cfab.V_DEFL),
                    ("reorthogonalised", cfab.V_REORTH)):
        r = oracle_one_step(inp, u0, g0, threads, variant=v)[:2]
        out["variants"][name] = {"vs_oracle": d(r, base), "vs_extended": d(r, (ue, ge)),
                                 "vs_gpu": d(r, gpu)}
        log(f"variant {name}: vs oracle {d(r, base):.2e}, vs extended {d(r, (ue, ge)):.2e}, "
            f"vs GPU {d(r, gpu):.2e}")
    for name, v in (("deflation", cfab.V_RDEFL), ("lanczos_vectors", cfab.V_RLANCZOS),
                    ("sparse_products", cfab.V_RSPMV)):
        r = oracle_one_step(inp, u0, g0, threads, variant=v, extended=True)[:2]
        out["extended_with_fp64_stage"][name] = {"vs_extended": d(r, (ue, ge)),
                                                 "vs_oracle": d(r, base)}
        log(f"long double with {name} rounded to double: moves {d(r, (ue, ge)):.2e}")
    m2 = 2 * inp.m
    from gpumatfun_amd import quadrature as q
    inp2 = q.make_inputs(inp.A, u0, g0, m=m2, n=inp.extra["n"], eta=inp.extra["eta"],
                         alpha=inp.extra["alpha"], dt=inp.dt, l1=inp.extra["l1"],
                         lN=inp.extra["lN"], degree=inp.gamma.size - 1,
                         prob_dim=None, Q=inp.Q, lam=inp.lam, mass=inp.mass)
    inp2.gamma = inp.gamma
    conv = oracle_one_step(inp2, u0, g0, threads, variant=cfab.V_REORTH)[:2]
    out["krylov_truncation"] = {"m": inp.m, "against": f"reorthogonalised Lanczos, m = {m2}",
                                "extended": d((ue, ge), conv), "gpu": d(gpu, conv),
                                "oracle": d(base, conv)}
    log(f"Krylov truncation of m={inp.m} against m={m2}: {d((ue, ge), conv):.2e}")
    result["parity_study"] = out


def cpu_leg(inp, u0, g0, U1, result, study=False):
    """Parity of the GPU's first step against the oracle, the oracle's own rounding
    floor (a row-permuted run), and the timed CPU baseline."""
    threads = min(16, os.cpu_count() or 1)
    log(f"CPU oracle (C restatement, {threads} threads): one step ...")
    u, gu, cpu_s, used = oracle_one_step(inp, u0, g0, threads)
    e_u, e_g = rel(U1[:, 0], u), rel(U1[:, 1], inp.dt * gu)
    up, gp, _, _ = oracle_one_step(inp, u0, g0, threads, perm_seed=1)
    floor = max(rel(up, u), rel(gp, gu))
    result["rel_err_vs_oracle"] = {"u": e_u, "dt_g": e_g, "steps": 1,
                                   "oracle_floor": floor, "tolerance": 1e-10}
    if floor > 1e-10:
        # past the loss of orthogonality: the oracle's spread over summation orders
        # (three row permutations and the original), and the GPU against each run
        runs = [(u, gu), (up, gp)] + [oracle_one_step(inp, u0, g0, threads, perm_seed=k)[:2]
                                      for k in (2, 3)]
        pair = [max(rel(a[0], b[0]), rel(a[1], b[1]))
                for i, a in enumerate(runs) for b in runs[i + 1:]]
        gpu = [max(rel(U1[:, 0], a[0]), rel(U1[:, 1], inp.dt * a[1])) for a in runs]
        result["rel_err_vs_oracle"]["oracle_spread"] = {
            "runs": "original and rows permuted with seeds 1, 2, 3",
            "pairwise_min": min(pair), "pairwise_max": max(pair),
            "gpu_vs_runs_min": min(gpu), "gpu_vs_runs_max": max(gpu)}
        log(f"oracle spread over 4 summation orders: {min(pair):.2e} .. {max(pair):.2e}; "
            f"GPU against them {min(gpu):.2e} .. {max(gpu):.2e}")
        # the same evaluation in long double: rounding errors 2^11 times smaller, the
        # reference point for every FP64 evaluation order, the GPU's included
        ue, ge, ext_s, _ = oracle_one_step(inp, u0, g0, threads, extended=True)
        ext = lambda a: max(rel(a[0], ue), rel(a[1], ge))
        result["rel_err_vs_oracle"]["vs_extended_precision"] = {
            "reference": "the oracle's C restatement in long double (64-bit significand)",
            "gpu": max(rel(U1[:, 0], ue), rel(U1[:, 1], inp.dt * ge)),
            "oracle_runs_min": min(ext(a) for a in runs),
            "oracle_runs_max": max(ext(a) for a in runs),
            "seconds": round(ext_s, 1)}
        log(f"vs the long-double evaluation: GPU {result['rel_err_vs_oracle']['vs_extended_precision']['gpu']:.2e}, "
            f"oracle runs {min(ext(a) for a in runs):.2e} .. {max(ext(a) for a in runs):.2e}")
        if study:
            parity_study(inp, u0, g0, U1, threads, runs, (ue, ge), result)
    reps = 1
    if cpu_s < 5.0:   # small workloads: repeat the step to a ~10 s sample
        reps = int(min(1000, max(2, np.ceil(10.0 / max(cpu_s, 1e-3)))))
        cpu_s = sum(oracle_one_step(inp, u0, g0, threads)[2] for _ in range(reps)) / reps
    result["cpu_baseline"] = {"value": round(2.0 / cpu_s, 6), "unit": "solves/s",
                              "cores": used, "kind": "port",
                              "sample": f"{reps} time step(s) (2 solves each) of one problem of "
                                        "the same workload by the oracle's C/OpenMP restatement "
                                        "(oracle/c/fab_oracle.c), mean per step",
                              "seconds_per_step": round(cpu_s, 4), "host_cpus": os.cpu_count()}
    log(f"oracle step {cpu_s:.2f}s, rel err u={e_u:.3e} dt*g={e_g:.3e}, floor {floor:.3e}")


def timed(step_fn, sync_fn, steps, warm):
    import torch
    step_fn(warm)
    sync_fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_fn(steps)
    sync_fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--config", default="fem1m",
                    choices=["fem1m", "dirichlet128", "batch3d", "uneven64", "adjoint256",
                             "ms512", "eigs512", "msadj512", "bfab512", "barrier"])
    ap.add_argument("--nodes", type=int, default=1_000_000)
    ap.add_argument("--side3d", type=int, default=128)
    ap.add_argument("--problems", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--eigs", default="native", choices=["native", "arpack"],
                    help="fem1m: deflation pairs and lambda_max from the library's "
                         "factorisation-based eigensolver (gm_eigs_shift_invert, "
                         "gm_lambda_max) or scipy's ARPACK")
    ap.add_argument("--rcm", default="native", choices=["native", "scipy"],
                    help="fem1m: reverse Cuthill-McKee order from the library's set-up code "
                         "(gm_rcm_order, pseudo-peripheral start as symrcm) or SciPy's")
    ap.add_argument("--node-order", default="natural", choices=["natural", "pencil", "locality"],
                    help="batch3d: grid nodes in natural (i + n j + n^2 k) or 16 x 16-pencil order, "
                         "or the library's locality order (gm_locality_order, leaves of 512)")
    ap.add_argument("--locality-order", default="on", choices=["on", "off"],
                    help="batch3d: the batch engine renumbers the nodes internally "
                         "(GM_FLAG_LOCALITY_ORDER); inputs and results in the caller's order")
    ap.add_argument("--column-slabs", action="store_true",
                    help="batch3d with --chunk-columns, ms512: the SpMM in column slabs of 32")
    ap.add_argument("--chunk-columns", type=int, default=0,
                    help="batch3d: columns per pass (0: chosen by the create-time autotuner)")
    ap.add_argument("--no-forms", action="store_true",
                    help="fem1m: skip the batched (uneven64) and adjoint (adjoint256) legs")
    ap.add_argument("--setup-cache", default=None,
                    help="directory to keep the FEM set-up (mesh, operator, eigenpairs) between runs")
    ap.add_argument("--schedule", default="auto", choices=["auto", "rows", "blocks", "blocks16", "blocks16e"],
                    help="SpMM schedule of the plain path: autotuned at create (default), "
                         "or forced (fixed kernels for counter runs; blocks16e: 16-bit "
                         "column offsets with the epilogue operands loaded early)")
    ap.add_argument("--adjoint-engine", default="batch", choices=["batch", "streams"],
                    help="msadj512: all stimuli in one multi-column engine, or one engine "
                         "per stimulus on concurrent host threads")
    ap.add_argument("--stimuli", type=int, default=16,
                    help="msadj512: number of stimulus patches (16 as in ms512)")
    ap.add_argument("--adj-m", type=int, default=40,
                    help="msadj512: Lanczos steps of the adjoint's solves (40 as in ms512)")
    ap.add_argument("--parity-study", action="store_true",
                    help="add the oracle's arithmetic variants, stage-by-stage rounding of "
                         "the long-double evaluation and the Krylov truncation error")
    ap.add_argument("--profile-only", action="store_true",
                    help="a few steps and the kernel timings only (for rocprofv3 runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    from gpumatfun_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from gpumatfun_amd import build as b
        b.build()
    from gpumatfun_amd.engine import LanczosEngine

    use_graph = not args.no_graph
    if args.config == "batch3d":
        steps = args.steps or 3
        warm = 1 if args.warmup is None else args.warmup
    else:
        steps = args.steps or 20
        warm = 3 if args.warmup is None else args.warmup
    if args.profile_only:
        steps, warm = 2, 1

    if args.config == "batch3d":
        run_batch(args, steps, warm, use_graph)
        return
    if args.config == "ms512":
        if args.steps is None and not args.profile_only:
            steps = 1000
        run_ms512(args, steps, warm, use_graph)
        return
    if args.config == "adjoint256":
        print(json.dumps(run_adjoint(args, steps, warm)), flush=True)
        return
    if args.config == "eigs512":
        run_eigs512(args)
        return
    if args.config == "msadj512":
        run_msadj512(args, args.steps or 1000)
        return
    if args.config == "barrier":
        run_barrier(args)
        return
    if args.config == "bfab512":
        run_bfab512(args, args.steps or 6)
        return
    if args.config == "uneven64":
        if args.problems == 256:
            args.problems = 64
        print(json.dumps(run_uneven(args, steps, warm, use_graph)), flush=True)
        return

    inp, cfg, setup_s = (build_fem1m if args.config == "fem1m" else build_dirichlet128)(args)
    log(f"setup {setup_s:.1f}s")
    t0 = time.time()
    if args.schedule == "auto":
        eng = LanczosEngine.from_inputs(inp, clamp=False, autotune=True)
    else:
        eng = LanczosEngine.from_inputs(inp, clamp=False, schedule=args.schedule[:6],
                                        index16=args.schedule.startswith("blocks16"),
                                        early_loads=args.schedule == "blocks16e")
    info = eng.info()
    log(f"create {time.time() - t0:.2f}s, schedule {info['spmm_schedule']}, "
        f"shifted solves {info['solve_form']}")

    # parity leg: one step from the initial state
    eng.step(1, use_graph=False)
    U1 = eng.get_U()
    eng.synchronize()

    elapsed = timed(lambda k: eng.step(k, use_graph=use_graph), eng.synchronize, steps, warm)
    ms_per_step = 1000.0 * elapsed / steps
    value = 2.0 * steps / elapsed
    log(f"{steps} steps in {elapsed:.4f}s: {ms_per_step:.3f} ms/step, {value:.2f} solves/s")

    # roofline of the dominant kernel (SpMM), HIP events on the solver's stream
    n, nnz = info["n"], info["nnz"]
    mode = 0 if inp.gamma.size > 1 else 1
    spmm_ms = eng.profile_spmm(mode=mode, reps=10 if args.profile_only else 200)
    bytes_per = spmm_bytes(info, 1 if mode == 0 else 2)
    achieved = bytes_per / (spmm_ms * 1e-3) / 1e9
    log(f"spmm mode {mode}: {spmm_ms * 1e3:.2f} us/launch, {achieved:.1f} GB/s algorithmic")
    copy_bytes = 2 * 16 * n * 8   # read + write of 8 column pairs
    copy_ms = eng.profile_copy(copy_bytes // 2, reps=5 if args.profile_only else 50)
    # the same launches with L2 and the Infinity Cache flushed before each (a 1 GiB
    # streaming copy), so nothing of the working set can be served from cache
    cold_ms = eng.profile_spmm_flushed(mode=mode, reps=5 if args.profile_only else 50)
    cold = bytes_per / (cold_ms * 1e-3) / 1e9
    log(f"spmm mode {mode} after a 1 GiB flush: {cold_ms * 1e3:.2f} us/launch, {cold:.1f} GB/s")

    result = {
        "metric": "f(A)b solves/sec (FP64, plain path)",
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mesh and initial condition)",
        "config": cfg,
        "graph": use_graph,
        "spmm_schedule": info["spmm_schedule"], "solve_form": info["solve_form"],
        "setup_s": round(setup_s, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(cfg["workload"], kernel_key(info, 1 if mode == 0 else 2)),
                     "kernel": kernel_symbol(info, 1 if mode == 0 else 2),
                     "avg_launch_us": round(spmm_ms * 1e3, 3),
                     "algorithmic_bytes_per_launch": int(bytes_per),
                     "cache_flushed": {
                         "what": "each launch after a 1 GiB streaming copy (cold L2 and "
                                 "Infinity Cache), timed alone with its launch latency",
                         "avg_launch_us": round(cold_ms * 1e3, 3),
                         "achieved": round(cold, 1), "frac": round(cold / HBM_PEAK_GBS, 4)}},
        "roofline_step": step_roofline(
            fab_step_bytes(n, info["nnz"], info["m"], info["gamma_size"] - 1, info["num_eigs"],
                           2, 2),
            ms_per_step * 1e-3, "every kernel of one plain step (the SpMM's share is above)"),
        "copy_kernel_gbs": round(copy_bytes / (copy_ms * 1e-3) / 1e9, 1),
        "cpu_baseline": None,
    }
    if not args.no_cpu and not args.profile_only:
        cpu_leg(inp, inp.U[:, 0], inp.U[:, 1], U1, result, study=args.parity_study)
    eng.delete()
    if args.config == "fem1m" and not args.no_forms and not args.profile_only:
        result.update(form_legs(args, use_graph))
    print(json.dumps(result), flush=True)


def run_batch(args, steps, warm, use_graph):
    from gpumatfun_amd.batch import BatchEngine
    from gpumatfun_amd.engine import LanczosEngine
    inp, U, cfg, setup_s = build_batch3d(args)
    log(f"setup {setup_s:.1f}s, U {U.shape}")
    P = U.shape[1] // 2
    t0 = time.time()
    loc = args.locality_order == "on"
    if args.chunk_columns:
        b = BatchEngine.from_inputs(inp, U, clamp=False, chunk_columns=args.chunk_columns,
                                    column_slabs=args.column_slabs, locality_order=loc)
    else:
        b = BatchEngine.from_inputs(inp, U, clamp=False, autotune=True, locality_order=loc)
    info = b.info()
    log(f"batch create {time.time() - t0:.1f}s, chunk columns {info['chunk_columns']}, "
        f"column slabs {info['column_slabs']}, locality order {loc}")
    # parity leg: problem 0 after one step
    b.step(1, use_graph=False)
    b.synchronize()
    U1 = b.get_U()[:, 0:2]
    elapsed = timed(lambda k: b.step(k, use_graph=use_graph), b.synchronize, steps, warm)
    value = 2.0 * P * steps / elapsed
    log(f"batched: {steps} steps in {elapsed:.3f}s -> {value:.1f} solves/s")
    b.delete()
    u0_cpu, g0_cpu = U[:, 0].copy(), U[:, 1].copy()   # problem 0, for the CPU leg
    del U

    # the plain path on the same matrix (one problem, 2 columns per step)
    if args.schedule == "auto":
        e = LanczosEngine.from_inputs(inp, clamp=False, autotune=True)
    else:
        e = LanczosEngine.from_inputs(inp, clamp=False, schedule=args.schedule[:6],
                                      index16=args.schedule.startswith("blocks16"),
                                      early_loads=args.schedule == "blocks16e")
    einfo = e.info()
    psteps = max(5, steps * 4)
    pel = timed(lambda k: e.step(k, use_graph=use_graph), e.synchronize, psteps, 2)
    plain = 2.0 * psteps / pel
    mode = 1
    spmm_ms = e.profile_spmm(mode=mode, reps=10 if args.profile_only else 100)
    cold_ms = e.profile_spmm_flushed(mode=mode, reps=5 if args.profile_only else 50)
    n, nnz = einfo["n"], einfo["nnz"]
    bytes_per = spmm_bytes(einfo, 2)
    e.delete()
    log(f"plain same matrix: {plain:.1f} solves/s; batched/plain = {value / plain:.2f}")
    result = {
        "metric": "f(A)b solves/sec (FP64, batched form)",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(1000.0 * elapsed / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded right-hand sides)",
        "config": cfg,
        "graph": use_graph,
        "chunk_columns": info["chunk_columns"], "column_slabs": info["column_slabs"],
        "locality_order": info["locality_order"],
        "setup_s": round(setup_s, 1),
        "plain_same_matrix": {"value": round(plain, 3), "unit": "solves/s",
                              "spmm_schedule": einfo["spmm_schedule"]},
        "batched_vs_plain": round(value / plain, 3),
        "roofline": {"bound": "hbm", "kernel": kernel_symbol(einfo, 2) + " (plain path, same matrix)",
                     "achieved": round(bytes_per / (spmm_ms * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_per / (spmm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("laplace3d_128_plain", kernel_key(einfo, 2)),
                     "avg_launch_us": round(spmm_ms * 1e3, 3),
                     "cache_flushed": {
                         "avg_launch_us": round(cold_ms * 1e3, 3),
                         "frac": round(bytes_per / (cold_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
        "cpu_baseline": None,
    }
    if not args.no_cpu and not args.profile_only:
        cpu_leg(inp, u0_cpu, g0_cpu, U1, result)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
