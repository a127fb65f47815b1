"""Benchmark of the f(A)b engine on one MI355X.

Default workload (``--config fem1m``): the plain method at full scale, FP64 --
P1 FEM stiffness K and lumped mass M on a seeded Delaunay triangulation of
the unit disc with 1,000,000 nodes (uniform random interior points plus
4096 equispaced boundary points), A = M^{-1/2} K M^{-1/2} in CSR (symrcm
order), f1 = exp(-tau kappa A^{alpha/2}), f2 = phi(...), alpha = 1.7,
tau = 1e-3, kappa = 1, degree-8 Jacobi polynomial preconditioner, Lanczos
m = 80, 64 contour shifts, 20 deflated eigenvectors.  One step is one time
step of the exponential integrator = 2 f(A)b solves (exp and phi columns);
inputs are resident in HBM when the timed region starts.

``--config dirichlet128``: 2D 5-point Dirichlet Laplacian 128 x 128 scaled by
1/h^2, b ~ N(0,1) seed 0, alpha = 1.5, m = 50, 32 shifts, no preconditioner,
no deflation.

Prints one JSON line.
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


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_fem1m(args):
    from gpumatfun_amd import problems, quadrature as q
    t0 = time.time()
    nodes, tri = problems.disc_mesh(args.nodes, 4096, seed=0)
    log(f"mesh {nodes.shape[0]} nodes, {tri.shape[0]} triangles ({time.time() - t0:.1f}s)")
    K, mdiag = problems.p1_stiffness_lumped_mass(nodes, tri)
    A, mass, perm = problems.fem_operator(K, mdiag)
    nodes = nodes[perm]
    log(f"A: n={A.shape[0]} nnz={A.nnz} ({time.time() - t0:.1f}s)")
    Kdefl = 20
    lam, Q = problems.smallest_eigenpairs(A, Kdefl + 1)
    lN = problems.largest_eigenvalue(A)
    log(f"eigenpairs: l1={lam[Kdefl]:.6g} lN={lN:.6g} ({time.time() - t0:.1f}s)")
    # initial condition of problem_startup_nonuniform.m:38-52
    x, y = nodes[:, 0], nodes[:, 1]
    r = np.sqrt((x - 0.1) ** 2 + (y + 0.1) ** 2)
    xh = np.zeros(A.shape[0])
    xh[r < 1 / np.sqrt(2900)] = 0.01
    if xh.max() == 0:
        xh[np.argmin(r)] = 0.01
    u0 = mass * xh
    g0 = mass * xh * (1 - xh)
    inp = q.make_inputs(A, u0, g0, m=80, n=64, eta=1e-3, alpha=1.7, dt=1e-3,
                        l1=float(lam[Kdefl]), lN=lN, degree=8, prob_dim=lN / 4.0,
                        Q=Q[:, :Kdefl], lam=lam[:Kdefl], mass=mass)
    cfg = {"workload": "fem_disc_1m_plain", "n": int(A.shape[0]), "nnz": int(A.nnz),
           "m": 80, "poly_degree": 8, "shifts": 64, "deflation": Kdefl, "alpha": 1.7,
           "tau": 1e-3, "kappa": 1.0, "columns_per_step": 2, "mesh_seed": 0}
    return inp, cfg, "nonuniform", time.time() - t0


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
    return inp, cfg, "fisher", time.time() - t0


def oracle_one_step(inp, threads):
    """One time step of the oracle's C restatement (oracle/c/fab_oracle.c) on `threads`
    host threads: the step loop of problem_startup_CPU_nonuniform.m (mass = 1 for the
    Dirichlet workload, which is then problem_startup_CPU.m's loop)."""
    from oracle import cfab
    S = {"A": inp.A, "Q": inp.Q if inp.Q.shape[1] else None,
         "exp_lambda": inp.exp_lambda, "phi_lambda": inp.phi_lambda, "dt": inp.dt,
         "n": inp.extra["n"], "m": inp.m, "s": inp.extra["s"], "gamma": inp.gamma,
         "gamma_s_real": inp.gamma_s_real, "gamma_s_imag": inp.gamma_s_imag,
         "etaxi": inp.extra["etaxi"], "wts1_cpu": inp.extra["wts1_cpu"],
         "wts2_cpu": inp.extra["wts2_cpu"], "mass": inp.mass}
    cfab.load()
    t0 = time.perf_counter()
    u, gu, used = cfab.cpu_steps_c(S, inp.U[:, 0], inp.U[:, 1], 1, threads=threads)
    dt = time.perf_counter() - t0
    return u, gu, dt, used


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="fem1m", choices=["fem1m", "dirichlet128"])
    ap.add_argument("--nodes", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg")
    ap.add_argument("--no-graph", action="store_true")
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

    inp, cfg, kind, setup_s = (build_fem1m if args.config == "fem1m" else build_dirichlet128)(args)
    log(f"setup {setup_s:.1f}s")
    t0 = time.time()
    eng = LanczosEngine.from_inputs(inp, clamp=False, autotune=True)
    info = eng.info()
    log(f"create {time.time() - t0:.2f}s, lanes={info['spmm_lanes']}")

    # parity leg: one step from the initial state against the oracle
    eng.step(1, use_graph=False)
    U1 = eng.get_U()
    eng.synchronize()

    # timed region
    use_graph = not args.no_graph
    steps = 2 if args.profile_only else args.steps
    warm = 1 if args.profile_only else args.warmup
    eng.step(warm, use_graph=use_graph)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(steps, use_graph=use_graph)
    eng.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / steps
    value = 2.0 * steps / elapsed
    log(f"{steps} steps in {elapsed:.4f}s: {ms_per_step:.3f} ms/step, {value:.2f} solves/s")

    # roofline of the dominant kernel (SpMM), HIP events on the solver's stream
    n, nnz = info["n"], info["nnz"]
    mode = 0 if inp.gamma.size > 1 else 1
    spmm_ms = eng.profile_spmm(mode=mode, reps=10 if args.profile_only else 200)
    bytes_per = 12 * nnz + 4 * (n + 1) + (48 if mode == 0 else 64) * n
    achieved = bytes_per / (spmm_ms * 1e-3) / 1e9
    log(f"spmm mode {mode}: {spmm_ms * 1e3:.2f} us/launch, {achieved:.1f} GB/s algorithmic")
    copy_bytes = 2 * 16 * n * 8   # read + write of an 8-column-pair vector block
    copy_ms = eng.profile_copy(copy_bytes // 2, reps=5 if args.profile_only else 50)
    copy_gbs = copy_bytes / (copy_ms * 1e-3) / 1e9

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
        "spmm_lanes": info["spmm_lanes"],
        "setup_s": round(setup_s, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "k_spmm<%d,%s>" % (info["spmm_lanes"],
                                                                   "kMid" if mode == 0 else "kLast"),
                     "avg_launch_us": round(spmm_ms * 1e3, 3),
                     "algorithmic_bytes_per_launch": int(bytes_per)},
        "copy_kernel_gbs": round(copy_gbs, 1),
    }

    if not args.no_cpu and not args.profile_only:
        threads = min(16, os.cpu_count() or 1)
        log(f"CPU oracle (C restatement, {threads} threads): one step ...")
        u, gu, cpu_s, used = oracle_one_step(inp, threads)
        e_u = float(np.linalg.norm(U1[:, 0] - u) / np.linalg.norm(u))
        e_g = float(np.linalg.norm(U1[:, 1] - inp.dt * gu) / np.linalg.norm(inp.dt * gu))
        result["rel_err_vs_oracle"] = {"u": e_u, "dt_g": e_g, "steps": 1}
        reps = 1
        if cpu_s < 5.0:   # small workloads: repeat the step to a ~10 s sample
            reps = int(min(1000, max(2, np.ceil(10.0 / max(cpu_s, 1e-3)))))
            cpu_s = sum(oracle_one_step(inp, threads)[2] for _ in range(reps)) / reps
        result["cpu_baseline"] = {"value": round(2.0 / cpu_s, 6), "unit": "solves/s",
                                  "cores": used, "kind": "port",
                                  "sample": f"{reps} time step(s) (2 solves each) of the same "
                                            "workload by the oracle's C/OpenMP restatement "
                                            "(oracle/c/fab_oracle.c), mean per step",
                                  "seconds_per_step": round(cpu_s, 4),
                                  "host_cpus": os.cpu_count()}
        log(f"oracle step {cpu_s:.1f}s, rel err u={e_u:.3e} dt*g={e_g:.3e}")
    else:
        result["cpu_baseline"] = None
    eng.delete()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
