"""Shared problem builders for the tests (same inputs go to the HIP path and the oracle)."""

import numpy as np

from gpumatfun_amd import problems, quadrature as q


def oracle_setup(inp):
    """The dict the oracle's step loops take, built from the same create inputs."""
    return {
        "A": inp.A, "Q": inp.Q if inp.Q.shape[1] else None,
        "exp_lambda": inp.exp_lambda, "phi_lambda": inp.phi_lambda, "dt": inp.dt,
        "n": inp.extra["n"], "m": inp.m, "s": inp.extra["s"], "gamma": inp.gamma,
        "gamma_s_real": inp.gamma_s_real, "gamma_s_imag": inp.gamma_s_imag,
        "etaxi": inp.extra["etaxi"], "wts1_cpu": inp.extra["wts1_cpu"],
        "wts2_cpu": inp.extra["wts2_cpu"], "mass": inp.mass,
    }


def dirichlet_case(side=128, m=50, n=32, degree=0, seed=0, tau=1e-3, kappa=1.0, alpha=1.5):
    """2D 5-point Dirichlet Laplacian scaled by 1/h^2, b ~ N(0,1)."""
    A = problems.laplacian_2d_dirichlet(side)
    lam = problems.dirichlet_eigs_2d(side)
    rng = np.random.default_rng(seed)
    u0 = rng.standard_normal(A.shape[0])
    return q.make_inputs(A, u0, u0 * (1 - u0), m=m, n=n, eta=tau * kappa, alpha=alpha,
                         dt=tau, l1=float(lam.min()), lN=float(lam.max()), degree=degree)


def fem_case(n_nodes=3000, n_boundary=256, K=6, m=40, n=32, degree=2, seed=3,
             eta=1e-3, alpha=1.7, dt=1e-3):
    """Small FEM disc with mass lumping, deflation of K eigenpairs, Jacobi polynomial.
    The deflation pairs and lambda_max come from the library's factorisation-based
    eigensolver, as in the full-scale set-up."""
    from gpumatfun_amd.eigs import eigs_shift_invert, lambda_max
    A, mass, nodes = problems.fem_disc(n_nodes, n_boundary, seed)
    lam, Q = eigs_shift_invert(A, K + 1)
    lN = lambda_max(A)
    l1 = float(lam[K])
    x = nodes[:, 0]
    y = nodes[:, 1]
    xh = 0.5 * np.exp(-((x - 0.3) ** 2 + (y + 0.1) ** 2) * 4.0)
    u0 = mass * xh
    g0 = mass * xh * (1 - xh)
    return q.make_inputs(A, u0, g0, m=m, n=n, eta=eta, alpha=alpha, dt=dt, l1=l1, lN=lN,
                         degree=degree, prob_dim=lN / 4.0, Q=Q[:, :K], lam=lam[:K], mass=mass)


FHN = dict(aode_cpu=0.1, epsil=0.01, beta=0.5, gamma=1.0, deltas=0.0)


def monodomain_case(side=48, m=30, n=24, degree=0, dt=0.05, eta=2e-3, alpha=1.7):
    """FitzHugh-Nagumo problem of problem_startup_CPU_monodomain.m:38-49 on a small grid."""
    A = problems.laplacian_2d_dirichlet(side)
    lam = problems.dirichlet_eigs_2d(side)
    N = A.shape[0]
    xs = np.linspace(0, 2.5, side)
    X, Y = np.meshgrid(xs, xs)
    ode = np.zeros_like(X)
    xh = np.zeros_like(X)
    ode[Y >= 1.25] = 0.1
    xh[(X <= 1.25) & (Y < 1.25)] = 1.0
    ode = ode.reshape(N, order="F")
    xh = xh.reshape(N, order="F")
    a = FHN["aode_cpu"]
    yh = xh * (1 - xh) * (xh - a) - ode
    # GPU-form parameters of update_U_monodomain.cu:11-22 matching the CPU model
    fhn = dict(aode=FHN["epsil"] * FHN["beta"], bode=-FHN["epsil"] * FHN["gamma"],
               code=-dt * FHN["epsil"] * FHN["deltas"], vrest=a, vamp=1.0, vpeak=1.0,
               c1ion=-1.0)
    inp = q.make_inputs(A, xh, yh, m=m, n=n, eta=eta, alpha=alpha, dt=dt,
                        l1=float(lam.min()), lN=float(lam.max()), degree=degree, ode=ode, fhn=fhn)
    return inp


def dirichlet_sine_eigenpairs(side, K):
    """The K smallest eigenpairs of ``laplacian_2d_dirichlet(side)`` in closed form:
    lambda = (4/h^2)(sin^2(j pi h/2) + sin^2(k pi h/2)), eigenvector
    sin(j pi x) sin(k pi y) normalised, in natural order (index i + side*j)."""
    h = 1.0 / (side + 1)
    js = np.arange(1, side + 1)
    lx = 4.0 * np.sin(js * np.pi * h / 2) ** 2 / h ** 2
    L = lx[:, None] + lx[None, :]
    order = np.argsort(L.ravel(), kind="stable")[:K]
    jj, kk = np.unravel_index(order, L.shape)
    xs = js * h
    S1 = np.sqrt(2 * h) * np.sin(np.pi * np.outer(xs, js))   # column j: mode j on the 1D grid
    Q = np.stack([np.kron(S1[:, k], S1[:, j]) for j, k in zip(jj, kk)], 1)
    return L.ravel()[order], Q


def headline_parameters_case(side=128, m=80, seed=0):
    """The full-scale plain workload's parameters on a grid whose answer FP64 fixes to
    ~1e-13: degree-8 Jacobi polynomial, m = 80, 64 shifts, K = 20 deflation vectors
    (closed-form sine eigenpairs), alpha = 1.7, tau = 1e-3, a non-unit mass and
    g = mass ubar (1 - ubar) (problem_startup_CPU_nonuniform.m's reaction)."""
    A = problems.laplacian_2d_dirichlet(side)
    lam_all = problems.dirichlet_eigs_2d(side)
    lam, Q = dirichlet_sine_eigenpairs(side, 21)
    rng = np.random.default_rng(seed)
    N = A.shape[0]
    u0 = rng.standard_normal(N)
    mass = 1.0 + 0.5 * rng.uniform(size=N)
    ub = u0 / mass
    return q.make_inputs(A, u0, mass * ub * (1 - ub), m=m, n=64, eta=1e-3, alpha=1.7, dt=1e-3,
                         l1=float(lam[20]), lN=float(lam_all.max()), degree=8,
                         Q=Q[:, :20], lam=lam[:20], mass=mass)


def laplace3d_case(side=20, m=30, n=64, seed=0, alpha=1.5, tau=1e-3):
    """3D 7-point Dirichlet Laplacian side^3 scaled by 1/h^2 (the multi-RHS workload's
    operator at a test size), closed-form spectrum bounds, no preconditioner."""
    A = problems.laplacian_3d_dirichlet(side)
    h = 1.0 / (side + 1)
    lo = 3 * 4.0 * np.sin(np.pi * h / 2) ** 2 / h ** 2
    hi = 3 * 4.0 * np.sin(side * np.pi * h / 2) ** 2 / h ** 2
    u0 = np.random.default_rng(seed).standard_normal(A.shape[0])
    return q.make_inputs(A, u0, u0 * (1 - u0), m=m, n=n, eta=tau, alpha=alpha, dt=tau,
                         l1=lo, lN=hi, degree=0)
