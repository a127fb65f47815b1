"""Exponential-integrator step loops (oracle).

* ``cpu_steps``             -- ``problem_startup_CPU.m:115-129`` (g(u) = u(1-u))
* ``cpu_steps_nonuniform``  -- ``problem_startup_CPU_nonuniform.m:103-128``
                               (g = mass*ubar*(1-ubar), ubar = u/mass)
* ``cpu_steps_monodomain``  -- ``problem_startup_CPU_monodomain.m:124-141``
                               (FitzHugh-Nagumo source)
* ``GpuSemantics``          -- the reference's GPU pipeline, entry point by entry
  point: constructor (``lanczos_impl.cu:77-91``), ``build_subspace``
  (``build_subspace.cu:5-91``), ``lin_sys_solves`` (``lin_sys_solves.cu:5-21``
  with the packing of ``lanczos_impl.cu:292-387``, solved as 2n systems of size
  m), ``apply_V`` (``apply_V.cu:5-95``), ``update_U`` (``update_U.cu:5-32``,
  with the u >= 0 clamp of ``update_U.cu:9`` switchable) and
  ``update_U_monodomain`` (``update_U_monodomain.cu:5-43``).
"""

import numpy as np
import scipy.sparse as sp

from .krylov import jacobi_poly_eval, nicksmethodcpu, lanczos, tridiag_solve_pivoted


def _deflate(Q, exp_lambda, phi_lambda, u, gu):
    """problem_startup_CPU.m:120-125."""
    if Q is None or Q.shape[1] == 0:
        z = np.zeros_like(u)
        return u.copy(), gu.copy(), z, z.copy()
    Qu = Q.T @ u
    Qgu = Q.T @ gu
    xbar = u - Q @ Qu
    ybar = gu - Q @ Qgu
    update1 = Q @ (exp_lambda * Qu)
    update2 = Q @ (phi_lambda * Qgu)
    return xbar, ybar, update1, update2


def cpu_steps(setup, u0, gu0, steps, g=None):
    """Run ``steps`` CPU time steps; ``problem_startup_CPU.m:119-129``.

    ``setup`` holds A, Q, exp_lambda, phi_lambda, dt, n, m, s, gamma,
    gamma_s_real, gamma_s_imag, etaxi, wts1_cpu, wts2_cpu.  ``g`` maps u to
    g(u) (default u(1-u)).  Returns (u, gu).
    """
    if g is None:
        g = lambda u: u * (1 - u)
    u = np.array(u0, dtype=np.float64)
    gu = np.array(gu0, dtype=np.float64)
    S = setup
    for _ in range(steps):
        xbar, ybar, up1, up2 = _deflate(S["Q"], S["exp_lambda"], S["phi_lambda"], u, gu)
        x1, x2 = nicksmethodcpu(S["wts1_cpu"], S["wts2_cpu"], S["A"], xbar, ybar,
                                S["n"], S["m"], S["s"], S["gamma"],
                                S["gamma_s_real"], S["gamma_s_imag"], S["etaxi"])
        u = x1 + S["dt"] * x2 + up1 + S["dt"] * up2       # :127
        gu = g(u)                                          # :128
    return u, gu


def cpu_steps_nonuniform(setup, u0, gu0, steps):
    """``problem_startup_CPU_nonuniform.m:110-128``: g = mass*ubar*(1-ubar)."""
    mass = np.asarray(setup["mass"], dtype=np.float64)

    def g(u):
        ubar = u / mass
        return mass * ubar * (1 - ubar)
    return cpu_steps(setup, u0, gu0, steps, g=g)


def cpu_steps_monodomain(setup, u0, gu0, ode0, steps, aode, epsil, beta, gamma_fhn, deltas):
    """``problem_startup_CPU_monodomain.m:128-141`` (FitzHugh-Nagumo).

    ode <- ode + dt*epsil*(beta*u - gamma*ode - deltas);
    gu  <- u(1-u)(u-aode) - ode.
    """
    u = np.array(u0, dtype=np.float64)
    gu = np.array(gu0, dtype=np.float64)
    ode = np.array(ode0, dtype=np.float64)
    S = setup
    dt = S["dt"]
    for _ in range(steps):
        xbar, ybar, up1, up2 = _deflate(S["Q"], S["exp_lambda"], S["phi_lambda"], u, gu)
        x1, x2 = nicksmethodcpu(S["wts1_cpu"], S["wts2_cpu"], S["A"], xbar, ybar,
                                S["n"], S["m"], S["s"], S["gamma"],
                                S["gamma_s_real"], S["gamma_s_imag"], S["etaxi"])
        u = x1 + dt * x2 + up1 + dt * up2                  # :138
        ode = ode + dt * epsil * (beta * u - gamma_fhn * ode - deltas)   # :139
        gu = u * (1 - u) * (u - aode) - ode                # :140
    return u, gu, ode


MS_DEFAULT = dict(tau_in=0.3, tau_out=6.0, tau_open=120.0, tau_close=150.0, v_gate=0.13)


def ms_reaction(u, h, dt, p=MS_DEFAULT):
    """Mitchell-Schaeffer gate update and current (Mitchell & Schaeffer 2003):

        dh/dt = (1 - h)/tau_open  (u < v_gate),   -h/tau_close  (u >= v_gate)
        J     = h u^2 (1 - u)/tau_in - u/tau_out

    One forward-Euler step of h, then J with the updated h (the order the FHN
    path uses, update_U_monodomain.cu:11-22).  No reference counterpart.
    """
    dh = np.where(u < p["v_gate"], (1.0 - h) / p["tau_open"], -h / p["tau_close"])
    hn = h + dt * dh
    return hn, (hn * (u * u) * (1.0 - u)) / p["tau_in"] - u / p["tau_out"]


def cpu_steps_ms(setup, u0, h0, steps, p=MS_DEFAULT):
    """Exponential-Euler steps with the Mitchell-Schaeffer reaction; returns (u, g, h).

    The step loop of problem_startup_CPU_monodomain.m:128-141 with the FHN
    source replaced by ``ms_reaction``; g(u0) uses h0 unchanged.
    """
    S = setup
    dt = S["dt"]
    u = np.array(u0, dtype=np.float64)
    h = np.array(h0, dtype=np.float64)
    gu = (h * (u * u) * (1.0 - u)) / p["tau_in"] - u / p["tau_out"]
    for _ in range(steps):
        xbar, ybar, up1, up2 = _deflate(S["Q"], S["exp_lambda"], S["phi_lambda"], u, gu)
        x1, x2 = nicksmethodcpu(S["wts1_cpu"], S["wts2_cpu"], S["A"], xbar, ybar,
                                S["n"], S["m"], S["s"], S["gamma"],
                                S["gamma_s_real"], S["gamma_s_imag"], S["etaxi"])
        u = x1 + dt * x2 + up1 + dt * up2
        h, gu = ms_reaction(u, h, dt, p)
    return u, gu, h


class GpuSemantics:
    """The reference GPU solver's state machine, restated on the CPU.

    Inputs follow ``lanczos_create`` (``lanczos_create.cpp:7-32``): U is
    N x 2 [u, g]; wts1/wts2 are n x 2 [Re, Im] *including* s; ``shifts`` is
    the packed 4*m*n vector.  Deviations from the reference code, all
    deliberate: the batched solve is 2n systems of size m (the reference
    passes size 2m, ``lin_sys_solves.cu:15``, which reads past the end of its
    buffers but decouples into the same systems); the clamp of
    ``update_U.cu:9`` is a switch.
    """

    def __init__(self, U, A, Q, m, gamma, gamma_s_real, gamma_s_imag, exp_lambda,
                 phi_lambda, dt, mass, wts1, wts2, shifts, ode=None, fhn=None,
                 clamp=True):
        self.A = sp.csr_matrix(A)
        self.A.sort_indices()
        self.N = self.A.shape[0]
        self.m = int(m)
        self.gamma = np.asarray(gamma, dtype=np.float64).ravel()
        self.gsr = np.asarray(gamma_s_real, dtype=np.float64).reshape(self.gamma.size, -1)
        self.gsi = np.asarray(gamma_s_imag, dtype=np.float64).reshape(self.gamma.size, -1)
        self.n = self.gsr.shape[1]
        self.Q = np.asarray(Q, dtype=np.float64).reshape(self.N, -1)
        self.el = np.asarray(exp_lambda, dtype=np.float64).ravel()
        self.pl = np.asarray(phi_lambda, dtype=np.float64).ravel()
        self.dt = float(dt)
        self.mass = np.asarray(mass, dtype=np.float64).ravel()
        w1 = np.asarray(wts1, dtype=np.float64).reshape(self.n, 2)
        w2 = np.asarray(wts2, dtype=np.float64).reshape(self.n, 2)
        self.w1 = w1[:, 0] + 1j * w1[:, 1]
        self.w2 = w2[:, 0] + 1j * w2[:, 1]
        sh = np.asarray(shifts, dtype=np.float64).ravel()
        shc = sh[0::2] + 1j * sh[1::2]               # 2n*m complex slots
        self.shift_c = shc.reshape(2 * self.n, self.m)   # system j, row i
        self.ode = np.zeros(self.N) if ode is None else np.array(ode, dtype=np.float64)
        self.fhn = fhn or {}
        self.clamp = clamp
        U = np.asarray(U, dtype=np.float64).reshape(2, self.N).T.copy() \
            if np.asarray(U).ndim == 1 else np.array(U, dtype=np.float64)
        # lanczos_impl.cu:79-89
        U[:, 1] *= self.dt
        self.Usol = U.copy()
        self.U, self.update = self._deflate(U)
        self.X = np.zeros((self.N, 2))
        self.aabb = np.zeros((self.m + 1, 4))
        self.V = None
        self.Yr = np.zeros((self.m, 2 * self.n))
        self.Yi = np.zeros((self.m, 2 * self.n))

    def _deflate(self, U):
        if self.Q.shape[1] == 0:
            return U.copy(), np.zeros_like(U)
        Z = self.Q.T @ U                               # matmatprod(Q,U,Z,T,..)
        Ud = U - self.Q @ Z                            # matmatprod(Q,Z,U,..,-1,1)
        Z = Z.copy()
        Z[:, 0] *= self.el                             # diag_mult_on_left_column
        Z[:, 1] *= self.pl
        return Ud, self.Q @ Z

    def build_subspace(self):
        """build_subspace.cu:5-91 (same recurrence as lanczos.m:21-49)."""
        V1, V2, (a1, b1, a2, b2), _ = lanczos(self.A, self.U[:, 0], self.U[:, 1],
                                              self.m, self.gamma)
        self.V = (V1, V2)
        self.aabb[:, 0] = np.append(a1, 0.0)
        self.aabb[:, 1] = np.append(a2, 0.0)
        self.aabb[:, 2] = b1
        self.aabb[:, 3] = b2

    def lin_sys_solves(self):
        """lin_sys_solves.cu:5-21: (T - etaxi_j I) y = -beta0 w_j e1 per system."""
        m, n = self.m, self.n
        for sys_j in range(2 * n):
            col = 0 if sys_j < n else 1
            node = sys_j % n
            alpha = self.aabb[:m, col]
            beta = self.aabb[:, 2 + col]
            w = (self.w1 if col == 0 else self.w2)[node]
            d = alpha.astype(np.complex128) - self.shift_c[sys_j]
            off = beta[1:m].astype(np.complex128)
            rhs = np.zeros(m, dtype=np.complex128)
            rhs[0] = -(beta[0] * w.real + 1j * beta[0] * w.imag)
            y = tridiag_solve_pivoted(off, d, off, rhs)
            self.Yr[:, sys_j] = y.real
            self.Yi[:, sys_j] = y.imag

    def apply_V(self):
        """apply_V.cu:5-95: X(:,c) += Im(q_j(A) V_c y_j) per node, exp then phi."""
        n = self.n
        for c in range(2):
            Vc = self.V[c]
            for i in range(n):
                sys_j = i + c * n
                ur = Vc @ self.Yr[:, sys_j]
                ui = Vc @ self.Yi[:, sys_j]
                _, yi = jacobi_poly_eval(self.gsr[:, i], self.gsi[:, i], self.A, ur, ui)
                self.X[:, c] += yi

    def get_xsol(self):
        return self.X.copy()

    def get_U(self):
        return self.Usol.copy()

    def update_U(self):
        """update_U.cu:5-32."""
        u = self.X[:, 0] + self.X[:, 1] + self.update[:, 0] + self.update[:, 1]
        if self.clamp:
            u = u * (u >= 0.0)
        ubar = u / self.mass
        g = self.mass * ubar * (1 - ubar)
        g = g * self.dt
        self.Usol = np.column_stack([u, g])
        self.U, self.update = self._deflate(self.Usol)
        self.X[:] = 0.0

    def update_U_monodomain(self):
        """update_U_monodomain.cu:5-43."""
        p = self.fhn
        dt = self.dt
        u = self.X[:, 0] + self.X[:, 1] + self.update[:, 0] + self.update[:, 1]
        odepart1 = u * (dt * p["aode"])
        odepart2 = self.ode * (dt * p["bode"])
        self.ode = self.ode + (odepart1 + odepart2 + p["code"])
        source2 = (u - p["vrest"]) * u * (u - p["vpeak"])
        source2 = source2 * (p["c1ion"] / p["vamp"] / p["vamp"])
        g = (source2 - self.ode) * dt
        self.Usol = np.column_stack([u, g])
        self.U, self.update = self._deflate(self.Usol)
        self.X[:] = 0.0

    def step(self, monodomain=False):
        """nicksmethodgpu_problem.m:7-16 / nicksmethodgpu_problem_monodomain.m:7-16."""
        self.build_subspace()
        self.lin_sys_solves()
        self.apply_V()
        if monodomain:
            self.update_U_monodomain()
        else:
            self.update_U()
