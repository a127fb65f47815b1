"""Pinning of the CPU oracle by independent mathematics (no GPU).

Each test checks an oracle function against something the reference's
algorithm must satisfy that is NOT a retyping of the oracle's own formula:
least-squares optimality, closed-form spectra, dense linear algebra,
brute-force matrix functions.
"""

from fractions import Fraction

import numpy as np
import pytest
import scipy.fft as sfft
import scipy.integrate as si
import scipy.sparse as sp

import oracle
from oracle import krylov, poly, contour
from conftest import rel
from gpumatfun_amd import problems


# ----------------------------------------------------------- polynomial_coeffs
def test_poly_degree0_is_one():
    # degree 0: R = 1 - t/3 on [0,4] -> s = 1/3, times (3+2d) = 3 -> [1]
    assert poly.polynomial_coeffs_exact(0) == [Fraction(1)]


@pytest.mark.parametrize("degree", [1, 2, 3, 5, 8])
def test_poly_is_chebyshev_weight_least_squares(degree):
    """gamma/(3+2d) must be the minimiser of int_0^4 (1 - t s(t))^2 w(t) dt with the
    Chebyshev weight w = t^(-1/2) (4-t)^(-1/2): solve the normal equations by quadrature."""
    d = degree
    g = np.array([float(x) for x in poly.polynomial_coeffs_exact(d)]) / (3 + 2 * d)
    # Gauss-Chebyshev nodes: int_0^4 f(t) w(t) dt = (pi/nq) sum f(t_i), exact for
    # polynomial f of degree < 2 nq.  Weighted least squares in the well-conditioned
    # basis t*T_j((t-2)/2), solved by QR.
    nq = 64
    th = (np.arange(nq) + 0.5) * np.pi / nq
    t = 2 - 2 * np.cos(th)
    basis = np.column_stack([t * np.cos(j * np.arccos((t - 2) / 2)) for j in range(d + 1)])
    c, *_ = np.linalg.lstsq(basis, np.ones(nq), rcond=None)
    tt = np.linspace(0, 4, 101)
    r_ls = 1 - np.column_stack([tt * np.cos(j * np.arccos(np.clip((tt - 2) / 2, -1, 1)))
                                for j in range(d + 1)]) @ c
    # evaluate the oracle's residual 1 - t s(t) exactly (rational arithmetic)
    ge = [x / (3 + 2 * d) for x in poly.polynomial_coeffs_exact(d)]
    r_or = []
    for x in tt:
        fx = Fraction(x)
        v = Fraction(0)
        for co in ge:
            v = v * fx + co
        r_or.append(float(1 - fx * v))
    assert np.max(np.abs(r_ls - np.array(r_or))) < 1e-12


def test_poly_scaling_by_prob_dim():
    g1 = poly.polynomial_coeffs(4, 1.0)
    g2 = poly.polynomial_coeffs(4, 2.5)
    for k, (a, b) in enumerate(zip(g1, g2)):
        p = 4 - k
        assert b == pytest.approx(a / 2.5 ** p, rel=1e-15)


def test_product_poly_matches_exact():
    from gpumatfun_amd.quadrature import polynomial_coeffs as prod
    for d in range(0, 12):
        ex = np.array(poly.polynomial_coeffs(d, 3.7))
        assert np.max(np.abs(prod(d, 3.7) - ex) / np.abs(ex)) < 1e-14


# ------------------------------------------------------------------- contour
@pytest.mark.parametrize("l1,lN,n,eta,alpha,tol", [
    (19.7, 133000.0, 64, 1e-3, 1.5, 1e-13),
    (2.0, 8.0, 16, 0.3, 1.0, 1e-13),
    (1.0, 1e4, 64, 1e-2, 1.7, 1e-9),
])
def test_contour_scalar_identity(l1, lN, n, eta, alpha, tol):
    """f(lam) = Im sum_j w_j/(xi_j - lam) on [l1, lN] (the HHT quadrature identity)."""
    c = contour.contour_setup(l1, lN, n, eta, alpha)
    for lam in np.geomspace(l1, lN, 25):
        for which, w in ((1, c["wts1"]), (2, c["wts2"])):
            approx = np.imag(np.sum(w / (c["xi"] - lam)))
            exact = contour.f_scalar(lam, eta, alpha, which).real
            assert abs(approx - exact) <= tol * max(1.0, abs(exact)) + 1e-15


def test_contour_nodes_jacobi_closed_form():
    """At t = i K'/2 (real part 0) sn = i/sqrt(k) (A&S 16.20), so xi has |xi| = sqrt(l1 lN)."""
    l1, lN = 3.0, 300.0
    c = contour.contour_setup(l1, lN, 2, 0.1, 1.0)   # nodes at -K/2 and +K/2 ... symmetric
    # the two nodes are mirror images: xi_0 = conj-inverse map of xi_1 about sqrt(l1 lN)
    prod = c["xi"][0] * np.conj(c["xi"][1])
    assert abs(abs(prod) - l1 * lN) / (l1 * lN) < 1e-13


def test_product_quadrature_matches_oracle():
    from gpumatfun_amd import quadrature as q
    for (l1, lN, n, eta, alpha, tol) in [(19.7, 133000.0, 32, 1e-3, 1.5, 1e-12),
                                         (2.0, 8.0, 8, 0.1, 1.0, 1e-13),
                                         (3.39, 1.19e10, 64, 1e-3, 1.7, 1e-9)]:
        a = q.contour_nodes(l1, lN, n, eta, alpha)
        b = contour.contour_setup(l1, lN, n, eta, alpha)
        for key in ("xi", "dxidt", "wts1", "wts2"):
            m = np.abs(b[key]) > 1e-300
            assert np.max(np.abs(a[key] - b[key])[m] / np.abs(b[key])[m]) < tol, key
        assert abs(a["s"] - b["s"]) / abs(b["s"]) < tol


def test_shifted_gamma_identity():
    """q_z(lam) (z - lam) = z p(z) - lam p(lam) for the synthetic-division coefficients."""
    gamma = np.array(poly.polynomial_coeffs(5, 2.0))
    z = np.array([3.0 + 2.0j, -1.0 + 0.5j, 10.0 + 7.0j])
    gs, etaxi = contour.shifted_gammas(gamma, z)
    for lam in (0.3, 1.7, 6.0):
        for j in range(z.size):
            qv = np.polyval(gs[:, j], lam)
            lhs = qv * (z[j] - lam)
            rhs = z[j] * np.polyval(gamma, z[j]) - lam * np.polyval(gamma, lam)
            assert abs(lhs - rhs) <= 1e-12 * abs(rhs)
    assert np.allclose(etaxi, z * np.polyval(gamma, z), rtol=1e-13)


def test_pack_shifts_layout():
    from gpumatfun_amd.quadrature import pack_shifts as prod_pack
    et = np.array([1 + 2j, 3 - 4j, 5 + 0.5j])
    m = 4
    a = contour.pack_shifts(et, m)
    assert a.size == 4 * m * 3
    c = a[0::2] + 1j * a[1::2]
    for sys_j in range(6):
        assert np.all(c[sys_j * m:(sys_j + 1) * m] == et[sys_j % 3])
    assert np.array_equal(a, prod_pack(et, m))


# ---------------------------------------------------------------- tridiagonal
@pytest.mark.parametrize("n", [1, 2, 3, 7, 40])
def test_tridiag_pivoted_vs_dense(n):
    rng = np.random.default_rng(n)
    for trial in range(5):
        dl = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
        du = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
        d = 0.1 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))  # forces pivoting
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        M = np.diag(d) + np.diag(dl, -1) + np.diag(du, 1)
        x = krylov.tridiag_solve_pivoted(dl, d, du, b)
        xr = np.linalg.solve(M, b)
        assert rel(x, xr) < 1e-11 * max(1.0, np.linalg.cond(M) * 1e-4)


def test_tridiag_singular_raises():
    with pytest.raises(ZeroDivisionError):
        krylov.tridiag_solve_pivoted([0.0], [0.0, 1.0], [1.0], [1.0, 1.0])


# ------------------------------------------------------------ poly evaluation
def test_poly_eval_dense():
    rng = np.random.default_rng(0)
    A = sp.random(30, 30, density=0.2, random_state=1, format="csr")
    A = A + A.T
    gr = rng.standard_normal(5)
    gi = rng.standard_normal(5)
    br = rng.standard_normal((30, 2))
    bi = rng.standard_normal((30, 2))
    xr, xi = krylov.jacobi_poly_eval(gr, gi, A, br, bi)
    Ad = A.toarray()
    P = np.zeros((30, 30), dtype=complex)
    for c in gr + 1j * gi:
        P = P @ Ad + c * np.eye(30)
    ref = P @ (br + 1j * bi)
    assert rel(xr + 1j * xi, ref) < 1e-13


def test_poly_eval_real_fast_path_is_exact():
    rng = np.random.default_rng(1)
    A = problems.laplacian_2d_dirichlet(6)
    g = rng.standard_normal(4)
    b = rng.standard_normal((36, 2))
    xr, _ = krylov.jacobi_poly_eval(g, np.zeros(4), A, b, np.zeros_like(b))
    # general path (complex lane live) must give the bit-identical real part
    xr2, _ = krylov.jacobi_poly_eval(g, np.full(4, 1e-300), A, b, np.zeros_like(b))
    assert np.allclose(xr, xr2, rtol=0, atol=1e-280)


# ------------------------------------------------------------------- lanczos
@pytest.mark.parametrize("gamma", [[1.0], [0.1, 1.0]])
def test_lanczos_full_krylov_reproduces_eigh(gamma):
    """m = N: the Ritz values of T equal the eigenvalues of A p(A) (distinct spectrum)."""
    A = problems.laplacian_1d_dirichlet(10)
    N = A.shape[0]
    gamma = np.array(gamma)
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal(N), rng.standard_normal(N)
    V1, V2, (al1, be1, al2, be2), (H1, H2) = krylov.lanczos(A, a, b, N, gamma)
    Ad = A.toarray()
    lam = np.linalg.eigvalsh(Ad)
    ev = np.sort(lam * np.polyval(gamma, lam))
    assert np.allclose(np.sort(np.linalg.eigvalsh(H1)), ev, rtol=1e-8, atol=1e-8)
    assert np.allclose(np.sort(np.linalg.eigvalsh(H2)), ev, rtol=1e-8, atol=1e-8)


def test_lanczos_invariants_small_m():
    A = problems.laplacian_2d_dirichlet(20)
    N = A.shape[0]
    gamma = np.array(poly.polynomial_coeffs(1, 2.0 * 21 ** 2))
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(N), rng.standard_normal(N)
    m = 12
    V1, V2, (al1, be1, al2, be2), (H1, H2) = krylov.lanczos(A, a, b, m, gamma)
    assert np.abs(V1.T @ V1 - np.eye(m)).max() < 1e-10
    Ap = (A @ (gamma[0] * A + gamma[1] * sp.identity(N)))
    T = V1.T @ (Ap @ V1)
    assert np.abs(T - H1).max() < 1e-8 * np.abs(H1).max()
    assert be1[0] == pytest.approx(np.linalg.norm(a), rel=1e-15)


# ----------------------------------------------------- f(A)b end to end
def _dense_f(A, v, eta, alpha, which):
    w, Q = np.linalg.eigh(A.toarray())
    z = -eta * w ** (alpha / 2)
    f = np.exp(z) if which == 1 else np.expm1(z) / z
    return Q @ (f * (Q.T @ v))


@pytest.mark.parametrize("degree", [0, 2])
def test_fab_dense_small(degree):
    """m = N: nicksmethodcpu reproduces dense f(A)b to rounding."""
    n = 7
    A = problems.laplacian_2d_dirichlet(n)
    N = A.shape[0]
    lam = problems.dirichlet_eigs_2d(n)
    eta, alpha = 1e-3, 1.5
    c = oracle.contour_setup(lam.min(), lam.max(), 40, eta, alpha)
    gamma = oracle.polynomial_coeffs(degree, lam.max() / 4)
    gs, etaxi = oracle.shifted_gammas(gamma, c["xi"])
    rng = np.random.default_rng(4)
    a, b = rng.standard_normal(N), rng.standard_normal(N)
    v1, v2 = oracle.nicksmethodcpu(c["wts1_cpu"], c["wts2_cpu"], A, a, b, 40, N, c["s"],
                                   gamma, gs.real, gs.imag, etaxi)
    assert rel(v1, _dense_f(A, a, eta, alpha, 1)) < 1e-12
    assert rel(v2, _dense_f(A, b, eta, alpha, 2)) < 1e-12


def test_fab_dst_closed_form_dirichlet():
    """2D Dirichlet: exact f(A)b by the sine transform (startup_nicksmethod_CPU.m:59-77)."""
    n = 48
    A = problems.laplacian_2d_dirichlet(n)
    lam = problems.dirichlet_eigs_2d(n)
    eta, alpha = 1e-3, 1.5
    rng = np.random.default_rng(0)
    b = rng.standard_normal(n * n)
    B = b.reshape(n, n, order="F")
    Bh = sfft.dstn(B, type=1, norm="ortho")
    exact = sfft.idstn(np.exp(-eta * lam ** (alpha / 2)) * Bh, type=1, norm="ortho")
    exact = exact.reshape(-1, order="F")
    c = oracle.contour_setup(lam.min(), lam.max(), 64, eta, alpha)
    gs, etaxi = oracle.shifted_gammas([1.0], c["xi"])
    v1, _ = oracle.nicksmethodcpu(c["wts1_cpu"], c["wts2_cpu"], A, b, b, 64, 120, c["s"],
                                  [1.0], gs.real, gs.imag, etaxi)
    assert rel(v1, exact) < 1e-8


def test_step_dense_recurrence():
    """cpu_steps = u' = f1(A)u + dt f2(A) g(u), with deflation, against dense f."""
    from cases import oracle_setup
    from gpumatfun_amd import quadrature as q
    n = 6
    A = problems.laplacian_2d_dirichlet(n)
    N = A.shape[0]
    lamd = problems.dirichlet_eigs_2d(n)
    w, Qf = np.linalg.eigh(A.toarray())
    K = 3
    rng = np.random.default_rng(5)
    u0 = rng.uniform(0, 1, N)
    inp = q.make_inputs(A, u0, u0 * (1 - u0), m=N, n=48, eta=1e-2, alpha=1.5, dt=0.05,
                        l1=float(w[K]), lN=float(lamd.max()), degree=1, prob_dim=lamd.max() / 4,
                        Q=Qf[:, :K], lam=w[:K])
    inp.phi_lambda[:] = (np.expm1(-1e-2 * w[:K] ** 0.75) / (-1e-2 * w[:K] ** 0.75))
    S = oracle_setup(inp)
    u, gu = oracle.cpu_steps(S, inp.U[:, 0], inp.U[:, 1], 2)
    ud = u0.copy()
    for _ in range(2):
        ud = _dense_f(A, ud, 1e-2, 1.5, 1) + 0.05 * _dense_f(A, ud * (1 - ud), 1e-2, 1.5, 2)
    assert rel(u, ud) < 1e-9


def test_gpu_semantics_equals_cpu_path_without_clamp():
    from cases import dirichlet_case, oracle_setup
    inp = dirichlet_case(side=16, m=20, n=16)
    S = oracle_setup(inp)
    u, gu = oracle.cpu_steps(S, inp.U[:, 0], inp.U[:, 1], 3)
    G = oracle.GpuSemantics(inp.U, inp.A, inp.Q, inp.m, inp.gamma, inp.gamma_s_real,
                            inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda, inp.dt,
                            inp.mass, inp.wts1, inp.wts2, inp.shifts, clamp=False)
    for _ in range(3):
        G.step()
    assert rel(G.get_U()[:, 0], u) < 1e-13
    assert rel(G.get_U()[:, 1], inp.dt * gu) < 1e-13


@pytest.mark.parametrize("which", ["dirichlet", "poly", "fem"])
def test_c_restatement_matches_numpy_oracle(which):
    """oracle/c/fab_oracle.c against the NumPy oracle, and thread-count independence."""
    from cases import dirichlet_case, fem_case, oracle_setup
    from oracle import cfab
    inp = {"dirichlet": lambda: dirichlet_case(side=48, m=30, n=16),
           "poly": lambda: dirichlet_case(side=32, m=24, n=16, degree=3),
           "fem": lambda: fem_case(m=30)}[which]()
    S = oracle_setup(inp)
    u, g = oracle.cpu_steps_nonuniform(S, inp.U[:, 0], inp.U[:, 1], 2)
    u1, g1, _ = cfab.cpu_steps_c(S, inp.U[:, 0], inp.U[:, 1], 2, threads=1)
    u4, g4, _ = cfab.cpu_steps_c(S, inp.U[:, 0], inp.U[:, 1], 2, threads=4)
    assert np.array_equal(u1, u4) and np.array_equal(g1, g4)
    assert rel(u1, u) < 1e-10 and rel(g1, g) < 1e-10


def test_gpu_semantics_clamp_zeroes_negatives():
    from cases import dirichlet_case
    inp = dirichlet_case(side=16, m=20, n=16)
    G = oracle.GpuSemantics(inp.U, inp.A, inp.Q, inp.m, inp.gamma, inp.gamma_s_real,
                            inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda, inp.dt,
                            inp.mass, inp.wts1, inp.wts2, inp.shifts, clamp=True)
    G.step()
    assert np.all(G.get_U()[:, 0] >= 0) and np.any(G.get_U()[:, 0] == 0)


# ----------------------------------------------------------- Mitchell-Schaeffer
def test_ms_reaction_gate_closed_form_and_fixed_point():
    p = oracle.stepping.MS_DEFAULT
    dt = 0.05
    u = np.full(4, 0.05)                    # below v_gate: gate opens
    h = np.array([0.0, 0.2, 0.7, 1.0])
    h0 = h.copy()
    for _ in range(40):
        h, _ = oracle.ms_reaction(u, h, dt)
    assert np.allclose(h, 1 - (1 - h0) * (1 - dt / p["tau_open"]) ** 40, rtol=1e-13)
    u = np.full(2, 0.5)                     # above v_gate: gate closes
    h = np.array([1.0, 0.4])
    hc = h.copy()
    for _ in range(30):
        h, J = oracle.ms_reaction(u, h, dt)
    assert np.allclose(h, hc * (1 - dt / p["tau_close"]) ** 30, rtol=1e-13)
    # resting state (u = 0, h = 1) is a fixed point with zero current
    hn, J = oracle.ms_reaction(np.zeros(3), np.ones(3), dt)
    assert np.array_equal(hn, np.ones(3)) and np.array_equal(J, np.zeros(3))
    # u = 1: inward current vanishes, only -1/tau_out remains
    _, J = oracle.ms_reaction(np.ones(1), np.ones(1), dt)
    assert J[0] == pytest.approx(-1.0 / p["tau_out"], rel=1e-15)
