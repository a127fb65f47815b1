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


def test_c_restatement_variants_and_long_double():
    """The measurement switches of oracle/c/fab_oracle.c: variant 0 is the literal
    restatement bit for bit; on a case whose answer FP64 fixes to ~1e-14 every
    arithmetic variant and the long-double build agree with it to 1e-12, and full
    reorthogonalisation with m = N reproduces dense f(A)b; the long-double build's
    rounding is visibly smaller (it agrees with a 40-digit dense evaluation more
    closely than the double build does)."""
    from cases import dirichlet_case, headline_parameters_case, oracle_setup
    from oracle import cfab
    inp = dirichlet_case(side=32, m=24, n=16, degree=3)
    S = oracle_setup(inp)
    u0, g0 = inp.U[:, 0], inp.U[:, 1]
    base = cfab.cpu_steps_c(S, u0, g0, 2, threads=2)[:2]
    again = cfab.cpu_steps_c(S, u0, g0, 2, threads=2, variant=0)[:2]
    assert np.array_equal(base[0], again[0]) and np.array_equal(base[1], again[1])
    for v in (cfab.V_SCALE, cfab.V_FMA, cfab.V_TREE, cfab.V_DEFL, cfab.V_REORTH,
              cfab.V_SCALE | cfab.V_FMA | cfab.V_TREE | cfab.V_DEFL):
        r = cfab.cpu_steps_c(S, u0, g0, 2, threads=2, variant=v)[:2]
        assert max(rel(r[0], base[0]), rel(r[1], base[1])) < 1e-12, v
    ext = cfab.cpu_steps_c(S, u0, g0, 2, threads=2, extended=True)[:2]
    assert max(rel(ext[0], base[0]), rel(ext[1], base[1])) < 1e-12
    for v in (cfab.V_RDEFL, cfab.V_RLANCZOS, cfab.V_RSPMV):
        r = cfab.cpu_steps_c(S, u0, g0, 2, threads=2, variant=v, extended=True)[:2]
        assert max(rel(r[0], ext[0]), rel(r[1], ext[1])) < 1e-12, v
    # deflation on: the headline parameters on a small grid
    hp = headline_parameters_case(side=32, m=40)
    Sh = oracle_setup(hp)
    b = cfab.cpu_steps_c(Sh, hp.U[:, 0], hp.U[:, 1], 1, threads=2)[:2]
    e = cfab.cpu_steps_c(Sh, hp.U[:, 0], hp.U[:, 1], 1, threads=2, extended=True)[:2]
    r = cfab.cpu_steps_c(Sh, hp.U[:, 0], hp.U[:, 1], 1, threads=2, variant=cfab.V_REORTH)[:2]
    assert rel(e[0], b[0]) < 1e-12 and rel(r[0], b[0]) < 1e-12
    # one f(A)b with the 1D Laplacian on 30 points (distinct eigenvalues, so m = N
    # Lanczos steps span the whole space): with full reorthogonalisation it is exact
    # up to the quadrature; the long-double build's rounding is far below the double
    # build's against the same 40-digit evaluation of the same quadrature
    import mpmath as mp
    from gpumatfun_amd import quadrature as q
    N = 30
    A = sp.csr_matrix(problems.laplacian_1d_dirichlet(N) * float((N + 1) ** 2))
    A.sort_indices()
    lam = 4.0 * np.sin(np.arange(1, N + 1) * np.pi / (2 * (N + 1))) ** 2 * (N + 1) ** 2
    rng = np.random.default_rng(12)
    a = rng.standard_normal(N)
    tiny = q.make_inputs(A, a, a, m=N, n=24, eta=1e-2, alpha=1.5, dt=1e-3,
                         l1=float(lam.min()), lN=float(lam.max()), degree=0)
    St = oracle_setup(tiny)
    # u = f1(A) a + dt f2(A) a (g = a, so neither Krylov run starts from zero)
    vd = cfab.cpu_steps_c(St, a, a, 1, threads=1, variant=cfab.V_REORTH)[0]
    vl = cfab.cpu_steps_c(St, a, a, 1, threads=1, variant=cfab.V_REORTH, extended=True)[0]
    mp.mp.dps = 40
    w, V = mp.eigsy(mp.matrix(A.toarray()))
    xi, s_ = tiny.extra["xi"], tiny.extra["s"]
    w1, w2 = tiny.extra["wts1_cpu"] * s_, tiny.extra["wts2_cpu"] * s_
    r = lambda wt, x: mp.im(sum(mp.mpc(wt[j]) / (mp.mpc(xi[j]) - x) for j in range(len(xi))))
    fl = [r(w1, w[k]) + mp.mpf(tiny.dt) * r(w2, w[k]) for k in range(N)]
    c = V.T * mp.matrix(a.tolist())
    ref = np.array([float(sum(V[i, k] * fl[k] * c[k] for k in range(N))) for i in range(N)])
    ed, el = rel(vd, ref), rel(vl, ref)
    assert ed < 1e-13 and el < 0.05 * ed, (ed, el)


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


# --------------------------------------------- monodomain step loops (FHN, MS)
# Each loop is pinned against the exponential-Euler recurrence of its reaction
# model evaluated with dense matrix functions (eigh) on a tiny grid.  With m = N
# Lanczos steps the Krylov approximation is exact, so the oracle's whole chain
# (deflation split, preconditioned Lanczos, contour solves, recombination,
# reaction update) must reproduce the dense recurrence to the quadrature's
# accuracy.  The planted variants show that the plausible slips -- the recovery
# variable or gate advanced with the previous u, the current evaluated with the
# previous gate -- move the result far beyond the tolerance, so the check
# discriminates them.

def _mono_problem(K=3, dt=0.05, eta=1e-2, alpha=1.5, seed=7, ode0=None):
    from gpumatfun_amd import quadrature as q
    n = 6
    A = problems.laplacian_2d_dirichlet(n)
    N = A.shape[0]
    lamd = problems.dirichlet_eigs_2d(n)
    w, Qf = np.linalg.eigh(A.toarray())
    rng = np.random.default_rng(seed)
    u0 = rng.uniform(0, 1, N)
    v0 = rng.uniform(0, 0.2, N) if ode0 is None else ode0
    inp = q.make_inputs(A, u0, u0, m=N, n=48, eta=eta, alpha=alpha, dt=dt, l1=float(w[K]),
                        lN=float(lamd.max()), degree=1, prob_dim=lamd.max() / 4,
                        Q=Qf[:, :K], lam=w[:K], ode=v0)
    z = -eta * w[:K] ** (alpha / 2)
    inp.phi_lambda[:] = np.expm1(z) / z      # exact phi (no phi(lambda_1) = 1 guard)
    return A, inp, u0, v0


def _dense_mono(A, u, v, steps, dt, eta, alpha, react):
    """u_{n+1} = exp(-eta A^(alpha/2)) u_n + dt phi(-eta A^(alpha/2)) J(u_n, v_n), with
    (v_{n+1}, J_{n+1}) = react(u_{n+1}, v_n)."""
    F1 = lambda x: _dense_f(A, x, eta, alpha, 1)
    F2 = lambda x: _dense_f(A, x, eta, alpha, 2)
    _, J = react(u, v, initial=True)
    for _ in range(steps):
        u = F1(u) + dt * F2(J)
        v, J = react(u, v)
    return u, J, v


def _fhn(eps, beta, gam, delta, a, dt, slip=None):
    """FitzHugh-Nagumo: du/dt = ... + u(1-u)(u-a) - v, dv/dt = eps(beta u - gam v - delta)."""
    def react(u, v, initial=False, u_prev=[None]):
        if initial:
            u_prev[0] = u
            return v, u * (1 - u) * (u - a) - v
        ua = u_prev[0] if slip == "old_u" else u
        vn = v + dt * eps * (beta * ua - gam * v - delta)
        u_prev[0] = u
        return vn, u * (1 - u) * (u - a) - vn
    return react


def test_fhn_step_loop_dense_recurrence():
    """cpu_steps_monodomain (problem_startup_CPU_monodomain.m:128-141) against the
    dense FitzHugh-Nagumo recurrence, with deflation and a degree-1 polynomial."""
    from cases import oracle_setup
    dt, eta, alpha = 0.05, 1e-2, 1.5
    eps, beta, gam, delta, a = 0.3, 0.5, 1.0, 0.02, 0.1
    A, inp, u0, v0 = _mono_problem(dt=dt, eta=eta, alpha=alpha)
    g0 = u0 * (1 - u0) * (u0 - a) - v0
    u, gu, v = oracle.cpu_steps_monodomain(oracle_setup(inp), u0, g0, v0, 3, a, eps, beta,
                                           gam, delta)
    ud, Jd, vd = _dense_mono(A, u0, v0, 3, dt, eta, alpha, _fhn(eps, beta, gam, delta, a, dt))
    e = max(rel(u, ud), rel(gu, Jd), rel(v, vd))
    assert e < 1e-9, e
    us, Js, vs = _dense_mono(A, u0, v0, 3, dt, eta, alpha,
                             _fhn(eps, beta, gam, delta, a, dt, slip="old_u"))
    assert max(rel(us, ud), rel(vs, vd)) > 1e3 * max(e, 1e-12)


def test_gpu_semantics_monodomain_dense_recurrence():
    """GpuSemantics.update_U_monodomain (update_U_monodomain.cu:5-43) in its GPU
    parametrisation (aode, bode, code, vrest, vamp, vpeak, c1ion) against the same
    dense FitzHugh-Nagumo recurrence: aode = eps beta, bode = -eps gam,
    code = -dt eps delta, vrest = a, vpeak = 1, c1ion / vamp^2 = -1."""
    dt, eta, alpha = 0.05, 1e-2, 1.5
    eps, beta, gam, delta, a = 0.3, 0.5, 1.0, 0.02, 0.1
    A, inp, u0, v0 = _mono_problem(dt=dt, eta=eta, alpha=alpha, seed=8)
    g0 = u0 * (1 - u0) * (u0 - a) - v0
    vamp = 2.0
    fhn = dict(aode=eps * beta, bode=-eps * gam, code=-dt * eps * delta, vrest=a, vamp=vamp,
               vpeak=1.0, c1ion=-vamp * vamp)
    G = oracle.GpuSemantics(np.column_stack([u0, g0]), inp.A, inp.Q, inp.m, inp.gamma,
                            inp.gamma_s_real, inp.gamma_s_imag, inp.exp_lambda, inp.phi_lambda,
                            dt, inp.mass, inp.wts1, inp.wts2, inp.shifts, ode=v0, fhn=fhn,
                            clamp=False)
    for _ in range(3):
        G.step(monodomain=True)
    ud, Jd, vd = _dense_mono(A, u0, v0, 3, dt, eta, alpha, _fhn(eps, beta, gam, delta, a, dt))
    U = G.get_U()
    e = max(rel(U[:, 0], ud), rel(U[:, 1], dt * Jd), rel(G.ode, vd))
    assert e < 1e-9, e


def _ms(p, dt, slip=None):
    """Mitchell-Schaeffer: dh/dt = (1-h)/tau_open (u < v_gate), -h/tau_close otherwise;
    J = h u^2 (1-u)/tau_in - u/tau_out."""
    def react(u, h, initial=False, u_prev=[None]):
        if initial:
            u_prev[0] = u
            return h, h * u * u * (1 - u) / p["tau_in"] - u / p["tau_out"]
        ug = u_prev[0] if slip == "old_u" else u
        rate = np.where(ug < p["v_gate"], (1 - h) / p["tau_open"], -h / p["tau_close"])
        hn = h + dt * rate
        hj = h if slip == "old_gate" else hn
        u_prev[0] = u
        return hn, hj * u * u * (1 - u) / p["tau_in"] - u / p["tau_out"]
    return react


def test_ms_step_loop_dense_recurrence():
    """cpu_steps_ms against the dense Mitchell-Schaeffer recurrence (gate advanced with
    the new u, current with the new gate), with u0 straddling v_gate so both gate
    branches are taken."""
    from cases import oracle_setup
    p = dict(oracle.stepping.MS_DEFAULT)
    p.update(tau_open=2.0, tau_close=3.0)     # gate moves visibly within 3 steps
    dt, eta, alpha = 0.05, 1e-2, 1.5
    A, inp, u0, _ = _mono_problem(dt=dt, eta=eta, alpha=alpha, seed=9)
    u0 = 0.26 * u0                             # values on both sides of v_gate = 0.13
    h0 = np.random.default_rng(3).uniform(0.3, 1.0, u0.size)
    u, gu, h = oracle.cpu_steps_ms(oracle_setup(inp), u0, h0, 3, p)
    ud, Jd, hd = _dense_mono(A, u0, h0, 3, dt, eta, alpha, _ms(p, dt))
    e = max(rel(u, ud), rel(gu, Jd), rel(h, hd))
    assert e < 1e-9, e
    assert np.any(u0 < p["v_gate"]) and np.any(u0 >= p["v_gate"])
    for slip in ("old_u", "old_gate"):
        us, Js, hs = _dense_mono(A, u0, h0, 3, dt, eta, alpha, _ms(p, dt, slip=slip))
        assert max(rel(us, ud), rel(Js, Jd), rel(hs, hd)) > 1e3 * max(e, 1e-12), slip
