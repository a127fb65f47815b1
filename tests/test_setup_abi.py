"""The C++ set-up generator (gm_polynomial_coeffs, gm_contour_setup,
gm_shift_setup, gm_eig_weights, gm_subspace_size) against the oracle.

Host code only: runs without a GPU.  The contour is checked against the
oracle's 40-digit mpmath restatement of problem_startup.m:80-103 (itself
pinned by the scalar quadrature identity), the polynomial against the exact
rational least-squares coefficients, the shifts against the oracle's
synthetic division and packing, and the subspace size against numpy roots of
(x p(x))'.
"""

import ctypes as C

import numpy as np
import pytest

import oracle
from oracle import contour as oc
from gpumatfun_amd import _lib

_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def lib():
    if not _lib.os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    return _lib.load()


def _p(a):
    return a.ctypes.data_as(_dp)


@pytest.mark.parametrize("degree,dim", [(0, 1.0), (1, 2.0), (3, 2.0), (8, 2.9e9), (16, 3.0)])
def test_polynomial_coeffs(lib, degree, dim):
    g = np.zeros(degree + 1)
    assert lib.gm_polynomial_coeffs(degree, dim, _p(g)) == 0
    ref = np.array(oracle.polynomial_coeffs(degree, dim))   # exact, rounded once
    np.testing.assert_allclose(g, ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("l1,lN,n", [(19.7, 78940.0, 32), (64.24, 1.185e10, 64), (2.5, 32.5, 16)])
def test_contour_setup_vs_mpmath(lib, l1, lN, n):
    eta, alpha = 1e-3, 1.7
    xi = np.zeros(2 * n)
    dx = np.zeros(2 * n)
    w1 = np.zeros(2 * n)
    w2 = np.zeros(2 * n)
    s = C.c_double()
    assert lib.gm_contour_setup(l1, lN, n, eta, alpha, _p(xi), _p(dx), _p(w1), _p(w2),
                                C.byref(s)) == 0
    ref = oc.contour_setup(l1, lN, n, eta, alpha)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    # FP64 evaluation of xi = sqrt(l1 lN)(1/k + sn)/(1/k - sn) cancels in
    # 1/k - sn as k -> 1: at lN/l1 = 1.8e8 (k = 0.99985) any FP64 evaluation
    # of this formula is good to ~2e-12, at lN/l1 = 4e3 to ~2e-14.
    tol = 1e-11 if lN / l1 > 1e6 else 1e-13
    assert rel(xi[0::2] + 1j * xi[1::2], ref["xi"]) < tol
    assert rel(dx[0::2] + 1j * dx[1::2], ref["dxidt"]) < 10 * tol
    assert rel(w1[:n] + 1j * w1[n:], ref["wts1"]) < 10 * tol
    assert rel(w2[:n] + 1j * w2[n:], ref["wts2"]) < 10 * tol
    assert abs(s.value - ref["s"]) <= 1e-13 * abs(ref["s"])


def test_contour_rejects_bad_interval(lib):
    z = np.zeros(8)
    assert lib.gm_contour_setup(2.0, 1.0, 4, 1e-3, 1.5, _p(z), None, _p(z), _p(z), None) == 1
    assert lib.gm_contour_setup(0.0, 1.0, 4, 1e-3, 1.5, _p(z), None, _p(z), _p(z), None) == 1


@pytest.mark.parametrize("degree", [0, 2, 5])
def test_shift_setup(lib, degree):
    n, m = 24, 7
    ref_c = oc.contour_setup(3.0, 900.0, n, 1e-3, 1.5)
    xi = ref_c["xi"]
    gamma = np.array(oracle.polynomial_coeffs(degree, 225.0))
    nd = gamma.size
    xin = np.empty(2 * n)
    xin[0::2], xin[1::2] = xi.real, xi.imag
    gr = np.zeros(nd * n)
    gi = np.zeros(nd * n)
    ez = np.zeros(2 * n)
    sh = np.zeros(4 * m * n)
    assert lib.gm_shift_setup(_p(np.ascontiguousarray(gamma)), nd, _p(xin), n, m, _p(gr),
                              _p(gi), _p(ez), _p(sh)) == 0
    gs, etaxi = oc.shifted_gammas(gamma, xi)
    G = (gr + 1j * gi).reshape(n, nd).T          # (d+1) x n column-major
    np.testing.assert_allclose(G, gs, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(ez[0::2] + 1j * ez[1::2], etaxi, rtol=1e-13)
    np.testing.assert_array_equal(sh, oc.pack_shifts(ez[0::2] + 1j * ez[1::2], m))


def test_eig_weights(lib):
    lam = np.array([0.5, 3.0, 70.0, 1e4])
    el = np.zeros(4)
    pl = np.zeros(4)
    assert lib.gm_eig_weights(_p(lam), 4, 1e-3, 1.7, _p(el), _p(pl)) == 0
    z = -1e-3 * lam ** 0.85
    np.testing.assert_allclose(el, np.exp(z), rtol=1e-15)
    np.testing.assert_allclose(pl[1:], np.expm1(z[1:]) / z[1:], rtol=1e-15)
    assert pl[0] == 1.0                      # problem_startup.m:77


@pytest.mark.parametrize("degree,l1,lN", [(0, 1.0, 1e4), (2, 0.3, 40.0), (8, 64.0, 1.2e10)])
def test_subspace_size(lib, degree, l1, lN):
    gamma = np.array(oracle.polynomial_coeffs(degree, lN / 4.0))
    m = C.c_int32()
    assert lib.gm_subspace_size(_p(np.ascontiguousarray(gamma)), gamma.size, l1, lN, 1e-10,
                                C.byref(m)) == 0
    q = np.append(gamma, 0.0)
    crit = np.roots(np.polyder(q)) if degree > 0 else np.array([])
    crit = crit[np.abs(crit.imag) < 1e-9 * np.abs(crit)].real if crit.size else crit
    crit = crit[(crit > l1) & (crit < lN)] if crit.size else crit
    vals = np.concatenate([np.polyval(q, crit), np.polyval(q, [l1, lN])])
    r = np.sqrt(vals.max() / vals.min())
    ref = int(np.ceil(np.log(1e-10 / 2) / np.log((r - 1) / (r + 1))))
    assert m.value == ref


def _bandwidth(A):
    A = A.tocoo()
    return int(np.abs(A.row - A.col).max()) if A.nnz else 0


def test_rcm_order_is_a_permutation_and_bands_like_symrcm():
    """gm_rcm_order (symrcm of process_nonuniform.m:15): a permutation, a path graph
    in scrambled order comes back with bandwidth 1, a scrambled 2D grid within 2n of
    the diagonal, and the FEM disc's bandwidth within 10% of SciPy's reverse
    Cuthill-McKee (or below it); components are handled separately."""
    import scipy.sparse as sp
    from gpumatfun_amd import problems
    rng = np.random.default_rng(5)
    n = 200
    P = problems.laplacian_1d_dirichlet(n)
    s = rng.permutation(n)
    Ps = sp.csr_matrix(P[s][:, s])
    p = problems.rcm_order(Ps)
    assert np.array_equal(np.sort(p), np.arange(n))
    assert _bandwidth(Ps[p][:, p]) == 1
    g = 30
    G = problems.laplacian_2d_dirichlet(g)
    s = rng.permutation(g * g)
    Gs = sp.csr_matrix(G[s][:, s])
    p = problems.rcm_order(Gs)
    assert _bandwidth(Gs[p][:, p]) <= 2 * g
    # two components plus an isolated node
    B = sp.block_diag([P, G, sp.csr_matrix(np.ones((1, 1)))]).tocsr()
    p = problems.rcm_order(B)
    assert np.array_equal(np.sort(p), np.arange(B.shape[0]))
    nodes, tri = problems.disc_mesh(3000, 256, seed=3)
    K, mdiag = problems.p1_stiffness_lumped_mass(nodes, tri)
    A_native, _, _ = problems.fem_operator(K, mdiag, reorder="native")
    A_scipy, _, _ = problems.fem_operator(K, mdiag, reorder=True)
    assert _bandwidth(A_native) <= 1.1 * _bandwidth(A_scipy)


def _far_gather_fraction(A, T):
    """Fraction of a CSR matrix's gathers x[col] whose previous gather of the same
    column lies more than T rows earlier (or that are its first): the share a
    cache holding the last T rows' operands would miss."""
    A = A.tocsr()
    r = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    o = np.lexsort((r, A.indices))
    cs, rs = A.indices[o], r[o]
    d = np.r_[0, np.diff(rs)]
    d[np.r_[True, cs[1:] != cs[:-1]]] = 1 << 40
    return float((d > T).mean())


def test_locality_order_is_a_permutation_and_local():
    """gm_locality_order (recursive breadth-first bisection): a permutation; on the
    24^3 7-point grid with leaves of 512 nodes, fewer than 25% of the gathers miss
    a window of 512 rows, against 3/7 in natural order (the k +- 1 neighbours,
    576 rows away, come back too late) and over 75% in a scrambled order, from
    which the order recovers the same locality; components and an isolated node
    are kept; with leaf >= n the order is the identity."""
    import scipy.sparse as sp
    from gpumatfun_amd import problems
    g = 24
    A = problems.laplacian_3d_dirichlet(g).tocsr()
    n = A.shape[0]
    p = problems.locality_order(A, 512)
    assert np.array_equal(np.sort(p), np.arange(n))
    nat = _far_gather_fraction(A, 512)
    loc = _far_gather_fraction(A[p][:, p], 512)
    assert nat > 0.4 and loc < 0.25, (nat, loc)
    s = np.random.default_rng(7).permutation(n)
    As = sp.csr_matrix(A[s][:, s])
    ps = problems.locality_order(As, 512)
    assert np.array_equal(np.sort(ps), np.arange(n))
    assert _far_gather_fraction(As, 512) > 0.75
    assert _far_gather_fraction(As[ps][:, ps], 512) < 0.25
    B = sp.block_diag([A, problems.laplacian_2d_dirichlet(30), sp.csr_matrix(np.ones((1, 1)))]).tocsr()
    pb = problems.locality_order(B, 256)
    assert np.array_equal(np.sort(pb), np.arange(B.shape[0]))
    assert np.array_equal(problems.locality_order(A, n), np.arange(n))
