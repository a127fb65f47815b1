"""Host sparse Cholesky and the factorisation-based deflation eigensolver (no
GPU): gm_nd_order, gm_chol_*, gm_eigs_shift_invert and gm_lambda_max.

They replace eigs(K, M, k, 'sm') and eigs(K, M, 1, 'lm') of
process_nonuniform.m:38-40,106-110.  Pinned by dense LAPACK (numpy), SciPy's
ARPACK shift-invert (the Python drivers' earlier path), the full-Krylov case
(a basis of all n vectors makes the Ritz values the exact eigenvalues), and a
deflated step that must not depend on which eigensolver deflated it."""

import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from conftest import ROOT, rel
from gpumatfun_amd import _lib, eigs, problems, quadrature as q
from gpumatfun_amd.eigs import SparseCholesky, eigs_shift_invert, lambda_max, nd_order

_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


def _random_spd(n, seed, deg=6):
    """Random sparse symmetric positive definite matrix (random graph, weights
    of both signs, diagonal dominant by a margin)."""
    rng = np.random.default_rng(seed)
    i = rng.integers(0, n, n * deg)
    j = rng.integers(0, n, n * deg)
    w = rng.standard_normal(n * deg)
    B = sp.coo_matrix((w, (i, j)), shape=(n, n)).tocsr()
    B = B + B.T
    B.setdiag(0)
    B.eliminate_zeros()
    d = np.asarray(abs(B).sum(axis=1)).ravel() + 0.5 + rng.random(n)
    return sp.csr_matrix(B + sp.diags(d))


@pytest.fixture(scope="module")
def disc3k():
    A, mass, nodes = problems.fem_disc(3000, 256, 3)
    return A


@pytest.mark.parametrize("which", ["grid", "disc", "random"])
def test_nd_order_is_a_permutation(which, disc3k):
    A = {"grid": lambda: problems.laplacian_2d_dirichlet(37), "disc": lambda: disc3k,
         "random": lambda: _random_spd(1500, 1)}[which]()
    p = nd_order(A)
    assert np.array_equal(np.sort(p), np.arange(A.shape[0]))


def test_nd_order_fill_below_natural_and_rcm():
    """Nested dissection's point: on a 2D grid the factor of the dissected
    matrix is much sparser than in natural (banded) order."""
    A = sp.csr_matrix(problems.laplacian_2d_dirichlet(60))
    F = SparseCholesky(A, sigma=-1.0)
    nnz_nd = F.info()["nnz_l"]
    L = np.linalg.cholesky((A + sp.identity(A.shape[0])).toarray())
    nnz_band = int(np.count_nonzero(np.abs(L) > 0))
    assert nnz_nd < 0.5 * nnz_band


@pytest.mark.parametrize("which,sigma", [("grid", -2.0), ("disc", None), ("random", 0.0)])
def test_solve_matches_dense(which, sigma, disc3k):
    A = {"grid": lambda: problems.laplacian_2d_dirichlet(33), "disc": lambda: disc3k,
         "random": lambda: _random_spd(1200, 2)}[which]()
    if sigma is None:
        sigma = eigs.default_shift(A)
    n = A.shape[0]
    M = (A - sigma * sp.identity(n)).toarray()
    B = np.random.default_rng(5).standard_normal((n, 7))
    F = SparseCholesky(A, sigma=sigma, threads=3)
    X = F.solve(B)
    Xd = np.linalg.solve(M, B)
    cond = np.linalg.cond(M)
    assert rel(X, Xd) <= 1e-15 * cond * 10
    # backward error at rounding: ||M X - B|| / (||M|| ||X||)
    be = np.linalg.norm(M @ X - B) / (np.linalg.norm(M, 2) * np.linalg.norm(X))
    assert be <= 1e-15
    assert np.array_equal(F.solve(B[:, 3]), X[:, 3]) or rel(F.solve(B[:, 3]), X[:, 3]) <= 1e-14


def test_symbolic_structure_covers_the_fill(disc3k):
    """The stored supernodal structure holds every nonzero of the dense
    Cholesky factor of the permuted matrix (the count may exceed it by the
    explicit zeros that merging supernodes adds)."""
    A = disc3k
    sigma = eigs.default_shift(A)
    F = SparseCholesky(A, sigma=sigma)
    p = F.perm()
    n = A.shape[0]
    Mp = (A - sigma * sp.identity(n)).toarray()[np.ix_(p, p)]
    L = np.linalg.cholesky(Mp)
    nnz_true = int(np.count_nonzero(np.abs(L) > 1e-300))
    inf = F.info()
    assert inf["nnz_l"] >= nnz_true
    assert inf["nnz_l"] <= 1.6 * nnz_true
    assert inf["supernodes"] > 1 and inf["flops"] > 0


def test_results_do_not_depend_on_thread_count(disc3k):
    A = disc3k
    B = np.random.default_rng(2).standard_normal((A.shape[0], 5))
    X1 = SparseCholesky(A, sigma=-1.0, threads=1).solve(B)
    X6 = SparseCholesky(A, sigma=-1.0, threads=6).solve(B)
    assert np.array_equal(X1, X6)
    l1, Q1 = eigs_shift_invert(A, 8, sigma=-1.0, threads=1)
    l5, Q5 = eigs_shift_invert(A, 8, sigma=-1.0, threads=5)
    assert np.array_equal(l1, l5) and np.array_equal(Q1, Q5)


_ISA_SCRIPT = r"""
import sys, numpy as np
sys.path.insert(0, {root!r})
from gpumatfun_amd import problems
from gpumatfun_amd.eigs import SparseCholesky
A = problems.laplacian_2d_dirichlet(41)
B = np.random.default_rng(4).standard_normal((A.shape[0], 9))
np.save(sys.argv[1], SparseCholesky(A, sigma=-3.0).solve(B))
"""


def test_instruction_set_variants_agree(tmp_path):
    """The dense kernel is chosen by the host CPU (AVX-512, AVX2 or plain
    x86-64); each variant gives the same solve to rounding."""
    script = tmp_path / "isa.py"
    script.write_text(_ISA_SCRIPT.format(root=ROOT))
    out = {}
    for isa in ("native", "avx2", "generic"):
        env = dict(os.environ)
        if isa != "native":
            env["GM_CHOL_ISA"] = isa
        f = tmp_path / f"{isa}.npy"
        subprocess.run([sys.executable, str(script), str(f)], check=True, env=env, timeout=120)
        out[isa] = np.load(f)
    assert rel(out["avx2"], out["native"]) <= 1e-13
    assert rel(out["generic"], out["native"]) <= 1e-13


def test_not_positive_definite_is_reported():
    A = problems.laplacian_2d_dirichlet(20)
    lam = problems.dirichlet_eigs_2d(20)
    with pytest.raises(_lib.GpuMatfunError) as e:
        SparseCholesky(A, sigma=float(lam.min()) * 1.5)
    assert "positive definite" in str(e.value)
    assert e.value.status == _lib.GM_EBREAKDOWN


def test_bad_arguments_are_rejected():
    lib = _lib.load()
    ptr = np.array([0, 1, 2], dtype=np.int32)
    ind = np.array([0, 5], dtype=np.int32)          # column out of range
    val = np.ones(2)
    h = C.c_void_p()
    assert lib.gm_chol_factor(2, ptr.ctypes.data_as(_ip), ind.ctypes.data_as(_ip),
                              val.ctypes.data_as(_dp), 0.0, 1, C.byref(h)) == _lib.GM_EINVAL
    assert lib.gm_chol_factor(0, ptr.ctypes.data_as(_ip), ind.ctypes.data_as(_ip),
                              val.ctypes.data_as(_dp), 0.0, 1, C.byref(h)) == _lib.GM_EINVAL
    assert lib.gm_chol_solve(None, 1, val.ctypes.data_as(_dp)) == _lib.GM_EBADHANDLE
    with pytest.raises(_lib.GpuMatfunError):
        eigs_shift_invert(problems.laplacian_2d_dirichlet(4), 9, block_size=8)


def test_full_krylov_space_gives_exact_ritz_values():
    """With the basis grown to all n vectors the Rayleigh-Ritz values are the
    eigenvalues of A itself: block Lanczos, the small symmetric
    eigensolver (Householder + implicit QL) and the shift inversion are all
    checked against LAPACK at rounding level."""
    A = _random_spd(96, 9)
    w = np.linalg.eigvalsh(A.toarray())
    lam, Q, info = eigs_shift_invert(A, 12, sigma=0.0, block_size=16, max_blocks=6,
                                     max_restarts=0, tol=1e-10, info=True)
    assert info["blocks"] <= 6
    assert np.abs(lam - w[:12]).max() <= 1e-12 * w[-1]
    assert np.abs(Q.T @ Q - np.eye(12)).max() <= 1e-12
    assert np.abs(A @ Q - Q * lam).max() <= 1e-11 * w[-1]


def _sin_angle(V, Q):
    """Sine of the largest principal angle between span(V) and span(Q)
    (orthonormal columns): ||(I - V V^T) Q||_2, free of the cancellation in
    sqrt(1 - cos^2)."""
    return float(np.linalg.norm(Q - V @ (V.T @ Q), 2))


def _first_gap(lam, k, rtol=1e-3):
    """Largest k' <= k with a relative gap after lambda_k' (subspaces are
    compared only where the spectrum separates them)."""
    for kk in range(k, 0, -1):
        if lam[kk] - lam[kk - 1] > rtol * abs(lam[kk]):
            return kk
    return 0


def test_smallest_eigenpairs_match_dense_on_a_sliver_mesh(disc3k):
    """3000-node disc (lambda_max / lambda_21 ~ 1e6): eigenvalues to rounding
    of lambda_max, the subspace up to a spectral gap, residuals and
    orthonormality, against LAPACK's dense eigh."""
    A = disc3k
    w, V = np.linalg.eigh(A.toarray())
    lam, Q, info = eigs_shift_invert(A, 21, info=True)
    assert info["max_residual"] <= 1e-12
    assert np.abs(lam - w[:21]).max() <= 1e-13 * w[-1]
    k = _first_gap(w, 20)
    assert k >= 10
    assert _sin_angle(V[:, :k], Q[:, :k]) <= 1e-9
    assert np.abs(Q.T @ Q - np.eye(21)).max() <= 1e-12
    assert np.linalg.norm(A @ Q - Q * lam, axis=0).max() <= 1e-13 * w[-1]
    assert lambda_max(A) == pytest.approx(w[-1], rel=1e-12)


def test_matches_arpack_on_a_20k_node_disc():
    A, mass, nodes = problems.fem_disc(20000, 1024, 0)
    lam, Q = eigs_shift_invert(A, 21)
    lam_a, Q_a = problems.smallest_eigenpairs(A, 21)
    lmax = lambda_max(A)
    assert lmax == pytest.approx(problems.largest_eigenvalue(A), rel=1e-11)
    assert np.abs(lam - lam_a).max() <= 1e-12 * lmax
    k = _first_gap(lam_a, 20)
    assert k >= 10
    assert _sin_angle(Q_a[:, :k], Q[:, :k]) <= 1e-8


def test_deflated_step_does_not_depend_on_the_eigensolver():
    """One oracle step of the FEM disc with K = 6 deflation pairs from this
    eigensolver and from ARPACK: the f(A)b outputs agree far below the step's
    own accuracy (the deflated subspace, not the basis, is what matters)."""
    A, mass, nodes = problems.fem_disc(3000, 256, 3)
    K = 6
    lN = lambda_max(A)
    x, y = nodes[:, 0], nodes[:, 1]
    xh = 0.5 * np.exp(-((x - 0.3) ** 2 + (y + 0.1) ** 2) * 4.0)
    u0, g0 = mass * xh, mass * xh * (1 - xh)
    outs = []
    for lam, Q in (eigs_shift_invert(A, K + 1), problems.smallest_eigenpairs(A, K + 1)):
        inp = q.make_inputs(A, u0, g0, m=30, n=32, eta=1e-3, alpha=1.7, dt=1e-3,
                            l1=float(lam[K]), lN=lN, degree=2, prob_dim=lN / 4.0,
                            Q=Q[:, :K], lam=lam[:K], mass=mass)
        from cases import oracle_setup
        outs.append(oracle.cpu_steps_nonuniform(oracle_setup(inp), inp.U[:, 0], inp.U[:, 1], 1))
    assert rel(outs[0][0], outs[1][0]) <= 1e-10
    assert rel(outs[0][1], outs[1][1]) <= 1e-10
