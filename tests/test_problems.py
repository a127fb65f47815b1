"""Problem generators and host set-up (no GPU)."""

import os

import numpy as np
import pytest
import scipy.fft as sfft
import scipy.sparse as sp

from conftest import ROOT, rel
from gpumatfun_amd import problems, quadrature as q

GOLD = os.path.join(ROOT, "tests", "golden", "circle1.npz")


def test_dirichlet_2d_spectrum_closed_form():
    n = 12
    A = problems.laplacian_2d_dirichlet(n)
    w = np.sort(np.linalg.eigvalsh(A.toarray()))
    lam = np.sort(problems.dirichlet_eigs_2d(n).ravel())
    assert np.allclose(w, lam, rtol=1e-12)
    # sine modes diagonalise A (DST-I), startup_nicksmethod_CPU.m:59-77
    rng = np.random.default_rng(0)
    B = rng.standard_normal((n, n))
    Ab = (A @ B.reshape(-1, order="F")).reshape(n, n, order="F")
    via = sfft.idstn(problems.dirichlet_eigs_2d(n) * sfft.dstn(B, type=1, norm="ortho"),
                     type=1, norm="ortho")
    assert rel(Ab, via) < 1e-12


def test_laplacian_3d_size_and_rowsums():
    A = problems.laplacian_3d_dirichlet(8)
    assert A.shape == (512, 512)
    assert A.nnz == 512 * 7 - 6 * 64
    h2 = 9.0 ** 2
    assert np.allclose(A.diagonal(), 6 * h2)


def test_fem_assembly_matches_reference_mesh_fixture():
    """P1 stiffness + lumped mass of the reference's circle1 mesh (468 nodes):
    same sparsity, K agrees to 1e-5 relative (the stored K was assembled from
    coordinates that differ in the 6th digit), M to 2e-5."""
    d = np.load(GOLD, allow_pickle=False)
    nodes = d["nodes"][:, :2]
    el = d["elements"].astype(np.int64) - 1
    K, mdiag = problems.p1_stiffness_lumped_mass(nodes, el)
    N = nodes.shape[0]
    Kr = sp.csr_matrix((d["K_val"], (d["K_row"], d["K_col"])), shape=(N, N))
    assert sp.linalg.norm(K - Kr) / sp.linalg.norm(Kr) < 1e-5
    assert set(zip(*K.nonzero())) == set(zip(*Kr.nonzero()))
    assert np.max(np.abs(mdiag - d["M_diag"]) / d["M_diag"]) < 2e-5
    assert np.abs(np.asarray(K.sum(axis=1))).max() < 1e-12


def test_fem_operator_symmetric_and_deflation_pairs():
    d = np.load(GOLD, allow_pickle=False)
    el = d["elements"].astype(np.int64) - 1
    K, mdiag = problems.p1_stiffness_lumped_mass(d["nodes"][:, :2], el)
    A, mass, perm = problems.fem_operator(K, mdiag)
    assert abs(A - A.T).max() == 0.0
    lam, Q = problems.smallest_eigenpairs(A, 5)
    assert abs(lam[0]) < 1e-8                       # Neumann: constant null vector
    assert np.allclose(Q.T @ Q, np.eye(5), atol=1e-10)
    assert rel(A @ Q, Q * lam) < 1e-8
    # A = M^-1/2 K M^-1/2: mass-weighted constant is the null vector
    assert rel(A @ mass, np.zeros_like(mass) + 0.0) < 1e-12 or np.linalg.norm(A @ mass) < 1e-9


def test_disc_mesh_is_seeded_and_on_the_disc():
    n1, t1 = problems.disc_mesh(2000, 128, seed=7)
    n2, t2 = problems.disc_mesh(2000, 128, seed=7)
    assert np.array_equal(n1, n2) and np.array_equal(t1, t2)
    r = np.hypot(n1[:, 0], n1[:, 1])
    assert r.max() <= 1.0 + 1e-15
    assert np.sum(np.isclose(r, 1.0)) == 128


def test_varcoef_laplacian_constant_kappa_and_kappa_model():
    n = 16
    A1 = problems.varcoef_laplacian_2d(np.ones(n * n), n)
    assert abs(A1 - problems.laplacian_2d_dirichlet(n)).max() == 0.0
    kappa = np.exp(0.5 * problems.gaussian_random_field(n, 0.1, seed=1))
    A = problems.varcoef_laplacian_2d(kappa, n)
    assert abs(A - A.T).max() == 0.0
    # linear in kappa: A(2 kappa) = 2 A(kappa); off-diagonals -(k_i + k_j)/(2h^2)
    assert abs(problems.varcoef_laplacian_2d(2 * kappa, n) - 2 * A).max() < 1e-9
    h2 = (n + 1) ** 2
    assert A[0, 1] == pytest.approx(-(kappa[0] + kappa[1]) * h2 / 2, rel=1e-15)
    assert np.all(np.linalg.eigvalsh(A.toarray()) > 0)


def test_grf_is_seeded_unit_variance():
    a = problems.gaussian_random_field(64, 0.1, seed=1)
    b = problems.gaussian_random_field(64, 0.1, seed=1)
    assert np.array_equal(a, b)
    assert abs(a.mean()) < 1e-12 and abs(a.std() - 1) < 1e-12
    # correlation length: neighbouring values are strongly correlated (h = 1/65 << 0.1)
    g = a.reshape(64, 64, order="F")
    assert np.corrcoef(g[:-1].ravel(), g[1:].ravel())[0, 1] > 0.9


def test_anisotropic_9pt_spd_and_isotropic_limit():
    A = problems.anisotropic_9pt(24, ratio=9.0)
    assert abs(A - A.T).max() == 0.0
    assert np.linalg.eigvalsh(A.toarray()).min() > 0
    assert A.getnnz(axis=1).max() == 9
    # ratio 1 (isotropic): the Q1 Laplacian annihilates linear functions in the interior
    A1 = problems.anisotropic_9pt(10, ratio=1.0)
    x = np.tile(np.arange(1, 11, dtype=float), 10)        # linear in x
    r = (A1 @ x).reshape(10, 10, order="F")
    assert np.abs(r[1:-1, 1:-1]).max() < 1e-9


def test_make_inputs_shapes_follow_lanczos_create():
    A = problems.laplacian_2d_dirichlet(10)
    lam = problems.dirichlet_eigs_2d(10)
    u = np.ones(100) * 0.3
    inp = q.make_inputs(A, u, u * 0.7, m=12, n=9, eta=1e-2, alpha=1.5, dt=0.1,
                        l1=float(lam.min()), lN=float(lam.max()), degree=3)
    assert inp.U.shape == (100, 2)
    assert inp.gamma.shape == (4,)
    assert inp.gamma_s_real.shape == (4, 9) and inp.gamma_s_imag.shape == (4, 9)
    assert inp.wts1.shape == (9, 2) and inp.wts2.shape == (9, 2)
    assert inp.shifts.size == 4 * 12 * 9
    # weights carry s (problem_startup.m:100-101)
    w = inp.wts1[:, 0] + 1j * inp.wts1[:, 1]
    assert np.allclose(w, inp.extra["s"] * inp.extra["wts1_cpu"], rtol=1e-15)
