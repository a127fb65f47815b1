"""The GPU deflation eigensolver (gm_eigs_smallest) against what the
mathematics fixes: the closed-form spectrum of the 2D Dirichlet Laplacian,
LAPACK's dense eigh on an anisotropic 9-point operator (eigenvalues, and the
spanned subspace through its orthogonal projector, which is unique even
inside a degenerate cluster), and, end to end, a deflated time step whose
result must not depend on which orthonormal eigenbasis deflated it."""

import numpy as np
import pytest

import oracle
from cases import oracle_setup
from conftest import rel
from gpumatfun_amd import problems, quadrature as q
from gpumatfun_amd.eigs import eigs_smallest
from gpumatfun_amd.engine import LanczosEngine

pytestmark = pytest.mark.gpu


def _check_pairs(A, lam, Q, tol):
    """Residuals, orthonormality; returns the max relative residual."""
    R = A @ Q - Q * lam
    res = np.linalg.norm(R, axis=0) / np.abs(lam)
    assert res.max() <= tol
    assert np.abs(Q.T @ Q - np.eye(Q.shape[1])).max() <= 1e-12
    return res.max()


@pytest.mark.parametrize("block,degree", [(16, 40), (32, 40), (64, 20)])
def test_dirichlet_closed_form(block, degree):
    side, k = 64, 7 if block == 16 else 12
    A = problems.laplacian_2d_dirichlet(side)
    exact = np.sort(problems.dirichlet_eigs_2d(side).ravel())[:k]
    print(f"case dirichlet side={side} k={k} block={block} degree={degree}", flush=True)
    lam, Q, info = eigs_smallest(A, k, block_size=block, degree=degree, tol=1e-10, info=True)
    print(f"passes {info['passes']} residual {info['max_residual']:.2e} "
          f"eig error {np.abs(lam - exact).max() / exact[0]:.2e}")
    assert np.abs(lam - exact).max() <= 1e-10 * exact[0]
    _check_pairs(A, lam, Q, 1e-10)


def test_anisotropic_vs_dense_eigh_subspace():
    """k = 10 at a gap of the spectrum (lambda_10 < lambda_11 by more than
    rounding), so the invariant subspace is well defined: ||P - P_ref|| is
    bounded by the residual over the gap."""
    A = problems.anisotropic_9pt(48)
    w, V = np.linalg.eigh(A.toarray())
    k = 10
    while w[k] - w[k - 1] < 1e-6 * w[k]:
        k -= 1
    print(f"case anisotropic 48^2 k={k}", flush=True)
    lam, Q, info = eigs_smallest(A, k, block_size=32, degree=40, tol=1e-11, seed=3, info=True)
    print(f"passes {info['passes']} residual {info['max_residual']:.2e}")
    assert np.abs(lam - w[:k]).max() <= 1e-10 * w[0]
    _check_pairs(A, lam, Q, 1e-11)
    P = Q @ Q.T
    Pr = V[:, :k] @ V[:, :k].T
    gap = (w[k] - w[k - 1]) / w[k - 1]
    assert np.abs(P - Pr).max() <= max(1e-9, 10 * 1e-11 / gap)


def test_deflated_step_independent_of_eigenbasis():
    """One deflated step of the plain engine with the GPU eigenpairs against
    the oracle deflated with ARPACK's: f(A) b does not depend on the basis,
    so the two agree to the Krylov floor of this case (about 5e-12)."""
    A = problems.laplacian_2d_dirichlet(40)
    lam_all = problems.dirichlet_eigs_2d(40)
    K = 6
    lam_g, Q_g = eigs_smallest(A, K, block_size=16, degree=40, tol=1e-12)
    lam_a, Q_a = problems.smallest_eigenpairs(A, K)
    assert np.abs(lam_g - lam_a).max() <= 1e-10 * lam_a[0]
    rng = np.random.default_rng(5)
    u0 = rng.random(A.shape[0])
    kw = dict(m=30, n=16, eta=1e-3, alpha=1.5, dt=1e-3, l1=float(lam_all.min()),
              lN=float(lam_all.max()))
    inp_g = q.make_inputs(A, u0, u0 * (1 - u0), Q=Q_g, lam=lam_g, **kw)
    inp_a = q.make_inputs(A, u0, u0 * (1 - u0), Q=Q_a, lam=lam_a, **kw)
    print(f"case deflated step n={A.shape[0]} K={K}", flush=True)
    eng = LanczosEngine.from_inputs(inp_g, clamp=False)
    eng.step(1)
    U = eng.get_U()
    eng.delete()
    u, gu = oracle.cpu_steps(oracle_setup(inp_a), inp_a.U[:, 0], inp_a.U[:, 1], 1)
    e0, e1 = rel(U[:, 0], u), rel(U[:, 1], inp_a.dt * gu)
    print(f"vs oracle with ARPACK basis: {e0:.2e} {e1:.2e}")
    assert e0 < 1e-9 and e1 < 1e-9


def test_not_converged_reports_breakdown():
    A = problems.laplacian_2d_dirichlet(32)
    from gpumatfun_amd._lib import GpuMatfunError
    with pytest.raises(GpuMatfunError, match="not converged"):
        eigs_smallest(A, 8, block_size=16, degree=4, max_passes=1, tol=1e-14)
