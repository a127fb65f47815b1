"""Test-problem matrices for the f(A)b engine (host set-up, NumPy/SciPy).

These build the sparse symmetric A and the deflation data that
``lanczos_create`` receives.  The reference takes them from ``laplacian.m``
(an unvendored File-Exchange helper called at ``problem_startup.m:42``) and
from ``process_nonuniform.m:12-111`` (FEM meshes: symrcm reordering,
A = M^{-1/2} K M^{-1/2}, mass = sqrt(diag(M)), eigenpairs via eigs).

All matrices are returned as int32-indexed CSR with sorted column indices.
"""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla
import scipy.spatial as sspatial
from scipy.sparse.csgraph import reverse_cuthill_mckee


def _csr(A):
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    A.indptr = A.indptr.astype(np.int32)
    A.indices = A.indices.astype(np.int32)
    return A


def laplacian_1d_dirichlet(n):
    """Tridiagonal [-1 2 -1], n interior points (unscaled)."""
    e = np.ones(n)
    return sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1], format="csr")


def laplacian_2d_dirichlet(n, scaled=True):
    """5-point Laplacian on an n x n interior grid of the unit square.

    Dirichlet boundary, natural ordering (index i + n*j).  With
    ``scaled=True`` the matrix is multiplied by 1/h^2, h = 1/(n+1).
    """
    L1 = laplacian_1d_dirichlet(n)
    I = sp.identity(n, format="csr")
    A = sp.kron(I, L1) + sp.kron(L1, I)
    if scaled:
        A = A * float((n + 1) ** 2)
    return _csr(A)


def dirichlet_eigs_2d(n, scaled=True):
    """Closed-form eigenvalues (4/h^2)(sin^2(j pi h/2) + sin^2(k pi h/2)).

    Same eigenvalues as ``startup_nicksmethod_CPU.m:60-61`` (there unscaled).
    Returned as the n x n grid lam[j, k], j, k = 1..n.
    """
    h = 1.0 / (n + 1)
    lx = 4.0 * np.sin(np.arange(1, n + 1) * np.pi * h / 2) ** 2
    lam = lx[:, None] + lx[None, :]
    if scaled:
        lam = lam / h ** 2
    return lam


def laplacian_3d_dirichlet(n, scaled=True):
    """7-point Laplacian on an n^3 interior grid of the unit cube."""
    L1 = laplacian_1d_dirichlet(n)
    I = sp.identity(n, format="csr")
    A = (sp.kron(sp.kron(I, I), L1) + sp.kron(sp.kron(I, L1), I)
         + sp.kron(sp.kron(L1, I), I))
    if scaled:
        A = A * float((n + 1) ** 2)
    return _csr(A)


def p1_stiffness_lumped_mass(nodes, elements):
    """P1 finite-element stiffness K and lumped (row-sum) mass diagonal.

    ``nodes`` n x 2, ``elements`` t x 3 (0-based).  K_ij = sum over
    triangles of (b_i b_j + c_i c_j)/(4|T|); M_ii = sum |T|/3.
    """
    nodes = np.asarray(nodes, dtype=np.float64)
    tri = np.asarray(elements, dtype=np.int64)
    p = nodes[tri]
    x = p[:, :, 0]
    y = p[:, :, 1]
    d1 = p[:, 1] - p[:, 0]
    d2 = p[:, 2] - p[:, 0]
    area = 0.5 * np.abs(d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0])
    b = np.stack([y[:, 1] - y[:, 2], y[:, 2] - y[:, 0], y[:, 0] - y[:, 1]], 1)
    c = np.stack([x[:, 2] - x[:, 1], x[:, 0] - x[:, 2], x[:, 1] - x[:, 0]], 1)
    Ke = (b[:, :, None] * b[:, None, :] + c[:, :, None] * c[:, None, :]) / (4 * area[:, None, None])
    N = nodes.shape[0]
    I = np.repeat(tri, 3, axis=1).ravel()
    J = np.tile(tri, (1, 3)).ravel()
    K = sp.csr_matrix((Ke.ravel(), (I, J)), shape=(N, N))
    K.sum_duplicates()
    mdiag = np.bincount(tri.ravel(), weights=np.repeat(area / 3.0, 3), minlength=N)
    return K, mdiag


def disc_mesh(n_nodes, n_boundary=4096, seed=0):
    """Seeded Delaunay triangulation of the unit disc.

    ``n_nodes - n_boundary`` uniform random interior points (radius < 1 -
    half the boundary spacing) plus ``n_boundary`` equispaced boundary
    points.  Returns (nodes, elements).
    """
    rng = np.random.default_rng(seed)
    n_int = n_nodes - n_boundary
    margin = 0.5 * 2 * np.pi / n_boundary
    r = np.sqrt(rng.uniform(0.0, 1.0, n_int)) * (1.0 - margin)
    th = rng.uniform(0.0, 2 * np.pi, n_int)
    tb = 2 * np.pi * np.arange(n_boundary) / n_boundary
    nodes = np.vstack([np.c_[r * np.cos(th), r * np.sin(th)], np.c_[np.cos(tb), np.sin(tb)]])
    tri = sspatial.Delaunay(nodes).simplices
    return nodes, tri


def fem_operator(K, mdiag, reorder=True):
    """A = M^{-1/2} K M^{-1/2} with symrcm reordering, ``process_nonuniform.m:15-20,90-91``.

    Returns (A, mass, perm) where mass = sqrt(diag(M)) in the new order and
    ``perm`` maps new index -> old index.
    """
    K = sp.csr_matrix(K)
    N = K.shape[0]
    if reorder:
        perm = reverse_cuthill_mckee(K, symmetric_mode=True).astype(np.int64)
    else:
        perm = np.arange(N)
    K = K[perm][:, perm]
    mass = np.sqrt(np.asarray(mdiag, dtype=np.float64)[perm])
    Dinv = sp.diags(1.0 / mass)
    A = Dinv @ K @ Dinv
    # symmetrise exactly: the product above is symmetric up to rounding order
    A = 0.5 * (A + A.T)
    return _csr(A), mass, perm


def fem_disc(n_nodes, n_boundary=4096, seed=0):
    """FEM disc problem: (A, mass, nodes) in symrcm order."""
    nodes, tri = disc_mesh(n_nodes, n_boundary, seed)
    K, mdiag = p1_stiffness_lumped_mass(nodes, tri)
    A, mass, perm = fem_operator(K, mdiag)
    return A, mass, nodes[perm]


def smallest_eigenpairs(A, k, sigma=None):
    """k smallest eigenpairs of symmetric A (shift-invert Lanczos via ARPACK).

    Host-side stand-in for ``eigs(K, M, k, 'sm')`` (``process_nonuniform.m:38,106``).
    Returns (lam ascending, Q with orthonormal columns).
    """
    if sigma is None:
        sigma = -1e-3 * max(1.0, abs(A.diagonal()).min())
    lam, Q = sla.eigsh(A, k=k, sigma=sigma, which="LM")
    order = np.argsort(lam)
    return lam[order], np.ascontiguousarray(Q[:, order])


def largest_eigenvalue(A, tol=1e-10):
    """lambda_max of symmetric A (``eigs(K, M, 1, 'lm')``, ``process_nonuniform.m:40``)."""
    return float(sla.eigsh(A, k=1, which="LA", return_eigenvectors=False, tol=tol)[0])
