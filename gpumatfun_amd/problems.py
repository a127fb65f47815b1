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


def pencil_order(n, tile=16):
    """Node order of an n^3 grid in pencils: the (i, j) plane is cut into tile x tile
    squares, each square's column of n planes is one contiguous range (k outer,
    (i, j) inner), so a node's k +- 1 neighbours are tile^2 positions away instead
    of n^2.  Returns old_of_new: new position q holds natural node i + n j + n^2 k.
    ``tile`` must divide n."""
    if n % tile:
        raise ValueError("tile must divide n")
    t = n // tile
    q = np.arange(n ** 3)
    tid, rem = np.divmod(q, tile * tile * n)
    k, w = np.divmod(rem, tile * tile)
    jw, iw = np.divmod(w, tile)
    tj, ti = np.divmod(tid, t)
    return (ti * tile + iw) + n * (tj * tile + jw) + n * n * k


def gaussian_random_field(n, corr_len=0.1, seed=1):
    """Zero-mean, unit-variance stationary GRF on the n x n interior grid of the unit square.

    Squared-exponential covariance exp(-r^2 / (2 l^2)), synthesised spectrally:
    white noise (``default_rng(seed).standard_normal``) on a periodic grid of
    twice the size (to suppress wrap-around), filtered by the square root of the
    covariance's spectral density, cropped and normalised to unit sample variance.
    Returned in natural ordering (index i + n*j).
    """
    h = 1.0 / (n + 1)
    P = 2 * n
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((P, P))
    k = 2 * np.pi * np.fft.fftfreq(P, d=h)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    spec = np.exp(-0.25 * corr_len ** 2 * (kx ** 2 + ky ** 2))   # sqrt of the SE density
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * spec))[:n, :n]
    f = (f - f.mean()) / f.std()
    return f.reshape(-1, order="F")


def varcoef_laplacian_2d(kappa, n):
    """-div(kappa grad) on the n x n interior grid (Dirichlet), 5-point, scaled 1/h^2.

    Edge conductance (kappa_i + kappa_k)/2 between grid neighbours; an edge to
    the boundary uses kappa_i.  With this model the matrix depends on kappa as
        A_ik = -(kappa_i + kappa_k) / (2 h^2),   row sum_k A_ik = nb_i kappa_i / h^2
    (nb_i = number of boundary edges of node i), which is what the adjoint's
    d A / d kappa uses.
    """
    kappa = np.asarray(kappa, dtype=np.float64).reshape(n, n, order="F")
    h2 = float((n + 1) ** 2)
    idx = np.arange(n * n).reshape(n, n, order="F")
    rows, cols, vals = [], [], []
    diag = np.zeros((n, n))
    for axis in (0, 1):
        a = [slice(None), slice(None)]
        b = [slice(None), slice(None)]
        a[axis] = slice(0, n - 1)
        b[axis] = slice(1, n)
        ke = 0.5 * (kappa[tuple(a)] + kappa[tuple(b)]) * h2
        ia = idx[tuple(a)].ravel(order="F")
        ib = idx[tuple(b)].ravel(order="F")
        kv = ke.ravel(order="F")
        rows += [ia, ib]
        cols += [ib, ia]
        vals += [-kv, -kv]
        diag[tuple(a)] += ke
        diag[tuple(b)] += ke
        # boundary edges (both ends of the axis)
        lo = [slice(None), slice(None)]
        hi = [slice(None), slice(None)]
        lo[axis] = 0
        hi[axis] = n - 1
        diag[tuple(lo)] += kappa[tuple(lo)] * h2
        diag[tuple(hi)] += kappa[tuple(hi)] * h2
    rows.append(idx.ravel(order="F"))
    cols.append(idx.ravel(order="F"))
    vals.append(diag.ravel(order="F"))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n * n, n * n))
    return _csr(A)


def anisotropic_9pt(n, ratio=9.0, theta_max=np.pi / 3, length=1.0):
    """-div(D grad) on an n x n Dirichlet grid with a rotating fibre tensor, 9-point stencil.

    D(x) = R(theta) diag(ratio, 1) R(theta)^T with theta = theta_max * x / length
    (fibre angle rotating linearly across x).  Assembled as bilinear (Q1)
    finite elements on the uniform grid with D evaluated at cell centres, which
    gives a symmetric positive definite 9-point matrix for any SPD D; scaled so
    that it is the stiffness matrix divided by h^2 (a lumped-mass-free
    finite-difference analogue).  Returns A (CSR).
    """
    h = length / (n + 1)
    # Q1 element stiffness of a constant tensor [[a, b], [b, c]] on a unit square
    # cell is a*Ka + b*Kb + c*Kc; local nodes (0,0), (1,0), (0,1), (1,1);
    # 2x2 Gauss quadrature (exact for the bilinear gradients).
    gp = np.array([-1, 1]) / np.sqrt(3.0)
    Ka = np.zeros((4, 4))
    Kb = np.zeros((4, 4))
    Kc = np.zeros((4, 4))
    for s in gp:
        for t in gp:
            xi, et = 0.5 * (s + 1), 0.5 * (t + 1)
            dx = np.array([-(1 - et), (1 - et), -et, et])
            dy = np.array([-(1 - xi), -xi, (1 - xi), xi])
            Ka += 0.25 * np.outer(dx, dx)
            Kb += 0.25 * (np.outer(dx, dy) + np.outer(dy, dx))
            Kc += 0.25 * np.outer(dy, dy)
    N1 = n + 2                                   # node grid including the boundary
    ci, cj = np.meshgrid(np.arange(N1 - 1), np.arange(N1 - 1), indexing="ij")
    ci = ci.ravel()
    cj = cj.ravel()
    th = theta_max * (ci + 0.5) * h / length
    c2, s2 = np.cos(th) ** 2, np.sin(th) ** 2
    a = ratio * c2 + s2
    b = (ratio - 1.0) * np.cos(th) * np.sin(th)
    c = ratio * s2 + c2
    Ke = a[:, None, None] * Ka + b[:, None, None] * Kb + c[:, None, None] * Kc
    lx = np.stack([ci, ci + 1, ci, ci + 1], 1)
    ly = np.stack([cj, cj, cj + 1, cj + 1], 1)
    inside = (lx >= 1) & (lx <= n) & (ly >= 1) & (ly <= n)
    gidx = np.where(inside, (lx - 1) + n * (ly - 1), -1)
    R = np.repeat(gidx[:, :, None], 4, axis=2)
    Cc = np.repeat(gidx[:, None, :], 4, axis=1)
    keep = (R >= 0) & (Cc >= 0)
    A = sp.csr_matrix((Ke[keep] / h ** 2, (R[keep], Cc[keep])), shape=(n * n, n * n))
    A = 0.5 * (A + A.T)
    return _csr(A)


def anisotropic_9pt_ratio_derivative(n, theta_max=np.pi / 3, length=1.0):
    """dA/d(ratio) of ``anisotropic_9pt``: the operator is linear in the
    conductivity ratio (D = ratio f f^T + s s^T), so this is
    anisotropic_9pt(ratio=1) - anisotropic_9pt(ratio=0), returned on A's
    sparsity pattern (same indptr and indices for every ratio)."""
    A1 = anisotropic_9pt(n, ratio=1.0, theta_max=theta_max, length=length)
    A0 = anisotropic_9pt(n, ratio=0.0, theta_max=theta_max, length=length)
    assert np.array_equal(A1.indptr, A0.indptr) and np.array_equal(A1.indices, A0.indices)
    E = A1.copy()
    E.data = A1.data - A0.data
    return E


def uneven_batch_sides(count=64, lo=32, hi=448, seed=0):
    """Side lengths drawn log-uniformly in [lo, hi] (integers), seeded."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)


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


def rcm_order(A):
    """Reverse Cuthill-McKee order of A's (symmetric) pattern from the library's host
    set-up code (``gm_rcm_order``, in place of symrcm at ``process_nonuniform.m:15``):
    ``perm[q]`` is the old index placed at q."""
    import ctypes as C
    from gpumatfun_amd import _lib
    A = sp.csr_matrix(A)
    n = A.shape[0]
    ptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ind = np.ascontiguousarray(A.indices, dtype=np.int32)
    perm = np.empty(n, dtype=np.int32)
    ip = C.POINTER(C.c_int32)
    _lib.check(_lib.load().gm_rcm_order(n, ptr.ctypes.data_as(ip), ind.ctypes.data_as(ip),
                                        perm.ctypes.data_as(ip)))
    return perm.astype(np.int64)


def locality_order(A, leaf=4096):
    """Cache-locality order of A's (symmetric) pattern from the library's host code
    (``gm_locality_order``: recursive breadth-first bisection down to parts of at
    most ``leaf`` nodes): ``perm[q]`` is the old index placed at q."""
    import ctypes as C
    from gpumatfun_amd import _lib
    A = sp.csr_matrix(A)
    n = A.shape[0]
    ptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ind = np.ascontiguousarray(A.indices, dtype=np.int32)
    perm = np.empty(n, dtype=np.int32)
    ip = C.POINTER(C.c_int32)
    _lib.check(_lib.load().gm_locality_order(n, ptr.ctypes.data_as(ip), ind.ctypes.data_as(ip),
                                             int(leaf), perm.ctypes.data_as(ip)))
    return perm.astype(np.int64)


def fem_operator(K, mdiag, reorder=True):
    """A = M^{-1/2} K M^{-1/2} with symrcm reordering, ``process_nonuniform.m:15-20,90-91``.

    ``reorder``: True (SciPy's reverse Cuthill-McKee), "native" (the library's
    ``gm_rcm_order``) or False.  Returns (A, mass, perm) where mass = sqrt(diag(M))
    in the new order and ``perm`` maps new index -> old index.
    """
    K = sp.csr_matrix(K)
    N = K.shape[0]
    if reorder == "native":
        perm = rcm_order(K)
    elif reorder:
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
