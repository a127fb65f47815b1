"""Exact-resolvent gradients of the contour-quadrature f(A)b (oracle, test infrastructure).

The reference has no adjoint; this is the FP64 CPU reference gradient the
HIP adjoint is checked against.  The forward map is the quadrature of
``problem_startup.m:95-103`` / ``nicksmethodcpu.m:17-43`` with resolvents
solved exactly (sparse LU) instead of by Lanczos:

    u(b, kappa, alpha) = Im sum_j c_j (xi_j I - A(kappa))^{-1} b,
    c_j = s f_alpha(xi_j) dxi/dt_j,   f_alpha(x) = exp(-tau x^(alpha/2)).

For an upstream gradient g = dL/du (A symmetric, so (xi - A)^T = xi - A):

    dL/db       = Im sum_j c_j z_j,                        z_j = (xi_j - A)^{-1} g
    dL/dkappa_i = Im sum_j c_j z_j^T (dA/dkappa_i) x_j,    x_j = (xi_j - A)^{-1} b
    dL/dalpha   = Im sum_j (dc_j/dalpha) g^T x_j,
    dc_j/dalpha = c_j * (-tau xi_j^(alpha/2) log(xi_j) / 2).

For a scalar parameter theta of A with derivative E = dA/dtheta (any
symmetric sparse matrix, e.g. the fibre part of an anisotropic conductivity
when theta is the conductivity ratio):

    dL/dtheta   = Im sum_j c_j z_j^T E x_j.

dA/dkappa_i follows the variable-coefficient model of
``gpumatfun_amd.problems.varcoef_laplacian_2d``: off-diagonals
-(kappa_i + kappa_k)/(2 h^2) and boundary terms linear in kappa_i, so
z^T (dA/dkappa_i) x = sum_{k != i} w_ik (z_i - z_k)(x_i - x_k) + s_i z_i x_i with
w_ik = -A_ik/(kappa_i + kappa_k) and s_i = (sum_k A_ik)/kappa_i.
"""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def resolvents(A, xi, rhs):
    """x_j = (xi_j I - A)^{-1} rhs for every node, as an n x len(xi) complex array."""
    A = sp.csc_matrix(A, dtype=np.complex128)
    n = A.shape[0]
    I = sp.identity(n, dtype=np.complex128, format="csc")
    out = np.empty((n, len(xi)), dtype=np.complex128)
    for j, x in enumerate(xi):
        out[:, j] = sla.splu(x * I - A).solve(np.asarray(rhs, dtype=np.complex128))
    return out


def forward(A, b, xi, c):
    X = resolvents(A, xi, b)
    return np.imag(X @ c), X


def kappa_edge_weights(A, kappa):
    """w (same pattern as A, zero diagonal) and s for z^T dA/dkappa_i x."""
    A = sp.csr_matrix(A)
    Ac = A.tocoo()
    off = Ac.row != Ac.col
    rows, cols = Ac.row[off], Ac.col[off]
    w = -Ac.data[off] / (kappa[rows] + kappa[cols])
    W = sp.csr_matrix((w, (rows, cols)), shape=A.shape)
    s = np.asarray(A.sum(axis=1)).ravel() / kappa
    return W, s


def krylov_resolvents(A, xi, rhs, m):
    """x_j ~= V ytilde_j, ytilde_j = beta0 (xi_j - T)^{-1} e1 from m Lanczos steps
    (``lanczos.m``, shifted solves as in ``nicksmethodcpu.m:18-21`` with unit
    weights): the Krylov approximation the HIP adjoint uses."""
    from .krylov import lanczos, shifted_solves
    V, _, (al, be, _, _), _ = lanczos(A, rhs, rhs, m, [1.0])
    Y = shifted_solves(al, be, np.asarray(xi), np.ones(len(xi)), 1.0)
    return V @ Y


def krylov_gradients(A, kappa, b, g, xi, c, dc_dalpha, m):
    """``gradients`` with the resolvents replaced by their m-step Krylov approximations."""
    return gradients(A, kappa, b, g, xi, c, dc_dalpha,
                     solver=lambda A_, xi_, r: krylov_resolvents(A_, xi_, r, m))


def gradients(A, kappa, b, g, xi, c, dc_dalpha, solver=None):
    """(dL/db, dL/dkappa, dL/dalpha) for upstream gradient g."""
    solver = solver or resolvents
    X = solver(A, xi, b)
    Z = solver(A, xi, g)
    grad_b = np.imag(Z @ c)
    W, s = kappa_edge_weights(A, kappa)
    Wc = W.tocoo()
    grad_k = np.zeros(A.shape[0])
    for j in range(len(xi)):
        x, z = X[:, j], Z[:, j]
        # sum_k w_ik (z_i - z_k)(x_i - x_k)
        e = Wc.data * (z[Wc.row] - z[Wc.col]) * (x[Wc.row] - x[Wc.col])
        per = np.bincount(Wc.row, weights=e.real, minlength=A.shape[0]) + \
            1j * np.bincount(Wc.row, weights=e.imag, minlength=A.shape[0])
        per = per + s * z * x
        grad_k += np.imag(c[j] * per)
    grad_a = float(np.imag(np.sum(dc_dalpha * (g @ X))))
    return grad_b, grad_k, grad_a


def param_gradient(A, E, b, g, xi, c, solver=None):
    """dL/dtheta = Im sum_j c_j z_j^T E x_j for E = dA/dtheta (see the module text)."""
    solver = solver or resolvents
    X = solver(A, xi, b)
    Z = solver(A, xi, g)
    E = sp.csr_matrix(E)
    return float(np.imag(sum(c[j] * (Z[:, j] @ (E @ X[:, j])) for j in range(len(xi)))))


def ms_forward(A, u0, h0, steps, xi, c1, c2, dt, p):
    """The monodomain Mitchell-Schaeffer steps with exact resolvents:
    u_{n+1} = Im sum_j c1_j x_j(u_n) + dt Im sum_j c2_j x_j(g_n) (the contour
    rule of problem_startup.m:95-103 for f1 and f2 without deflation), then the
    gate and g from ``stepping.ms_reaction``.  Returns (u_T, h_T)."""
    from .stepping import ms_reaction
    u = np.array(u0, dtype=np.float64)
    h = np.array(h0, dtype=np.float64)
    g = (h * (u * u) * (1.0 - u)) / p["tau_in"] - u / p["tau_out"]
    for _ in range(steps):
        u = forward(A, u, xi, c1)[0] + dt * forward(A, g, xi, c2)[0]
        h, g = ms_reaction(u, h, dt, p)
    return u, h

