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


def deflated_forward(A, b, xi, c, Q, f_lambda):
    """f(A) b = Q (f(Lambda) Q^T b) + Im sum_j c_j (xi_j - A)^{-1} (b - Q Q^T b): the
    deflated split the GPU path uses (``lanczos_impl.cu:79-89``, ``update_U.cu:18-27``),
    here with exact resolvents."""
    a = Q.T @ b
    u, _ = forward(A, b - Q @ a, xi, c)
    return u + Q @ (f_lambda * a)


def deflated_param_gradient(A, E, b, g, xi, c, Q, lam, f_lambda, f_divdiff, parts=False):
    """dL/dtheta = g^T L_f(A, E) b for the deflated forward, E = dA/dtheta symmetric.

    The Frechet derivative of f in A's eigenbasis is sum_kl f[l_k, l_l] u_k u_k^T E
    u_l u_l^T; split over Q (the K deflated pairs) and P = I - QQ^T:
      Q-Q:  (Q^T g)^T (F o Q^T E Q) (Q^T b),  F = f_divdiff;
      Q-P:  sum_k (Q^T g)_k q_k^T E g_k(A) P b + sum_k (Q^T b)_k q_k^T E g_k(A) P g,
            g_k(x) = (f(l_k) - f_c(x)) / (l_k - x);
      P-P:  Im sum_j c_j z_j^T E x_j, x_j = (xi_j - A)^{-1} P b, z_j likewise for g,
    with f_c(x) = Im sum_j c_j / (xi_j - x) the contour rule on the P part.  Dense
    (small grids): g_k(A) P is formed from the full eigendecomposition.  With parts,
    returns the three blocks (pp, qq, cross) instead of their sum."""
    E = sp.csr_matrix(E)
    K = Q.shape[1]
    ab, ag = Q.T @ b, Q.T @ g
    Pb, Pg = b - Q @ ab, g - Q @ ag
    X = resolvents(A, xi, Pb)
    Z = resolvents(A, xi, Pg)
    pp = float(np.imag(sum(c[j] * (Z[:, j] @ (E @ X[:, j])) for j in range(len(xi)))))
    EQ = E @ Q
    qq = float(ag @ ((f_divdiff * (Q.T @ EQ)) @ ab))
    w, U = np.linalg.eigh(A.toarray() if sp.issparse(A) else A)
    Up, wp = U[:, K:], w[K:]
    fc = np.imag((c[None, :] / (np.asarray(xi)[None, :] - wp[:, None])).sum(axis=1))
    cross = 0.0
    for k in range(K):
        gk = (f_lambda[k] - fc) / (lam[k] - wp)
        cross += ag[k] * (EQ[:, k] @ (Up @ (gk * (Up.T @ Pb))))
        cross += ab[k] * (EQ[:, k] @ (Up @ (gk * (Up.T @ Pg))))
    return (pp, qq, cross) if parts else pp + qq + cross


def ms_forward(A, u0, h0, steps, xi, c1, c2, dt, p, defl=None, states=False):
    """The monodomain Mitchell-Schaeffer steps with exact resolvents:
    u_{n+1} = Im sum_j c1_j x_j(u_n) + dt Im sum_j c2_j x_j(g_n) (the contour
    rule of problem_startup.m:95-103 for f1 and f2), then the gate and g from
    ``stepping.ms_reaction``; with defl = (Q, f1(Lambda), f2(Lambda)) each
    product is the deflated split of ``deflated_forward``.  Returns (u_T, h_T), and
    with states also the list u_1 .. u_T."""
    from .stepping import ms_reaction
    u = np.array(u0, dtype=np.float64)
    h = np.array(h0, dtype=np.float64)
    g = (h * (u * u) * (1.0 - u)) / p["tau_in"] - u / p["tau_out"]
    us = []
    for _ in range(steps):
        if defl is None:
            u = forward(A, u, xi, c1)[0] + dt * forward(A, g, xi, c2)[0]
        else:   # defl = (Q, f1(lambda), f2(lambda))
            Q, fl1, fl2 = defl
            u = deflated_forward(A, u, xi, c1, Q, fl1) + dt * deflated_forward(A, g, xi, c2, Q, fl2)
        h, g = ms_reaction(u, h, dt, p)
        us.append(u)
    return (u, h, us) if states else (u, h)

