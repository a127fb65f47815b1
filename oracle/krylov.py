"""Lanczos, polynomial evaluation, shifted solves and the f(A)b assembly (oracle).

Literal NumPy restatements of the reference's MATLAB CPU path:

* ``lanczos``            -- ``lanczos.m:1-54``
* ``jacobi_poly_eval``   -- ``jacobi_poly_eval.m:1-19``
* ``tridiag_solve_pivoted`` -- MATLAB's sparse ``\\`` on a tridiagonal matrix
  (``nicksmethodcpu.m:19-20``), which dispatches to LAPACK's ?GTSV: Gaussian
  elimination with partial pivoting (row interchange when
  |dl(k)|_1 > |d(k)|_1, with |z|_1 = |Re z| + |Im z|), restated from the
  published LAPACK algorithm (LAPACK 3.x ZGTSV).
* ``nicksmethodcpu``     -- ``nicksmethodcpu.m:1-45``

Sparse products use scipy CSR with sorted column indices, so each row is
summed in increasing column order as MATLAB's column-oriented sparse product
does.
"""

import numpy as np
import scipy.sparse as sp


def jacobi_poly_eval(gammareal, gammaimag, A, breal, bimag):
    """Complex Horner p(A)b with p = sum gamma_i A^(n-1-i), ``jacobi_poly_eval.m:1-19``.

    ``breal``/``bimag`` are N x z real arrays (z columns).  Returns
    (xreal, ximag).  When ``gammaimag`` and ``bimag`` are identically zero the
    imaginary lane is skipped: it would stay exactly zero, and
    ``Ax + b*g - 0*0`` rounds exactly as ``Ax + b*g``.
    """
    gr = np.asarray(gammareal, dtype=np.float64).ravel()
    gi = np.asarray(gammaimag, dtype=np.float64).ravel()
    br = np.asarray(breal, dtype=np.float64)
    bi = np.asarray(bimag, dtype=np.float64)
    one_col = br.ndim == 1
    if one_col:
        br = br[:, None]
        bi = bi[:, None]
    real_only = not gi.any() and not bi.any()
    # jacobi_poly_eval.m:12-13
    xr = br * gr[0] - bi * gi[0]
    xi = br * gi[0] + bi * gr[0]
    for i in range(1, gr.size):                       # jacobi_poly_eval.m:15
        if real_only:
            xr = A @ xr + br * gr[i]                  # :16-17 with bimag = 0
        else:
            Ax = A @ np.hstack([xr, xi])              # :16
            z = br.shape[1]
            xr = Ax[:, :z] + br * gr[i] - bi * gi[i]  # :17
            xi = Ax[:, z:] + br * gi[i] + bi * gr[i]  # :18
    if one_col:
        return xr[:, 0], xi[:, 0]
    return xr, xi


def lanczos(A, a, b, N, gamma):
    """Two-column preconditioned Lanczos, ``lanczos.m:1-54``.

    Returns V1, V2 (n x N), and the tridiagonal data as
    (alpha1, beta1, alpha2, beta2) with beta arrays of length N+1
    (beta[0] = ||a||, beta[k] = beta_k), plus dense H1, H2.
    """
    A = sp.csr_matrix(A)
    A.sort_indices()
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    gamma = np.asarray(gamma, dtype=np.float64).ravel()
    n = a.size
    beta1 = np.zeros(N + 1)
    beta2 = np.zeros(N + 1)
    beta1[0] = np.linalg.norm(a)                      # lanczos.m:9
    beta2[0] = np.linalg.norm(b)                      # lanczos.m:10
    V1 = np.zeros((n, N))
    V2 = np.zeros((n, N))
    alpha1 = np.zeros(N)
    alpha2 = np.zeros(N)
    V1[:, 0] = a / beta1[0]                           # lanczos.m:17
    V2[:, 0] = b / beta2[0]                           # lanczos.m:18
    zg = np.zeros_like(gamma)
    for m in range(N):                                # lanczos.m:21
        w = np.column_stack([V1[:, m], V2[:, m]])     # :22
        w, _ = jacobi_poly_eval(gamma, zg, A, w, np.zeros_like(w))   # :24
        w = A @ w                                     # :27
        if m != 0:                                    # :30-32
            w = w - np.column_stack([beta1[m] * V1[:, m - 1], beta2[m] * V2[:, m - 1]])
        alpha1[m] = np.dot(w[:, 0], V1[:, m])         # :35
        alpha2[m] = np.dot(w[:, 1], V2[:, m])         # :36
        w = w - np.column_stack([alpha1[m] * V1[:, m], alpha2[m] * V2[:, m]])  # :38
        beta1[m + 1] = np.linalg.norm(w[:, 0])        # :40
        beta2[m + 1] = np.linalg.norm(w[:, 1])        # :41
        if m < N - 1:                                 # :44-47
            V1[:, m + 1] = w[:, 0] / beta1[m + 1]
            V2[:, m + 1] = w[:, 1] / beta2[m + 1]
    H1 = tridiag_dense(alpha1, beta1)
    H2 = tridiag_dense(alpha2, beta2)
    return V1, V2, (alpha1, beta1, alpha2, beta2), (H1, H2)


def tridiag_dense(alpha, beta):
    """Dense H of ``lanczos.m:52-53``: diag alpha, off-diagonals beta[1..N-1]."""
    N = alpha.size
    H = np.diag(alpha.astype(np.float64))
    off = np.asarray(beta[1:N], dtype=np.float64)
    H += np.diag(off, -1) + np.diag(off, 1)
    return H


def _cabs1(z):
    return abs(z.real) + abs(z.imag)


def tridiag_solve_pivoted(dl, d, du, b):
    """Solve a complex tridiagonal system by Gaussian elimination with partial pivoting.

    ``dl`` (n-1 sub-diagonal), ``d`` (n diagonal), ``du`` (n-1
    super-diagonal), ``b`` (n right-hand side).  The elimination and the
    back substitution follow LAPACK ZGTSV step for step, which is what
    MATLAB's sparse backslash runs on a tridiagonal matrix
    (``nicksmethodcpu.m:19-20``).  Raises ZeroDivisionError on an exactly
    singular pivot, as ZGTSV reports INFO > 0.
    """
    dl = np.array(dl, dtype=np.complex128)
    d = np.array(d, dtype=np.complex128)
    du = np.array(du, dtype=np.complex128)
    x = np.array(b, dtype=np.complex128)
    n = d.size
    if n == 0:
        return x
    du2 = np.zeros(max(n - 2, 0), dtype=np.complex128)   # second super-diagonal (fill-in)
    for k in range(n - 1):
        if dl[k] == 0:
            if d[k] == 0:
                raise ZeroDivisionError("singular tridiagonal pivot")
            continue
        if _cabs1(d[k]) >= _cabs1(dl[k]):
            mult = dl[k] / d[k]
            d[k + 1] = d[k + 1] - mult * du[k]
            x[k + 1] = x[k + 1] - mult * x[k]
        else:
            mult = d[k] / dl[k]
            d[k] = dl[k]
            temp = d[k + 1]
            d[k + 1] = du[k] - mult * temp
            if k < n - 2:
                du2[k] = du[k + 1]
                du[k + 1] = -mult * du2[k]
            du[k] = temp
            t = x[k]
            x[k] = x[k + 1]
            x[k + 1] = t - mult * x[k + 1]
    if d[n - 1] == 0:
        raise ZeroDivisionError("singular tridiagonal pivot")
    x[n - 1] = x[n - 1] / d[n - 1]
    if n > 1:
        x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / d[n - 2]
    for k in range(n - 3, -1, -1):
        x[k] = (x[k] - du[k] * x[k + 1] - du2[k] * x[k + 2]) / d[k]
    return x


def shifted_solves(alpha, beta, etaxi, wts, scaling_factor):
    """y(:,j) = s*w_j*((etaxi_j I - H) \\ beta_0 e1), ``nicksmethodcpu.m:17-21``.

    Returns an N x n complex array (column j for node j).
    """
    N = alpha.size
    rhs = np.zeros(N, dtype=np.complex128)
    rhs[0] = beta[0]
    off = np.asarray(beta[1:N], dtype=np.float64)
    y = np.zeros((N, len(etaxi)), dtype=np.complex128)
    for j, ex in enumerate(etaxi):
        d = ex - alpha.astype(np.complex128)
        sol = tridiag_solve_pivoted(-off, d, -off, rhs)
        y[:, j] = (scaling_factor * wts[j]) * sol
    return y


def nicksmethodcpu(wts1, wts2, A, a, b, n, N, scaling_factor, gamma,
                   gammasreal, gammasimag, etaxi, return_parts=False):
    """[v1, v2] = nicksmethodcpu(...), ``nicksmethodcpu.m:1-45``.

    ``wts1``/``wts2`` are complex node weights without s (CPU form),
    ``scaling_factor`` is s.  Returns v1 ~ f1(A) a and v2 ~ f2(A) b.
    """
    A = sp.csr_matrix(A)
    A.sort_indices()
    V1, V2, (al1, be1, al2, be2), _ = lanczos(A, a, b, N, gamma)   # :7
    etaxi = np.asarray(etaxi, dtype=np.complex128)
    y1 = shifted_solves(al1, be1, etaxi[:n], np.asarray(wts1)[:n], scaling_factor)  # :18-21
    y2 = shifted_solves(al2, be2, etaxi[:n], np.asarray(wts2)[:n], scaling_factor)
    gsr = np.asarray(gammasreal, dtype=np.float64).reshape(len(gamma), -1)
    gsi = np.asarray(gammasimag, dtype=np.float64).reshape(len(gamma), -1)
    v1 = np.zeros(A.shape[0])
    v2 = np.zeros(A.shape[0])
    for i in range(n):                                # :33
        Y1r = V1 @ y1[:, i].real                      # :34-37
        Y1i = V1 @ y1[:, i].imag
        Y2r = V2 @ y2[:, i].real
        Y2i = V2 @ y2[:, i].imag
        _, Y1i = jacobi_poly_eval(gsr[:, i], gsi[:, i], A, Y1r, Y1i)   # :39
        _, Y2i = jacobi_poly_eval(gsr[:, i], gsi[:, i], A, Y2r, Y2i)   # :40
        v1 = v1 + Y1i                                 # :41
        v2 = v2 + Y2i                                 # :42
    if return_parts:
        return v1, v2, {"V1": V1, "V2": V2, "aabb": (al1, be1, al2, be2),
                        "y1": y1, "y2": y2}
    return v1, v2
