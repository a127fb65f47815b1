"""Pin the exact-resolvent adjoint oracle by central finite differences (no GPU)."""

import numpy as np
import pytest

from oracle import adjoint, contour
from conftest import rel
from gpumatfun_amd import problems

TAU, ALPHA = 1e-3, 1.5


def setup(n=12, seed=1):
    kappa = np.exp(0.5 * problems.gaussian_random_field(n, 0.1, seed=seed))
    A = problems.varcoef_laplacian_2d(kappa, n)
    w = np.linalg.eigvalsh(A.toarray())
    rng = np.random.default_rng(7)
    b = rng.standard_normal(n * n)
    y = rng.standard_normal(n * n)
    return kappa, A, w, b, y


def weights(l1, lN, alpha, ns=40):
    c = contour.contour_setup(l1, lN, ns, TAU, alpha)
    xi = c["xi"]
    cw = c["wts1"]                         # s * f(xi) * dxi/dt
    dcw = cw * (-TAU * xi ** (alpha / 2) * np.log(xi) / 2)
    return xi, cw, dcw


def loss(A, b, y, xi, cw):
    u, _ = adjoint.forward(A, b, xi, cw)
    return float(np.sum((u - y) ** 2)), u


def test_forward_matches_dense_fA_b():
    kappa, A, w, b, y = setup()
    xi, cw, _ = weights(w.min(), w.max(), ALPHA)
    u, _ = adjoint.forward(A, b, xi, cw)
    W, Q = np.linalg.eigh(A.toarray())
    ref = Q @ (np.exp(-TAU * W ** (ALPHA / 2)) * (Q.T @ b))
    assert rel(u, ref) < 1e-9


def test_kappa_gradient_central_differences():
    n = 12
    kappa, A, w, b, y = setup(n)
    # contour on bounds that stay valid under the small kappa perturbations
    xi, cw, dcw = weights(0.9 * w.min(), 1.1 * w.max(), ALPHA)
    L0, u = loss(A, b, y, xi, cw)
    g = 2 * (u - y)
    gb, gk, ga = adjoint.gradients(A, kappa, b, g, xi, cw, dcw)
    rng = np.random.default_rng(3)
    for i in rng.choice(n * n, 8, replace=False):
        d = 1e-6 * kappa[i]
        kp = kappa.copy(); kp[i] += d
        km = kappa.copy(); km[i] -= d
        Lp, _ = loss(problems.varcoef_laplacian_2d(kp, n), b, y, xi, cw)
        Lm, _ = loss(problems.varcoef_laplacian_2d(km, n), b, y, xi, cw)
        fd = (Lp - Lm) / (2 * d)
        assert abs(fd - gk[i]) <= 1e-5 * max(abs(gk[i]), 1e-3 * np.abs(gk).max())


def test_b_and_alpha_gradients_central_differences():
    kappa, A, w, b, y = setup()
    l1, lN = 0.9 * w.min(), 1.1 * w.max()
    xi, cw, dcw = weights(l1, lN, ALPHA)
    L0, u = loss(A, b, y, xi, cw)
    g = 2 * (u - y)
    gb, gk, ga = adjoint.gradients(A, kappa, b, g, xi, cw, dcw)
    v = np.random.default_rng(5).standard_normal(b.size)
    h = 1e-6
    fd = (loss(A, b + h * v, y, xi, cw)[0] - loss(A, b - h * v, y, xi, cw)[0]) / (2 * h)
    assert abs(fd - gb @ v) <= 1e-6 * abs(fd)
    ha = 1e-6
    xp, cp, _ = weights(l1, lN, ALPHA + ha)
    xm, cm, _ = weights(l1, lN, ALPHA - ha)
    fda = (loss(A, b, y, xp, cp)[0] - loss(A, b, y, xm, cm)[0]) / (2 * ha)
    assert abs(fda - ga) <= 1e-5 * abs(fda)


def test_param_gradient_conductivity_ratio_central_differences():
    """dL/dr for the conductivity ratio r of the anisotropic operator (A linear in r)."""
    n, r, h = 12, 4.0, 1e-5
    A = problems.anisotropic_9pt(n, ratio=r)
    E = problems.anisotropic_9pt_ratio_derivative(n)
    # linear in r: A(r) = A(0) + r E, to rounding
    assert rel((problems.anisotropic_9pt(n, ratio=0.0) + r * E).toarray(), A.toarray()) < 1e-14
    w = np.linalg.eigvalsh(A.toarray())
    xi, cw, _ = weights(0.9 * w.min(), 1.1 * w.max(), ALPHA)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n * n)
    y = rng.standard_normal(n * n)
    L0, u = loss(A, b, y, xi, cw)
    g = adjoint.param_gradient(A, E, b, 2 * (u - y), xi, cw)
    Lp, _ = loss(problems.anisotropic_9pt(n, ratio=r + h), b, y, xi, cw)
    Lm, _ = loss(problems.anisotropic_9pt(n, ratio=r - h), b, y, xi, cw)
    fd = (Lp - Lm) / (2 * h)
    assert abs(fd - g) <= 1e-6 * abs(g), (fd, g)


def test_deflated_param_gradient_central_differences():
    """The deflated forward Q f(Lambda) Q^T b + f_c(A)(b - QQ^T b), with its K = 6
    eigenpairs recomputed at every ratio, differentiated through the split of the
    Frechet derivative (Q-Q divided differences, cross-block real-shift
    resolvents, contour on the rest), against central differences.  Dropping any
    one of the three blocks moves the result by more than the tolerance (checked
    below), so each is pinned."""
    from gpumatfun_amd import quadrature as q
    n, r, h, K = 12, 4.0, 1e-5, 6
    E = problems.anisotropic_9pt_ratio_derivative(n)
    A = problems.anisotropic_9pt(n, ratio=r)
    w = np.linalg.eigvalsh(A.toarray())
    xi, cw, _ = weights(0.9 * w[K], 1.1 * w.max(), ALPHA)
    f = (lambda x: q.f1(x, TAU, ALPHA))
    df = (lambda x: q.df1(x, TAU, ALPHA))
    rng = np.random.default_rng(4)
    Q0 = np.linalg.eigh(A.toarray())[1][:, :K]
    b = rng.standard_normal(n * n) + 5.0 * Q0 @ rng.standard_normal(K)   # strong low modes
    y = rng.standard_normal(n * n)

    def defl(rr):
        Ar = problems.anisotropic_9pt(n, ratio=rr)
        wr, Ur = np.linalg.eigh(Ar.toarray())
        return Ar, Ur[:, :K], wr[:K]

    def L(rr):
        Ar, Qr, lr = defl(rr)
        u = adjoint.deflated_forward(Ar, b, xi, cw, Qr, f(lr))
        return float(np.sum((u - y) ** 2)), u

    L0, u = L(r)
    _, Q, lam = defl(r)
    F = q.divided_differences(f, df, lam)
    g = adjoint.deflated_param_gradient(A, E, b, 2 * (u - y), xi, cw, Q, lam, f(lam), F)
    fd = (L(r + h)[0] - L(r - h)[0]) / (2 * h)
    assert abs(fd - g) <= 1e-8 * abs(g), (fd, g)
    # the deflated forward is f(A) b to the contour's accuracy
    u_full, _ = adjoint.forward(A, b, *weights(0.9 * w.min(), 1.1 * w.max(), ALPHA)[:2])
    assert rel(u, u_full) < 1e-9
    # each block is far above the agreement, so dropping any one would fail it
    parts = adjoint.deflated_param_gradient(A, E, b, 2 * (u - y), xi, cw, Q, lam, f(lam), F,
                                            parts=True)
    print("blocks", parts, "fd", fd, "rel", abs(fd - g) / abs(g))
    assert min(abs(v) for v in parts) > 1e-4 * abs(g)


def test_divided_differences_and_derivatives():
    """f1', f2' against central differences in extended steps; the divided
    difference matrix is symmetric, its diagonal the derivative, and an
    off-diagonal entry the secant slope."""
    from gpumatfun_amd import quadrature as q
    eta, al = 5e-5, 1.7
    x = np.array([3.0, 80.0, 1014.0, 2.5e4])
    for fn, dfn in ((q.f1, q.df1), (q.f2, q.df2)):
        for xv in x:
            hh = 1e-4 * xv
            fd = (fn(xv + hh, eta, al) - fn(xv - hh, eta, al)) / (2 * hh)
            assert abs(fd - dfn(xv, eta, al)) <= 1e-7 * abs(dfn(xv, eta, al))
        F = q.divided_differences(lambda t: fn(t, eta, al), lambda t: dfn(t, eta, al), x)
        assert np.allclose(F, F.T, rtol=1e-14, atol=0)
        assert np.allclose(np.diag(F), dfn(x, eta, al), rtol=1e-14, atol=0)
        assert abs(F[0, 2] - (fn(x[0], eta, al) - fn(x[2], eta, al)) / (x[0] - x[2])) < 1e-18


def test_activation_time_sources_match_finite_differences():
    """The smooth activation-time map t = dt sum_n s(u_n) and its direct sources
    dL/du_n for L = ||t - t_obs||^2, against central differences per entry."""
    from gpumatfun_amd.monodomain import activation_times
    rng = np.random.default_rng(2)
    us = [rng.random(30) for _ in range(4)]
    t_obs = rng.random(30) * 0.2
    dt = 0.05
    t, sources = activation_times(us, dt, v_thr=0.5, eps=0.05)
    srcs = sources(2 * (t - t_obs))
    for n in range(4):
        for i in (0, 7, 19):
            h = 1e-6
            up = [u.copy() for u in us]
            um = [u.copy() for u in us]
            up[n][i] += h
            um[n][i] -= h
            Lp = np.sum((activation_times(up, dt)[0] - t_obs) ** 2)
            Lm = np.sum((activation_times(um, dt)[0] - t_obs) ** 2)
            fd = (Lp - Lm) / (2 * h)
            assert abs(fd - srcs[n][i]) <= 1e-6 * abs(fd) + 1e-10, (n, i, fd, srcs[n][i])
