"""Boundary-input set-up for ``lanczos_create`` (host, double precision).

Computes what ``problem_startup.m:71-144`` computes before it creates the
engine:

* the Jacobi least-squares preconditioner coefficients gamma
  (``polynomial_coeffs.m:1-28``), highest power first;
* the Hale-Higham-Trefethen nodes xi, dxi/dt and weights on [l1, lN]
  (``problem_startup.m:85-103``); the Jacobi elliptic functions of complex
  argument are built from real-argument ones with the addition formulas of
  Abramowitz & Stegun 16.21.1-3, and the imaginary part is exactly K'/2, where
  sn, cn, dn have the closed forms 1/sqrt(1+k), sqrt(k/(1+k)), sqrt(k)
  (A&S 16.5, with the complementary parameter);
* shifted coefficients gammas and etaxi (``problem_startup.m:113-123``);
* the packed ``shifts`` vector (``problem_startup.m:140-144``).

The number of nodes n is an input (the reference picks it with ``fminbnd``,
``problem_startup.m:92``).
"""

from dataclasses import dataclass, field
from math import comb

import numpy as np
import scipy.special as ss


def polynomial_coeffs(degree, prob_dim):
    """Least-squares polynomial coefficients, highest power first (``polynomial_coeffs.m``)."""
    d = int(degree)
    R = np.zeros(d + 2)
    for j in range(1, d + 3):
        n_bin = d - j + 3
        binc = np.array([comb(d - j + 2, i) for i in range(n_bin)], dtype=np.float64)
        kap = float(comb(d + 1, j - 1))
        for t in range(j - 1):
            kap *= (d + 0.5 - t) / (t + 1.5)
        sc = (0.25 ** (j - 1)) * (-1.0) ** (j - 1) * kap
        R[j - 1:] += sc * binc * (-4.0) ** (-np.arange(n_bin, dtype=np.float64))
    rev = R[::-1][:-1]
    powers = np.arange(d, -1, -1, dtype=np.float64)
    return -rev * (3 + 2 * d) / float(prob_dim) ** powers


def _ellipj_agm(u, m, m1):
    """sn, cn, dn of real u by the arithmetic-geometric mean (A&S 16.4).

    Takes the complementary parameter m1 = 1 - m directly so that
    b0 = sqrt(m1) carries no cancellation when m is close to 1.
    """
    u = np.asarray(u, dtype=np.float64)
    a = [1.0]
    c = [np.sqrt(m)]
    b = np.sqrt(m1)
    while abs(c[-1]) > 1e-17 * a[-1] and len(a) < 64:
        an = 0.5 * (a[-1] + b)
        cn = 0.5 * (a[-1] - b)
        b = np.sqrt(a[-1] * b)
        a.append(an)
        c.append(cn)
    nlev = len(a) - 1
    phi = (2.0 ** nlev) * a[-1] * u
    phis = [phi]
    for lev in range(nlev, 0, -1):
        phi = 0.5 * (np.arcsin(c[lev] / a[lev] * np.sin(phi)) + phi)
        phis.append(phi)
    phi0 = phis[-1]
    phi1 = phis[-2] if nlev > 0 else phi0
    sn = np.sin(phi0)
    cn = np.cos(phi0)
    dn = cn / np.cos(phi1 - phi0)
    return sn, cn, dn


def _jacobi_complex(x, m, k, m1):
    """sn, cn, dn at x + i K'(m)/2 for real x (A&S 16.21.1-3)."""
    s, c, d = _ellipj_agm(x, m, m1)
    s1 = 1.0 / np.sqrt(1.0 + k)
    c1 = np.sqrt(k / (1.0 + k))
    d1 = np.sqrt(k)
    den = c1 * c1 + m * s * s * s1 * s1
    sn = (s * d1 + 1j * c * d * s1 * c1) / den
    cn = (c * c1 - 1j * s * d * s1 * d1) / den
    dn = (d * c1 * d1 - 1j * m * s * c * s1) / den
    return sn, cn, dn


def f1(x, eta, alpha):
    """exp(-eta x^(alpha/2)), ``problem_startup.m:73``."""
    return np.exp(-eta * np.power(x, alpha / 2.0))


def f2(x, eta, alpha):
    """phi(-eta x^(alpha/2)), phi(z) = (e^z - 1)/z, ``problem_startup.m:72,74``."""
    z = -eta * np.power(x, alpha / 2.0)
    return np.expm1(z) / z


def _dphi(z):
    """phi'(z) for phi(z) = (e^z - 1)/z: sum_{n>=1} n z^(n-1)/(n+1)! near 0, else
    (e^z (z - 1) + 1)/z^2."""
    z = np.asarray(z, dtype=np.float64)
    small = np.abs(z) < 0.5
    out = np.empty_like(z)
    zs = z[small]
    acc = np.zeros_like(zs)
    term_fact = 1.0
    for n in range(1, 24):
        term_fact *= (n + 1)
        acc += n * zs ** (n - 1) / term_fact
    out[small] = acc
    zb = z[~small]
    out[~small] = (np.exp(zb) * (zb - 1.0) + 1.0) / (zb * zb)
    return out


def df1(x, eta, alpha):
    """d/dx exp(-eta x^(alpha/2))."""
    x = np.asarray(x, dtype=np.float64)
    a = alpha / 2.0
    return -eta * a * np.power(x, a - 1.0) * f1(x, eta, alpha)


def df2(x, eta, alpha):
    """d/dx phi(-eta x^(alpha/2))."""
    x = np.asarray(x, dtype=np.float64)
    a = alpha / 2.0
    return _dphi(-eta * np.power(x, a)) * (-eta * a * np.power(x, a - 1.0))


def divided_differences(fun, dfun, lam, rel=1e-6):
    """K x K matrix f[lam_k, lam_l] = (f(lam_k) - f(lam_l)) / (lam_k - lam_l), with
    f' at the midpoint where |lam_k - lam_l| <= rel |lam_k| (the diagonal and
    near-degenerate pairs)."""
    lam = np.asarray(lam, dtype=np.float64)
    fl = fun(lam)
    dl = lam[:, None] - lam[None, :]
    close = np.abs(dl) <= rel * np.maximum(np.abs(lam[:, None]), np.abs(lam[None, :]))
    with np.errstate(divide="ignore", invalid="ignore"):
        F = (fl[:, None] - fl[None, :]) / dl
    mid = 0.5 * (lam[:, None] + lam[None, :])
    F[close] = dfun(mid[close])
    return F


def contour_nodes(l1, lN, n, eta, alpha):
    """HHT quadrature on [l1, lN]: dict of xi, dxidt, wts1, wts2 (GPU form, with s), s."""
    r = np.sqrt(lN / l1)
    k = (r - 1.0) / (r + 1.0)
    one_minus_k = 2.0 / (r + 1.0)
    m = k * k
    m1 = one_minus_k * (1.0 + k)          # 1 - k^2 without cancellation
    K = float(ss.ellipkm1(m1))           # K(m)
    Kp = float(ss.ellipk(m1))            # K(1 - m)
    j = np.arange(n, dtype=np.float64)
    x = -K + (n - 0.5 - j) * 2.0 * K / n
    sn, cn, dn = _jacobi_complex(x, m, k, m1)
    scale = np.sqrt(l1 * lN)
    xi = scale * (1.0 / k + sn) / (1.0 / k - sn)
    dxidt = cn * dn / (1.0 / k - sn) ** 2
    s = -4.0 * K * scale / (k * np.pi * n)
    w1 = f1(xi, eta, alpha) * dxidt
    w2 = f2(xi, eta, alpha) * dxidt
    return {"xi": xi, "dxidt": dxidt, "wts1": s * w1, "wts2": s * w2,
            "wts1_cpu": w1, "wts2_cpu": w2, "s": s, "k": k, "K": K, "Kp": Kp}


def shifted_gammas(gamma, xi):
    gamma = np.asarray(gamma, dtype=np.float64)
    gs = np.zeros((gamma.size, np.size(xi)), dtype=np.complex128)
    gs[0, :] = gamma[0]
    for i in range(1, gamma.size):
        gs[i, :] = gamma[i] + xi * gs[i - 1, :]
    return gs, xi * gs[-1, :]


def pack_shifts(etaxi, m):
    """4*m*n doubles; system j and system n+j hold etaxi[j] in all m complex slots."""
    etaxi = np.asarray(etaxi, dtype=np.complex128)
    blk = np.repeat(etaxi, m)                  # n*m complex
    both = np.concatenate([blk, blk])          # exp systems then phi systems
    out = np.empty(2 * both.size)
    out[0::2] = both.real
    out[1::2] = both.imag
    return out


def f_of_eigs(lam, eta, alpha):
    """exp_lambda and phi_lambda of deflated eigenvalues, ``problem_startup.m:75-77``.

    A negative eigenvalue is the rounding of a Neumann operator's zero mode
    (the FEM disc's constant vector) and is evaluated as 0 (its real power
    would be NaN); phi(0) = 1.
    """
    lam = np.maximum(np.asarray(lam, dtype=np.float64), 0.0)
    el = f1(lam, eta, alpha)
    with np.errstate(divide="ignore", invalid="ignore"):
        pl = f2(lam.astype(np.float64), eta, alpha)
    pl[lam == 0.0] = 1.0
    if pl.size:
        pl[0] = 1.0                               # phi_lambda(1) = 1 guard
    return el, pl


@dataclass
class CreateInputs:
    """The 26 positional inputs of ``lanczos_create`` (``lanczos_create.cpp:7-32``)."""
    U: np.ndarray                     # N x 2 [u, g(u)]
    A: object                         # scipy CSR, N x N
    Q: np.ndarray                     # N x K
    lam: np.ndarray                   # K
    m: int                            # subspace size ("N")
    gamma: np.ndarray                 # d+1
    gamma_s_real: np.ndarray          # (d+1) x n
    gamma_s_imag: np.ndarray
    exp_lambda: np.ndarray            # K
    phi_lambda: np.ndarray            # K
    dt: float
    ode: np.ndarray                   # N
    aode: float = 0.0
    bode: float = 0.0
    code: float = 0.0
    vrest: float = 0.0
    vamp: float = 0.0
    vth: float = 0.0
    vpeak: float = 0.0
    c1ion: float = 0.0
    c2ion: float = 0.0
    mass: np.ndarray = None           # N
    wts1: np.ndarray = None           # n x 2 [Re Im], includes s
    wts2: np.ndarray = None
    shifts: np.ndarray = None         # 4*m*n
    extra: dict = field(default_factory=dict)


def make_inputs(A, u0, g0, *, m, n, eta, alpha, dt, l1, lN, degree=0, prob_dim=None,
                Q=None, lam=None, mass=None, ode=None, fhn=None):
    """Assemble ``lanczos_create`` inputs the way ``problem_startup.m:71-147`` does."""
    N = A.shape[0]
    if prob_dim is None:
        prob_dim = lN / 4.0
    gamma = polynomial_coeffs(degree, prob_dim)
    c = contour_nodes(l1, lN, n, eta, alpha)
    gs, etaxi = shifted_gammas(gamma, c["xi"])
    if Q is None:
        Q = np.zeros((N, 0))
        lam = np.zeros(0)
    el, pl = f_of_eigs(lam, eta, alpha)
    U = np.column_stack([np.asarray(u0, dtype=np.float64), np.asarray(g0, dtype=np.float64)])
    inp = CreateInputs(
        U=U, A=A, Q=np.asarray(Q, dtype=np.float64), lam=np.asarray(lam, dtype=np.float64),
        m=int(m), gamma=gamma, gamma_s_real=gs.real.copy(), gamma_s_imag=gs.imag.copy(),
        exp_lambda=el, phi_lambda=pl, dt=float(dt),
        ode=np.zeros(N) if ode is None else np.asarray(ode, dtype=np.float64),
        mass=np.ones(N) if mass is None else np.asarray(mass, dtype=np.float64),
        wts1=np.column_stack([c["wts1"].real, c["wts1"].imag]),
        wts2=np.column_stack([c["wts2"].real, c["wts2"].imag]),
        shifts=pack_shifts(etaxi, m),
        extra={"xi": c["xi"], "etaxi": etaxi, "s": c["s"], "wts1_cpu": c["wts1_cpu"],
               "wts2_cpu": c["wts2_cpu"], "l1": l1, "lN": lN, "eta": eta, "alpha": alpha,
               "n": n},
    )
    if fhn:
        for key, v in fhn.items():
            setattr(inp, key, float(v))
    return inp
