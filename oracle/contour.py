"""Hale-Higham-Trefethen contour quadrature set-up (oracle, mpmath).

Restates ``problem_startup.m:71-144`` (identical text in
``problem_startup_CPU.m:49-112`` and ``problem_startup_nonuniform.m:54-119``):

* f1(x) = exp(-eta x^(alpha/2)), f2(x) = phi(-eta x^(alpha/2)),
  phi(z) = (e^z - 1)/z                                  (``problem_startup.m:72-74``)
* k = (sqrt(lN/l1) - 1)/(sqrt(lN/l1) + 1), [K, K'] = complete elliptic integrals
  of parameter m = k^2 (``ellipkkp(-log(k)/pi)``, ``problem_startup.m:85-86``)
* t_j = i K'/2 - K + (n - 1/2 - j) 2K/n, j = 0..n-1    (``problem_startup.m:95``)
* xi = sqrt(l1 lN)(1/k + sn)/(1/k - sn), dxi/dt = cn dn/(1/k - sn)^2,
  s = -4 K sqrt(l1 lN)/(k pi n)                          (``problem_startup.m:96-99``)
* GPU weights wts = s f(xi) dxi/dt (``problem_startup.m:100-103``); CPU weights
  omit s and pass it separately (``problem_startup_CPU.m:77-79``).
* shifted polynomial coefficients gammas(1,:) = gamma(1),
  gammas(i,:) = gamma(i) + xi gammas(i-1,:), etaxi = xi gammas(end,:)
                                                        (``problem_startup.m:113-123``)
* shift packing for the batched solve                  (``problem_startup.m:140-144``)

``ellipkkp``/``ellipjc`` (Driscoll's SC toolbox, not vendored in the
reference) are replaced by mpmath's ``ellipk`` and ``ellipfun`` at 40 digits,
which accept complex arguments.  The fminbnd choice of n
(``problem_startup.m:92``) is not restated: n is an input.
"""

import mpmath as mp
import numpy as np

_DPS = 40


def _f1(x, eta, alpha):
    return mp.exp(-eta * mp.power(x, mp.mpf(alpha) / 2))


def _f2(x, eta, alpha):
    z = -eta * mp.power(x, mp.mpf(alpha) / 2)
    return (mp.exp(z) - 1) / z


def contour_setup(l1, lN, n, eta, alpha):
    """Nodes, weights and scale factor of the HHT rule on [l1, lN].

    Returns a dict with complex128 arrays ``xi``, ``dxidt``, ``wts1_cpu``,
    ``wts2_cpu`` (without s), ``wts1``, ``wts2`` (with s, GPU form) and the
    float ``s``, ``k``, ``K``, ``Kp``.
    """
    with mp.workdps(_DPS):
        l1m, lNm = mp.mpf(l1), mp.mpf(lN)
        eta_m, alpha_m = mp.mpf(eta), mp.mpf(alpha)
        r = mp.sqrt(lNm / l1m)
        k = (r - 1) / (r + 1)
        m = k * k
        K = mp.ellipk(m)
        Kp = mp.ellipk(1 - m)
        scale = mp.sqrt(l1m * lNm)
        s = -4 * K * scale / (k * mp.pi * n)
        out = {k_: [] for k_ in ("xi", "dxidt", "w1", "w2")}
        for j in range(n):
            t = mp.mpc(0, Kp / 2) - K + (n - mp.mpf(1) / 2 - j) * 2 * K / n
            sn = mp.ellipfun("sn", t, m=m)
            cn = mp.ellipfun("cn", t, m=m)
            dn = mp.ellipfun("dn", t, m=m)
            xi = scale * (1 / k + sn) / (1 / k - sn)
            dxidt = cn * dn / (1 / k - sn) ** 2
            out["xi"].append(xi)
            out["dxidt"].append(dxidt)
            out["w1"].append(_f1(xi, eta_m, alpha_m) * dxidt)
            out["w2"].append(_f2(xi, eta_m, alpha_m) * dxidt)
        c = lambda v: np.array([complex(x) for x in v], dtype=np.complex128)
        return {
            "xi": c(out["xi"]),
            "dxidt": c(out["dxidt"]),
            "wts1_cpu": c(out["w1"]),
            "wts2_cpu": c(out["w2"]),
            "wts1": c([s * w for w in out["w1"]]),
            "wts2": c([s * w for w in out["w2"]]),
            "s": float(s), "k": float(k), "K": float(K), "Kp": float(Kp),
        }


def f_scalar(x, eta, alpha, which):
    """f1 (which=1) or f2 (which=2) at real or complex x, to 40 digits."""
    with mp.workdps(_DPS):
        f = _f1 if which == 1 else _f2
        v = f(mp.mpmathify(x), mp.mpf(eta), mp.mpf(alpha))
        return complex(v)


def shifted_gammas(gamma, xi):
    """gammas (size(gamma) x n complex) and etaxi, ``problem_startup.m:113-123``."""
    gamma = np.asarray(gamma, dtype=np.float64)
    xi = np.asarray(xi, dtype=np.complex128)
    gs = np.zeros((gamma.size, xi.size), dtype=np.complex128)
    gs[0, :] = gamma[0]
    for i in range(1, gamma.size):
        gs[i, :] = gamma[i] + xi * gs[i - 1, :]
    etaxi = xi * gs[-1, :]
    return gs, etaxi


def pack_shifts(etaxi, m):
    """``shiftsvec`` of ``problem_startup.m:140-144``: 4*m*n doubles.

    Complex-interleaved; system j (0-based) occupies complex slots j*m..j*m+m-1
    for the exp systems and n*m + j*m.. for the phi systems, all equal to
    etaxi[j].
    """
    etaxi = np.asarray(etaxi, dtype=np.complex128)
    n = etaxi.size
    out = np.zeros(4 * m * n, dtype=np.float64)
    for i in range(1, n + 1):  # 1-based as in the MATLAB loop
        re_idx = list(range(2 * (i - 1) * m + 1, 2 * i * m + 1, 2))
        re_idx += [q + 2 * n * m for q in re_idx]
        im_idx = [q + 1 for q in re_idx]
        out[np.array(re_idx) - 1] = etaxi[i - 1].real
        out[np.array(im_idx) - 1] = etaxi[i - 1].imag
    return out
