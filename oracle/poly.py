"""Jacobi least-squares polynomial preconditioner coefficients (oracle).

Restates ``polynomial_coeffs.m:1-28``: the residual polynomial R of degree
``degree+1`` for the Jacobi weight with mu = 1/2, nu = -1/2 is accumulated term
by term (``polynomial_coeffs.m:15-24``), then the least-squares polynomial
s(lam) = (1 - R(lam)) / lam is returned highest power first and rescaled from
[0, 4] to [0, 4*prob_dim] (``polynomial_coeffs.m:28``).

The exact version works in rationals, so it is free of the rounding order of
the MATLAB loop; the float version rounds the exact result once per
coefficient.
"""

from fractions import Fraction
from math import comb


def _residual_exact(degree):
    """R(1..degree+2) of ``polynomial_coeffs.m:11-24`` as exact rationals."""
    d = int(degree)
    R = [Fraction(0)] * (d + 2)
    for j in range(1, d + 3):                      # polynomial_coeffs.m:15
        n_bin = d - j + 3
        bin_coeff = [comb(d - j + 2, i - 1) for i in range(1, n_bin + 1)]  # :17-20
        kappaj = Fraction(comb(d + 1, j - 1))      # :22
        for t in range(0, j - 1):
            kappaj *= (Fraction(2 * d + 1, 2) - t) / (Fraction(t) + Fraction(3, 2))
        coef = Fraction(1, 4 ** (j - 1)) * (-1) ** (j - 1) * kappaj   # :23
        for i in range(n_bin):
            R[j - 1 + i] += coef * bin_coeff[i] * Fraction(1, (-4) ** i)
    return R


def polynomial_coeffs_exact(degree, prob_dim=1):
    """Coefficients gamma (highest power first) as Fractions.

    ``prob_dim`` must be rational here (int or Fraction).
    """
    d = int(degree)
    R = _residual_exact(d)
    pd = Fraction(prob_dim)
    # polynomial_coeffs.m:28  gamma = -R(end:-1:2)*(3+2d) ./ prob_dim.^(d:-1:0)
    rev = R[::-1][:-1]                            # R(end:-1:2)
    return [-r * (3 + 2 * d) / pd ** p for r, p in zip(rev, range(d, -1, -1))]


def polynomial_coeffs(degree, prob_dim):
    """Float coefficients gamma (highest power first), ``polynomial_coeffs.m:1-28``.

    ``prob_dim`` may be any positive float (e.g. lN/4 as in
    ``problem_startup_nonuniform.m:87``); the scaling is applied in floating
    point to the exactly computed unscaled coefficients.
    """
    d = int(degree)
    unscaled = polynomial_coeffs_exact(d, 1)
    return [float(g) / float(prob_dim) ** p for g, p in zip(unscaled, range(d, -1, -1))]
