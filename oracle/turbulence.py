"""Von Karman phase structure function, closed form (test infrastructure only).

The reference's ``rodconan`` (shesha/shesha/util/iterkolmo.py:313-351) computes
    D(r) = k1 * L0^(5/3) * (k2 - (2 pi r / L0)^(5/6) K_{5/6}(2 pi r / L0))
for r0 = 1 m with k1 = 0.1716613621245709486 and k2 the r -> 0 limit of the
Bessel term, through a series / asymptotic expansion.  Here the same function
is evaluated with scipy's modified Bessel function K_nu, and scaled by
r0^(-5/3).  Units: rad^2 at the wavelength r0 is given at.
"""
import numpy as np
from scipy.special import gamma, kv

K1 = 0.1716613621245709486


def structure_function(r, r0, L0):
    r = np.asarray(r, dtype=np.float64)
    x = 2 * np.pi * r / L0
    k2 = gamma(5.0 / 6.0) * 2.0 ** (-1.0 / 6.0)   # lim_{x->0} x^(5/6) K_{5/6}(x)
    with np.errstate(invalid="ignore"):
        bessel = np.where(x > 0, x ** (5.0 / 6.0) * kv(5.0 / 6.0, np.maximum(x, 1e-300)), k2)
    return K1 * L0 ** (5.0 / 3.0) * (k2 - bessel) * r0 ** (-5.0 / 3.0)


def kolmogorov(r, r0):
    """Infinite outer scale limit 6.88 (r/r0)^(5/3)."""
    return 6.88 * (np.asarray(r, dtype=np.float64) / r0) ** (5.0 / 3.0)


def variance(r0, L0):
    """Phase variance: half the saturation value of D for r >> L0."""
    k2 = gamma(5.0 / 6.0) * 2.0 ** (-1.0 / 6.0)
    return 0.5 * K1 * L0 ** (5.0 / 3.0) * k2 * r0 ** (-5.0 / 3.0)
