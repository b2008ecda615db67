"""Von Karman structure function (the closed form behind the reference's
rodconan) against its textbook limits, and the synthetic screens against it."""
import numpy as np

from oracle import turbulence as ot
from shwfs import screens


def test_small_separation_expansion():
    # D(r) = 6.88 (r/r0)^(5/3) [1 - 1.485 (r/L0)^(1/3) + O((r/L0)^2)] for r << L0
    r0, L0 = 0.16, 25.0
    r = np.array([1e-5, 1e-4, 1e-3]) * L0
    expect = ot.kolmogorov(r, r0) * (1 - 1.485 * (r / L0) ** (1.0 / 3.0))
    np.testing.assert_allclose(ot.structure_function(r, r0, L0), expect, rtol=3e-3)


def test_large_separation_saturates_at_twice_the_von_karman_variance():
    r0, L0 = 0.12, 25.0
    # sigma^2 = 0.0863 (L0 / r0)^(5/3) rad^2 (von Karman phase variance)
    sig2 = 0.0863 * (L0 / r0) ** (5.0 / 3.0)
    np.testing.assert_allclose(ot.structure_function(1e4 * L0, r0, L0), 2 * sig2, rtol=2e-3)
    np.testing.assert_allclose(ot.variance(r0, L0), sig2, rtol=2e-3)


def test_structure_function_is_zero_at_zero_and_increasing():
    r = np.linspace(0.0, 50.0, 200)
    d = ot.structure_function(r, 0.16, 25.0)
    assert d[0] == 0.0 and np.all(np.diff(d) > 0)


def test_fft_screens_follow_the_structure_function():
    n, dx, r0, L0 = 256, 0.05, 0.16, 25.0
    scr = screens.vonkarman_many(n, dx, r0, L0, seeds=range(1, 9))
    lags = np.array([1, 2, 4, 8])
    emp = screens.structure_function(scr, 8)[lags - 1]
    ref = ot.structure_function(lags * dx, r0, L0)
    np.testing.assert_allclose(emp, ref, rtol=0.15)
