"""Pins of the geometric-slope oracle (oracle/slopes.py) and of the tip/tilt
filter restatement (oracle/centroid.py: tt_filter) by closed forms:

* a phase ramp phi = gx x + gy y gives the wavefront tilt
  0.206265 (gx, gy) nphase / subapd arcsec through the edge differences
  (type 0), the central differences (type 1, full windows) and the tip/tilt
  projection matrix built by the product geometry (type 2, every window);
* a quadratic phi = c x^2 gives c (n-1)^2 alpha along x for types 0 and 1
  (the central differences of x^2 sum to the same as the edge difference)
  and 0 along y: a transposed index or a dropped edge term fails it;
* a constant phase gives 0;
* the TT filter on a hand-computed two-element case with non-orthogonal
  tip/tilt vectors (both projections on the unfiltered slopes).
"""
import numpy as np
import pytest

from oracle import centroid as oc
from oracle import slopes as osl
from shwfs import configs, geom, system


def _setup(name):
    c = configs.get(name)
    system.geometry(c)
    w, g = c["p_wfs"], c["p_geom"]
    return c, w, g


def _ramp_screen(n, gx, gy, c0=0.3):
    x = np.arange(n, dtype=np.float64)[:, None]
    y = np.arange(n, dtype=np.float64)[None, :]
    return c0 + gx * x + gy * y  # [x, y]


@pytest.mark.parametrize("name", ["small", "large"])
def test_ramp_gives_the_tilt_for_every_type(name):
    c, w, g = _setup(name)
    n, sd = w._pdiam, w._subapd
    gx, gy = 0.013, -0.007
    phase = _ramp_screen(g._n, gx, gy)
    ex, ey = osl.ramp_slopes(gx, gy, n, sd)
    nv = w._nvalid
    flux = geom.flux_per_valid_subap(w)
    full = flux == 1.0
    s0 = osl.slopes_geom(0, phase, w._phasemap, n, subapd=sd)
    np.testing.assert_allclose(s0[:nv], ex, rtol=1e-12)
    np.testing.assert_allclose(s0[nv:], ey, rtol=1e-12)
    s1 = osl.slopes_geom(1, phase, w._phasemap, n, subapd=sd, pupil=g._mpupil, flux=flux)
    np.testing.assert_allclose(s1[:nv][full], ex, rtol=1e-12)
    np.testing.assert_allclose(s1[nv:][full], ey, rtol=1e-12)
    s2 = osl.slopes_geom(2, phase, w._phasemap, n, ttprojmat=w._ttprojmat)
    # ttprojmat is float32 and uses RAD2ARCSEC = 206264.806 (not 0.206265e6)
    np.testing.assert_allclose(s2[:nv], ex, rtol=2e-6)
    np.testing.assert_allclose(s2[nv:], ey, rtol=2e-6)


def test_quadratic_phase_pins_the_difference_stencils():
    c, w, g = _setup("small")
    n, sd, nv = w._pdiam, w._subapd, w._nvalid
    cq = 0.002
    # phi = cq * (x - x0)^2 inside one window (x0 its corner), zero elsewhere:
    # one window at a time, since the windows of neighbours touch
    flat = np.zeros(g._n * g._n)
    pm = np.asarray(w._phasemap).reshape(n * n, nv, order="F")
    a = osl.alpha(sd, n)
    flux = geom.flux_per_valid_subap(w)
    xs = (np.arange(n * n) % n).astype(np.float64)
    for s in np.where(flux == 1.0)[0][:4]:
        flat[:] = 0.0
        flat[pm[:, s]] = cq * xs ** 2
        phase = flat.reshape(g._n, g._n, order="F")
        for kind in (0, 1):
            out = osl.slopes_geom(kind, phase, w._phasemap, n, subapd=sd, pupil=g._mpupil,
                                  flux=flux)
            assert out[s] == pytest.approx(cq * (n - 1) ** 2 * a, rel=1e-12)
            assert abs(out[nv + s]) < 1e-15


def test_constant_phase_gives_zero():
    c, w, g = _setup("small")
    n, sd = w._pdiam, w._subapd
    phase = np.full((g._n, g._n), 1.7)
    flux = geom.flux_per_valid_subap(w)
    for kind in (0, 1, 2):
        out = osl.slopes_geom(kind, phase, w._phasemap, n, subapd=sd, pupil=g._mpupil, flux=flux,
                              ttprojmat=w._ttprojmat)
        assert np.abs(out).max() < 1e-6


def test_tt_filter_projects_on_the_unfiltered_slopes():
    # non-orthogonal reference vectors: a = <s, tip> = 1, b = <s, tilt> = 3
    out, (a, b) = oc.tt_filter([1.0, 2.0], [1.0, 0.0], [1.0, 1.0])
    assert (a, b) == (1.0, 3.0)
    np.testing.assert_array_equal(out, [-3.0, -1.0])
    # orthonormal references: the result is orthogonal to both, and a second
    # pass removes nothing
    rng = np.random.default_rng(0)
    q, _ = np.linalg.qr(rng.standard_normal((40, 2)))
    s = rng.standard_normal(40)
    f, _ = oc.tt_filter(s, q[:, 0], q[:, 1])
    assert abs(f @ q[:, 0]) < 1e-12 and abs(f @ q[:, 1]) < 1e-12
    f2, (a2, b2) = oc.tt_filter(f, q[:, 0], q[:, 1])
    assert abs(a2) < 1e-12 and abs(b2) < 1e-12
    np.testing.assert_allclose(f2, f, atol=1e-12)
