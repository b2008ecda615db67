"""The CoG-family oracle against scipy.ndimage.center_of_mass (the check the
reference's own centroider tests use, test_rtcFFF.py:154-164, 255-286,
324-363), plus the closed-form special cases of the threshold / epsilon
semantics and the slope orderings."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import centroid as oc


def _case(seed, npix=8, nxsub=5):
    rng = np.random.default_rng(seed)
    size = npix * nxsub
    img = rng.gamma(2.0, 50.0, size=(size, size))
    sub = [(i, j) for i in range(nxsub) for j in range(nxsub) if (i + j) % 3]
    vx = np.array([i * npix for i, _ in sub])
    vy = np.array([j * npix for _, j in sub])
    return img, vx, vy, npix


def _com_slopes(img, vx, vy, npix, offset, scale, f=lambda w, s: w):
    out = np.zeros(2 * len(vx))
    for s, (x, y) in enumerate(zip(vx, vy)):
        w = f(img[x:x + npix, y:y + npix], s)
        cx, cy = ndimage.center_of_mass(w)
        out[s], out[s + len(vx)] = (cx - offset) * scale, (cy - offset) * scale
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_cog_matches_center_of_mass(seed):
    img, vx, vy, npix = _case(seed)
    off, sc = npix / 2 - 0.5, 0.37
    got, inten = oc.cog(img, vx, vy, npix, off, sc)
    ref = _com_slopes(img, vx, vy, npix, off, sc)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-7)
    np.testing.assert_allclose(inten, [img[x:x + npix, y:y + npix].sum() for x, y in zip(vx, vy)])


def test_tcog_matches_center_of_mass_of_thresholded_pixels_and_writes_back():
    img, vx, vy, npix = _case(3)
    off, sc, t = npix / 2 - 0.5, 0.5, 120.0
    got, _, out = oc.tcog(img, vx, vy, npix, off, sc, t)
    ref = _com_slopes(img, vx, vy, npix, off, sc,
                      lambda w, s: np.where(w > t, w - t, 1e-6))
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)
    # the windows now hold d' = (img > t ? img - t : 1e-6); the rest is untouched
    mask = np.zeros(img.shape, bool)
    for x, y in zip(vx, vy):
        mask[x:x + npix, y:y + npix] = True
    np.testing.assert_array_equal(out[~mask], img[~mask])
    np.testing.assert_allclose(out[mask], np.where(img[mask] > t, img[mask] - t, 1e-6))


def test_tcog_with_everything_below_threshold_keeps_the_epsilon_terms():
    # every pixel contributes 1e-6 and the denominator adds 1e-6:
    # Sx / (I + 1e-6) = npix * sum(x) * 1e-6 / (npix^2 * 1e-6 + 1e-6), not the
    # window centre
    img, vx, vy, npix = _case(4)
    off = npix / 2 - 0.5
    got, inten, _ = oc.tcog(img, vx, vy, npix, off, 1.0, threshold=1e9)
    sx = npix * (npix * (npix - 1) / 2) * 1e-6
    expect = sx / (npix * npix * 1e-6 + 1e-6) - off
    np.testing.assert_allclose(got, expect, rtol=1e-12)
    np.testing.assert_allclose(inten, npix * npix * 1e-6)


def test_wcog_matches_center_of_mass_of_weighted_pixels():
    img, vx, vy, npix = _case(5)
    rng = np.random.default_rng(9)
    w = rng.uniform(0.1, 1.0, size=(npix, npix, len(vx)))
    off, sc, t = npix / 2 - 0.5, 0.25, 30.0
    got, _ = oc.wcog(img, vx, vy, npix, off, sc, w, threshold=t)
    ref = _com_slopes(img, vx, vy, npix, off, sc,
                      lambda win, s: np.where(win > t, win - t, 1e-6) * w[:, :, s])
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


def test_interlaced_order_and_reference_slopes():
    img, vx, vy, npix = _case(6)
    n = len(vx)
    untied, _ = oc.cog(img, vx, vy, npix, 3.5, 1.0)
    ref = np.linspace(-1, 1, 2 * n)
    inter, _ = oc.cog(img, vx, vy, npix, 3.5, 1.0, ref=ref, order=1)
    np.testing.assert_allclose(inter[0::2] + ref[0::2], untied[:n])
    np.testing.assert_allclose(inter[1::2] + ref[1::2], untied[n:])


def test_gaussian_weights_peak_at_the_spot_centre():
    w = oc.gaussian_weights(8, 4.0)
    # symmetric about npix/2 - 0.5 and maximal on the four central pixels
    np.testing.assert_allclose(w, w[::-1, :])
    np.testing.assert_allclose(w, w.T)
    assert w.argmax() in {3 * 8 + 3, 3 * 8 + 4, 4 * 8 + 3, 4 * 8 + 4}
    # exp(-(r / (width / 1.66))^2) at r = 0.5 * sqrt(2)
    np.testing.assert_allclose(w[3, 3], np.exp(-((np.sqrt(0.5) / (4.0 / 1.66)) ** 2)), rtol=1e-6)
