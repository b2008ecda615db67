"""GPU parity of the SH WFS path against the float64 oracle (noise off).

Bar: spots / image / flux within 1e-5 of the oracle relative to the maximum
(the reference tests' relative_array_error), CoG-family slopes within 1e-5
relative to the largest |slope|, index maps bit-exact.
"""
import numpy as np
import pytest

from conftest import relative_array_error
from oracle import centroid as oc
from oracle import wfs as ow
from shwfs import configs, system

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _oracle(sysm, phase):
    t = sysm.tables()
    b, inten, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                                  t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                                  t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                                  t["nxsub"])
    return b, inten, img


def _check_image(sysm, phase):
    sysm.set_phase(phase)
    sysm.comp_image(noise=False)
    b, inten, img = _oracle(sysm, phase)
    eb = relative_array_error(sysm.bincube(), b)
    ei = relative_array_error(sysm.binimg(), img)
    ef = relative_array_error(sysm.intensities(), inten)
    print("spot err %.2e image err %.2e flux err %.2e" % (eb, ei, ef), flush=True)
    assert eb < TOL and ei < TOL and ef < TOL
    return img


@pytest.mark.parametrize("name", ["small", "large"])
def test_spots_match_oracle(name):
    c = configs.get(name)
    c["p_centroider"].type = "cog"
    s = system.ShSystem(c)
    phase = system.make_screens(c, 1)[0]
    _check_image(s, phase)


# Plain-path FFT variants: N = 64, one-pixel binning, the binmap gather, and
# windows smaller than the FFT (the shapes of the MFMA variant tests); b74:
# the 74x74 WFS of the uneven batch (8 phase points in a 16-point transform,
# 2x2 binning to 6x6 pixels: the central 12x12 of the 16x16 grid, block
# origin 10).
FFT_SHAPES = {
    "b74": dict(diam=8.0, nxsub=74, pdiam=8, npix=6, nfft=16, nrebin=2),
    "n64": dict(diam=8.0, nxsub=8, pdiam=32, npix=16, nfft=64, nrebin=4),
    "rebin1": dict(diam=4.0, nxsub=10, pdiam=16, npix=12, nfft=16, nrebin=1),
    "gather": dict(diam=8.0, nxsub=24, pdiam=16, npix=7, nfft=32, nrebin=4),
    "masked32": dict(diam=8.0, nxsub=20, pdiam=24, npix=8, nfft=32, nrebin=4),
}


@pytest.mark.parametrize("name", sorted(FFT_SHAPES))
def test_fft_variants_match_oracle(name):
    c = configs.sh_config(**FFT_SHAPES[name], r0=0.16, seeds=(3,), name=name)
    s = system.ShSystem(c)
    phase = system.make_screens(c, 1)[0]
    _check_image(s, phase)


def test_small_cog_slopes_match_oracle():
    c = configs.get("small")
    s = system.ShSystem(c)
    phase = system.make_screens(c, 1)[0]
    img = _check_image(s, phase)
    s.do_centroids()
    w = c["p_wfs"]
    ref, _ = oc.cog(img, w._validsubsx, w._validsubsy, w.npix, s.centro.offset, s.centro.scale)
    got = s.slopes()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("slope err %.2e" % err, flush=True)
    assert err < TOL


def test_large_tcog_fp64_matches_oracle():
    c = configs.get("large")
    s = system.ShSystem(c, fp64_moments=True)
    phase = system.make_screens(c, 1)[0]
    img = _check_image(s, phase)
    thr = 0.1 * float(img.max())
    s.centro.set_threshold(thr)
    s.do_centroids()
    w = c["p_wfs"]
    ref, _, img_out = oc.tcog(img, w._validsubsx, w._validsubsy, w.npix, s.centro.offset,
                              s.centro.scale, thr)
    got = s.slopes()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("tcog slope err %.2e" % err, flush=True)
    assert err < TOL
    # the reference's side effect: thresholded pixels written back into the image
    assert relative_array_error(s.binimg(), img_out) < TOL


def test_unit_pupil_flags_match_the_pupil():
    c = configs.get("large")
    s = system.ShSystem(c)
    t = s.tables()
    pup, pd = t["pupil"], t["pdiam"]
    expect = np.array([np.all(pup[i:i + pd, j:j + pd] == 1.0)
                       for i, j in zip(t["istart"], t["jstart"])], dtype=np.int32)
    got = np.array(s.wfs.d_pupil_unit).ravel()
    print("unit-pupil sub-apertures: %d of %d" % (got.sum(), got.size), flush=True)
    np.testing.assert_array_equal(got, expect)
    assert 0 < got.sum() < got.size
