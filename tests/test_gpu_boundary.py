"""Boundary behaviours of SutraWfsSH that are not the fused frame itself:

* roket: the detector image before the noise (d_binimg_notnoisy, filled by
  the reference before the noise, libsutra/src.cpp/sutra_wfs_sh.cpp:572-578)
  and the CoG of it through get_cog(noise=False);
* fill_binimage after set_bincube re-scatters the spot cube into the
  detector image (sutra_wfs_sh.cpp:599-621);
* set_bincube / set_binimg with a wrong size raise (a defined error of the
  C ABI, shwfs_upload) instead of leaving stale data.
"""
import numpy as np
import pytest

from conftest import relative_array_error
from oracle import centroid as oc
from oracle import wfs as ow
from shwfs import configs, system

pytestmark = pytest.mark.gpu


def _noisy(roket, seed=21):
    c = configs.get("noisy")
    c["p_centroider"].type = "cog"
    c["roket"] = roket
    s = system.ShSystem(c)
    s.wfs.set_noise(0.5, seed)
    return c, s


def test_roket_keeps_the_image_before_the_noise():
    c, plain = _noisy(False)
    _, rok = _noisy(True)
    phase = system.make_screens(c, 1)[0]
    for s in (plain, rok):
        s.set_phase(phase)
        s.comp_image(noise=True)
    # the noisy outputs are those of the fused launch, bit for bit
    assert np.array_equal(rok.bincube(), plain.bincube())
    assert np.array_equal(rok.binimg(), plain.binimg())
    # the clean image: the noise-free spots of the same phase (zeros elsewhere)
    t = rok.tables()
    _, _, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                              t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                              t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                              t["nxsub"])
    clean = np.array(rok.wfs.d_binimg_notnoisy)
    err = relative_array_error(clean, img)
    print("roket clean image err %.2e" % err, flush=True)
    assert err < 1e-5
    # get_cog(noise=False) reads the clean image, get_cog(noise=True) the noisy one
    w, cen = rok.p_wfs, rok.centro
    rok.rtc.do_centroids(0, False)
    ref_clean, _ = oc.cog(img, w._validsubsx, w._validsubsy, w.npix, cen.offset, cen.scale)
    e_clean = np.abs(rok.slopes() - ref_clean).max() / np.abs(ref_clean).max()
    rok.rtc.do_centroids(0, True)
    ref_noisy, _ = oc.cog(rok.binimg(), w._validsubsx, w._validsubsy, w.npix, cen.offset,
                          cen.scale)
    e_noisy = np.abs(rok.slopes() - ref_noisy).max() / np.abs(ref_noisy).max()
    print("roket cog: clean err %.2e, noisy err %.2e" % (e_clean, e_noisy), flush=True)
    assert e_clean < 1e-5 and e_noisy < 1e-5


def test_fill_binimage_rescatters_the_spot_cube():
    c = configs.get("small")
    s = system.ShSystem(c)
    s.set_phase(system.make_screens(c, 1)[0])
    s.comp_image(noise=False)
    w = s.p_wfs
    before = s.binimg()
    rng = np.random.default_rng(3)
    cube = rng.uniform(0, 100, (w.npix, w.npix, w._nvalid)).astype(np.float32)
    s.wfs.set_bincube(np.asfortranarray(cube), cube.size)
    assert np.array_equal(s.bincube(), cube)
    s.wfs.fill_binimage()
    got = s.binimg()
    exp = ow.binimg_from_cube(cube.astype(np.float64), w._validsubsx, w._validsubsy,
                              w.npix * w.nxsub, base=before)
    assert np.array_equal(got, exp.astype(np.float32))


def test_set_bincube_and_binimg_reject_a_wrong_size():
    c = configs.get("small")
    s = system.ShSystem(c)
    w = s.p_wfs
    bad = np.zeros(w.npix * w.npix * w._nvalid - 1, dtype=np.float32)
    with pytest.raises(RuntimeError, match="elements given"):
        s.wfs.set_bincube(bad, bad.size)
    img = np.zeros(10, dtype=np.float32)
    with pytest.raises(RuntimeError, match="elements given"):
        s.wfs.set_binimg(img, img.size)
