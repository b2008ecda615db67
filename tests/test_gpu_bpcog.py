"""Brightest-pixel CoG (CentroiderBPCOG_FF / _UF) against the oracle
(oracle/centroid.py: bpcog), on the detector image of a WFS frame loaded into
a standalone centroider, as the reference's RTC tests load images
(shesha/tests/pytest/rtc/test_rtcUFF.py:45-52)."""
import numpy as np
import pytest

from oracle import camera
from oracle import centroid as oc
from shwfs import configs, system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("uint16", [False, True])
def test_bpcog_matches_oracle(uint16):
    import sutra
    c = configs.get("small")
    c["nphot"] = 1e4
    s = system.ShSystem(c)
    s.set_phase(system.make_screens(c, 1)[0])
    s.comp_image(noise=False)
    img = s.binimg()
    if uint16:  # a camera frame: the image digitised on the host by the oracle
        img = camera.digitalize(img, np.zeros(img.shape), np.ones(img.shape), 2 ** 15, 2 ** 15)
    p = c["p_wfs"]
    nvalid = int(s.wfs.nvalid)
    vx = np.array(s.centro.d_validx).astype(np.int32)
    vy = np.array(s.centro.d_validy).astype(np.int32)
    rtc = sutra.Rtc_UFF() if uint16 else sutra.Rtc_FFF()
    rtc.add_centroider(s.ctx, nvalid, p.npix / 2 - 0.5, p.pixsize, False, 0, "bpcog")
    rtc.add_controller(s.ctx, "generic", 0, 0.0, 2 * nvalid, 1,
                       idx_centro=np.zeros(1, np.int32), ncentro=1)
    centro = rtc.d_centro[0]
    assert type(centro).__name__ == ("CentroiderBPCOG_UF" if uint16 else "CentroiderBPCOG_FF")
    assert centro.nmax == 10  # the reference factory's default
    centro.set_npix(p.npix)
    centro.load_validpos(vx, vy, nvalid)
    n = img.shape[0]
    for nmax in (10, 3, p.npix * p.npix - 1):
        centro.set_nmax(nmax)
        centro.load_img(np.asfortranarray(img), n)
        centro.calibrate_img()
        print("bpcog uint16=%s nmax=%d nvalid=%d" % (uint16, nmax, nvalid), flush=True)
        rtc.do_centroids(0)
        slopes = np.array(rtc.d_control[0].d_centroids)
        ref, inten, out = oc.bpcog(img.astype(np.float64), vx, vy, p.npix, nmax,
                                   p.npix / 2 - 0.5, p.pixsize)
        err = np.abs(slopes - ref).max() / np.abs(ref).max()
        img_err = np.abs(np.array(centro.d_img) - out).max() / np.abs(out).max()
        print("  slope err %.2e, thresholded image err %.2e" % (err, img_err), flush=True)
        assert err < 1e-5
        assert img_err < 1e-6
        assert np.allclose(np.array(centro.d_intensities), inten, rtol=1e-5)
