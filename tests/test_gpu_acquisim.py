"""Spots from an external detector image (Acquisim, the reference's
SutraAcquisim: libsutra/src.cpp/sutra_acquisim.cpp:52-76).  Gathering a WFS's
own noise-free detector image back must give its bincube bit for bit (the
detector image is the spots scattered at the valid positions), through the
valid positions and through the sub-aperture number map."""
import numpy as np
import pytest

from shwfs import configs, system

pytestmark = pytest.mark.gpu


def test_acquisim_gathers_the_spots_back():
    import sutra
    c = configs.get("small")
    s = system.ShSystem(c)
    screen = system.make_screens(c, 1)[0]
    s.set_phase(screen)
    s.comp_image(noise=False)
    spots, img = s.bincube().copy(), np.asfortranarray(s.binimg())
    acq = sutra.Acquisim(s.sensors, 0)
    npix, nxsub, nvalid = acq.npix, acq.nxsub, acq.nvalid
    assert img.shape == (npix * nxsub, npix * nxsub)
    vx = np.array(s.wfs.d_validsubsx).astype(np.int64)
    vy = np.array(s.wfs.d_validsubsy).astype(np.int64)
    num = np.zeros(nxsub * nxsub, np.int32)
    num[vx // npix + nxsub * (vy // npix)] = np.arange(1, nvalid + 1)
    for form in ("valid positions", "number map"):
        s.set_phase(np.roll(screen, 37, axis=0))  # another frame in the bincube first
        s.comp_image(noise=False)
        assert not np.array_equal(s.bincube(), spots)
        print("acquisim by %s: nvalid %d, npix %d" % (form, nvalid, npix), flush=True)
        if form == "valid positions":
            acq.comp_image(img)
        else:
            acq.comp_image_2D(img, num)
        np.testing.assert_array_equal(s.bincube(), spots)
