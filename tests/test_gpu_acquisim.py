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


def test_fill_bincube_known_answer():
    """shwfs_fill_bincube against the oracle's restatement of bcube_krnl
    (oracle/corr.py: acquisim_bincube) on a random image and random window
    positions, bit for bit.  The row stride is the image side npix * nxsub
    (the stride the reference's correlation centroider passes); the
    reference's acquisim passes npix * nxsub as `nsub`, which the kernel
    multiplies by npix again."""
    import torch

    import capi
    from oracle import corr as oc
    lib = capi.load()
    rng = np.random.default_rng(21)
    npix, nxsub = 6, 9
    size = npix * nxsub
    img = rng.standard_normal((size, size)).astype(np.float32)  # [x, y]
    nvalid = 31
    vx = rng.integers(0, size - npix + 1, nvalid).astype(np.int32)
    vy = rng.integers(0, size - npix + 1, nvalid).astype(np.int32)
    dev = torch.device("cuda:0")
    d_img = torch.tensor(np.ascontiguousarray(img.T.ravel()), device=dev)  # x fastest
    d_vx = torch.tensor(vx, device=dev)
    d_vy = torch.tensor(vy, device=dev)
    d_cube = torch.full((npix * npix * nvalid,), np.nan, dtype=torch.float32, device=dev)
    print("fill_bincube known answer: npix %d nxsub %d nvalid %d" % (npix, nxsub, nvalid), flush=True)
    rc = lib.shwfs_fill_bincube(capi.ptr(d_img), capi.ptr(d_cube), capi.ptr(d_vx), capi.ptr(d_vy),
                                npix, nvalid, size, None)
    assert rc == 0, capi.last_error(lib)
    got = d_cube.cpu().numpy().reshape(nvalid, npix, npix)  # [s, y, x]
    want = oc.acquisim_bincube(img, vx, vy, npix, size)      # [x, y, s]
    np.testing.assert_array_equal(got, want.transpose(2, 1, 0))
