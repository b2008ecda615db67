"""Simulated camera image (fakecam) and centroids of uint16 camera frames
(Rtc_UFF), checked against the oracle (oracle/camera.py, oracle/centroid.py).
The flow mirrors the reference's own test of the uint16 RTC
(shesha/tests/pytest/rtc/test_rtcUFF.py:15-60): a fakecam WFS produces the
frame, a standalone uint16 centroider loads, calibrates and centroids it."""
import numpy as np
import pytest

import capi
from oracle import camera
from oracle import centroid as oc
from shwfs import configs, system

pytestmark = pytest.mark.gpu


def test_digitalize_and_calibrate_u16_match_oracle():
    import torch
    lib = capi.load()
    rng = np.random.default_rng(5)
    n = 200_000
    # negatives, values past max_flux_per_pix (saturation) and a random dark / flat
    img = rng.uniform(-50, 3000, n).astype(np.float32)
    dark = rng.uniform(-5, 5, n).astype(np.float32)
    flat = rng.uniform(0.5, 1.5, n).astype(np.float32)
    mfpp, mpv = 2500, 65535
    d_img, d_dark, d_flat = (torch.from_numpy(a).cuda() for a in (img, dark, flat))
    d_cam = torch.zeros(n, dtype=torch.int16, device="cuda")  # uint16 bits
    print("digitalize n=%d" % n, flush=True)
    assert lib.shwfs_digitalize(capi.ptr(d_cam), capi.ptr(d_img), capi.ptr(d_dark),
                                capi.ptr(d_flat), mfpp, mpv, n, None) == 0
    torch.cuda.synchronize()
    cam = d_cam.cpu().numpy().view(np.uint16)
    ref = camera.digitalize(img, dark, flat, mfpp, mpv)
    print("saturated %d, zero %d" % ((ref == mpv).sum(), (ref == 0).sum()), flush=True)
    assert (ref == mpv).any() and (ref == 0).any()
    np.testing.assert_array_equal(cam, ref)
    # calibration of the uint16 frame through a reversed lookup table
    lut = np.ascontiguousarray(np.arange(n, dtype=np.int32)[::-1])
    d_lut = torch.from_numpy(lut).cuda()
    d_cal = torch.zeros(n, dtype=torch.float32, device="cuda")
    print("calibrate_u16 n=%d" % n, flush=True)
    assert lib.shwfs_calibrate_u16(capi.ptr(d_cam), capi.ptr(d_cal), capi.ptr(d_dark),
                                   capi.ptr(d_flat), capi.ptr(d_lut), n, None) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_cal.cpu().numpy(), camera.calibrate(ref, dark, flat, lut))


def _fakecam_system():
    c = configs.get("small")
    c["nphot"] = 1e5  # photons per sub-aperture: pixel values well inside int32
    s = system.ShSystem(c)
    s.set_phase(system.make_screens(c, 1)[0])
    w = s.wfs
    s.comp_image(noise=False)
    mfpp = max(1, int(s.binimg().max() // 2))  # the brightest pixels saturate
    w.set_max_flux_per_pix(mfpp)
    w.set_max_pix_value(2 ** 16 - 1)
    w.set_fakecam(True)  # allocates the camera image, dark 0 and flat 1
    return c, s, mfpp


def test_fakecam_camera_image_matches_oracle():
    c, s, mfpp = _fakecam_system()
    w = s.wfs
    shape = s.binimg().shape
    rng = np.random.default_rng(7)
    dark = np.asfortranarray(rng.uniform(0, 20, shape).astype(np.float32))
    flat = np.asfortranarray(rng.uniform(0.8, 1.2, shape).astype(np.float32))
    for dk, fl in ((np.zeros(shape, np.float32), np.ones(shape, np.float32)), (dark, flat)):
        w.set_dark(np.asfortranarray(dk), dk.size)
        w.set_flat(np.asfortranarray(fl), fl.size)
        s.comp_image(noise=False)
        img = s.binimg()
        cam = np.array(w.d_camimg)
        ref = camera.digitalize(img, dk, fl, mfpp, 2 ** 16 - 1)
        print("camera image %s, saturated %d" % (cam.shape, (ref == 2 ** 16 - 1).sum()), flush=True)
        assert cam.dtype == np.uint16
        np.testing.assert_array_equal(cam, ref)


def test_rtc_uff_centroids_of_the_camera_frame():
    import sutra
    c, s, mfpp = _fakecam_system()
    s.comp_image(noise=False)
    frame = np.array(s.wfs.d_camimg)
    p = c["p_wfs"]
    nvalid = int(s.wfs.nvalid)
    vx = np.array(s.centro.d_validx).astype(np.int32)
    vy = np.array(s.centro.d_validy).astype(np.int32)
    rtc = sutra.Rtc_UFF()
    rtc.add_centroider(s.ctx, nvalid, p.npix / 2 - 0.5, p.pixsize, False, 0, "cog")
    rtc.add_controller(s.ctx, "generic", 0, 0.0, 2 * nvalid, 1,
                       idx_centro=np.zeros(1, np.int32), ncentro=1)
    centro = rtc.d_centro[0]
    assert type(centro).__name__ == "CentroiderCOG_UF"
    centro.set_npix(p.npix)
    centro.load_validpos(vx, vy, nvalid)
    n = frame.shape[0]
    rng = np.random.default_rng(9)
    dark = np.asfortranarray(rng.uniform(0, 3, frame.shape).astype(np.float32))
    flat = np.asfortranarray(rng.uniform(0.9, 1.1, frame.shape).astype(np.float32))
    for dk, fl in ((None, None), (dark, flat)):
        centro.load_img(np.asfortranarray(frame), n)
        if dk is not None:
            centro.set_dark(dk, n)
            centro.set_flat(fl, n)
        centro.calibrate_img()
        rtc.do_centroids(0)
        slopes = np.array(rtc.d_control[0].d_centroids)
        dk0 = np.zeros(frame.shape, np.float32) if dk is None else dk
        fl0 = np.ones(frame.shape, np.float32) if fl is None else fl
        cal = camera.calibrate(frame, dk0, fl0, np.arange(frame.size)).reshape(frame.shape,
                                                                              order="F")
        np.testing.assert_array_equal(np.array(centro.d_img), cal)
        ref, _ = oc.cog(cal, vx, vy, p.npix, p.npix / 2 - 0.5, p.pixsize)
        err = np.abs(slopes - ref).max() / np.abs(ref).max()
        print("uint16 frame cog: slope err %.2e" % err, flush=True)
        assert err < 1e-5
