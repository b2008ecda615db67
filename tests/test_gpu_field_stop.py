"""Field stop (shwfs_field_stop, Sensors.set_field_stop) against the float64
oracle (oracle/wfs.py: field_stop, comp_image with submask): the stopped
pupil field at the SH sizes and at the 80x80 size (n = 2048), spots, image
and CoG slopes of WFS frames with round and square stops on every method,
and an all-pass stop reproducing the frame without a stop."""
import numpy as np
import pytest

import capi
from conftest import relative_array_error
from oracle import centroid as oc
from oracle import geometry as og
from oracle import wfs as ow
from shwfs import configs, system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("npup,n,fstop", [(68, 128, "round"), (132, 256, "square"),
                                          (1284, 2048, "round")])
def test_field_stop_matches_oracle(npup, n, fstop):
    import torch
    lib = capi.load()
    rng = np.random.default_rng(npup)
    phase = (3.0 * rng.standard_normal((npup, npup))).astype(np.float32)  # [x, y]
    pupil = og.padded_pupil(npup - 4, 0.12).astype(np.float32)
    mask = og.field_stop_mask(npup, npup - 4, 0.5, 8.0, 4.0 if n < 2048 else 1.0, fstop)
    mask = mask.astype(np.float32)
    scale = np.float32(2 * np.pi / 0.5)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a.T.ravel()), device=dev)  # x fastest
    d_ph, d_pu, d_m = t(phase), t(pupil), t(mask)
    d_work = torch.zeros(2 * n * npup, dtype=torch.float32, device=dev)
    d_oph = torch.zeros(npup * npup, dtype=torch.float32, device=dev)
    d_oam = torch.zeros(npup * npup, dtype=torch.float32, device=dev)
    print("field stop npup=%d n=%d %s (mask keeps %d of %d)" % (npup, n, fstop, int(mask.sum()),
                                                               n * n), flush=True)
    rc = lib.shwfs_field_stop(capi.ptr(d_ph), capi.ptr(d_pu), npup, float(scale), capi.ptr(d_m), n,
                              capi.ptr(d_work), capi.ptr(d_oph), capi.ptr(d_oam), None)
    assert rc == 0, capi.last_error(lib)
    torch.cuda.synchronize()
    amp = d_oam.cpu().numpy().reshape(npup, npup).T.astype(np.float64)
    ph = d_oph.cpu().numpy().reshape(npup, npup).T.astype(np.float64)
    got = amp * np.exp(1j * ph * float(scale))
    want = ow.field_stop(phase, pupil, scale, mask)
    err = np.abs(got - want).max() / np.abs(want).max()
    print("  field error %.2e of max |F| %.3f" % (err, np.abs(want).max()), flush=True)
    assert err < 5e-6


def _frame(name, method, fssize, fstop):
    cfg = configs.get(name)
    cfg["p_wfs"].noise = -1.0
    cfg["p_wfs"].fssize, cfg["p_wfs"].fstop = fssize, fstop
    s = system.ShSystem(cfg, method=method)
    phase = system.make_screens(s.cfg, 1)[0]
    s.set_phase(phase)
    s.comp_image(noise=False)
    return s, phase


@pytest.mark.parametrize("name,method,fssize,fstop", [
    ("small", 0, 1.0, "round"), ("small", 1, 1.0, "square"), ("small", 2, 0.6, "round"),
    ("noisy", 0, 1.5, "square"), ("large", 1, 1.5, "round")])
def test_frame_with_field_stop_matches_oracle(name, method, fssize, fstop):
    s, phase = _frame(name, method, fssize, fstop)
    t, w = s.tables(), s.p_wfs
    assert w._submask is not None and s.wfs.d_submask is not None
    print("%s method %d: fssize %.2f %s, mask %d of %d, nvalid %d" %
          (name, method, fssize, fstop, int(w._submask.sum()), w._submask.size, w._nvalid),
          flush=True)
    b, _, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                              t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                              t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                              t["nxsub"], submask=w._submask)
    b_open, _, _ = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                                 t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                                 t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                                 t["nxsub"])
    eb = relative_array_error(s.bincube(), b)
    ei = relative_array_error(s.binimg(), img)
    moved = relative_array_error(b_open, b)  # the stop changes the spots
    print("  spots %.2e image %.2e (stop changes the spots by %.2e)" % (eb, ei, moved), flush=True)
    bar = 1e-5 if method == 0 else (1e-4 if method == 1 else 5e-4)
    assert moved > 10 * bar
    assert eb < bar and ei < bar
    if method == 0 and s.cfg["p_centroider"].type == "cog":
        s.centro.get_cog()
        got = np.array(s.wfs.d_slopes).ravel()
        ref, _ = oc.cog(img, w._validsubsx, w._validsubsy, w.npix, s.centro.offset, s.centro.scale)
        assert relative_array_error(got, ref) < 1e-5


def test_all_pass_stop_equals_no_stop():
    s0, phase = _frame("small", 0, 0.0, "round")
    spots_open, img_open = s0.bincube().copy(), s0.binimg().copy()
    n = 128
    s0.sensors.set_field_stop(0, np.ones((n, n), np.float32), n)
    s0.set_phase(phase)
    s0.comp_image(noise=False)
    print("all-pass stop: spots %.2e image %.2e" % (relative_array_error(s0.bincube(), spots_open),
                                                    relative_array_error(s0.binimg(), img_open)),
          flush=True)
    assert relative_array_error(s0.bincube(), spots_open) < 1e-5
    assert relative_array_error(s0.binimg(), img_open) < 1e-5
