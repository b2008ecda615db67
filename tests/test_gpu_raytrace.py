"""Guide-star raytrace: shwfs_raytrace against the oracle (float32
coordinates, float64 interpolation) and its closed forms, and the spot kernel
that traces its windows through the layers (Source.raytrace keeps the layers
pending; comp_image traces them in its field stage) against rendering the
screen first, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import capi
from conftest import relative_array_error
from oracle import raytrace as ort
from shwfs import configs, system

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).ravel(order="F"))).cuda()


def _layers(specs, keep):
    arr = (capi.Layer * len(specs))()
    for k, sp in enumerate(specs):
        t = _dev(sp["screen"])
        keep.append(t)
        arr[k] = capi.Layer(capi.ptr(t), sp["screen"].shape[0], sp.get("xoff", 0.0),
                            sp.get("yoff", 0.0), sp.get("G", 1.0), sp.get("thetaML", 0.0),
                            sp.get("dx", 0.0), sp.get("dy", 0.0), sp.get("delta", 1.0))
    return arr


def _trace(specs, nx, ny, accumulate=0, init=None):
    import torch
    lib = capi.load()
    keep = []
    arr = _layers(specs, keep)
    out = _dev(np.zeros((nx, ny)) if init is None else init)
    rc = lib.shwfs_raytrace(capi.ptr(out), nx, ny, arr, len(specs), accumulate, None)
    assert rc == 0, capi.last_error(lib)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(nx, ny, order="F")


def test_raytrace_matches_oracle():
    rng = np.random.default_rng(3)
    specs = [dict(screen=rng.normal(size=(160, 160)), xoff=21.37, yoff=9.83),
             dict(screen=rng.normal(size=(200, 200)), xoff=40.5, yoff=33.25, G=1.07,
                  thetaML=0.12, dx=0.3, dy=-0.45, delta=0.93)]
    got = _trace(specs, 132, 132)
    ref = ort.raytrace(specs, 132, 132)
    err = relative_array_error(got, ref)
    print("raytrace: rel err %.2e" % err, flush=True)
    assert err < 1e-5
    # accumulate onto an existing screen
    base = rng.normal(size=(132, 132)).astype(np.float32)
    got2 = _trace(specs, 132, 132, accumulate=1, init=base)
    assert relative_array_error(got2, ref + base) < 1e-5


def test_integer_shift_is_an_exact_copy():
    rng = np.random.default_rng(4)
    L = rng.normal(size=(96, 96)).astype(np.float32)
    got = _trace([dict(screen=L, xoff=7.0, yoff=12.0)], 68, 68)
    assert np.array_equal(got, L[7:75, 12:80])


def _atmos(s, seeds=(5, 6), altitude=None):
    n = s.p_geom._n
    dims = (n + 100, n + 160)
    carma, sutra = s.carma, s.sutra
    n = len(seeds)
    alt = np.zeros(n, np.float32) if altitude is None else np.asarray(altitude, np.float32)
    at = sutra.Atmos(s.ctx, n, 0.16, np.full(n, 10.0, np.float32), np.array(dims, np.int64),
                     np.full(n, 4, np.int64), alt, np.zeros(n, np.float32),
                     np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32), 0)
    screens = []
    for k, (sd, d) in enumerate(zip(seeds, dims)):
        scr = system.screens.vonkarman(d, s.p_tel.diam / s.p_geom.pupdiam, 0.2, 25.0, sd)
        scr = np.asfortranarray(scr.astype(np.float32))
        at.d_screens[k].d_screen.host2device(scr)
        screens.append(scr)
    return at, screens


@pytest.mark.parametrize("name", ["small", "large"])
def test_traced_field_stage_equals_rendered_screen(name):
    s = system.ShSystem(configs.get(name))
    at, screens = _atmos(s)
    gs = s.wfs.d_gs
    offs = [(13.4, 27.8), (50.25, 4.5)]
    for k, (xo, yo) in enumerate(offs):
        gs.add_layer("atmos", k, xo, yo)
    gs.raytrace(at)
    assert gs.npending == 2
    s.comp_image(noise=False)
    s.do_centroids()
    traced = (s.bincube(), s.binimg(), s.slopes())
    assert gs.npending == 2   # still pending: the screen was never rendered
    gs.lazy_raytrace = False
    gs.raytrace(at)           # renders the screen, as the reference
    assert gs.npending == 0
    n = s.p_geom._n
    phase = np.array(gs.d_phase)
    ref = ort.raytrace([dict(screen=screens[k], xoff=o[0], yoff=o[1]) for k, o in enumerate(offs)],
                       n, n)
    err = relative_array_error(phase, ref)
    print("%s: rendered screen rel err %.2e" % (name, err), flush=True)
    assert err < 1e-5
    s.comp_image(noise=False)
    s.do_centroids()
    rendered = (s.bincube(), s.binimg(), s.slopes())
    for a, b in zip(traced, rendered):
        assert np.array_equal(a, b)


def test_lgs_source_compresses_the_layers():
    """An LGS WFS's source sees each layer through the cone: delta = 1 -
    altitude / hg (sutra_source.cpp:250-252), hg the sodium layer's centre of
    gravity set by lgs_init."""
    s = system.ShSystem(configs.lgs())
    alts = (5000.0, 12000.0)
    at, screens = _atmos(s, altitude=alts)
    gs = s.wfs.d_gs
    hg = gs.d_lgs.hg
    assert gs.lgs and abs(hg - 90000.0) < 1.0
    offs = [(13.4, 27.8), (50.25, 4.5)]
    for k, (xo, yo) in enumerate(offs):
        gs.add_layer("atmos", k, xo, yo)
    gs.lazy_raytrace = False
    gs.raytrace(at)
    n = s.p_geom._n
    ref = ort.raytrace([dict(screen=screens[k], xoff=o[0], yoff=o[1],
                             delta=float(np.float32(1.0 - np.float32(alts[k]) / np.float32(hg))))
                        for k, o in enumerate(offs)], n, n)
    err = relative_array_error(np.array(gs.d_phase), ref)
    print("lgs raytrace (hg %.1f): rel err %.2e" % (hg, err), flush=True)
    assert err < 1e-5
