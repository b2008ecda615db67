"""Geometric slopes (the labels of the learned centroider) and the
centroider-base features around the CoG family on the GPU, against the
float64 oracle:

* slopes_geom types 0, 1, 2 (SutraWfs::slopes_geom, libsutra/src.cpp/
  sutra_wfs.cpp:199-245) on von Karman screens and on a phase ramp (closed
  form, oracle/slopes.py: ramp_slopes);
* reference slopes and the interlaced slope order (libsutra/src.cu/
  sutra_centroider_utils.cuh:31-51) for CoG / TCoG / WCoG, standalone
  (get_cog) and fused into the spot kernel (comp_frame);
* the tip/tilt filter (libsutra/src.cpp/sutra_centroider.cpp:322-338) with
  non-orthogonal reference vectors.

Bar: 1e-5 of the largest |slope| (float sums against float64).
"""
import numpy as np
import pytest

from oracle import centroid as oc
from oracle import slopes as osl
from oracle import wfs as ow
from shwfs import configs, geom, system

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _geom_slopes(s, kind):
    s.wfs.slopes_geom(kind)
    return np.array(s.wfs.d_slopes)


@pytest.mark.parametrize("name", ["small", "large"])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_slopes_geom_matches_oracle(name, kind):
    c = configs.get(name)
    s = system.ShSystem(c)
    phase = system.make_screens(c, 1)[0]
    s.set_phase(phase)
    w, g = s.p_wfs, s.p_geom
    got = _geom_slopes(s, kind)
    ref = osl.slopes_geom(kind, phase.astype(np.float32), w._phasemap, w._pdiam, subapd=w._subapd,
                          ttprojmat=w._ttprojmat, pupil=g._mpupil,
                          flux=geom.flux_per_valid_subap(w))
    err = _rel(got, ref)
    print("%s type %d: %d slopes, err %.2e" % (name, kind, got.size, err), flush=True)
    assert err < TOL


def test_slopes_geom_ramp_closed_form():
    c = configs.get("large")
    s = system.ShSystem(c)
    w, g = s.p_wfs, s.p_geom
    n = g._n
    gx, gy = 0.011, -0.004
    x = np.arange(n, dtype=np.float64)[:, None]
    y = np.arange(n, dtype=np.float64)[None, :]
    s.set_phase((gx * x + gy * y).astype(np.float32))
    ex, ey = osl.ramp_slopes(gx, gy, w._pdiam, w._subapd)
    nv = w._nvalid
    full = geom.flux_per_valid_subap(w) == 1.0
    for kind in (0, 1, 2):
        got = _geom_slopes(s, kind)
        gxs, gys = got[:nv], got[nv:]
        if kind == 1:
            gxs, gys = gxs[full], gys[full]
        ex_err = np.abs(gxs - ex).max() / abs(ex)
        ey_err = np.abs(gys - ey).max() / abs(ey)
        print("ramp type %d: x err %.2e, y err %.2e" % (kind, ex_err, ey_err), flush=True)
        # float32 screen values up to ~14 um, differences of neighbours
        assert ex_err < 1e-4 and ey_err < 2e-4


def _oracle_image(s, phase):
    t = s.tables()
    _, _, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                              t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                              t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                              t["nxsub"])
    return img


def _oracle_slopes(kind, img, w, c, thr, ref, order, weights=None):
    if kind == "cog":
        return oc.cog(img, w._validsubsx, w._validsubsy, w.npix, c.offset, c.scale, ref, order)[0]
    if kind == "tcog":
        return oc.tcog(img, w._validsubsx, w._validsubsy, w.npix, c.offset, c.scale, thr, ref,
                       order)[0]
    return oc.wcog(img, w._validsubsx, w._validsubsy, w.npix, c.offset, c.scale, weights, thr,
                   ref, order)[0]


@pytest.mark.parametrize("kind", ["cog", "tcog", "wcog"])
@pytest.mark.parametrize("fused", [False, True])
def test_reference_slopes_and_interlaced_order(kind, fused):
    import sutra
    cfg = configs.get("noisy")
    cfg["p_wfs"].noise = -1.0
    cfg["p_centroider"].type = kind
    cfg["p_centroider"].thresh = 5.0
    s = system.ShSystem(cfg)
    phase = system.make_screens(cfg, 1)[0]
    s.set_phase(phase)
    w, c = s.p_wfs, s.centro
    nv = w._nvalid
    rng = np.random.default_rng(4)
    ref = rng.uniform(-0.3, 0.3, 2 * nv).astype(np.float32)
    c.set_centroids_ref(ref)
    c.slope_order = 1
    assert c.slope_order == 1
    out = s.rtc.d_control[0].d_centroids
    if fused:
        sutra.comp_frame(s.wfs, c, out, False)
    else:
        s.comp_image(noise=False)
        s.do_centroids()
    got = np.array(out)
    img = _oracle_image(s, phase)
    weights = np.array(c.d_weights) if kind == "wcog" else None
    exp = _oracle_slopes(kind, img, w, c, 5.0, ref.astype(np.float64), 1, weights)
    err = _rel(got, exp)
    print("%s fused=%s interlaced + ref: err %.2e" % (kind, fused, err), flush=True)
    assert err < TOL
    # the x slope of sub-aperture 0 sits at 0, its y slope at 1 (interlaced)
    untied = _oracle_slopes(kind, img, w, c, 5.0, None, 0, weights)
    assert got[1] == pytest.approx(untied[nv] - ref[1], rel=1e-4, abs=1e-5)


def test_tt_filter_matches_oracle():
    cfg = configs.get("small")
    cfg["p_centroider"].filter_TT = True
    s = system.ShSystem(cfg)
    phase = system.make_screens(cfg, 1)[0]
    s.set_phase(phase)
    w, c = s.p_wfs, s.centro
    nv = w._nvalid
    assert c.filter_TT
    # non-orthogonal references (a tip and tilt with a shared component)
    tip = np.concatenate([np.ones(nv), 0.3 * np.ones(nv)]).astype(np.float32) / np.sqrt(nv)
    tilt = np.concatenate([0.2 * np.ones(nv), np.ones(nv)]).astype(np.float32) / np.sqrt(nv)
    c.d_ref_Tip.host2device(tip)
    c.d_ref_Tilt.host2device(tilt)
    s.comp_image(noise=False)
    s.do_centroids()
    got = s.slopes()
    img = _oracle_image(s, phase)
    raw, _ = oc.cog(img, w._validsubsx, w._validsubsy, w.npix, c.offset, c.scale)
    exp, (a, b) = oc.tt_filter(raw, tip.astype(np.float64), tilt.astype(np.float64))
    tt = np.array(c.d_TT_slopes).ravel()
    err = np.abs(got - exp).max() / np.abs(raw).max()
    print("tt filter: err %.2e, tip %.4f (%.4f) tilt %.4f (%.4f)" % (err, tt[0], a, tt[1], b),
          flush=True)
    assert err < TOL
    assert tt[0] == pytest.approx(a, rel=1e-4) and tt[1] == pytest.approx(b, rel=1e-4)
