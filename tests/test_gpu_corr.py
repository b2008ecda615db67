"""Correlation centroider (shwfs_centroids_corr, CentroiderCORR_FF) against
the oracle (oracle/corr.py), which restates SutraCentroiderCorr::get_cog
(libsutra/src.cpp/sutra_centroider_corr.cpp:190-262).  Seeded spots with
sub-pixel offsets, shared and per-sub-aperture reference functions, a
non-uniform corrnorm, both window branches (2 npix - 1 <= 20 and > 20),
3x3 and 5x5 fit patches, both slope orders; a whole-pixel shift measured
exactly; and the RTC path the reference's rtc_init sets up
(shesha/shesha/init/rtc_init.py:349-377) on a WFS frame."""
import numpy as np
import pytest

import capi
from oracle import corr as oc
from shwfs import configs, system

pytestmark = pytest.mark.gpu


def _spots(rng, npix, nxsub, positions, width=1.4):
    size = npix * nxsub
    img = 0.05 * rng.random((size, size))
    x = np.arange(npix)[:, None]
    y = np.arange(npix)[None, :]
    for vx, vy in positions:
        cx, cy = npix / 2 + rng.uniform(-1.8, 1.8), npix / 2 + rng.uniform(-1.8, 1.8)
        img[vx:vx + npix, vy:vy + npix] += 3.0 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) /
                                                         (2 * width ** 2))
    return img  # indexed [x, y]


def _run(lib, img, vx, vy, npix, ref_fn, corrnorm, sx, sy, offset, scale, ref, order):
    import torch
    dev = torch.device("cuda:0")
    nv = len(vx)
    N = 2 * npix - 1
    t = lambda a, dt=torch.float32: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    d_img = t(img.T.ravel())  # x fastest
    per = ref_fn.ndim == 3
    d_rf = t(np.concatenate([ref_fn[:, :, s].T.ravel() for s in range(nv)]) if per
             else ref_fn.T.ravel())
    d_norm = None if corrnorm is None else t(corrnorm.T.ravel())
    d_mat = t(oc.interp_mat(sx, sy).ravel())  # (6, np) C order = (np, 6) column-major
    d_vx, d_vy = t(vx, torch.int32), t(vy, torch.int32)
    if ref is not None and order == 1:  # the reference slopes follow the slope order
        ref = np.stack([ref[:nv], ref[nv:]], 1).ravel()
    d_ref = None if ref is None else t(ref)
    d_sl = torch.zeros(2 * nv, dtype=torch.float32, device=dev)
    d_corr = torch.zeros(N * N * nv, dtype=torch.float32, device=dev)
    d_max = torch.full((nv,), -1, dtype=torch.int32, device=dev)
    rc = lib.shwfs_centroids_corr(capi.ptr(d_img), img.shape[0], capi.ptr(d_vx), capi.ptr(d_vy),
                                  nv, npix, capi.ptr(d_rf), int(per), capi.ptr(d_norm),
                                  capi.ptr(d_mat), sx, sy, offset, scale, capi.ptr(d_ref), order,
                                  capi.ptr(d_sl), capi.ptr(d_corr), capi.ptr(d_max), None)
    assert rc == 0, capi.last_error(lib)
    torch.cuda.synchronize()
    sl = d_sl.cpu().numpy().astype(np.float64)
    if order == 1:  # interlaced x0 y0 x1 y1 ...
        sl = np.concatenate([sl[0::2], sl[1::2]])
    return sl, d_corr.cpu().numpy().reshape(nv, N, N), d_max.cpu().numpy()


@pytest.mark.parametrize("npix,sx,per_subap,order", [
    (8, 3, False, 0), (6, 3, True, 1), (12, 3, False, 0), (8, 5, True, 0), (16, 3, True, 1)])
def test_corr_matches_oracle(npix, sx, per_subap, order):
    lib = capi.load()
    rng = np.random.default_rng(100 + npix + sx)
    nxsub = 6
    pos = [(i * npix, j * npix) for i in range(nxsub) for j in range(nxsub) if (i + j) % 3 != 1]
    vx = np.array([p[0] for p in pos], np.int32)
    vy = np.array([p[1] for p in pos], np.int32)
    nv = len(pos)
    img = _spots(rng, npix, nxsub, pos)
    x = np.arange(npix)[:, None]
    y = np.arange(npix)[None, :]
    g = np.exp(-((x - npix // 2) ** 2 + (y - npix // 2) ** 2) / (2 * 1.6 ** 2))
    if per_subap:
        ref_fn = np.stack([g * (1 + 0.2 * rng.random((npix, npix))) for _ in range(nv)], 2)
    else:
        ref_fn = g
    N = 2 * npix - 1
    # smooth and asymmetric: a transposed or shifted norm changes every map
    kx = np.arange(N)[:, None] / N
    ky = np.arange(N)[None, :] / N
    corrnorm = 1.0 + 0.3 * kx + 0.15 * ky ** 2
    ref = 0.01 * rng.standard_normal(2 * nv)
    offset, scale = npix / 2 - 0.5, 0.3
    print("corr npix=%d patch=%dx%d per_subap=%s order=%d nvalid=%d" %
          (npix, sx, sx, per_subap, order, nv), flush=True)
    sl, cmap, cmax = _run(lib, img, vx, vy, npix, ref_fn.astype(np.float32),
                          corrnorm.astype(np.float32), sx, sx, offset, scale,
                          ref.astype(np.float32), order)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    want = oc.centroids(f32(img), vx, vy, npix, f32(ref_fn), offset, scale, f32(corrnorm), sx, sx,
                        f32(ref))
    # the maps: each sub-aperture's correlation over its norm, [x, y] per map
    nbmax = npix if N > 20 else N
    off = N // 2 - nbmax // 2
    checked = 0
    refv = f32(ref)
    for s in range(nv):
        r = ref_fn[:, :, s] if per_subap else ref_fn
        m = oc.corr_map(f32(img)[vx[s]:vx[s] + npix, vy[s]:vy[s] + npix], f32(r)) / f32(corrnorm)
        got = cmap[s].T  # device map is x fastest
        assert np.abs(got - m).max() <= 1e-5 * np.abs(m).max()
        # the fit stage alone, in float64 from the device's own map
        fx, fy, imax = oc.fit_map(got.astype(np.float64), npix, offset, scale, sx, sx)
        assert cmax[s] == imax
        assert abs(sl[s] - (fx - refv[s])) <= 2e-5 * scale
        assert abs(sl[s + nv] - (fy - refv[s + nv])) <= 2e-5 * scale
        w = np.sort(m[off:off + nbmax, off:off + nbmax].ravel())
        if w[-1] - w[-2] > 1e-4 * abs(w[-1]):  # a clear maximum: its index is fixed
            k = int(np.argmax(m[off:off + nbmax, off:off + nbmax].T.ravel()))
            assert cmax[s] == (off + k % nbmax) + (off + k // nbmax) * N
            checked += 1
    assert checked >= nv // 2
    err = np.abs(sl - want).max()
    print("  slope max err %.2e (scale %.2f)" % (err, scale), flush=True)
    assert err <= 2e-4 * scale


def test_corr_whole_pixel_shift_is_exact():
    lib = capi.load()
    npix = 8
    pat = np.random.default_rng(5).random((3, 3)) + 0.5
    r = np.zeros((npix, npix))
    r[2:5, 2:5] = pat
    shifts = [(0, 0), (1, -2), (-2, 1), (3, 3)]
    img = np.zeros((npix * 4, npix * 4))
    pos = []
    for i, (dx, dy) in enumerate(shifts):
        vx, vy = npix * i, 0
        img[vx + 2 + dx:vx + 5 + dx, vy + 2 + dy:vy + 5 + dy] = pat
        pos.append((vx, vy))
    vx = np.array([p[0] for p in pos], np.int32)
    vy = np.array([p[1] for p in pos], np.int32)
    print("corr whole-pixel shifts %s" % (shifts,), flush=True)
    sl, _, _ = _run(lib, img, vx, vy, npix, r.astype(np.float32), None, 3, 3, npix / 2 - 0.5,
                    0.25, None, 0)
    want = np.array([0.25 * d[0] for d in shifts] + [0.25 * d[1] for d in shifts])
    assert np.abs(sl - want).max() < 1e-5


def test_corr_centroider_through_the_rtc():
    """The reference's rtc_init set-up for a correlation centroider (Gaussian
    reference function, ones for corrnorm, 3x3 fit) on a WFS frame, then
    do_centroids: the slopes equal the oracle's on the same detector image."""
    c = configs.get("small")
    c["nphot"] = 1e4
    c["p_centroider"].type = "corr"
    c["p_centroider"].width = 2.0
    s = system.ShSystem(c)
    centro = s.centro
    assert type(centro).__name__ == "CentroiderCORR_FF"
    assert centro.interp_sizex == 3 and centro.interp_sizey == 3
    assert centro.d_corrspot is None
    s.set_phase(system.make_screens(c, 1)[0])
    s.comp_image(noise=False)
    img = s.binimg()
    p = c["p_wfs"]
    vx = np.array(centro.d_validx).astype(np.int32)
    vy = np.array(centro.d_validy).astype(np.int32)
    s.rtc.do_centroids(0)
    slopes = np.array(s.rtc.d_control[0].d_centroids).astype(np.float64)
    w = np.array(centro.d_corrfnct)[:, :, 0].astype(np.float64)  # shared: the first npix^2
    assert np.allclose(w, c["p_centroider"].weights)
    want = oc.centroids(img.astype(np.float64), vx, vy, p.npix, w, p.npix // 2 - 0.5, p.pixsize)
    err = np.abs(slopes - want).max()
    print("corr rtc nvalid=%d npix=%d slope max err %.2e arcsec" % (len(vx), p.npix, err),
          flush=True)
    assert err <= 2e-4 * p.pixsize
    cm = np.array(centro.d_corrmax)
    assert cm.shape == (len(vx),) and cm.min() >= 0 and cm.max() < (2 * p.npix - 1) ** 2

