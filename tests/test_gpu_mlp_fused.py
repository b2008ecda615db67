"""The learned centroider in the spot launch (fused after binning) and its
stand-alone forms, through the C++ layer and the C ABI:

* one fused frame (spots + network + CoG fallback in one launch) gives the
  same slopes, intensities and fallback flags, bit for bit, as the same
  frame's spots followed by shwfs_mlp_centroids over the spot cube, and as
  get_cog over the detector image windows;
* the image-window path without fallback equals the network's forward
  entry on the same spots times the pixel scale, bit for bit.

Workload: 40x40 SH (pdiam 16, Nfft 32, 8x8-pixel spots), 50 photons per
sub-aperture, Poisson + 3 e- readout noise; FFT, split-BF16x3 and FP16 spots."""
import ctypes as C

import numpy as np
import pytest

import capi
from oracle import philox as op
from shwfs import configs, learned, system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    s = system.ShSystem(configs.get("lowflux"))
    c = learned.LearnedCentroider(s, seed=3)
    c.mlp.set_input_scale(1.0 / 50.0)
    s.set_phase(system.make_screens(s.cfg, 1)[0])
    return c


def _obj(lc, a, dtype=np.float32):
    return lc.carma.obj_float(lc.sys.ctx, np.asfortranarray(np.asarray(a, dtype=dtype)))


def _standalone_on_cube(lc, cube_ptr, nvalid, lo, hi):
    """shwfs_mlp_centroids over a spot cube (no valid positions)."""
    lib = capi.load()
    lib.shwfs_mlp_work_bytes.restype = C.c_size_t
    f32, i32, vp = C.c_float, C.c_int32, C.c_void_p
    lib.shwfs_mlp_centroids.argtypes = [vp, vp, vp, i32, vp, vp, i32, i32, f32, f32, f32, f32,
                                        f32, f32, vp, i32, vp, vp, vp, vp]
    work = _obj(lc, np.zeros((lib.shwfs_mlp_work_bytes() + 3) // 4))
    slopes = _obj(lc, np.zeros(2 * nvalid))
    inten = _obj(lc, np.zeros(nvalid))
    fb = lc.carma.obj_int32_t(lc.sys.ctx, np.zeros(nvalid, dtype=np.int32))
    m, pw = lc.mlp, lc.sys.p_wfs
    rc = lib.shwfs_mlp_centroids(m.d_params.data_ptr, work.data_ptr, cube_ptr, 0, None, None,
                                 nvalid, 8, m.x_scale, lo, hi, pw.npix / 2.0 - 0.5, pw.pixsize,
                                 pw.pixsize, None, 0, slopes.data_ptr, inten.data_ptr,
                                 fb.data_ptr, None)
    assert rc == 0, capi.last_error(lib)
    return np.array(slopes).ravel(), np.array(inten).ravel(), np.array(fb).ravel()


@pytest.mark.parametrize("method", [0, 1, 2])
def test_fused_frame_equals_spots_then_network(lc, method):
    """FFT spots (method 0) or MFMA DFT spots (1: split-BF16x3, 2: FP16; the
    eight-sub-aperture kernel of shwfs_dft_mlp.hip) with the network after
    them, in one launch."""
    s, m = lc.sys, lc.mlp
    s.wfs.set_method(method)
    try:
        _fused_frame_checks(lc, s, m)
    finally:
        s.wfs.set_method(0)


def _fused_frame_checks(lc, s, m):
    nv = s.wfs.nvalid
    lo, hi = 20.0, 80.0  # some spots outside: their slopes are the CoG of the same pixels
    m.set_flux_range(lo, hi)
    slopes = _obj(lc, np.zeros(2 * nv))
    seed = 777
    s.wfs.set_noise(3.0, seed)                    # call counter 0
    lc.sutra.comp_frame(s.wfs, m, slopes, True)   # spots + network in one launch
    got = np.array(slopes).ravel()
    # the readout background drawn by the same launch: 3 N(0, 1) (domain 2) on
    # the pixels outside the spot windows, the spots inside them
    p = s.p_wfs
    img = s.binimg()
    mask = np.zeros(img.shape, bool)
    for x, y in zip(p._validsubsx, p._validsubsy):
        mask[x:x + p.npix, y:y + p.npix] = True
    idx = np.nonzero(~mask.ravel(order="F"))[0]
    np.testing.assert_allclose(img.ravel(order="F")[idx], 3.0 * op.normal(idx, 0, 2, seed),
                               rtol=1e-5, atol=1e-6)
    cube = s.bincube().reshape(p.npix, p.npix, -1, order="F")
    for k, (x, y) in enumerate(zip(p._validsubsx, p._validsubsy)):
        np.testing.assert_array_equal(img[x:x + p.npix, y:y + p.npix], cube[:, :, k])
    got_i = np.array(m.d_intensities).ravel()
    got_f = np.array(m.d_fallback).ravel()
    print("fused frame: %d of %d sub-apertures fell back to CoG" % (got_f.sum(), nv), flush=True)
    assert 0 < got_f.sum() < nv
    # the same frame's spot cube through the stand-alone kernel
    ref, ref_i, ref_f = _standalone_on_cube(lc, s.wfs.d_bincube.data_ptr, nv, lo, hi)
    assert np.array_equal(got_f, ref_f)
    assert np.array_equal(got_i, ref_i)
    assert np.array_equal(got, ref)
    # and through get_cog over the detector image windows
    m.get_cog()
    img = np.array(s.wfs.d_slopes).ravel()
    assert np.array_equal(img, ref)


def test_image_windows_equal_forward(lc):
    s, m = lc.sys, lc.mlp
    nv = s.wfs.nvalid
    m.set_flux_range(-np.inf, np.inf)   # no fallback
    s.comp_image(noise=True)
    m.get_cog()                          # network on the image windows
    got = np.array(s.wfs.d_slopes).ravel()
    assert not np.any(np.array(m.d_fallback))
    cube = np.array(s.wfs.d_bincube).reshape(64, nv, order="F")
    X = _obj(lc, cube)
    Y = _obj(lc, np.zeros((2, nv)))
    m.forward(X, nv, Y)
    y = np.array(Y).reshape(2, nv, order="F")
    ref = np.concatenate([y[0] * np.float32(s.p_wfs.pixsize), y[1] * np.float32(s.p_wfs.pixsize)])
    assert np.array_equal(got, ref.astype(np.float32))
