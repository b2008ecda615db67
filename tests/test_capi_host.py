"""C ABI checks that need no GPU: the library loads, exports every declared
symbol, validates its arguments on the host and reports errors."""
import ctypes as C

import numpy as np
import pytest

import capi
from oracle import geometry as og


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def test_exports_every_declared_symbol(lib):
    names = capi.declared_functions()
    assert {"shwfs_sh_comp_image", "shwfs_centroids", "shwfs_prng", "shwfs_last_error"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n


def test_version_and_error_reporting(lib):
    assert b"gfx950" in lib.shwfs_version()
    g = capi.ShGeom()
    g.nvalid = 4
    g.nphase, g.npup, g.nfft, g.ntot, g.npix, g.nrebin, g.nxsub = 8, 68, 12, 12, 6, 2, 8
    assert lib.shwfs_sh_check_geom(C.byref(g), 0) == capi.SHWFS_ERR_UNSUPPORTED
    assert "nfft" in capi.last_error(lib)
    g.nfft = g.ntot = 16
    assert lib.shwfs_sh_check_geom(C.byref(g), 0) == capi.SHWFS_ERR_STATE   # tables not loaded
    g.ntot = 32   # enlarged HR grid: needs its hrmap
    assert lib.shwfs_sh_check_geom(C.byref(g), 0) == capi.SHWFS_ERR_STATE
    assert "hrmap" in capi.last_error(lib)
    g.ntot = 128  # beyond the enlarged-grid kernel
    assert lib.shwfs_sh_check_geom(C.byref(g), 0) == capi.SHWFS_ERR_UNSUPPORTED
    assert lib.shwfs_sh_check_geom(None, 0) == capi.SHWFS_ERR_ARG
    # the background pixel list: its length within [0, (npix * nxsub)^2]
    # (the pointer is not dereferenced by the check)
    g.ntot = 16
    g.d_bgpix = 256
    for n, rc in ((-1, capi.SHWFS_ERR_ARG), ((6 * 8) ** 2 + 1, capi.SHWFS_ERR_ARG),
                  (0, capi.SHWFS_ERR_STATE), ((6 * 8) ** 2, capi.SHWFS_ERR_STATE)):
        g.nbgpix = n
        assert lib.shwfs_sh_check_geom(C.byref(g), 0) == rc, n
        if rc == capi.SHWFS_ERR_ARG:
            assert "background" in capi.last_error(lib)


def test_null_and_size_arguments_rejected_without_launch(lib):
    assert lib.shwfs_sh_comp_image(None, None, None) == capi.SHWFS_ERR_ARG
    a = capi.CentroidArgs()
    assert lib.shwfs_centroids(None, None, C.byref(a), None) == capi.SHWFS_ERR_ARG
    a.nvalid, a.npix, a.size = 4, 8, 64
    assert lib.shwfs_centroids(C.c_void_p(16), None, C.byref(a), None) == capi.SHWFS_ERR_STATE
    assert lib.shwfs_prng(None, 10, b"P", 1.0, 0.0, 1, 0, 0, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_prng(C.c_void_p(16), 10, b"X", 1.0, 0.0, 1, 0, 0, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_fill_weights(None, None, 8, 64, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_calibrate(None, None, None, None, None, 4, None) == capi.SHWFS_ERR_ARG
    dev = C.c_void_p(16)  # never dereferenced: every call below fails its checks first
    # unknown type; type 1 without pupil / flux; type 2 without ttprojmat
    assert lib.shwfs_slopes_geom(3, dev, dev, dev, dev, dev, dev, 8, 4, 1.0, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_slopes_geom(1, dev, None, None, dev, dev, None, 8, 4, 1.0, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_slopes_geom(2, dev, None, None, dev, dev, None, 8, 4, 1.0, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_tt_filter(None, dev, dev, None, 8, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_fill_binimg(dev, None, dev, dev, 8, 4, 64, 0.0, 1, 0, 2, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_mfma_probe(None, None, None, 1, None) == capi.SHWFS_ERR_ARG


def test_upload_with_a_wrong_size_is_a_defined_error(lib):
    """set_bincube / set_binimg go through shwfs_upload: a size that differs
    from the device array's is SHWFS_ERR_ARG, checked before any copy (the
    reference only prints a message and returns success)."""
    host = np.zeros(10, dtype=np.float32)
    rc = lib.shwfs_upload(C.c_void_p(16), 12, host.ctypes.data_as(C.c_void_p), 10, 4, None)
    assert rc == capi.SHWFS_ERR_ARG
    assert "10 elements given" in capi.last_error(lib) and "holds 12" in capi.last_error(lib)
    assert lib.shwfs_upload(None, 10, host.ctypes.data_as(C.c_void_p), 10, 4, None) == capi.SHWFS_ERR_ARG


def test_camera_bpcog_and_acquisim_arguments_rejected_without_launch(lib):
    dev = C.c_void_p(16)  # never dereferenced: every call below fails its checks first
    assert lib.shwfs_calibrate_u16(None, None, None, None, None, 4, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_digitalize(None, dev, dev, dev, 100, 65535, 4, None) == capi.SHWFS_ERR_ARG
    # max_flux_per_pix must be positive (it divides)
    assert lib.shwfs_digitalize(dev, dev, dev, dev, 0, 65535, 4, None) == capi.SHWFS_ERR_ARG
    a = capi.CentroidArgs()
    a.nvalid, a.npix, a.size = 4, 8, 64
    assert lib.shwfs_centroids_bpcog(None, 3, C.byref(a), None) == capi.SHWFS_ERR_ARG
    # nbpix outside [0, npix^2)
    assert lib.shwfs_centroids_bpcog(dev, 64, C.byref(a), None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_centroids_bpcog(dev, -1, C.byref(a), None) == capi.SHWFS_ERR_ARG
    # windows above 1024 pixels
    a.npix = 33
    assert lib.shwfs_centroids_bpcog(dev, 3, C.byref(a), None) == capi.SHWFS_ERR_UNSUPPORTED
    a.npix = 8
    # valid positions and outputs must be set
    assert lib.shwfs_centroids_bpcog(dev, 3, C.byref(a), None) == capi.SHWFS_ERR_STATE
    assert lib.shwfs_fill_bincube(None, dev, dev, dev, 8, 4, 64, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_fill_bincube(dev, dev, dev, dev, 0, 4, 64, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_fill_bincube_2d(dev, dev, None, 8, 8, 4, None) == capi.SHWFS_ERR_ARG


@pytest.mark.parametrize("Ntot,npix,nrebin", [(16, 8, 2), (32, 8, 4), (16, 6, 2), (64, 16, 4), (16, 5, 3)])
def test_binmap_block_analysis_recognises_the_ngs_binmap(lib, Ntot, npix, nrebin):
    bm = np.ascontiguousarray(og.binmap_table(Ntot, npix, nrebin).T)  # [p][i] = F-order of [i, p]
    o = C.c_int32(-1)
    rc = lib.shwfs_binmap_block_origin(bm.ctypes.data_as(C.c_void_p), Ntot, npix, nrebin, C.byref(o))
    assert rc == 1
    # the origin reproduces the table: block of pixel 0 starts at (o, o)
    w = nrebin * npix
    indi = (Ntot - w) // 2 + (1 if (w % 2) != (Ntot % 2) else 0)
    assert o.value == (indi + Ntot // 2) % Ntot
    # a permuted table is not block form
    bad = bm.copy()
    bad[[0, 1]] = bad[[1, 0]]
    assert lib.shwfs_binmap_block_origin(bad.ctypes.data_as(C.c_void_p), Ntot, npix, nrebin,
                                         C.byref(o)) == 0


def test_batch_arguments_rejected_without_launch(lib):
    out = C.c_void_p(123)
    assert lib.shwfs_batch_create(None, None, 2, C.byref(out)) == capi.SHWFS_ERR_ARG
    assert out.value is None  # cleared on failure
    g, f = capi.ShGeom(), capi.FrameArgs()
    assert lib.shwfs_batch_create(C.byref(g), C.byref(f), 0, C.byref(out)) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_batch_create(C.byref(g), C.byref(f), 1, None) == capi.SHWFS_ERR_ARG
    assert lib.shwfs_batch_destroy(None) == capi.SHWFS_OK
    for rc in (lib.shwfs_batch_set_tiles(None, None, 0, 0),
               lib.shwfs_batch_get_tiles(None, None, None, None),
               lib.shwfs_batch_set_frame(None, 0, None),
               lib.shwfs_batch_get_frame(None, None),
               lib.shwfs_batch_launch(None, None),
               lib.shwfs_batch_capture(None, 4, None),
               lib.shwfs_batch_replay(None, 1, None),
               lib.shwfs_batch_autotune(None, 1, 10, None, None)):
        assert rc == capi.SHWFS_ERR_ARG
    assert "batch" in capi.last_error(lib)


def test_corr_arguments_rejected_without_launch(lib):
    dev = C.c_void_p(16)  # never dereferenced: every call below fails its checks first

    def call(img=dev, size=64, npix=8, nvalid=4, ref_fn=dev, interp=dev, sx=3, sy=3, slopes=dev):
        return lib.shwfs_centroids_corr(img, size, dev, dev, nvalid, npix, ref_fn, 0, None, interp,
                                        sx, sy, 3.5, 1.0, None, 0, slopes, None, None, None)

    assert call(img=None) == capi.SHWFS_ERR_ARG
    assert call(ref_fn=None) == capi.SHWFS_ERR_ARG
    assert call(interp=None) == capi.SHWFS_ERR_ARG
    assert call(slopes=None) == capi.SHWFS_ERR_ARG
    assert call(nvalid=0) == capi.SHWFS_ERR_ARG
    assert call(size=4) == capi.SHWFS_ERR_ARG  # the image is smaller than one spot
    assert call(npix=17, size=200) == capi.SHWFS_ERR_UNSUPPORTED
    # the parabola patch: at least 6 points, at most 49, inside the 2 npix - 1 map
    assert call(sx=2, sy=2) == capi.SHWFS_ERR_ARG
    assert call(sx=7, sy=9) == capi.SHWFS_ERR_ARG
    assert call(npix=2, size=8, sx=5, sy=3) == capi.SHWFS_ERR_ARG
    assert "centroids_corr" in capi.last_error(lib)


def _header_struct_members(name):
    """Member names of `typedef struct name { ... } name;` in include/shwfs.h."""
    import os
    import re
    text = open(os.path.join(os.path.dirname(__file__), "..", "include", "shwfs.h")).read()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(","):  # "float xoff, yoff"
            names.append(re.findall(r"[A-Za-z_][A-Za-z_0-9]*", part)[-1])
    return names


@pytest.mark.parametrize("cname,pyname", [("shwfs_sh_geom", "ShGeom"),
                                          ("shwfs_centroid_args", "CentroidArgs"),
                                          ("shwfs_layer", "Layer"),
                                          ("shwfs_frame_args", "FrameArgs")])
def test_ctypes_structs_follow_the_header(cname, pyname):
    """The ctypes mirrors list the header's members in order (a missing
    trailing member would leave the library reading past the Python struct)."""
    want = _header_struct_members(cname)
    got = [f[0] for f in getattr(capi, pyname)._fields_]
    assert got == want
