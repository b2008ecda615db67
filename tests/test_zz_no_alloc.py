"""shwfs_sh_comp_image_noalloc (no device allocation: the DFT operand table is
rebuilt per call into the library's static device scratch array) against
shwfs_sh_comp_image, both called through the C ABI on the same device
tables: two calls in a row, the 80x80 frame (4964 sub-apertures, Nfft 32)
then the smaller 8x8 frame (52 sub-apertures, Nfft 16, another operand
table), for every method, with the thresholded CoG fused.  The bar is bit
equality: the two entries run the same kernels on the same table values."""
import ctypes as C

import numpy as np
import pytest

import capi
from shwfs import configs, system

pytestmark = pytest.mark.gpu


def _dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))).cuda()


class _Frame:
    """Device tables, inputs and outputs of one WFS frame, as plain pointers."""

    def __init__(self, lib, name, method):
        import torch
        cfg = system.geometry(configs.get(name))
        w, g = cfg["p_wfs"], cfg["p_geom"]
        phase = system.make_screens(cfg, 1)[0]
        from shwfs import geom
        self.keep = dict(
            istart=_dev(w._validpuppixx, np.int32), jstart=_dev(w._validpuppixy, np.int32),
            vx=_dev(w._validsubsx, np.int32), vy=_dev(w._validsubsy, np.int32),
            offsets=_dev(w._halfxy, np.float32),
            flux=_dev(geom.flux_per_valid_subap(w), np.float32),
            binmap=_dev(w._binmap, np.int32), phase=_dev(phase, np.float32),
            pupil=_dev(g._mpupil, np.float32))
        self.nvalid, self.npix, self.nsub = w._nvalid, w.npix, w.npix * w.nxsub
        origin = C.c_int32(0)
        bm = np.ascontiguousarray(np.asarray(w._binmap, np.int32).ravel(order="F"))
        blk = lib.shwfs_binmap_block_origin(bm.ctypes.data_as(C.c_void_p), w._Nfft, w.npix,
                                            w._nrebin, C.byref(origin))
        assert blk >= 0
        k = self.keep
        self.geom = capi.ShGeom(
            w._nvalid, w._pdiam, g._n, w._Nfft, w._Nfft, w.npix, w._nrebin, w.nxsub,
            capi.ptr(k["istart"]), capi.ptr(k["jstart"]), capi.ptr(k["vx"]), capi.ptr(k["vy"]),
            capi.ptr(k["offsets"]), capi.ptr(k["flux"]), capi.ptr(k["binmap"]), int(blk),
            origin.value)
        self.scale = float(2 * np.pi / w.Lambda)
        self.nphot = float(w._nphotons)
        self.method = method
        self.outputs = [self._alloc(torch) for _ in range(2)]

    def _alloc(self, torch):
        n = self.nvalid
        return dict(bincube=torch.zeros(self.npix * self.npix * n, device="cuda"),
                    binimg=torch.zeros(self.nsub * self.nsub, device="cuda"),
                    inten=torch.zeros(n, device="cuda"),
                    slopes=torch.zeros(2 * n, device="cuda"),
                    cinten=torch.zeros(n, device="cuda"))

    def args(self, out):
        k = self.keep
        cen = capi.CentroidArgs(
            capi.CENTROID_TCOG, self.nvalid, self.npix, self.nsub, capi.ptr(k["vx"]),
            capi.ptr(k["vy"]), self.npix / 2 - 0.5, 0.5, 0.3, None, None,
            capi.ptr(out["slopes"]), capi.ptr(out["cinten"]), 0, 1)
        self._cen = cen
        return capi.FrameArgs(
            capi.ptr(k["phase"]), capi.ptr(k["pupil"]), self.scale, self.nphot, -1.0, 0.0, 0,
            0, self.method, 0, capi.ptr(out["bincube"]), capi.ptr(out["binimg"]),
            capi.ptr(out["inten"]), C.pointer(cen), None)


@pytest.mark.parametrize("method", [capi.METHOD_DFT_FP16, capi.METHOD_DFT_BF16X3,
                                    capi.METHOD_FFT_FP32])
def test_noalloc_entry_equals_comp_image(method):
    import torch
    lib = capi.load()
    for name in ("large", "small"):  # the second call: a smaller frame, another table
        f = _Frame(lib, name, method)
        print("noalloc: %s frame, method %d, nvalid %d" % (name, method, f.nvalid), flush=True)
        a_ref, a_new = f.args(f.outputs[0]), f.args(f.outputs[1])
        rc = lib.shwfs_sh_comp_image(C.byref(f.geom), C.byref(a_ref), None)
        assert rc == 0, capi.last_error(lib)
        rc = lib.shwfs_sh_comp_image_noalloc(C.byref(f.geom), C.byref(a_new), None)
        assert rc == 0, capi.last_error(lib)
        torch.cuda.synchronize()
        for key in ("bincube", "binimg", "inten", "slopes", "cinten"):
            ref, got = f.outputs[0][key].cpu().numpy(), f.outputs[1][key].cpu().numpy()
            assert np.abs(ref).max() > 0, key
            assert np.array_equal(ref, got), "%s differs (%s frame)" % (key, name)
