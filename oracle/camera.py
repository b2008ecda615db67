"""Oracle for the simulated camera image and its calibration (test
infrastructure only).

* digitalize  libsutra/src.cu/sutra_wfs.cu:2243-2255 (digitalize_krnl), called
              after fill_binimage when the WFS has fakecam
              (libsutra/src.cpp/sutra_wfs_sh.cpp:635-640):
                  t   = (img * flat + dark) / max_flux_per_pix
                  t   = min(t, 1)
                  cam = uint16(t * max_pix_value)
              The product and sum are contracted into one fused multiply-add
              (the reference's compiler default), the float-to-integer
              conversion goes toward zero and saturates to [0, 65535]
              (NaN -> 0).  Every float32 operation is evaluated in float64 and
              rounded once to float32: exact for the product (24 x 24 bits)
              and correctly rounded for the division and the final product; the
              fused sum is exact in float64 for operands of similar magnitude.
* calibrate   libsutra/src.cu/sutra_centroider.cu:107-114 (calib_krnl, float
              and uint16 inputs):  cal[i] = (float(raw[lut[i]]) - dark[i]) * flat[i]
"""
import numpy as np


def _f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32)


def digitalize(img, dark, flat, max_flux_per_pix, max_pix_value):
    """uint16 camera image of the float detector image img (same shape)."""
    img = _f32(img).astype(np.float64)
    dark = _f32(dark).astype(np.float64)
    flat = _f32(flat).astype(np.float64)
    s = _f32(img * flat + dark).astype(np.float64)              # fused multiply-add
    t = _f32(s / np.float64(np.float32(max_flux_per_pix))).astype(np.float64)
    t = np.minimum(t, 1.0)
    v = _f32(t * np.float64(np.float32(max_pix_value))).astype(np.float64)
    v = np.trunc(v)
    v = np.where(np.isnan(v), 0.0, v)
    return np.clip(v, 0.0, 65535.0).astype(np.uint16)


def calibrate(raw, dark, flat, lut):
    """Calibrated float image: (float(raw[lut]) - dark) * flat (flat arrays)."""
    raw = np.asarray(raw).ravel(order="F")
    lut = np.asarray(lut, dtype=np.int64).ravel(order="F")
    dark = _f32(dark).ravel(order="F").astype(np.float64)
    flat = _f32(flat).ravel(order="F").astype(np.float64)
    x = _f32(raw[lut].astype(np.float64))
    d = _f32(x.astype(np.float64) - dark)
    return _f32(d.astype(np.float64) * flat)
