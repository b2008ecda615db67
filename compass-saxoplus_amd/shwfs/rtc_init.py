"""Centroider construction for SH WFS (shesha/shesha/init/rtc_init.py:272-433)."""
import numpy as np

from . import geom, native

RAD2ARCSEC = geom.RAD2ARCSEC


def comp_weights(p_centroider, p_wfs, npix):
    """Gaussian WCoG / correlation weights (the GAUSS branch of comp_weights,
    shesha/shesha/init/rtc_init.py:379-434): centred on the middle pixel, or
    between the two middle pixels of an even window for WCoG."""
    if p_centroider.width is None:
        p_centroider.width = npix
    if p_wfs.npix % 2 == 1 or p_centroider.type == "corr":
        c = p_wfs.npix // 2
    else:
        c = p_wfs.npix // 2 - 0.5
    p_centroider.weights = geom.makegaussian(p_wfs.npix, p_centroider.width, c, c).astype(np.float32)


def create_interp_mat(dimx, dimy):
    """Least-squares matrix [dimx*dimy, 6] of the parabola a x^2 + b y^2 +
    c xy + d x + e y + f over a dimx x dimy patch centred on its middle
    sample (shesha/shesha/util/rtc_util.py:23-45): coefficients = M.T @ patch."""
    n = max(dimx, dimy)
    rows, cols = np.indices((n, n))
    x = (cols[:dimx, :dimy] - dimx // 2).ravel()
    y = (rows[:dimx, :dimy] - dimy // 2).ravel()
    A = np.stack([x * x, y * y, x * y, x, y, np.ones_like(x)], 1).astype(np.float64)
    return (np.linalg.inv(A.T @ A) @ A.T).T


def init_centroider(context, nwfs, p_wfs, p_centroider, p_tel, r0, sensors, rtc):
    """Adds a cog / tcog / wcog / bpcog / corr centroider fed by
    sensors.d_wfs[nwfs] (shesha/shesha/init/rtc_init.py:272-377)."""
    s_offset = p_wfs.npix // 2.0 - 0.5
    s_scale = p_wfs.pixsize
    rtc.add_centroider(context, p_wfs._nvalid, s_offset, s_scale, p_centroider.filter_TT,
                       context.active_device, p_centroider.type, sensors.d_wfs[nwfs])
    c = rtc.d_centro[-1]
    c.load_validpos(p_wfs._validsubsx, p_wfs._validsubsy, p_wfs._nvalid * p_wfs.nPupils)
    c.set_npix(p_wfs.npix)
    p_centroider._nslope = 2 * p_wfs._nvalid
    if p_centroider.type in ("tcog", "wcog"):
        c.set_threshold(p_centroider.thresh)
    if p_centroider.type == "bpcog":
        c.set_nmax(p_centroider.nmax)
    if p_centroider.type in ("wcog", "corr"):
        r0w = r0 * (p_wfs.Lambda / 0.5) ** (6 / 5.0)
        seeing = RAD2ARCSEC * (p_wfs.Lambda * 1.0e-6) / r0w
        npix = seeing // p_wfs.pixsize
        if p_centroider.weights is None:
            comp_weights(p_centroider, p_wfs, npix)
        if p_centroider.type == "wcog":
            c.init_weights()
            c.load_weights(p_centroider.weights, p_centroider.weights.ndim)
        else:
            corrnorm = np.ones((2 * p_wfs.npix, 2 * p_wfs.npix), dtype=np.float32)
            p_centroider.sizex = 3
            p_centroider.sizey = 3
            p_centroider.interpmat = create_interp_mat(p_centroider.sizex,
                                                       p_centroider.sizey).astype(np.float32)
            c.init_corr(p_centroider.sizex, p_centroider.sizey, p_centroider.interpmat)
            # the first (2 npix - 1)^2 values of the ones array normalise the map
            c.load_corr(p_centroider.weights, corrnorm, p_centroider.weights.ndim)
    return c


def rtc_init(context, sensors, p_wfss, p_centroiders, p_tel, r0):
    """Rtc_FFF with one centroider per WFS and one slope-holding controller."""
    _, sutra = native()
    rtc = sutra.Rtc_FFF()
    for i, pc in enumerate(p_centroiders):
        init_centroider(context, pc.nwfs, p_wfss[pc.nwfs], pc, p_tel, r0, sensors, rtc)
    nslope = sum(2 * p_wfss[pc.nwfs]._nvalid for pc in p_centroiders)
    rtc.add_controller(context, "generic", context.active_device, 0.0, nslope, 1,
                       idx_centro=np.arange(len(p_centroiders), dtype=np.int32),
                       ncentro=len(p_centroiders))
    return rtc
