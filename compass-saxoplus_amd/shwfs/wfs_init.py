"""Telescope and Sensors construction, table upload.

Follows the SH branch of shesha/shesha/init/geom_init.py:30-107 (tel_init)
and shesha/shesha/init/wfs_init.py:29-233 (wfs_init): the Sensors object is
built from per-WFS arrays, initgs attaches a guide-star source (noise seed
1234 + i), load_arrays uploads the geometry tables in sub-aperture order.
"""
import numpy as np

from . import geom, lgs_init, native


def tel_init(context, p_geom, p_tel, r0, p_wfss, ittime=1.0):
    """Geometry of every WFS (largest nxsub first), then the Telescope object."""
    _, sutra = native()
    order = np.argsort([w.nxsub for w in p_wfss])[::-1]
    for i in order:
        geom.init_wfs_geom(p_wfss[i], r0, p_tel, p_geom, ittime=ittime)
    return sutra.Telescope(context, p_geom._spupil.shape[0], int(np.where(p_geom._spupil > 0)[0].size),
                           p_geom._spupil.astype(np.float32), p_geom._mpupil.shape[0], p_geom._mpupil)


def wfs_init(context, telescope, p_wfss, p_tel, p_geom, seeds=None, roket=False):
    _, sutra = native()
    n = len(p_wfss)
    i64 = lambda f: np.array([f(w) for w in p_wfss], dtype=np.int64)  # noqa: E731
    f32 = lambda f: np.array([f(w) for w in p_wfss], dtype=np.float32)  # noqa: E731
    sensors = sutra.Sensors(
        context, telescope, ["sh"] * n, n,
        i64(lambda w: w.nxsub), i64(lambda w: w._nvalid), i64(lambda w: w.nPupils),
        i64(lambda w: w.npix), i64(lambda w: w._pdiam), i64(lambda w: w._nrebin),
        i64(lambda w: w._Nfft), i64(lambda w: w._Ntot), np.full(n, p_geom._n, dtype=np.int64),
        f32(lambda w: w._subapd), f32(lambda w: w._nphotons), f32(lambda w: w.nphotons4imat),
        np.array([w.gsalt > 0 for w in p_wfss], dtype=np.int32),
        np.array([w.fakecam for w in p_wfss], dtype=bool),
        np.array([w.max_flux_per_pix for w in p_wfss], dtype=np.int32),
        np.array([w.max_pix_value for w in p_wfss], dtype=np.int32),
        context.active_device, bool(roket))
    if seeds is None:
        seeds = np.arange(n, dtype=np.int64) + 1234
    sensors.initgs(f32(lambda w: w.xpos), f32(lambda w: w.ypos), f32(lambda w: w.Lambda),
                   f32(lambda w: w.gsmag), float(p_wfss[0].zerop),
                   np.full(n, p_geom._n, dtype=np.int64), f32(lambda w: w.noise),
                   np.asarray(seeds, dtype=np.int64), f32(lambda w: w.G),
                   f32(lambda w: w.thetaML), f32(lambda w: w.dx), f32(lambda w: w.dy))
    for i, w in enumerate(p_wfss):
        sensors.d_wfs[i].load_arrays(w._phasemap, w._hrmap, w._binmap, w._halfxy,
                                     geom.flux_per_valid_subap(w), w._validsubsx, w._validsubsy,
                                     w._validpuppixx, w._validpuppixy, w._ttprojmat, w._ftkernel)
        if w._submask is not None:  # shesha/shesha/init/wfs_init.py:234-235
            sensors.set_field_stop(i, w._submask, w._submask.shape[0])
        if w.gsalt > 0:  # shesha/shesha/init/wfs_init.py:237-242
            lgs_init.prep_lgs_prof(w, i, p_tel, sensors)
    return sensors
