"""One SH WFS + its centroider, built the way shesha builds them.

``geometry(cfg)`` runs the host-side geometry only (no GPU needed);
``ShSystem(cfg)`` also creates the native Telescope / Sensors / Rtc_FFF.
"""
import numpy as np

from . import geom, native, rtc_init, screens, wfs_init


def geometry(cfg, ittime=1.0):
    """Fills the derived fields of cfg['p_wfs'] / cfg['p_geom'] (host only)."""
    p_wfs, p_tel, p_geom = cfg["p_wfs"], cfg["p_tel"], cfg["p_geom"]
    if not p_geom.is_init:
        geom.init_wfs_geom(p_wfs, cfg["r0"], p_tel, p_geom, ittime=ittime,
                           nfft=cfg.get("nfft"), nrebin=cfg.get("nrebin"))
        if cfg.get("nphot") is not None:
            p_wfs._nphotons = float(cfg["nphot"])
    return cfg


def make_screens(cfg, n_screens=None):
    p_tel, p_geom = cfg["p_tel"], cfg["p_geom"]
    dx = p_tel.diam / p_geom.pupdiam
    seeds = cfg["seeds"] if n_screens is None else cfg["seeds"][:n_screens]
    return screens.vonkarman_many(p_geom._n, dx, cfg["r0"], cfg["L0"], seeds)


class ShSystem:
    def __init__(self, cfg, method=0, fp64_moments=False, threshold=None):
        self.carma, self.sutra = native()
        self.cfg = geometry(cfg)
        self.ctx = self.carma.context.get_instance()
        p_wfs, p_tel, p_geom = cfg["p_wfs"], cfg["p_tel"], cfg["p_geom"]
        self.p_wfs, self.p_tel, self.p_geom = p_wfs, p_tel, p_geom
        self.tel = self.sutra.Telescope(
            self.ctx, p_geom._spupil.shape[0], int(np.count_nonzero(p_geom._spupil > 0)),
            p_geom._spupil, p_geom._mpupil.shape[0], p_geom._mpupil)
        self.sensors = wfs_init.wfs_init(self.ctx, self.tel, [p_wfs], p_tel, p_geom,
                                         roket=cfg.get("roket", False))
        self.wfs = self.sensors.d_wfs[0]
        if method:
            self.wfs.set_method(method)
        pc = cfg["p_centroider"]
        if threshold is not None:
            pc.thresh = float(threshold)
        self.rtc = rtc_init.rtc_init(self.ctx, self.sensors, [p_wfs], [pc], p_tel, cfg["r0"])
        self.centro = self.rtc.d_centro[0]
        if fp64_moments:
            self.centro.set_fp64_moments(True)

    # -- inputs
    def set_phase(self, phase_um):
        self.wfs.d_gs.set_phase(np.asarray(phase_um, dtype=np.float32))

    # -- the path
    def comp_image(self, noise=True):
        self.wfs.comp_image(noise)

    def do_centroids(self):
        self.rtc.do_centroids(0)

    # -- outputs (synchronising copies)
    def bincube(self):
        return np.array(self.wfs.d_bincube)

    def binimg(self):
        return np.array(self.wfs.d_binimg)

    def intensities(self):
        return np.array(self.wfs.d_intensities)

    def slopes(self):
        return np.array(self.rtc.d_control[0].d_centroids)

    def tables(self):
        """The tables as uploaded, in the oracle's argument order."""
        w = self.p_wfs
        return dict(pupil=self.p_geom._mpupil, scale=np.float32(self.wfs.d_gs.scale),
                    offsets=w._halfxy, istart=w._validpuppixx, jstart=w._validpuppixy,
                    validsubsx=w._validsubsx, validsubsy=w._validsubsy, binmap=w._binmap,
                    fluxPerSub=geom.flux_per_valid_subap(w), nphot=np.float32(w._nphotons),
                    pdiam=w._pdiam, Nfft=w._Nfft, npix=w.npix, nxsub=w.nxsub)


class BatchSystem:
    """Several independent SH WFS (one ShSystem each: own pupil, screen and
    centroider) computed by one launch per frame through ``sutra.WfsBatch``."""

    def __init__(self, cfgs, method=0, noise=False):
        self.systems = [ShSystem(c, method=method) for c in cfgs]
        self.carma, self.sutra = native()
        ctx = self.systems[0].ctx
        self.batch = self.sutra.WfsBatch(ctx, [s.wfs for s in self.systems],
                                         [s.centro for s in self.systems], noise)

    def set_phases(self, phases):
        for s, ph in zip(self.systems, phases):
            s.set_phase(ph)

    def slopes(self, i):
        return np.array(self.batch.d_slopes[i])
