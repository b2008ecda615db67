"""Learned centroider (MLP on 8x8-pixel spots) trained on pairs the WFS itself
generates: noisy spots of shifted / rescaled screen windows, labelled with
the geometric slopes of the same phase (mean gradient, arcsec).

The network predicts the spot displacement in pixels; slopes are that times
the pixel scale.  The CoG of the same pixels is both the fallback (spot flux
outside the training range) and the baseline its slope error is compared with.
"""
import numpy as np

from . import native, screens

GEOM_TYPE = 2  # slopes_geom: projection on the tip-tilt of each sub-aperture


class LearnedCentroider:
    def __init__(self, system, seed=1):
        self.carma, self.sutra = native()
        self.sys = system
        w, pw = system.wfs, system.p_wfs
        self.rtc = self.sutra.Rtc_FFF()
        self.rtc.add_centroider(system.ctx, w.nvalid, pw.npix / 2.0 - 0.5, pw.pixsize, False,
                                system.ctx.active_device, "mlp", w)
        self.mlp = self.rtc.d_centro[0]
        self.mlp.load_validpos(pw._validsubsx, pw._validsubsy, len(pw._validsubsx))
        self.mlp.set_npix(pw.npix)
        self.mlp.init_params(seed)
        self.mlp.set_input_scale(1.0 / float(pw._nphotons))
        self.pixsize = float(pw.pixsize)

    # -- data
    def base_screen(self, side, r0_base, seed):
        p_tel, p_geom = self.sys.p_tel, self.sys.p_geom
        dx = p_tel.diam / p_geom.pupdiam
        scr = screens.vonkarman(side, dx, r0_base, self.sys.cfg["L0"], seed)
        return self.carma.obj_float(self.sys.ctx, np.asfortranarray(scr))

    def make_data(self, base, r0_base, nframes, r0_range=(0.08, 0.25), seed=1):
        nv = self.sys.wfs.nvalid
        n = nv * nframes
        X = self.carma.obj_float(self.sys.ctx, np.zeros((64, n), dtype=np.float32, order="F"))
        T = self.carma.obj_float(self.sys.ctx, np.zeros((2, n), dtype=np.float32, order="F"))
        got = self.sutra.make_mlp_training_set(self.sys.wfs, base, r0_base, r0_range[0],
                                               r0_range[1], nframes, seed, GEOM_TYPE, X, T)
        assert got == n
        return X, T, n

    def set_flux_range_from(self, X, n, q=(0.001, 0.999)):
        flux = np.array(X).reshape(64, n, order="F").sum(0)
        lo, hi = np.quantile(flux, q)
        self.mlp.set_flux_range(float(lo), float(hi))
        return lo, hi

    # -- training
    def train(self, X, T, n, epochs=4, batch=4096, lr=2e-3, seed=1):
        return self.mlp.train(X, T, n, 1.0 / self.pixsize, batch, epochs, lr, seed)

    # -- evaluation (slopes in arcsec against the geometric slopes)
    def predict(self, X, n):
        Y = self.carma.obj_float(self.sys.ctx, np.zeros((2, n), dtype=np.float32, order="F"))
        self.mlp.forward(X, n, Y)
        return np.array(Y).reshape(2, n, order="F").T * self.pixsize

    def cog(self, X, n):
        """CoG of the same spots (x fastest), (Sx/(I+1e-6) - offset) * pixsize."""
        sp = np.array(X).reshape(64, n, order="F").T.reshape(n, 8, 8).astype(np.float64)
        I = sp.sum((1, 2))
        xs = np.arange(8.0)
        sx = (sp * xs[None, None, :]).sum((1, 2))
        sy = (sp * xs[None, :, None]).sum((1, 2))
        off = 3.5
        return np.stack([(sx / (I + 1e-6) - off), (sy / (I + 1e-6) - off)], 1) * self.pixsize

    def errors(self, X, T, n, min_flux=None):
        """Slope RMS errors (arcsec) of the network and of CoG against the
        geometric slopes; with min_flux, only over spots whose flux is at least
        that (CoG's error blows up as the flux nears zero)."""
        t = np.array(T).reshape(2, n, order="F").T
        keep = np.ones(n, bool)
        if min_flux is not None:
            keep = np.array(X).reshape(64, n, order="F").sum(0) >= min_flux
        e_mlp = float(np.sqrt(np.mean((self.predict(X, n) - t)[keep] ** 2)))
        e_cog = float(np.sqrt(np.mean((self.cog(X, n) - t)[keep] ** 2)))
        return e_mlp, e_cog
