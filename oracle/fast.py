"""Vectorised CPU restatement of one SH frame (test infrastructure only).

The same computation as ``oracle.wfs.comp_image`` followed by
``oracle.centroid.tcog`` / ``cog`` (libsutra/src.cpp/sutra_wfs_sh.cpp:337-643,
libsutra/src.cu/sutra_centroider_{cog,tcog}.cu), written with whole-array
NumPy operations and ``scipy.fft.fft2(..., workers=...)`` instead of
per-sub-aperture loops, in float64 or float32.  It is the CPU baseline that
``bench.py`` times beside the GPU (the reference has no CPU path of its own);
tests/test_oracle_fast.py checks it against the loop oracle.
"""
import numpy as np
import scipy.fft


class Frame:
    """Precomputed index tables of one WFS (the tables the GPU path uploads)."""

    def __init__(self, tables, dtype=np.float64):
        t = tables
        self.dtype = np.dtype(dtype)
        self.cdtype = np.complex128 if self.dtype == np.float64 else np.complex64
        pd, self.N, self.npix = int(t["pdiam"]), int(t["Nfft"]), int(t["npix"])
        self.pd = pd
        self.nsub = self.npix * int(t["nxsub"])
        ist = np.asarray(t["istart"], np.int64)
        jst = np.asarray(t["jstart"], np.int64)
        self.nv = len(ist)
        ax = np.arange(pd)
        # window point (x, y) of sub-aperture s: screen element [istart+x, jstart+y]
        self.wx = (ist[:, None, None] + ax[None, :, None])
        self.wy = (jst[:, None, None] + ax[None, None, :])
        self.pupil = np.asarray(t["pupil"], self.dtype)[self.wx, self.wy]
        self.off = np.asarray(t["offsets"], self.dtype)
        self.scale = self.dtype.type(np.float32(t["scale"]))
        self.binmap = np.asarray(t["binmap"], np.int64)          # [nrebin^2, npix^2]
        self.flux = np.asarray(t["fluxPerSub"], self.dtype)
        self.nphot = self.dtype.type(t["nphot"])
        p = np.arange(self.npix * self.npix)
        self.px, self.py = p % self.npix, p // self.npix
        vx = np.asarray(t["validsubsx"], np.int64)
        vy = np.asarray(t["validsubsy"], np.int64)
        self.ix = vx[:, None] + self.px[None, :]                 # [s, p]
        self.iy = vy[:, None] + self.py[None, :]

    def spots(self, phase, workers=1):
        """bincube [s, p] (normalised), raw intensities [s], binimg [x, y]."""
        ph = np.asarray(phase, self.dtype)[self.wx, self.wy]
        E = np.zeros((self.nv, self.N, self.N), dtype=self.cdtype)
        E[:, :self.pd, :self.pd] = self.pupil * np.exp(1j * (ph * self.scale - self.off))
        F = scipy.fft.fft2(E, axes=(1, 2), workers=workers)
        hr = (F.real * F.real + F.imag * F.imag).transpose(0, 2, 1).reshape(self.nv, -1)
        b = hr[:, self.binmap].sum(axis=1)                       # [s, p]
        inten = b.sum(axis=1)
        nz = inten != 0
        b[nz] *= (self.flux[nz] / inten[nz] * self.nphot)[:, None]
        img = np.zeros((self.nsub, self.nsub), dtype=self.dtype)
        img[self.ix, self.iy] = b
        return b, inten, img

    def centroids(self, img, offset, scale, threshold=None):
        """CoG (threshold None) or TCoG of the detector image; returns the
        slopes [2 nvalid] (x then y) and, for TCoG, the written-back image."""
        d = img[self.ix, self.iy]
        if threshold is not None:
            d = np.where(d > threshold, d - threshold, self.dtype.type(1e-6))
            img = img.copy()
            img[self.ix, self.iy] = d
        I = d.sum(axis=1)
        sx = (d * self.px).sum(axis=1)
        sy = (d * self.py).sum(axis=1)
        out = np.concatenate([(sx / (I + 1e-6) - offset) * scale,
                              (sy / (I + 1e-6) - offset) * scale])
        return out, img

    def frame(self, phase, offset, scale, threshold=None, workers=1):
        b, inten, img = self.spots(phase, workers)
        sl, img = self.centroids(img, offset, scale, threshold)
        return b, sl, img
