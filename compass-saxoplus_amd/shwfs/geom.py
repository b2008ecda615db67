"""Pupil and Shack-Hartmann geometry tables (host side, NumPy).

Produces, for one SH WFS, every table ``SutraWfsSH::load_arrays`` receives,
with the reference's numerical conventions (float32 tables, column-major
device layout, sub-aperture order = np.where(isvalid.T)):

* init_wfs_size  - shesha/shesha/init/geom_init.py:176-391 (SH branch)
* geom_init      - geom_init.py:998-1054
* make_pupil     - shesha/shesha/util/make_pupil.py:144-190 (annulus, no spiders)
* init_sh_geom   - geom_init.py:735-996
* compute_nphotons - geom_init.py:394-471

All arrays indexed [x, y] with x the fast device axis (numpy axis 0), as the
reference's F-ordered pybind boundary.
"""
import math

import numpy as np

RAD2ARCSEC = 3600.0 * 360.0 / (2 * np.pi)
DEG2RAD = np.pi / 180.0


def fft_goodsize(s):
    """Strictly next power of two (shesha/shesha/util/utilities.py:33-40)."""
    return 2 ** (int(np.log2(s)) + 1)


def dist(dim, xc=-1, yc=-1):
    """Distance of every pixel of a dim x dim grid to (xc, yc); xc on axis 1."""
    if xc < 0:
        xc = int(dim / 2.0)
    if yc < 0:
        yc = int(dim / 2.0)
    ax = np.arange(dim)
    return np.hypot(ax[None, :] - xc, ax[:, None] - yc)


def makegaussian(size, fwhm, xc=-1, yc=-1, norm=0):
    """exp(-(r / (fwhm/1.66))^2) (shesha/shesha/util/utilities.py:125-140)."""
    g = np.exp(-((dist(size, xc, yc) / (fwhm / 1.66)) ** 2.0))
    if norm > 0:
        g = g / (fwhm ** 2.0 * 1.140075)
    return g


def pad_array(A, N):
    out = np.zeros((N, N))
    d1 = (N - A.shape[0]) // 2
    d2 = (N - A.shape[1]) // 2
    out[d1:d1 + A.shape[0], d2:d2 + A.shape[1]] = A
    return out


def make_pupil(dim, pupd, cobs, xc, yc):
    """Generic annular pupil: (r < (pupd+1)/2) - (r < (pupd*cobs+1)/2)."""
    r = dist(dim, xc, yc)
    pup = (r < (pupd + 1.0) / 2).astype(np.float32)
    if cobs > 0:
        pup -= (r < (pupd * cobs + 1.0) * 0.5).astype(np.float32)
    return pup


def init_wfs_size(p_wfs, r0, p_tel, nfft=None, nrebin=None):
    """pdiam / Nfft / Ntot / nrebin / pixel sizes of an SH WFS.

    With ``nfft`` and ``nrebin`` given, they are imposed (the pixel size
    becomes nrebin quantum pixels); otherwise shesha's sizing rule picks them
    from the requested pixel size.  Note that the rule's strictly-next power
    of two never yields nfft == 2*pdiam for a power-of-two pdiam.
    """
    if nfft is not None and nrebin is not None:
        subapd = p_tel.diam / float(p_wfs.nxsub)
        pdiam = p_wfs._pdiam
        qpixsize = (pdiam * p_wfs.Lambda * 1.0e-6 / subapd * RAD2ARCSEC) / nfft
        p_wfs._pdiam, p_wfs._Nfft, p_wfs._nrebin = pdiam, int(nfft), int(nrebin)
        p_wfs._qpixsize, p_wfs.pixsize, p_wfs._subapd = qpixsize, nrebin * qpixsize, subapd
        Ntot = nfft
        if p_wfs.pixsize * p_wfs.npix > qpixsize * nfft:
            Ntot = fft_goodsize(int(p_wfs.pixsize * p_wfs.npix / qpixsize) + 1)
            if Ntot % 2 != nfft % 2:
                Ntot += 1
        p_wfs._Ntot = Ntot
        return
    r0w = r0 * (p_wfs.Lambda * 2) ** (6.0 / 5) if r0 else 0.0
    subapd = p_tel.diam / float(p_wfs.nxsub)
    lam = p_wfs.Lambda * 1.0e-6
    if p_wfs._pdiam <= 0:
        if r0w <= 0:
            raise ValueError("init_wfs_size: r0 is needed when pdiam is not imposed")
        pdiam = max(16, int(6 * subapd / r0w))
        if (pdiam * p_wfs.nxsub) % 2:
            pdiam += 1
    else:
        pdiam = p_wfs._pdiam
    nrebin = max(2, int(2 * subapd * p_wfs.pixsize / lam / RAD2ARCSEC) + 1)
    Nfft = fft_goodsize(int(pdiam / subapd * nrebin / p_wfs.pixsize * RAD2ARCSEC * lam))
    qpixsize = (pdiam * lam / subapd * RAD2ARCSEC) / Nfft
    frac = p_wfs.pixsize / qpixsize
    nrebin = int(frac) + 1 if frac - int(frac) > 0.5 else int(frac)
    pixsize = nrebin * qpixsize
    if pixsize * p_wfs.npix > qpixsize * Nfft:
        Ntot = fft_goodsize(int(pixsize * p_wfs.npix / qpixsize) + 1)
    else:
        Ntot = Nfft
    if Ntot % 2 != Nfft % 2:
        Ntot += 1
    p_wfs._pdiam = pdiam
    p_wfs.pixsize = pixsize
    p_wfs._qpixsize = qpixsize
    p_wfs._Nfft = Nfft
    p_wfs._Ntot = Ntot
    p_wfs._nrebin = nrebin
    p_wfs._subapd = subapd


def geom_init(p_geom, p_tel, padding=2):
    """Useful pupil (pupdiam^2) and its padded copy (pupdiam + 2*padding)^2."""
    p_geom.ssize = int(2 ** np.ceil(np.log2(p_geom.pupdiam) + 1))
    p_geom.cent = p_geom.ssize / 2 - 0.5
    p_geom._p1 = int(np.ceil(p_geom.cent - p_geom.pupdiam / 2.0))
    p_geom._p2 = int(np.floor(p_geom.cent + p_geom.pupdiam / 2.0))
    p_geom.pupdiam = p_geom._p2 - p_geom._p1 + 1
    p_geom._n = p_geom.pupdiam + 2 * padding
    p_geom._n1 = p_geom._p1 - padding
    p_geom._n2 = p_geom._p2 + padding
    cent = p_geom.pupdiam / 2.0 - 0.5
    if p_tel.spiders_type is not None:
        raise NotImplementedError("spiders are not part of this build (annular pupils only)")
    p_geom._spupil = make_pupil(p_geom.pupdiam, p_geom.pupdiam, p_tel.cobs, cent, cent).astype(np.float32)
    p_geom._mpupil = pad_array(p_geom._spupil, p_geom._n).astype(np.float32)
    p_geom._pixsize = p_tel.diam / p_geom.pupdiam
    p_geom.is_init = True


def init_wfs_geom(p_wfs, r0, p_tel, p_geom, ittime=1.0, nfft=None, nrebin=None):
    """pdiam from the pupil sampling, sizes, pupil, then the SH tables."""
    if p_geom.pupdiam:
        pdiam = p_geom.pupdiam // p_wfs.nxsub
        if (pdiam * p_wfs.nxsub) % 2:
            pdiam += 1
    else:
        pdiam = -1
    p_wfs._pdiam = pdiam
    init_wfs_size(p_wfs, r0, p_tel, nfft=nfft, nrebin=nrebin)
    if not p_geom.is_init:
        p_geom.pupdiam = p_wfs._pdiam * p_wfs.nxsub
        geom_init(p_geom, p_tel)
    init_sh_geom(p_wfs, p_tel, p_geom, ittime=ittime, r0=r0)


def compute_nphotons(ittime, optthroughput, diam, nxsub, zerop=0.0, gsmag=0.0,
                     lgsreturnperwatt=0.0, laserpower=0.0):
    """Photons per fully illuminated sub-aperture per frame: from the guide
    star magnitude, or for an LGS from its return per watt and laser power
    (shesha/shesha/init/geom_init.py:396-458)."""
    if laserpower > 0:
        return lgsreturnperwatt * laserpower * optthroughput * (diam / nxsub) ** 2.0 * 1e4 * ittime
    if zerop == 0:
        zerop = 1e11
    return zerop * 10.0 ** (-0.4 * gsmag) * ittime * optthroughput * (diam / nxsub) ** 2.0


def _valid_subaps(p_wfs, p_geom):
    pd, nx = p_wfs._pdiam, p_wfs.nxsub
    starts = (np.linspace(0, p_geom.pupdiam, nx + 1)[:-1]).astype(np.int64) + 2
    mp = p_geom._mpupil
    blocks = np.stack([np.stack([mp[a:a + pd, b:b + pd] for b in starts]) for a in starts])
    flux = blocks.sum(axis=(2, 3), dtype=np.float32) / np.float32(pd ** 2.0)
    return starts, flux.astype(np.float32)


def _ttprojmat(p_wfs, p_tel, p_geom, starts):
    """Least-squares tip/tilt projection of each sub-aperture phase (geom_init.py:800-830)."""
    pd, nv = p_wfs._pdiam, p_wfs._nvalid
    mT = p_geom._mpupil.T
    out = np.zeros((pd * pd, 2 * nv), dtype=np.float32)
    for i in range(nv):
        r0 = starts[p_wfs._validsubsy[i]]
        c0 = starts[p_wfs._validsubsx[i]]
        m = mT[r0:r0 + pd, c0:c0 + pd].astype(np.float32)
        left = np.zeros_like(m)
        left[:, 1:] = m[:, :-1]
        right = np.zeros_like(m)
        right[:, :-1] = m[:, 1:]
        up = np.zeros_like(m)
        up[1:, :] = m[:-1, :]
        down = np.zeros_like(m)
        down[:-1, :] = m[1:, :]
        sx = (left - right) * m
        sy = (up - down) * m
        dx = -sx.cumsum(axis=1).sum()
        dy = -sy.cumsum(axis=0).sum()
        if dx == 0 or dy == 0:
            continue
        out[:, i] = (sx / dx).ravel()
        out[:, i + nv] = (sy / dy).ravel()
    return out * p_geom.pupdiam / p_tel.diam * RAD2ARCSEC * 1e-6


def _binmap(p_wfs):
    """HR flat indices (u + v*Ntot) of the nrebin^2 pixels of each binned pixel,
    window centred in the Ntot grid and fftshifted for an NGS."""
    Ntot, w, nreb, npix = p_wfs._Ntot, p_wfs._nrebin * p_wfs.npix, p_wfs._nrebin, p_wfs.npix
    indi = (Ntot - w) // 2 + (1 if (w % 2) != (Ntot % 2) else 0)
    lab = np.zeros((Ntot, Ntot), dtype=np.int64)
    r, c = np.indices((w, w))
    lab[indi:indi + w, indi:indi + w] = (c // nreb) + (r // nreb) * npix + 1
    if p_wfs.gsalt <= 0:
        lab = np.roll(np.roll(lab, Ntot // 2, axis=0), Ntot // 2, axis=1)
    flat = lab.ravel()                       # C order: index = c + r*Ntot
    order = np.argsort(flat, kind="stable")  # stable: increasing flat index within a label
    counts = np.bincount(flat, minlength=npix * npix + 1)
    start = np.cumsum(counts)[:-1]
    out = np.zeros((nreb * nreb, npix * npix), dtype=np.int32)
    for p in range(npix * npix):
        s = start[p]
        out[:, p] = order[s:s + nreb * nreb]
    return out


def _hrmap(p_wfs):
    """Ntot-grid flat index (u + v*Ntot) of each Nfft-grid HR element
    (u + v*Nfft), geom_init.py:847-866: the Nfft block at (Ntot - Nfft)/2,
    rolled by Ntot//2 on both axes, its cells in increasing flat index; a
    (2, 2) placeholder when Ntot == Nfft (the reference uploads it unused)."""
    Ntot, Nfft = p_wfs._Ntot, p_wfs._Nfft
    if Ntot == Nfft:
        return np.zeros((2, 2), dtype=np.int32)
    indi = int((Ntot - Nfft) / 2.0)
    block = np.zeros((Ntot, Ntot), dtype=bool)
    block[indi:indi + Nfft, indi:indi + Nfft] = True
    block = np.roll(np.roll(block, Ntot // 2, axis=0), Ntot // 2, axis=1)
    return np.flatnonzero(block.ravel()).astype(np.int32).reshape(-1, 1)


def init_sh_geom(p_wfs, p_tel, p_geom, ittime=1.0, r0=None):
    pd = p_wfs._pdiam
    starts, flux = _valid_subaps(p_wfs, p_geom)
    p_wfs.nPupils = 1
    p_wfs._isvalid = (flux >= p_wfs.fracsub).astype(np.int32)
    p_wfs._nvalid = int(p_wfs._isvalid.sum())
    p_wfs._fluxPerSub = flux.copy()
    wy, wx = np.where(p_wfs._isvalid.T)
    p_wfs._validsubsx = wx.astype(np.int32)
    p_wfs._validsubsy = wy.astype(np.int32)
    p_wfs._validpuppixx = (p_wfs._validsubsx * pd + 2).astype(np.int32)
    p_wfs._validpuppixy = (p_wfs._validsubsy * pd + 2).astype(np.int32)

    n = p_geom._n
    xx, yy = np.meshgrid(np.arange(pd), np.arange(pd), indexing="xy")   # [yy, xx]
    local = (xx + yy * n).ravel()                                       # k = xx + yy*pd order
    x0 = starts[p_wfs._validsubsx]
    y0 = starts[p_wfs._validsubsy]
    p_wfs._phasemap = (local[:, None] + (x0 + y0 * n)[None, :]).astype(np.int32)
    p_wfs._ttprojmat = _ttprojmat(p_wfs, p_tel, p_geom, starts)

    p_wfs._validsubsx = (p_wfs._validsubsx * p_wfs.npix).astype(np.int32)
    p_wfs._validsubsy = (p_wfs._validsubsy * p_wfs.npix).astype(np.int32)

    if p_wfs.npix % 2 == 1 and p_wfs._nrebin % 2 == 1:
        p_wfs._halfxy = np.zeros((pd, pd), dtype=np.float32)
    else:
        h = np.linspace(0, 2 * np.pi, p_wfs._Nfft + 1)[0:pd] / 2.0
        p_wfs._halfxy = (h[None, :] + h[:, None]).astype(np.float32)

    p_wfs._hrmap = _hrmap(p_wfs)
    p_wfs._binmap = _binmap(p_wfs)

    # seeing + detector blur kernel (only read with kernel convolution; kept
    # because load_arrays takes it), geom_init.py:901-926
    Ntot = p_wfs._Ntot
    fwhm = p_wfs.kernel
    if r0:
        dr0 = (p_tel.diam / r0 * (0.5 / p_wfs.Lambda) ** 1.2 /
               np.cos(p_geom.zenithangle * DEG2RAD) ** 0.6)
        seeing = p_wfs.Lambda / (p_tel.diam / np.sqrt(p_wfs.nxsub ** 2.0 + (dr0 / 1.5) ** 2.0)) / 4.848
        seeing = min(seeing, 2 * p_wfs.pixsize)
        fwhm = np.sqrt(seeing ** 2.0 + p_wfs.kernel ** 2.0)
    k = makegaussian(Ntot, max(fwhm, 1e-6) / p_wfs._qpixsize, Ntot // 2, Ntot // 2)
    k = np.roll(np.roll(k.astype(np.float32), Ntot // 2, axis=0), Ntot // 2, axis=1)
    k[0, 0] = 1.0
    k = k / np.sum(k)
    p_wfs._ftkernel = (np.fft.fft2(k) / (Ntot * Ntot)).astype(np.complex64)

    lgs = p_wfs.gsalt > 0
    p_wfs._nphotons = compute_nphotons(ittime, p_wfs.optthroughput, p_tel.diam, p_wfs.nxsub,
                                       zerop=p_wfs.zerop, gsmag=p_wfs.gsmag,
                                       lgsreturnperwatt=p_wfs.lgsreturnperwatt if lgs else 0.0,
                                       laserpower=p_wfs.laserpower if lgs else 0.0)
    p_wfs._submask = field_stop_mask(p_wfs, p_tel, p_geom)


def field_stop_mask(p_wfs, p_tel, p_geom):
    """Focal-plane stop of the whole pupil's transform, fftshifted
    (shesha/shesha/init/geom_init.py:973-995): on the fft_goodsize of the
    padded pupil, pixels of fspixsize = pupdiam lambda / D / fftsize arcsec,
    a disc (round) or square of side fssize, centred between the middle
    pixels.  None when fssize is 0."""
    if p_wfs.fssize == 0:
        return None
    fftsize = fft_goodsize(p_geom._mpupil.shape[0])
    fspixsize = (p_geom.pupdiam * (p_wfs.Lambda * 1.0e-6) / p_tel.diam * RAD2ARCSEC) / fftsize
    r = int(p_wfs.fssize / fspixsize / 2.0)
    if p_wfs.fstop == "round":
        focmask = dist(fftsize, xc=fftsize / 2.0 - 0.5, yc=fftsize / 2.0 - 0.5) < r
    elif p_wfs.fstop == "square":
        X = np.indices((fftsize, fftsize)) + 1
        x = X[1] - (fftsize + 1.0) / 2.0
        y = X[0] - (fftsize + 1.0) / 2.0
        focmask = (np.abs(x) <= r) * (np.abs(y) <= r)
    else:
        raise ValueError("wfs fstop must be 'round' or 'square'")
    return np.fft.fftshift(focmask * 1.0).astype(np.float32)


def flux_per_valid_subap(p_wfs):
    """fluxPerSub in sub-aperture order, as wfs_init passes it to load_arrays."""
    return p_wfs._fluxPerSub.T[np.where(p_wfs._isvalid.T > 0)].copy()
