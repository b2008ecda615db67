"""Oracle restatement of the SH WFS geometry (test infrastructure only).

Restates, in plain Python / NumPy, what the reference's shesha init computes
for an NGS Shack-Hartmann WFS with a fixed number of phase points per
sub-aperture:

* ``wfs_size``     - shesha/shesha/init/geom_init.py:176-391 (``init_wfs_size``,
                     the "fixed pdiam" SH branch 259-279 and 287-312)
* ``annulus``      - shesha/shesha/util/make_pupil.py:144-190 (generic pupil,
                     no spiders) with ``dist`` from shesha/shesha/util/utilities.py
* ``sh_geometry``  - shesha/shesha/init/geom_init.py:735-899 (``init_sh_geom``)
                     and 998-1054 (``geom_init``)

The restatement is written differently from the product's geometry module on
purpose (per-pixel loops, exact rational flux fractions) so the two check each
other.

Conventions: arrays are indexed ``[x, y]`` with x the fast (contiguous) axis of
the device layout, exactly as the reference's F-ordered pybind arrays.
"""
from fractions import Fraction
import math

import numpy as np

RAD2ARCSEC = 3600.0 * 360.0 / (2 * math.pi)   # shesha/shesha/constants.py:25


def fft_goodsize(s):
    """Strictly next power of two (shesha/shesha/util/utilities.py:33-40)."""
    return 2 ** (int(math.log2(s)) + 1)


def wfs_size(diam, nxsub, pdiam, npix, pixsize, lam_um):
    """Array sizes of an SH WFS with ``pdiam`` phase points per sub-aperture.

    geom_init.py:259-279 (first nrebin / Nfft guess) then 287-312 (actual
    quantum pixel, rebin factor, pixel size and HR support).
    """
    subapd = diam / float(nxsub)
    lam = lam_um * 1e-6
    nrebin0 = max(2, int(2 * subapd * pixsize / lam / RAD2ARCSEC) + 1)
    Nfft = fft_goodsize(int(pdiam / subapd * nrebin0 / pixsize * RAD2ARCSEC * lam))
    qpixsize = (pdiam * lam / subapd * RAD2ARCSEC) / Nfft
    ratio = pixsize / qpixsize
    nrebin = int(ratio) + 1 if ratio - int(ratio) > 0.5 else int(ratio)
    pix = nrebin * qpixsize
    if pix * npix > qpixsize * Nfft:
        Ntot = fft_goodsize(int(pix * npix / qpixsize) + 1)
    else:
        Ntot = Nfft
    if Ntot % 2 != Nfft % 2:
        Ntot += 1
    return dict(pdiam=pdiam, Nfft=Nfft, Ntot=Ntot, nrebin=nrebin, qpixsize=qpixsize,
                pixsize=pix, subapd=subapd)


def pupil_extent(pupdiam):
    """geom_init.py:1009-1024: support size, useful pupil diameter, padded size."""
    ssize = int(2 ** math.ceil(math.log2(pupdiam) + 1))
    cent = ssize / 2 - 0.5
    p1 = int(math.ceil(cent - pupdiam / 2.0))
    p2 = int(math.floor(cent + pupdiam / 2.0))
    pupd = p2 - p1 + 1
    return pupd, pupd + 4


def annulus(pupd, cobs):
    """Per-pixel annulus (make_pupil.py:183-190, centre pupd/2-0.5, no spiders).

    A pixel is inside if its distance to the centre is < (pupd+1)/2 and, when
    cobs > 0, it is not < (pupd*cobs+1)/2.
    """
    c = pupd / 2.0 - 0.5
    out = np.zeros((pupd, pupd), dtype=np.float32)
    rout = (pupd + 1.0) / 2
    rin = (pupd * cobs + 1.0) * 0.5
    for i in range(pupd):
        for j in range(pupd):
            d = math.sqrt((j - c) ** 2 + (i - c) ** 2)
            v = 1.0 if d < rout else 0.0
            if cobs > 0 and d < rin:
                v -= 1.0
            out[i, j] = v
    return out


def padded_pupil(pupd, cobs):
    """Useful pupil centred in a (pupd+4)^2 support (utilities.pad_array)."""
    sp = annulus(pupd, cobs)
    n = pupd + 4
    mp = np.zeros((n, n), dtype=np.float32)
    d = (n - pupd) // 2
    mp[d:d + pupd, d:d + pupd] = sp
    return mp


def binmap_table(Ntot, npix, nrebin, shift=True):
    """binmap[i, p]: the i-th smallest flat HR index (u + v*Ntot) of binned pixel p.

    Restates geom_init.py:872-899: the central (nrebin*npix)^2 window of the
    Ntot^2 HR grid (offset by one pixel when the parities differ) is cut in
    nrebin x nrebin blocks; for an NGS (gsalt <= 0) the whole grid is then
    rolled by Ntot//2 on both axes (an fftshift of the spot).
    """
    w = nrebin * npix
    if w % 2 != Ntot % 2:
        indi = int((Ntot - w) / 2.0) + 1
    else:
        indi = int((Ntot - w) / 2.0)
    lists = [[] for _ in range(npix * npix)]
    for r in range(indi, indi + w):
        for c in range(indi, indi + w):
            p = (c - indi) // nrebin + ((r - indi) // nrebin) * npix
            if shift:
                rr, cc = (r + Ntot // 2) % Ntot, (c + Ntot // 2) % Ntot
            else:
                rr, cc = r, c
            lists[p].append(cc + rr * Ntot)
    out = np.zeros((nrebin * nrebin, npix * npix), dtype=np.int32)
    for p, l in enumerate(lists):
        out[:, p] = sorted(l)
    return out


def hrmap_table(Nfft, Ntot):
    """hrmap[k]: the Ntot-grid flat index (u + v*Ntot) receiving element k
    (u + v*Nfft) of an Nfft x Nfft HR image, geom_init.py:847-866: the
    Nfft x Nfft block at (Ntot - Nfft)/2 of an Ntot grid, rolled by Ntot//2 on
    both axes; its cells in increasing flat index, the k-th receiving element
    k (indexfill_krnl, libsutra/src.cu/sutra_wfs.cu:462-479)."""
    indi = int((Ntot - Nfft) / 2.0)
    cells = []
    for r in range(Ntot):
        for c in range(Ntot):
            rr, cc = (r - Ntot // 2) % Ntot, (c - Ntot // 2) % Ntot  # before the roll
            if indi <= rr < indi + Nfft and indi <= cc < indi + Nfft:
                cells.append(c + r * Ntot)
    return np.array(sorted(cells), dtype=np.int32).reshape(-1, 1)


def halfxy_table(pdiam, Nfft, npix, nrebin):
    """Half-pixel tip-tilt phase, geom_init.py:833-845: pi*(a+b)/Nfft (zero if
    npix and nrebin are both odd)."""
    h = np.zeros((pdiam, pdiam), dtype=np.float64)
    if not (npix % 2 == 1 and nrebin % 2 == 1):
        for a in range(pdiam):
            for b in range(pdiam):
                h[a, b] = math.pi * (a + b) / Nfft
    return h.astype(np.float32)


def sh_geometry(pupdiam, nxsub, pdiam, npix, nrebin, Nfft, Ntot, fracsub, cobs):
    """Valid sub-apertures and index tables of one SH WFS (init_sh_geom).

    Returns a dict with:
      mpupil [x,y] float32 (npup^2), isvalid [i,j] (i = x index),
      validsubsx/y (sub-aperture units, order: y slow, x fast),
      istart/jstart (validpuppix = valid*pdiam+2), fluxPerSub (valid order),
      halfxy [x,y], binmap [i,p], nvalid, npup.
    """
    pupd, npup = pupil_extent(pupdiam)
    assert pupd == pupdiam
    mp = padded_pupil(pupd, cobs)
    starts = [int(k * pupdiam / nxsub) + 2 for k in range(nxsub)]
    flux = [[None] * nxsub for _ in range(nxsub)]
    for i in range(nxsub):
        for j in range(nxsub):
            cnt = 0
            for a in range(pdiam):
                for b in range(pdiam):
                    cnt += int(mp[starts[i] + a, starts[j] + b])
            flux[i][j] = Fraction(cnt, pdiam * pdiam)
    isvalid = np.zeros((nxsub, nxsub), dtype=np.int32)
    for i in range(nxsub):
        for j in range(nxsub):
            isvalid[i, j] = 1 if flux[i][j] >= Fraction(fracsub).limit_denominator(10 ** 9) else 0
    vx, vy, fl = [], [], []
    for j in range(nxsub):          # y slow
        for i in range(nxsub):      # x fast
            if isvalid[i, j]:
                vx.append(i)
                vy.append(j)
                fl.append(float(flux[i][j]))
    vx = np.array(vx, dtype=np.int32)
    vy = np.array(vy, dtype=np.int32)
    return dict(
        mpupil=mp, npup=npup, isvalid=isvalid, nvalid=len(vx),
        validsubsx=vx, validsubsy=vy,
        istart=(vx * pdiam + 2).astype(np.int32), jstart=(vy * pdiam + 2).astype(np.int32),
        fluxPerSub=np.array(fl, dtype=np.float32),
        halfxy=halfxy_table(pdiam, Nfft, npix, nrebin),
        binmap=binmap_table(Ntot, npix, nrebin),
    )


def field_stop_mask(mpupil_side, pupdiam, lam_um, diam, fssize, fstop):
    """Field-stop mask of geom_init (shesha/shesha/init/geom_init.py:973-995):
    fftsize = fft_goodsize(mpupil_side), pixel fspixsize = pupdiam lambda / D
    / fftsize (arcsec), radius r = int(fssize / fspixsize / 2); round: the
    distance to the centre (fftsize/2 - 0.5 on both axes) below r; square:
    |x|, |y| <= r about (fftsize + 1) / 2 of the 1-based indices; fftshifted."""
    n = fft_goodsize(mpupil_side)
    fspix = pupdiam * lam_um * 1e-6 / diam * RAD2ARCSEC / n
    r = int(fssize / fspix / 2.0)
    i = np.arange(n, dtype=np.float64)
    if fstop == "round":
        c = n / 2.0 - 0.5
        m = np.sqrt((i[:, None] - c) ** 2 + (i[None, :] - c) ** 2) < r
    else:
        c = (n + 1.0) / 2.0
        a = np.abs(i + 1 - c)
        m = (a[:, None] <= r) & (a[None, :] <= r)
    return np.fft.fftshift(m.astype(np.float64))
