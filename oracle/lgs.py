"""Oracle of the LGS elongation (test infrastructure only).

Restates SutraLGS (libsutra/src.cpp/sutra_lgs.cpp:160-212) and its kernels
(libsutra/src.cu/sutra_lgs.cu), in float64, arrays indexed [x, y] (x = fast
device axis):

* iprof_krnl / interp_transf (sutra_lgs.cu:25-84): for sub-aperture s and HR
  pixel x, norm = 206265 doffaxis[s] / hg^2; the pixel position t = x - c
  (c = ntot/2 + 0.5 for even ntot, ntot/2 for odd), times pixsize, divided by
  norm, plus hg - h0, divided by deltah, is an index into the profile;
  linear interpolation with index 0 and weight 0 below 0, index nprof - 2 and
  weight 1 from nprof - 2 on.  When pixsize > deltah * norm the profile is
  resampled through its cumulative sum: profcum at t(x + 1) minus at t(x).
* CarmaFFT (1-D, forward), tftbeam_krnl (86-104: product with ftbeam),
  CarmaFFT (inverse, unnormalised): prof2d[x, s].
* rollbeamexp_krnl (106-129): image[x, y] = |Re prof2d[(x + ntot/2) % ntot]|
  * beam[y].
* rotate_krnl (155-226) by azimuth[s] about (ntot/2, ntot/2): the output
  pixel (x, y) samples the image at tu = u cos + v sin + c, tv = -u sin +
  v cos + c (u = x - c, v = y - c) bilinearly, each axis clamped as
  interp_transf clamps (width - 2, weight 1 at the top).
* CarmaFFT 2-D forward of each kernel: ftlgskern.
* comp_generic (sutra_wfs_sh.cpp:420-451, 483-512): the HR image of each
  sub-aperture is convolved with its kernel: unnormalised inverse FFT of
  FFT(hr) * ftlgskern[s], real part.
"""
import numpy as np


def interp_transf(idata, t, npix, pixsize, norm, shiftx, deltah, hmax):
    t = t - (npix / 2.0 + 0.5 if npix % 2 == 0 else npix / 2.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = ((t * pixsize) / norm + shiftx) / deltah
    if t < 0:  # NaN compares false twice and lands on the top sample, as on the device
        xref, w = 0, 0.0
    elif t < hmax - 2:
        xref = int(t)
        w = t - xref
    else:
        xref, w = hmax - 2, 1.0
    return idata[xref] * (1 - w) + idata[xref + 1] * w


def profiles(prof1d, profcum, nprof, doffaxis, hg, pixsize, h0, deltah, ntot):
    """The interpolated profile of every sub-aperture, [ntot, nvalid] (real)."""
    nv = len(doffaxis)
    out = np.zeros((ntot, nv))
    for s in range(nv):
        norm = np.float64(206265.0) * np.float64(doffaxis[s]) / (hg * hg)
        for x in range(ntot):
            if pixsize > deltah * norm:
                out[x, s] = (interp_transf(profcum, x + 1.0, ntot, pixsize, norm, hg - h0, deltah, nprof)
                             - interp_transf(profcum, float(x), ntot, pixsize, norm, hg - h0, deltah, nprof))
            else:
                out[x, s] = interp_transf(prof1d, float(x), ntot, pixsize, norm, hg - h0, deltah, nprof)
    return out


def prof2d(prof1d, profcum, nprof, doffaxis, hg, pixsize, h0, deltah, ftbeam, ntot):
    """lgs_update: profiles convolved with the beam, [ntot, nvalid] complex."""
    p = profiles(prof1d, profcum, nprof, doffaxis, hg, pixsize, h0, deltah, ntot)
    F = np.fft.fft(p, axis=0) * np.asarray(ftbeam, dtype=np.complex128)[:, None]
    return np.fft.ifft(F, axis=0) * ntot


def _axis(t, n):
    if t < 0:
        return 0, 0.0
    if t < n - 2:
        r = int(t)
        return r, t - r
    return n - 2, 1.0


def kernels(p2d, beam, azimuth, ntot):
    """lgs_makespot: (lgskern [x, y, s], ftlgskern [x, y, s])."""
    N = ntot
    nv = p2d.shape[1]
    lk = np.zeros((N, N, nv))
    c = float(N // 2)
    for s in range(nv):
        img = np.abs(p2d[(np.arange(N) + N // 2) % N, s].real)[:, None] * np.asarray(beam)[None, :]
        cs, sn = np.cos(azimuth[s]), np.sin(azimuth[s])
        for y in range(N):
            for x in range(N):
                u, v = x - c, y - c
                ur, wu = _axis(u * cs + v * sn + c, N)
                vr, wv = _axis(-u * sn + v * cs + c, N)
                lk[x, y, s] = ((img[ur, vr] * (1 - wu) + img[ur + 1, vr] * wu) * (1 - wv)
                               + (img[ur, vr + 1] * (1 - wu) + img[ur + 1, vr + 1] * wu) * wv)
    return lk, np.fft.fft2(lk, axes=(0, 1))


def convolve(hr, ftlgskern):
    """Each HR image hr[s] ([s, u, v]) convolved with its kernel ftlgskern[:, :, s]."""
    N = hr.shape[1]
    K = np.moveaxis(np.asarray(ftlgskern, dtype=np.complex128), 2, 0)
    return (np.fft.ifft2(np.fft.fft2(hr, axes=(1, 2)) * K, axes=(1, 2)) * (N * N)).real
