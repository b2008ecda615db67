"""Oracle for SutraWfsSH::comp_image in float64 (test infrastructure only).

Restates libsutra/src.cpp/sutra_wfs_sh.cpp:337-643 for the NGS, no-field-stop
case, kernel by kernel (the enlarged grid Ntot != Nfft and the kernel
convolution kernconv included):

* camplipup_krnl   libsutra/src.cu/sutra_wfs.cu:27-53
      E[x + y*Nfft + s*Nfft^2] = mask[p] * exp(i*(phase[p]*scale - offset[x + y*pdiam]))
      with p = x + y*npup + istart[s] + jstart[s]*npup, x, y < pdiam, zero elsewhere.
* CarmaFFT(dir=1)  libcarma/src.cpp/carma_fft.cpp:164-180 -> forward, unnormalised,
      exp(-2*pi*i*k*n/N)  (numpy.fft.fft2 convention).
* abs2c_krnl       libsutra/src.cu/sutra_utils.cu:144-168
* indexfill_krnl   libsutra/src.cu/sutra_wfs.cu:462-479 (Ntot != Nfft)
      hr_tot[hrmap[k] + s*Ntot^2] = hr[k + s*Nfft^2] on a zero Ntot^2 grid
* kernconv         libsutra/src.cpp/sutra_wfs_sh.cpp:453-464, 514-525:
      forward FFT, conv_krnl (sutra_wfs.cu:491-513: product with ftkernel),
      inverse FFT (CarmaFFT dir -1: unnormalised), real part read by binning
* fillbincube_krnl libsutra/src.cu/sutra_wfs.cu:369-414
      bincube[p + s*npix^2] = sum_i hr[binmap[i + p*nrebin^2] + s*Nfft^2]
* reduce2          libsutra/src.cu/sutra_wfs.cu:556-590  (per-subap flux)
* subapnorm_krnl   libsutra/src.cu/sutra_utils.cu:170-193
      b = b * fluxPerSub[s] / I[s] * nphot, skipped when I[s] == 0
* bimg_krnl        libsutra/src.cu/sutra_wfs.cu:183-229 (alpha = 0)
      binimg[x + y*nsub + vx[s] + vy[s]*nsub] = bincube[x + y*npix + s*npix^2]
* field stop       libsutra/src.cpp/sutra_wfs_sh.cpp:366-391:
      fsamplipup_krnl (sutra_wfs.cu:130-152): f[x, y] = mask * exp(i phase*scale)
      on the npup corner of an n x n zero grid; CarmaFFT forward; submask_krnl
      (1868-1888): product with the fftshifted stop; CarmaFFT inverse
      (unnormalised); camplipup_krnl's field-stop form (sutra_wfs.cu:87-128)
      then takes E = F[p] * exp(-i offset) / n^2 (no pupil product: the mask
      is already inside F).

Arrays are indexed [x, y] (x = fast device axis).
"""
import numpy as np


def fields(phase, pupil, scale, offsets, istart, jstart, pdiam, Nfft, field=None):
    """Zero-padded complex fields, shape (nvalid, Nfft, Nfft) indexed [s, x, y].
    field: the complex pupil field [npup, npup] of a field stop (F / n^2,
    from field_stop), which replaces pupil * exp(i phase * scale)."""
    nvalid = len(istart)
    E = np.zeros((nvalid, Nfft, Nfft), dtype=np.complex128)
    off = np.asarray(offsets, dtype=np.float64)
    if field is not None:
        for s in range(nvalid):
            i0, j0 = int(istart[s]), int(jstart[s])
            E[s, :pdiam, :pdiam] = field[i0:i0 + pdiam, j0:j0 + pdiam] * np.exp(-1j * off)
        return E
    ph = np.asarray(phase, dtype=np.float64)
    pu = np.asarray(pupil, dtype=np.float64)
    sc = float(np.float32(scale))
    for s in range(nvalid):
        i0, j0 = int(istart[s]), int(jstart[s])
        sub = ph[i0:i0 + pdiam, j0:j0 + pdiam]
        msk = pu[i0:i0 + pdiam, j0:j0 + pdiam]
        E[s, :pdiam, :pdiam] = msk * np.exp(1j * (sub * sc - off))
    return E


def field_stop(phase, pupil, scale, submask):
    """The field-stopped pupil field F / n^2 on the npup x npup support
    (arrays indexed [x, y]; submask [n, n] indexed [kx, ky])."""
    ph = np.asarray(phase, dtype=np.float64)
    npup = ph.shape[0]
    n = submask.shape[0]
    f = np.zeros((n, n), dtype=np.complex128)
    f[:npup, :npup] = np.asarray(pupil, dtype=np.float64) * np.exp(1j * ph * float(np.float32(scale)))
    # np.fft.ifft2 divides by n^2: the unnormalised inverse over n^2
    F = np.fft.ifft2(np.fft.fft2(f) * np.asarray(submask, dtype=np.float64))
    return F[:npup, :npup]


def hr_images(E):
    """|FFT2|^2 of each field (forward, unnormalised)."""
    F = np.fft.fft2(E, axes=(1, 2))
    return (F.real ** 2 + F.imag ** 2)


def extend_grid(hr, hrmap, Ntot):
    """indexfill: each Nfft x Nfft image onto a zero Ntot x Ntot grid, element
    k = u + v*Nfft (F-order flat) at flat index hrmap[k]; returns [s, u, v]."""
    nvalid, N, _ = hr.shape
    flat = hr.transpose(0, 2, 1).reshape(nvalid, N * N)
    out = np.zeros((nvalid, Ntot * Ntot))
    out[:, np.asarray(hrmap).ravel()] = flat
    return out.reshape(nvalid, Ntot, Ntot).transpose(0, 2, 1)


def convolve_kernel(hr, ftkernel):
    """kernconv: real part of the unnormalised inverse FFT of FFT(hr) * ftkernel;
    ftkernel [u, v] (F-order view of the device array) already holds the
    1/N^2 of the inverse transform."""
    N = hr.shape[1]
    K = np.asarray(ftkernel, dtype=np.complex128).reshape(N, N, order="F") \
        if np.ndim(ftkernel) == 1 else np.asarray(ftkernel, dtype=np.complex128)
    F = np.fft.fft2(hr, axes=(1, 2)) * K[None]
    return (np.fft.ifft2(F, axes=(1, 2)) * (N * N)).real


def bin_images(hr, binmap, npix):
    """Gather-sum of the HR pixels listed by binmap, returns [px, py, s]."""
    nvalid, N, _ = hr.shape
    flat = hr.transpose(0, 2, 1).reshape(nvalid, N * N)   # F-order flat index u + v*N
    bm = np.asarray(binmap)
    b = np.zeros((nvalid, npix * npix))
    for i in range(bm.shape[0]):
        b += flat[:, bm[i, :]]
    return b.reshape(nvalid, npix, npix).transpose(2, 1, 0)  # p = px + py*npix -> [px, py, s]


def normalise(bincube, fluxPerSub, nphot):
    """Per-subap flux (intensities) and photometric normalisation."""
    inten = bincube.sum(axis=(0, 1))
    out = bincube.copy()
    nz = inten != 0
    out[:, :, nz] = out[:, :, nz] * (np.asarray(fluxPerSub, np.float64)[nz] / inten[nz] * float(nphot))
    return out, inten


def binimg_from_cube(bincube, validsubsx, validsubsy, nsub, base=None):
    """Scatter each sub-aperture spot into the detector image at (vx, vy)."""
    npix = bincube.shape[0]
    img = np.zeros((nsub, nsub)) if base is None else np.array(base, dtype=np.float64)
    for s in range(bincube.shape[2]):
        vx, vy = int(validsubsx[s]), int(validsubsy[s])
        img[vx:vx + npix, vy:vy + npix] = bincube[:, :, s]
    return img


def comp_image(phase, pupil, scale, offsets, istart, jstart, validsubsx, validsubsy,
               binmap, fluxPerSub, nphot, pdiam, Nfft, npix, nxsub, Ntot=None, hrmap=None,
               ftkernel=None, submask=None, ftlgskern=None):
    """Noise-free SH image: returns (bincube [px,py,s], intensities_raw, binimg [x,y]).

    ``intensities_raw`` is what the reference leaves in d_intensities: the
    per-subap sum *before* normalisation.  Ntot / hrmap: the enlarged grid
    (binmap then indexes it); ftkernel: the kernel convolution (kernconv);
    submask: the field stop; ftlgskern [Ntot, Ntot, nvalid]: the LGS kernels
    (with Ntot == Nfft the kernel convolution is then skipped, as the
    reference's comp_generic does).
    """
    field = None if submask is None else field_stop(phase, pupil, scale, submask)
    E = fields(phase, pupil, scale, offsets, istart, jstart, pdiam, Nfft, field)
    hr = hr_images(E)
    if Ntot is not None and Ntot != Nfft:
        hr = extend_grid(hr, hrmap, Ntot)
    if ftlgskern is not None:
        from oracle import lgs
        hr = lgs.convolve(hr, ftlgskern)
    if ftkernel is not None and (ftlgskern is None or (Ntot is not None and Ntot != Nfft)):
        hr = convolve_kernel(hr, ftkernel)
    b = bin_images(hr, binmap, npix)
    bn, inten = normalise(b, fluxPerSub, nphot)
    img = binimg_from_cube(bn, validsubsx, validsubsy, npix * nxsub)
    return bn, inten, img


def brute_dft2(E):
    """O(N^4) textbook forward DFT of one 2-D field (pins the FFT convention)."""
    N = E.shape[0]
    out = np.zeros((N, N), dtype=np.complex128)
    for k in range(N):
        for l in range(N):
            acc = 0j
            for n in range(N):
                for m in range(N):
                    acc += E[n, m] * np.exp(-2j * np.pi * (k * n + l * m) / N)
            out[k, l] = acc
    return out
