"""Oracle of the correlation centroider (test infrastructure only).

Restates SutraCentroiderCorr::get_cog (libsutra/src.cpp/
sutra_centroider_corr.cpp:190-262) and its kernels
(libsutra/src.cu/sutra_centroider_corr.cu), per sub-aperture s:

* fill_bincube (bcube_krnl, libsutra/src.cu/sutra_acquisim.cu:65-87, called
  with the image side as row stride): spot b[x, y] = img[vx + x, vy + y];
* fillcorr / correl / CarmaFFT: the spot and the reference function r, each
  in the corner of a zero 2n x 2n grid (n = npix), C = IFFT(conj(FFT(R)) *
  FFT(B)) unnormalised: C[k] = sum_m b[m + k] r[m] (circular on 2n points,
  equal to the linear correlation for |k| < n);
* roll2real: the lags -(n-1) .. n-1 on both axes, out[kx + n - 1, ky + n - 1];
* corr_norm: out / corrnorm (elementwise, (2n-1)^2);
* subap_sortmaxi (nmax 1): position of the maximum over the central nbmax x
  nbmax window (nbmax = n if 2n - 1 > 20, else 2n - 1); on equal values the
  lowest flat index (x + y (2n-1)) here;
* interp_parab: the sizex x sizey patch centred on the maximum, coefficients
  c = interpmat^T patch of a*x^2 + b*y^2 + c*x*y + d*x + e*y + f, the vertex
  x0 = (2 c1 c3 - c4 c2) / D, y0 = (2 c0 c4 - c3 c2) / D, D = c2^2 - 4 c0 c1
  (0 when D == 0), plus the patch position xm - 0.5 - (2n)/4 with
  xm = ceil((2 (ix - sizex/2) + sizex) / 2), then (. - offset) * scale;
* minus the reference slopes (carma_axpy).

The patch must lie inside the (2n-1)^2 map: the reference reads past it
(other rows or the next sub-aperture's map) when the maximum sits on the
border; here the patch centre is moved inside (clamped to [sizex/2,
2n - 2 - sizex/2]), as the kernels this oracle checks do.
"""
import numpy as np


def interp_mat(dimx, dimy):
    """The least-squares parabola fit matrix (dimx*dimy, 6) of
    shesha/shesha/util/rtc_util.py:23-45 (sample t = col + row*dimy, col the x
    offset, row the y offset, both centred)."""
    n = max(dimx, dimy)
    rows, cols = np.indices((n, n))
    x = cols[:dimx, :dimy] - dimx // 2
    y = rows[:dimx, :dimy] - dimy // 2
    A = np.stack([(x ** 2).ravel(), (y ** 2).ravel(), (x * y).ravel(), x.ravel(), y.ravel(),
                  np.ones(x.size)], 1).astype(np.float64)
    return np.linalg.inv(A.T @ A) @ A.T  # (6, dimx*dimy): coefficients = M @ patch


def corr_map(spot, ref):
    """Correlation map [(2n-1), (2n-1)] indexed [x, y] of one spot [n, n]
    (indexed [x, y]) with the reference function [n, n], before corrnorm:
    (2n)^2 sum_m b[m + k] r[m]."""
    n = spot.shape[0]
    B = np.zeros((2 * n, 2 * n))
    R = np.zeros((2 * n, 2 * n))
    B[:n, :n] = spot
    R[:n, :n] = ref
    # unnormalised inverse transform, as the reference's: (2n)^2 times the
    # linear correlation
    C = np.fft.ifft2(np.conj(np.fft.fft2(R)) * np.fft.fft2(B)).real * (2 * n) ** 2
    # lags -(n-1) .. n-1 on both axes
    lag = np.arange(-(n - 1), n)
    return C[np.ix_(lag % (2 * n), lag % (2 * n))]


def fit_map(C, npix, offset, scale, sizex=3, sizey=3):
    """subap_sortmaxi + interp_parab on one normalised map C [(2n-1), (2n-1)]
    indexed [x, y]: (slope x, slope y, flat index of the maximum x + y (2n-1))."""
    n = npix
    N = 2 * n - 1
    M = interp_mat(sizex, sizey)
    nbmax = n if N > 20 else N
    off = N // 2 - nbmax // 2
    win = C[off:off + nbmax, off:off + nbmax]
    flat = win.T.ravel()  # index x + y * nbmax
    k = int(np.argmax(flat))  # first maximum
    ix, iy = off + k % nbmax, off + k // nbmax
    imax = ix + iy * N
    ix = min(max(ix, sizex // 2), N - 1 - sizex // 2)
    iy = min(max(iy, sizey // 2), N - 1 - sizey // 2)
    ox, oy = ix - sizex // 2, iy - sizey // 2
    patch = C[ox:ox + sizex, oy:oy + sizey].T.ravel()  # t = col + row * sizex
    c = M @ patch
    D = c[2] ** 2 - 4.0 * c[1] * c[0]
    if D == 0:
        return 0.0, 0.0, imax
    x0 = (2.0 * c[1] * c[3] - c[4] * c[2]) / D
    y0 = (2.0 * c[0] * c[4] - c[3] * c[2]) / D
    xm = 2 * ox + sizex
    xm = xm // 2 if xm % 2 == 0 else xm // 2 + 1
    ym = 2 * oy + sizey
    ym = ym // 2 if ym % 2 == 0 else ym // 2 + 1
    sx = ((x0 + xm - 0.5 - (N + 1) // 4) - offset) * scale
    sy = ((y0 + ym - 0.5 - (N + 1) // 4) - offset) * scale
    return sx, sy, imax


def centroids(img, validx, validy, npix, ref_fn, offset, scale, corrnorm=None, sizex=3, sizey=3,
              ref=None):
    """Slopes [2*nvalid] (x then y) of every sub-aperture.  ref_fn: [n, n]
    (shared) or [n, n, nvalid]; corrnorm [(2n-1), (2n-1)] or None (ones)."""
    img = np.asarray(img, dtype=np.float64)
    n = npix
    nv = len(validx)
    out = np.zeros(2 * nv)
    for s in range(nv):
        vx, vy = int(validx[s]), int(validy[s])
        spot = img[vx:vx + n, vy:vy + n]
        r = ref_fn if np.ndim(ref_fn) == 2 else ref_fn[:, :, s]
        C = corr_map(spot, np.asarray(r, dtype=np.float64))
        if corrnorm is not None:
            C = C / np.asarray(corrnorm, dtype=np.float64)
        out[s], out[s + nv], _ = fit_map(C, n, offset, scale, sizex, sizey)
    if ref is not None:
        out = out - np.asarray(ref, dtype=np.float64)
    return out


def acquisim_bincube(image, validx, validy, npix, stride):
    """bcube_krnl (libsutra/src.cu/sutra_acquisim.cu:65-87): bincube[x, y, s] =
    image_flat[x + validx[s] + (y + validy[s]) * stride] (image_flat: the
    x-fastest flat image)."""
    flat = np.asarray(image).ravel(order="F")
    nv = len(validx)
    out = np.zeros((npix, npix, nv), dtype=flat.dtype)
    for s in range(nv):
        for y in range(npix):
            for x in range(npix):
                out[x, y, s] = flat[x + validx[s] + (y + validy[s]) * stride]
    return out
