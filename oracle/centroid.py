"""Oracle for the CoG-family centroiders in float64 (test infrastructure only).

* cog   libsutra/src.cu/sutra_centroider_cog.cu:28-81
* wcog  libsutra/src.cu/sutra_centroider_wcog.cu:58-116 (+ fill_weights 27-50)
* tcog  libsutra/src.cu/sutra_centroider_tcog.cu:27-76 (writes d' back to the image)
* bpcog libsutra/src.cu/sutra_centroider_bpcog.cu:28-78: t = element nbpix of the
        window sorted in decreasing order (pixels outside the image count as 0),
        d = img > t ? img - t : 0 (written back), float moments
* slope ordering: libsutra/src.cu/sutra_centroider_utils.cuh:31-51
      untied:     x slopes at [0, nvalid), y slopes at [nvalid, 2*nvalid)
      interlaced: x at 2*s, y at 2*s+1

Per sub-aperture s with window img[vx+x, vy+y], x, y < npix:
      I  = sum d,  Sx = sum d*x,  Sy = sum d*y
      slope_x = (Sx / (I + 1e-6) - offset) * scale - ref_x
where d = img for CoG, d = (img > t ? img - t : 1e-6) for TCoG, and
d = (img > t ? img - t : 1e-6) * w for WCoG.  d_intensities[s] = I.
"""
import numpy as np


def _slot(s, nvalid, order):
    if order == 0:
        return s, s + nvalid
    return 2 * s, 2 * s + 1


def _moments(d):
    npix = d.shape[0]
    x = np.arange(npix, dtype=np.float64)[:, None]
    y = np.arange(npix, dtype=np.float64)[None, :]
    return d.sum(), (d * x).sum(), (d * y).sum()


def _finish(I, Sx, Sy, offset, scale):
    cx = (Sx / (I + 1e-6) - offset) * scale
    cy = (Sy / (I + 1e-6) - offset) * scale
    return cx, cy


def cog(img, validx, validy, npix, offset, scale, ref=None, order=0):
    img = np.asarray(img, dtype=np.float64)
    nvalid = len(validx)
    slopes = np.zeros(2 * nvalid)
    inten = np.zeros(nvalid)
    for s in range(nvalid):
        d = img[validx[s]:validx[s] + npix, validy[s]:validy[s] + npix]
        I, Sx, Sy = _moments(d)
        cx, cy = _finish(I, Sx, Sy, offset, scale)
        ix, iy = _slot(s, nvalid, order)
        slopes[ix] = cx - (0 if ref is None else ref[ix])
        slopes[iy] = cy - (0 if ref is None else ref[iy])
        inten[s] = I
    return slopes, inten


def threshold_pixels(d, threshold):
    d = np.asarray(d, dtype=np.float64)
    return np.where(d > threshold, d - threshold, 1e-6)


def tcog(img, validx, validy, npix, offset, scale, threshold, ref=None, order=0):
    """Returns slopes, intensities and the image with d' written back."""
    img = np.array(img, dtype=np.float64)
    nvalid = len(validx)
    slopes = np.zeros(2 * nvalid)
    inten = np.zeros(nvalid)
    for s in range(nvalid):
        win = img[validx[s]:validx[s] + npix, validy[s]:validy[s] + npix]
        d = threshold_pixels(win, threshold)
        img[validx[s]:validx[s] + npix, validy[s]:validy[s] + npix] = d
        I, Sx, Sy = _moments(d)
        cx, cy = _finish(I, Sx, Sy, offset, scale)
        ix, iy = _slot(s, nvalid, order)
        slopes[ix] = cx - (0 if ref is None else ref[ix])
        slopes[iy] = cy - (0 if ref is None else ref[iy])
        inten[s] = I
    return slopes, inten, img


def wcog(img, validx, validy, npix, offset, scale, weights, threshold=0.0, ref=None, order=0):
    """weights: [px, py, s] (3-D) or [px, py] broadcast to every sub-aperture."""
    img = np.asarray(img, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    nvalid = len(validx)
    slopes = np.zeros(2 * nvalid)
    inten = np.zeros(nvalid)
    for s in range(nvalid):
        win = img[validx[s]:validx[s] + npix, validy[s]:validy[s] + npix]
        ws = w if w.ndim == 2 else w[:, :, s]
        d = threshold_pixels(win, threshold) * ws
        I, Sx, Sy = _moments(d)
        cx, cy = _finish(I, Sx, Sy, offset, scale)
        ix, iy = _slot(s, nvalid, order)
        slopes[ix] = cx - (0 if ref is None else ref[ix])
        slopes[iy] = cy - (0 if ref is None else ref[iy])
        inten[s] = I
    return slopes, inten


def gaussian_weights(npix, width):
    """comp_weights GAUSS branch for even npix (shesha/shesha/init/rtc_init.py:420-433)
    with makegaussian (utilities.py:125-140): exp(-(r/(fwhm/1.66))^2), centre npix//2-0.5."""
    c = npix // 2 - 0.5
    out = np.zeros((npix, npix))
    for i in range(npix):
        for j in range(npix):
            r = np.sqrt((j - c) ** 2 + (i - c) ** 2)
            out[i, j] = np.exp(-((r / (width / 1.66)) ** 2.0))
    return out.astype(np.float32)


def bpcog(img, validx, validy, npix, nbpix, offset, scale, ref=None, order=0):
    """Brightest-pixel CoG.  Pixel (x, y) of sub-aperture s is the flat
    (x fastest) image element (x + validx[s]) + (y + validy[s]) * size, read
    only below size^2, as the reference indexes it.  Returns (slopes,
    intensities, thresholded image)."""
    img = np.array(img, dtype=np.float64)
    size = img.shape[0]
    flat = img.ravel(order="F")
    nvalid = len(validx)
    slopes = np.zeros(2 * nvalid)
    inten = np.zeros(nvalid)
    p = np.arange(npix * npix)
    x, y = p % npix, p // npix
    for s in range(nvalid):
        idim = (x + validx[s]) + (y + validy[s]) * size
        inside = idim < flat.size
        v = np.where(inside, flat[np.minimum(idim, flat.size - 1)], 0.0)
        t = np.sort(v)[::-1][nbpix]
        d = np.where(v > t, v - t, 0.0)
        I, Sx, Sy = d[inside].sum(), (d * x)[inside].sum(), (d * y)[inside].sum()
        flat[idim[inside]] = d[inside]
        cx, cy = _finish(I, Sx, Sy, offset, scale)
        ix, iy = _slot(s, nvalid, order)
        slopes[ix] = cx - (0 if ref is None else ref[ix])
        slopes[iy] = cy - (0 if ref is None else ref[iy])
        inten[s] = I
    return slopes, inten, flat.reshape(img.shape, order="F")


def tt_filter(slopes, tip, tilt):
    """Tip/tilt removal, SutraCentroider::apply_TT_filter
    (libsutra/src.cpp/sutra_centroider.cpp:322-338): tip = <s, ref_tip> and
    tilt = <s, ref_tilt> are both taken on the unfiltered slopes, then
    s - tip * ref_tip - tilt * ref_tilt.  Returns (filtered, (tip, tilt))."""
    s = np.asarray(slopes, dtype=np.float64)
    tp = np.asarray(tip, dtype=np.float64)
    tl = np.asarray(tilt, dtype=np.float64)
    a, b = float(s @ tp), float(s @ tl)
    return s - a * tp - b * tl, (a, b)
