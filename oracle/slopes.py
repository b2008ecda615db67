"""Oracle for the geometric slopes of an SH WFS in float64 (test infrastructure only).

Restates SutraWfs::slopes_geom (libsutra/src.cpp/sutra_wfs.cpp:199-245), the
mean phase gradient per sub-aperture in arcsec, for its three types:

* type 0  reduce_phasex / reduce_phasey (libsutra/src.cu/sutra_wfs.cu:868-913):
      slope_x[s] = alpha / n * sum_y (phi(n-1, y) - phi(0, y))
      slope_y[s] = alpha / n * sum_x (phi(x, n-1) - phi(x, 0))
* type 1  derive_phasex / derive_phasey (sutra_wfs.cu:916-1033): per point a
      central difference (one-sided at the window edges, not halved there),
      times the pupil value at the point, summed over the window, then
      / n * alpha / fluxPerSub[s] / 2
* type 2  phase_project (sutra_wfs.cu:1036-1090): the window gathered through
      phasemap, then a dot product with column s (x slope) and s + nvalid
      (y slope) of ttprojmat
with alpha = 0.206265 / subapd * n / (n - 1) (sutra_wfs.cpp:221, 229) and
n = nphase.  phi(x, y) = phase[phasemap[x + y*n, s]] (flat screen index,
x fastest); slopes [2*nvalid], x slopes first.
"""
import numpy as np


def alpha(subapd, nphase):
    """Radians-of-micrometres to arcsec factor of types 0 and 1."""
    return 0.206265 / subapd * nphase / (nphase - 1)


def _windows(phase, phasemap, nphase):
    """phi[s, y, x] of every sub-aperture (float64)."""
    flat = np.asarray(phase, dtype=np.float64).ravel(order="F")
    pm = np.asarray(phasemap).reshape(nphase * nphase, -1, order="F")
    return flat[pm.T].reshape(-1, nphase, nphase)


def slopes_geom(kind, phase, phasemap, nphase, subapd=None, ttprojmat=None, pupil=None,
                flux=None):
    """Geometric slopes of type `kind` (0, 1 or 2)."""
    n = nphase
    w = _windows(phase, phasemap, n)       # [s, y, x]
    nvalid = w.shape[0]
    out = np.zeros(2 * nvalid)
    if kind == 0:
        a = alpha(subapd, n)
        out[:nvalid] = (w[:, :, n - 1] - w[:, :, 0]).sum(axis=1) / n * a
        out[nvalid:] = (w[:, n - 1, :] - w[:, 0, :]).sum(axis=1) / n * a
    elif kind == 1:
        a = alpha(subapd, n)
        m = _windows(pupil, phasemap, n)
        dx = np.empty_like(w)
        dx[:, :, 1:-1] = w[:, :, 2:] - w[:, :, :-2]
        dx[:, :, 0] = w[:, :, 1] - w[:, :, 0]
        dx[:, :, -1] = w[:, :, -1] - w[:, :, -2]
        dy = np.empty_like(w)
        dy[:, 1:-1, :] = w[:, 2:, :] - w[:, :-2, :]
        dy[:, 0, :] = w[:, 1, :] - w[:, 0, :]
        dy[:, -1, :] = w[:, -1, :] - w[:, -2, :]
        fl = np.asarray(flux, dtype=np.float64)
        out[:nvalid] = (dx * m).sum(axis=(1, 2)) / n * a / fl / 2
        out[nvalid:] = (dy * m).sum(axis=(1, 2)) / n / 2 * a / fl
    elif kind == 2:
        tt = np.asarray(ttprojmat, dtype=np.float64).reshape(n * n, 2 * nvalid, order="F")
        v = w.reshape(nvalid, n * n)        # k = x + y*n
        out[:nvalid] = np.einsum("sk,ks->s", v, tt[:, :nvalid])
        out[nvalid:] = np.einsum("sk,ks->s", v, tt[:, nvalid:])
    else:
        raise ValueError("slopes_geom: type must be 0, 1 or 2")
    return out


def ramp_slopes(gx, gy, nphase, subapd):
    """Closed form for a phase ramp phi = gx*x + gy*y (micrometres per phase
    pixel) over fully illuminated windows: the tilt of the wavefront,
    (gx, gy) * nphase / subapd micrometres per metre, in arcsec."""
    return 0.206265 * gx * nphase / subapd, 0.206265 * gy * nphase / subapd
