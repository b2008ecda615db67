"""Seeded von Karman phase screens by the FFT method (synthetic inputs).

phase PSD (rad^2 m^2 at 500 nm):  0.023 r0^(-5/3) (f^2 + 1/L0^2)^(-11/6)
A complex white Gaussian field is coloured by sqrt(PSD)*df on an N x N grid,
transformed back, and the real part is kept (variance = sum PSD df^2); the grid side N is the power
of two >= max(n + 1, L0 / dx).  The central n x n window is
returned in micrometres of optical path (phase * 0.5 um / 2 pi), the unit of
SutraSource's d_screen.  The piston term (f = 0) is zero.
"""
import numpy as np
import scipy.fft


def grid_size(n, dx, L0):
    need = max(n + 1, int(np.ceil(L0 / dx)))
    return 1 << int(np.ceil(np.log2(need)))


def _psd_sqrt(N, dx, r0, L0):
    f = np.fft.fftfreq(N, d=dx)
    f2 = f[:, None] ** 2 + f[None, :] ** 2
    psd = 0.023 * r0 ** (-5.0 / 3.0) * (f2 + 1.0 / L0 ** 2) ** (-11.0 / 6.0)
    psd[0, 0] = 0.0
    df = 1.0 / (N * dx)
    return np.sqrt(psd) * df


def vonkarman(n, dx, r0, L0, seed, N=None):
    """One n x n screen in micrometres (float32, [x, y] with x fast)."""
    N = N or grid_size(n, dx, L0)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    c *= _psd_sqrt(N, dx, r0, L0)
    scr = scipy.fft.ifft2(c, workers=-1, overwrite_x=True).real * (N * N)
    o = (N - n) // 2
    rad = scr[o:o + n, o:o + n]
    return (rad * (0.5 / (2 * np.pi))).astype(np.float32)


def vonkarman_many(n, dx, r0, L0, seeds, N=None):
    return np.stack([vonkarman(n, dx, r0, L0, s, N) for s in seeds])


def structure_function(screens_um, max_lag):
    """Empirical D(r) along both axes, in rad^2 at 500 nm, for lags 1..max_lag."""
    ph = np.asarray(screens_um, dtype=np.float64) * (2 * np.pi / 0.5)
    if ph.ndim == 2:
        ph = ph[None]
    out = []
    for k in range(1, max_lag + 1):
        dxs = ph[:, k:, :] - ph[:, :-k, :]
        dys = ph[:, :, k:] - ph[:, :, :-k]
        out.append(0.5 * (np.mean(dxs ** 2) + np.mean(dys ** 2)))
    return np.array(out)
