"""LGS set-up of an SH WFS (shesha/shesha/init/lgs_init.py:204-343,
prep_lgs_prof): the sodium profile, the per-sub-aperture distance and
azimuth to the launch telescope, the beam, handed to the WFS's SutraLGS
(lgs_init), which builds the elongated spot kernels on the GPU."""
import numpy as np


def gauss_sodium_profile(h0=90000.0, sigma=4000.0, dh=23.984375, n=4965):
    """A one-Gaussian sodium layer: altitudes 0, dh, ... (m) and the density
    at each.  The reference reads this profile from its data file
    allProfileNa_withAltitude_1Gaussian.npy (ProfType.GAUSS1: a Gaussian at
    90 km of 4 km standard deviation sampled every 23.984375 m, 4965 samples);
    it is generated here from those parameters (the scale is irrelevant: the
    spots are normalised)."""
    alt = (np.arange(n) * dh).astype(np.float32)
    prof = np.exp(-((alt.astype(np.float64) - h0) ** 2) / (2 * sigma ** 2)).astype(np.float32)
    return alt, prof


def prep_lgs_prof(p_wfs, nsensors, p_tel, sensors, imat=0):
    """Sodium profile p_wfs._profna at altitudes p_wfs._altna (default: the
    one-Gaussian layer), sub-aperture distances and azimuths to (lltx, llty),
    Gaussian beam of FWHM beamsize centred on the image (as center="image"),
    then sensors.d_wfs[nsensors].d_gs.d_lgs.lgs_init."""
    if p_wfs._altna is None or p_wfs._profna is None:
        p_wfs._altna, p_wfs._profna = gauss_sodium_profile()
    alt = np.asarray(p_wfs._altna, dtype=np.float32)
    prof = np.asarray(p_wfs._profna, dtype=np.float32)
    p_wfs._prof1d = prof
    p_wfs._profcum = np.zeros(prof.size + 1, dtype=np.float32)
    p_wfs._profcum[1:] = prof.cumsum()
    subapdiam = p_tel.diam / p_wfs.nxsub
    if p_wfs.nxsub > 1:
        xsubs = np.linspace((subapdiam - p_tel.diam) / 2, (p_tel.diam - subapdiam) / 2,
                            p_wfs.nxsub).astype(np.float32)
    else:
        xsubs = np.zeros(1, dtype=np.float32)
    ysubs = xsubs.copy()
    hG = np.sum(alt * prof) / np.sum(prof)  # centre of gravity of the layer
    x = (np.arange(p_wfs._Ntot).astype(np.float32) - p_wfs._Ntot / 2) * p_wfs._qpixsize
    dh = alt[1] - alt[0]
    if p_wfs.nxsub > 1:
        ix, iy = p_wfs._validsubsx // p_wfs.npix, p_wfs._validsubsy // p_wfs.npix
        doff = np.sqrt((xsubs[ix] - p_wfs.lltx) ** 2 + (ysubs[iy] - p_wfs.llty) ** 2)
        azimuth = np.arctan2(ysubs[iy] - p_wfs.llty, xsubs[ix] - p_wfs.lltx)
    else:
        doff = np.sqrt((xsubs - p_wfs.lltx) ** 2 + (ysubs - p_wfs.llty) ** 2)
        azimuth = np.arctan2(ysubs - p_wfs.llty, xsubs - p_wfs.lltx)
    if imat > 0:
        doff = doff * 0.0
    w = p_wfs.beamsize / 2.35482005
    if w == 0:
        raise ValueError("prep_lgs_prof: beamsize must be positive")
    g = np.exp(-((x + p_wfs._qpixsize / 2) ** 2) / (2 * w ** 2.0))
    p_wfs._ftbeam = np.fft.fft(g, axis=0).astype(np.complex64)
    p_wfs._beam = g.astype(np.float32)
    p_wfs._azimuth = azimuth.astype(np.float32)
    p_wfs._doffaxis = doff.astype(np.float32)
    sensors.d_wfs[nsensors].d_gs.d_lgs.lgs_init(
        p_wfs._prof1d.size, float(hG), float(alt[0]), float(dh), float(p_wfs._qpixsize),
        p_wfs._doffaxis, p_wfs._prof1d, p_wfs._profcum, p_wfs._beam, p_wfs._ftbeam,
        p_wfs._azimuth)
