"""Pins the float64 comp_image oracle with closed forms and textbook checks."""
import numpy as np
import pytest

from oracle import centroid as oc
from oracle import geometry as og
from oracle import wfs as ow


def test_fft_convention_matches_brute_force_dft():
    rng = np.random.default_rng(3)
    E = np.zeros((8, 8), dtype=np.complex128)
    E[:4, :4] = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    F = np.fft.fft2(E)
    assert np.allclose(F, ow.brute_dft2(E), atol=1e-12)


def _system(pdiam=8, nfft=16, npix=8, nrebin=2, nxsub=4, cobs=0.0, fracsub=0.5):
    pup = pdiam * nxsub
    g = og.sh_geometry(pup, nxsub, pdiam, npix, nrebin, nfft, nfft, fracsub, cobs)
    return g


def _run(g, phase, pdiam=8, nfft=16, npix=8, nxsub=4, nphot=1000.0):
    return ow.comp_image(phase, g["mpupil"], np.float32(2 * np.pi / 0.5), g["halfxy"],
                         g["istart"], g["jstart"], g["validsubsx"] * npix,
                         g["validsubsy"] * npix, g["binmap"], g["fluxPerSub"], nphot, pdiam,
                         nfft, npix, nxsub)


def test_parseval_and_flux_normalisation():
    g = _system()
    rng = np.random.default_rng(1)
    phase = rng.standard_normal((g["npup"], g["npup"])).astype(np.float32) * 0.1
    E = ow.fields(phase, g["mpupil"], np.float32(4 * np.pi), g["halfxy"], g["istart"],
                  g["jstart"], 8, 16)
    hr = ow.hr_images(E)
    assert np.allclose(hr.sum(axis=(1, 2)), 256 * (np.abs(E) ** 2).sum(axis=(1, 2)), rtol=1e-12)
    b, inten, img = _run(g, phase)
    # after normalisation each spot holds nphot * fluxPerSub photons (full-window binning)
    assert np.allclose(b.sum(axis=(0, 1)), 1000.0 * g["fluxPerSub"], rtol=1e-12)
    # the image holds exactly the spots
    assert np.isclose(img.sum(), b.sum(), rtol=1e-12)


def test_flat_wavefront_spot_is_centred():
    g = _system(cobs=0.0, fracsub=1.0)
    phase = np.zeros((g["npup"], g["npup"]), dtype=np.float32)
    b, inten, img = _run(g, phase)
    sl, inten = oc.cog(img, g["validsubsx"] * 8, g["validsubsy"] * 8, 8, 3.5, 1.0)
    # centred on npix/2 - 0.5 up to the float32 rounding of the half-pixel
    # offsets (a ~1e-7 relative tilt) and the +1e-6 of the CoG denominator
    assert np.abs(sl).max() < 1e-7


@pytest.mark.parametrize("k0,axis", [(1, 0), (2, 0), (-3, 1), (5, 1)])
def test_tilt_shifts_spot_by_fourier_shift_theorem(k0, axis):
    # scale*phase = 2*pi*k0*x/Nfft multiplies each field by exp(2i pi k0 x/N):
    # the HR spot is the untilted one circularly shifted by +k0 along that axis
    # (a constant phase per sub-aperture from the global ramp does not matter).
    pdiam, nfft = 8, 16
    g = _system(pdiam=pdiam, nfft=nfft, nrebin=2, npix=8, cobs=0.12, fracsub=0.5)
    n = g["npup"]
    scale = np.float32(2 * np.pi / 0.5)
    rng = np.random.default_rng(7)
    base = rng.standard_normal((n, n)) * 0.05
    coord = np.arange(n)[:, None] * np.ones((1, n)) if axis == 0 else np.ones((n, 1)) * np.arange(n)[None, :]
    tilt = 2 * np.pi * k0 * coord / nfft / float(scale)
    off = np.zeros_like(g["halfxy"])
    E0 = ow.fields(base, g["mpupil"], scale, off, g["istart"], g["jstart"], pdiam, nfft)
    E1 = ow.fields(base + tilt, g["mpupil"], scale, off, g["istart"], g["jstart"], pdiam, nfft)
    h0, h1 = ow.hr_images(E0), ow.hr_images(E1)
    assert np.allclose(h1, np.roll(h0, k0, axis=1 + axis), rtol=1e-9, atol=1e-9 * h0.max())


def test_tilt_moves_cog_linearly_for_small_tilts():
    # in the linear regime the CoG slope follows the tilt: tilt of a quarter
    # FFT pixel -> slope of 0.25/nrebin detector pixels (within 5%)
    pdiam, nfft, nrebin = 8, 16, 2
    g = _system(pdiam=pdiam, nfft=nfft, nrebin=nrebin, npix=8, cobs=0.0, fracsub=1.0)
    n = g["npup"]
    scale = np.float32(2 * np.pi / 0.5)
    x = np.arange(n)[:, None] * np.ones((1, n))
    k0 = 0.25
    b, inten, img = _run(g, (2 * np.pi * k0 * x / nfft / float(scale)).astype(np.float32))
    sl, _ = oc.cog(img, g["validsubsx"] * 8, g["validsubsy"] * 8, 8, 3.5, 1.0)
    nv = len(g["validsubsx"])
    assert np.allclose(sl[:nv], k0 / nrebin, rtol=0.05)
    assert np.abs(sl[nv:]).max() < 1e-6


def test_binning_gathers_the_listed_pixels():
    rng = np.random.default_rng(0)
    hr = rng.random((3, 16, 16))
    bm = og.binmap_table(16, 8, 2)
    b = ow.bin_images(hr, bm, 8)
    flat = hr.transpose(0, 2, 1).reshape(3, 256)
    for s in range(3):
        for p in range(64):
            assert np.isclose(b[p % 8, p // 8, s], flat[s, bm[:, p]].sum())
    assert np.isclose(b.sum(), hr.sum())   # full-window binning keeps every photon
