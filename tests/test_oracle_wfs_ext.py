"""Pins of the enlarged-grid and kernel-convolution stages of the oracle
(oracle/wfs.py extend_grid / convolve_kernel, oracle/geometry.py hrmap_table):

* hrmap + extend_grid is the frequency-preserving zero padding of the HR
  image (fftshift, pad centred, ifftshift), computed here with numpy's shifts;
* the convolution with a Dirac kernel is the identity, and with the seeing
  kernel of the geometry it equals the circular convolution summed directly
  from its definition; both keep the flux;
* a full frame on the enlarged grid keeps the flux normalisation."""
import numpy as np
import pytest

from oracle import geometry as og
from oracle import wfs as ow


@pytest.mark.parametrize("Nfft,Ntot", [(16, 32), (16, 64), (32, 64), (8, 16)])
def test_hrmap_is_frequency_preserving_zero_padding(Nfft, Ntot):
    rng = np.random.default_rng(Nfft + Ntot)
    hr = rng.random((3, Nfft, Nfft))
    got = ow.extend_grid(hr, og.hrmap_table(Nfft, Ntot), Ntot)
    pad = (Ntot - Nfft) // 2
    c = np.fft.fftshift(hr, axes=(1, 2))
    c = np.pad(c, ((0, 0), (pad, pad), (pad, pad)))
    ref = np.fft.ifftshift(c, axes=(1, 2))
    assert np.array_equal(got, ref)
    assert np.isclose(got.sum(), hr.sum())


def test_dirac_kernel_is_the_identity():
    N = 16
    hr = np.random.default_rng(1).random((2, N, N))
    ft = np.fft.fft2(np.pad([[1.0]], ((0, N - 1), (0, N - 1)))) / (N * N)
    assert np.allclose(ow.convolve_kernel(hr, ft), hr, rtol=0, atol=1e-12)


def test_gaussian_kernel_is_the_circular_convolution():
    N = 16
    rng = np.random.default_rng(2)
    hr = rng.random((2, N, N))
    # the geometry's kernel: a normalised Gaussian centred on (0, 0), as
    # ftkernel = fft2(k) / N^2 (geom_init.py:901-926)
    x = np.arange(N)
    d = np.minimum(x, N - x).astype(np.float64)
    k = np.exp(-(d[:, None] ** 2 + d[None, :] ** 2) / (2 * 1.3 ** 2))
    k /= k.sum()
    ft = np.fft.fft2(k) / (N * N)
    got = ow.convolve_kernel(hr, ft)
    ref = np.zeros_like(hr)
    for a in range(N):
        for b in range(N):
            ref += k[a, b] * np.roll(np.roll(hr, a, axis=1), b, axis=2)
    assert np.allclose(got, ref, rtol=0, atol=1e-12)
    assert np.allclose(got.sum(axis=(1, 2)), hr.sum(axis=(1, 2)))


def test_enlarged_grid_frame_keeps_the_flux_normalisation():
    # 8x8 sub-apertures of 8 phase points, Nfft 16, 3x3 binning to 8x8 pixels:
    # the 24-point window needs the 32-point grid
    g = og.sh_geometry(64, 8, 8, 8, 3, 16, 32, 0.5, 0.12)
    rng = np.random.default_rng(3)
    phase = rng.normal(0, 0.05, (g["npup"], g["npup"]))
    b, inten, img = ow.comp_image(phase, g["mpupil"], 2 * np.pi / 0.5, g["halfxy"], g["istart"],
                                  g["jstart"], g["validsubsx"] * 8, g["validsubsy"] * 8,
                                  g["binmap"], g["fluxPerSub"], 1000.0, 8, 16, 8, 8, Ntot=32,
                                  hrmap=og.hrmap_table(16, 32))
    # the 24 x 24 window holds the whole 16 x 16 image: all the flux is binned
    assert np.allclose(b.sum(axis=(0, 1)), 1000.0 * g["fluxPerSub"], rtol=1e-12)
