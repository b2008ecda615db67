"""Pins of the LGS oracle (oracle/lgs.py): the profile interpolation is
exact on constant and linear profiles and clamps at both ends, the beam
convolution equals the circular convolution summed from its definition, the
kernel rotation by 0 is the identity up to the reference's clamp of the last
two samples, a Dirac kernel leaves the HR images unchanged (up to the n^2 of
the unnormalised transforms), and the product's
sodium layer has the one-Gaussian parameters."""
import numpy as np
import pytest

from oracle import lgs as ol


def test_interpolation_is_exact_on_linear_profiles_and_clamps():
    hmax = 50
    lin = 3.0 + 0.25 * np.arange(hmax + 1)
    args = dict(npix=16, pixsize=0.1, norm=0.02, shiftx=100.0, deltah=10.0, hmax=hmax)
    for x in np.arange(0.0, 16.0, 0.5):
        t = ((x - 8.5) * 0.1 / 0.02 + 100.0) / 10.0
        got = ol.interp_transf(lin, x, **args)
        want = lin[0] if t < 0 else (3.0 + 0.25 * t if t < hmax - 2 else lin[hmax - 1])
        assert np.isclose(got, want, rtol=0, atol=1e-12)
        assert ol.interp_transf(np.full(hmax + 1, 2.5), x, **args) == 2.5
    assert ol.interp_transf(lin, 0.0, 16, 1.0, 1e-3, -1e6, 1.0, hmax) == lin[0]
    assert ol.interp_transf(lin, 0.0, 16, 1.0, 1e-3, 1e6, 1.0, hmax) == lin[hmax - 1]


def test_beam_convolution_is_the_circular_convolution():
    rng = np.random.default_rng(4)
    n, nv = 16, 3
    prof = np.exp(-0.5 * ((np.arange(400) - 200) / 30.0) ** 2)
    cum = np.concatenate([[0.0], np.cumsum(prof)])
    doff = np.array([0.5, 2.0, 5.0])
    args = (prof, cum, prof.size, doff, 90000.0, 0.05, 85000.0, 25.0)
    p = ol.profiles(*args, n)
    g = rng.random(n)
    got = ol.prof2d(*args, np.fft.fft(g), n)
    want = np.zeros((n, nv))
    for s in range(nv):
        for x in range(n):
            want[x, s] = n * sum(p[m, s] * g[(x - m) % n] for m in range(n))
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()
    assert np.allclose(ol.prof2d(*args, np.ones(n), n), n * p, rtol=0, atol=1e-9)


def test_zero_azimuth_rotation_is_the_identity_up_to_the_edge_clamp():
    rng = np.random.default_rng(5)
    n = 16
    p2d = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    beam = rng.random(n)
    lk, ft = ol.kernels(p2d, beam, np.zeros(2), n)
    idx = np.minimum(np.arange(n), n - 3)  # samples n-2 and n-1 both read n-1
    idx[n - 2:] = n - 1
    for s in range(2):
        img = np.abs(p2d[(np.arange(n) + n // 2) % n, s].real)[:, None] * beam[None, :]
        assert np.allclose(lk[:, :, s], img[np.ix_(idx, idx)], rtol=0, atol=1e-12)
    assert np.allclose(ft, np.fft.fft2(lk, axes=(0, 1)))


def test_dirac_kernel_leaves_the_images_unchanged():
    rng = np.random.default_rng(6)
    hr = rng.random((3, 16, 16))
    assert np.allclose(ol.convolve(hr, np.ones((16, 16, 3))), 16 * 16 * hr)


def test_product_sodium_layer_has_the_one_gaussian_parameters():
    from shwfs import lgs_init
    alt, prof = lgs_init.gauss_sodium_profile()
    p = prof.astype(np.float64)
    hg = (alt * p).sum() / p.sum()
    sigma = np.sqrt(((alt - hg) ** 2 * p).sum() / p.sum())
    assert alt.size == 4965 and alt[1] - alt[0] == 23.984375
    assert abs(hg - 90000.0) < 1.0 and abs(sigma - 4000.0) < 1.0
