"""Noise path: the device Philox samplers against the oracle's on the same
counters, a noisy frame of the 40x40 WFS (1000 photons, 0.5 e- readout)
against the oracle's noise applied to the device's noise-free spots, and the
weighted CoG of the noisy image."""
import ctypes as C

import numpy as np
import pytest

import capi
from conftest import relative_array_error
from oracle import centroid as oc
from oracle import philox as op
from shwfs import configs, system

pytestmark = pytest.mark.gpu


def test_prng_streams_match_oracle():
    import torch
    lib = capi.load()
    rng = np.random.default_rng(0)
    lam = np.concatenate([rng.uniform(0, 10, 3000), rng.uniform(10, 1500, 3000)]).astype(np.float32)
    d = torch.from_numpy(lam.copy()).cuda()
    assert lib.shwfs_prng(capi.ptr(d), lam.size, b"P", 0.0, 0.0, 77, 5, 0, None) == 0
    got = d.cpu().numpy()
    ref = op.poisson(lam.astype(np.float64), 0, 5, 0, 77)
    mism = np.mean(got != ref)
    print("Poisson mismatches: %d of %d" % (np.sum(got != ref), lam.size), flush=True)
    assert mism <= 1e-3 and np.all(np.abs(got - ref)[got != ref] <= 1)
    z = torch.zeros(5000, device="cuda")
    assert lib.shwfs_prng(capi.ptr(z), 5000, b"N", 0.5, 1.0, 9, 2, 1, None) == 0
    zr = 0.5 * op.normal(np.arange(5000), 2, 1, 9)
    np.testing.assert_allclose(z.cpu().numpy(), zr, rtol=1e-5, atol=1e-6)


def test_poisson_float_decisions_equal_the_float64_sampler():
    """The device draws Poisson variates with float32 decisions checked
    against their rounding bounds (a float64 redraw where a decision is too
    close to call), so every draw must equal the float64 sampler's: around
    the inversion / PTRS switch at 10, at low and high flux, and above 1e6
    (always the float64 draw)."""
    import torch
    lib = capi.load()
    rng = np.random.default_rng(11)
    lam = np.concatenate([
        rng.uniform(0.0, 10.0, 4000), 10.0 + rng.uniform(-0.05, 0.05, 2000),
        rng.uniform(10.0, 300.0, 4000), rng.uniform(300.0, 5000.0, 2000),
        rng.uniform(5e4, 2e6, 400), [0.0, -1.0, 1e-30, 9.999999, 10.0, 1e6, 3e6],
    ]).astype(np.float32)
    d = torch.from_numpy(lam.copy()).cuda()
    assert lib.shwfs_prng(capi.ptr(d), lam.size, b"P", 0.0, 0.0, 2024, 3, 0, None) == 0
    got = d.cpu().numpy()
    ref = op.poisson(lam.astype(np.float64), 0, 3, 0, 2024)
    bad = np.nonzero(got != ref)[0]
    print("Poisson draws differing from the float64 sampler: %d of %d" % (bad.size, lam.size),
          flush=True)
    assert bad.size == 0, list(zip(lam[bad[:5]], got[bad[:5]], ref[bad[:5]]))


def _noisy_system(method=0):
    c = configs.get("noisy")
    s = system.ShSystem(c, method=method)
    s.set_phase(system.make_screens(c, 1)[0])
    return c, s


@pytest.mark.parametrize("method", [0, 1, 2])
def test_noisy_frame_matches_oracle_noise_on_device_spots(method):
    """Per method (the readout background drawn by extra workgroups of the
    spot launch on the pixels outside the spot windows): the noisy spots, the
    background draws, and the spot windows of the image equal to the spots."""
    c, s = _noisy_system(method)
    w = s.wfs
    seed = 4321
    w.set_noise(0.5, seed)           # call counter 0
    s.comp_image(noise=True)         # call 0: Poisson + readout on spots, background on binimg
    noisy = s.bincube()
    img = s.binimg()
    s.comp_image(noise=False)        # call 1: the noise-free spots
    clean = s.bincube().astype(np.float64)
    npix2 = w.npix * w.npix
    flat = clean.ravel(order="F")
    pois = op.poisson(flat, 0, 0, 0, seed)
    ref = (pois.astype(np.float32) + np.float32(0.5) *
           op.normal(np.arange(flat.size), 0, 1, seed).astype(np.float32))
    got = noisy.ravel(order="F")
    mism = np.abs(got - ref) > 1e-4 * np.maximum(1.0, np.abs(ref))
    print("noisy pixels differing from the oracle sampler: %d of %d" % (mism.sum(), got.size),
          flush=True)
    assert mism.mean() < 1e-3
    # background of the detector image outside the spots: 0.5 * N(0, 1) on domain 2
    p = s.p_wfs
    mask = np.zeros(img.shape, bool)
    for x, y in zip(p._validsubsx, p._validsubsy):
        mask[x:x + p.npix, y:y + p.npix] = True
    idx = np.nonzero(~mask.ravel(order="F"))[0]
    bg = 0.5 * op.normal(idx, 0, 2, seed)
    np.testing.assert_allclose(img.ravel(order="F")[idx], bg, rtol=1e-5, atol=1e-6)
    assert npix2 * w.nvalid == got.size
    cube = noisy.reshape(p.npix, p.npix, -1, order="F")
    for k, (x, y) in enumerate(zip(p._validsubsx, p._validsubsy)):
        np.testing.assert_array_equal(img[x:x + p.npix, y:y + p.npix], cube[:, :, k])


def test_wcog_of_the_noisy_image_matches_oracle():
    c, s = _noisy_system()
    s.wfs.set_noise(0.5, 99)
    s.comp_image(noise=True)
    img = s.binimg()
    s.do_centroids()
    p = s.p_wfs
    wts = np.array(s.centro.d_weights).reshape(p.npix, p.npix, -1, order="F")
    ref, _ = oc.wcog(img, p._validsubsx, p._validsubsy, p.npix, s.centro.offset, s.centro.scale,
                     wts, threshold=s.centro.threshold)
    got = s.slopes()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("wcog slope err %.2e" % err, flush=True)
    assert err < 1e-5
