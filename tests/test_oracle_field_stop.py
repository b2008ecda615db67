"""Pins of the field-stop oracle (oracle/wfs.py: field_stop, oracle/geometry.py:
field_stop_mask): an all-pass stop returns the pupil field, a stop that keeps
only the zero frequency returns the field's mean everywhere, the transform
pair equals the brute-force DFT on a small grid, and the product's mask
(shwfs/geom.py) equals the oracle's."""
import numpy as np
import pytest

from oracle import geometry as og
from oracle import wfs as ow


def _field(rng, npup):
    phase = rng.standard_normal((npup, npup))
    pupil = (rng.random((npup, npup)) > 0.3).astype(float)
    return phase, pupil


def test_all_pass_stop_is_the_identity():
    rng = np.random.default_rng(1)
    phase, pupil = _field(rng, 12)
    F = ow.field_stop(phase, pupil, 2.0, np.ones((32, 32)))
    want = pupil * np.exp(1j * phase * float(np.float32(2.0)))
    assert np.abs(F - want).max() < 1e-12


def test_zero_frequency_stop_gives_the_mean():
    rng = np.random.default_rng(2)
    phase, pupil = _field(rng, 10)
    n = 16
    m = np.zeros((n, n))
    m[0, 0] = 1.0  # the fftshifted centre
    F = ow.field_stop(phase, pupil, 1.5, m)
    f = pupil * np.exp(1j * phase * float(np.float32(1.5)))
    assert np.allclose(F, f.sum() / n ** 2, rtol=0, atol=1e-13)


def test_transform_pair_equals_the_brute_force_dft():
    rng = np.random.default_rng(3)
    phase, pupil = _field(rng, 5)
    n = 8
    m = (rng.random((n, n)) > 0.4).astype(float)
    f = np.zeros((n, n), complex)
    f[:5, :5] = pupil * np.exp(1j * phase * float(np.float32(0.7)))
    spec = ow.brute_dft2(f) * m
    # inverse = conj(DFT(conj(.))) / n^2
    back = np.conj(ow.brute_dft2(np.conj(spec))) / n ** 2
    assert np.abs(ow.field_stop(phase, pupil, 0.7, m) - back[:5, :5]).max() < 1e-12


@pytest.mark.parametrize("fstop", ["round", "square"])
def test_mask_is_centred_and_symmetric(fstop):
    m = og.field_stop_mask(68, 64, 0.5, 4.0, 1.0, fstop)
    n = m.shape[0]
    assert n == 128
    c = np.fft.ifftshift(m)  # centre between pixels n/2 - 1 and n/2
    assert np.array_equal(c, c[::-1, :]) and np.array_equal(c, c[:, ::-1]) and np.array_equal(c, c.T)
    assert m[0, 0] == 1.0 and 0 < m.sum() < n * n


@pytest.mark.parametrize("fstop", ["round", "square"])
def test_product_mask_equals_the_oracle(fstop):
    from shwfs import configs, geom
    c = configs.get("small")
    p_wfs, p_tel, p_geom = c["p_wfs"], c["p_tel"], c["p_geom"]
    geom.geom_init(p_geom, p_tel)
    geom.init_wfs_geom(p_wfs, c["r0"], p_tel, p_geom, nfft=c["nfft"], nrebin=c["nrebin"])
    p_wfs.fssize, p_wfs.fstop = 1.2, fstop
    got = geom.field_stop_mask(p_wfs, p_tel, p_geom)
    want = og.field_stop_mask(p_geom._mpupil.shape[0], p_geom.pupdiam, p_wfs.Lambda, p_tel.diam,
                              1.2, fstop)
    assert np.array_equal(got, want.astype(np.float32))
    p_wfs.fssize = 0
    assert geom.field_stop_mask(p_wfs, p_tel, p_geom) is None
