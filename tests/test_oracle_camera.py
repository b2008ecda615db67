"""The camera-image oracle pinned by known answers (CPU).

The reference holds no camera vectors; these cases fix the formula's order of
operations (flat multiplies the image before dark is added), the clamp at
max_flux_per_pix, the conversion toward zero and its saturation, and the
calibration's gather and order (dark subtracted before flat multiplies).
Powers of two keep every intermediate exact, so the expected values are exact.
"""
import numpy as np

from oracle import camera


def test_identity_when_scales_match():
    # dark 0, flat 1, max_flux_per_pix = max_pix_value = 4096: integer pixels
    # below 4096 come back unchanged
    x = np.arange(0, 4096, 7, dtype=np.float32)
    cam = camera.digitalize(x, np.zeros_like(x), np.ones_like(x), 4096, 4096)
    assert cam.dtype == np.uint16
    np.testing.assert_array_equal(cam, x.astype(np.uint16))


def test_flat_multiplies_before_dark_adds():
    x = np.array([0, 1, 5, 100], np.float32)
    flat = np.full_like(x, 2.0)
    dark = np.full_like(x, 3.0)
    cam = camera.digitalize(x, dark, flat, 4096, 4096)
    np.testing.assert_array_equal(cam, (2 * x + 3).astype(np.uint16))  # not 2 * (x + 3)


def test_clamp_conversion_and_saturation():
    mpv = 1000
    x = np.array([0.75, 1.5, 511.999, 512.0, 600.0, 1e9, -5.0, np.nan], np.float32)
    cam = camera.digitalize(x, np.zeros_like(x), np.ones_like(x), 512, mpv)
    # 0.75/512*1000 = 1.46 -> 1; 1.5 -> 2.93 -> 2; 511.999 -> 999.998 -> 999;
    # at or above max_flux_per_pix the pixel saturates at max_pix_value;
    # negative and NaN values convert to 0
    np.testing.assert_array_equal(cam, [1, 2, 999, 1000, 1000, 1000, 0, 0])
    # a max_pix_value beyond the uint16 range saturates at 65535
    big = camera.digitalize(np.array([10.0], np.float32), [0.0], [1.0], 10, 70000)
    assert big[0] == 65535


def test_calibration_gathers_then_subtracts_then_scales():
    raw = np.arange(8, dtype=np.uint16) * 10
    lut = np.arange(8)[::-1]
    dark = np.ones(8, np.float32)
    flat = np.full(8, 2.0, np.float32)
    cal = camera.calibrate(raw, dark, flat, lut)
    np.testing.assert_array_equal(cal, 2.0 * (raw[::-1].astype(np.float32) - 1.0))


def test_calibrate_inverts_digitalize_on_exact_scales():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 4095, 500).astype(np.float32)
    cam = camera.digitalize(x, np.zeros_like(x), np.ones_like(x), 4096, 4096)
    cal = camera.calibrate(cam, np.zeros(500), np.ones(500), np.arange(500))
    np.testing.assert_array_equal(cal, np.trunc(x))
