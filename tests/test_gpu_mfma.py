"""MFMA path: lane-layout probe through the C ABI and DFT-method accuracy.

Accuracy bar of the matrix-core methods (stated, measured against the float64
oracle on the 80x80 workload): split-BF16x3 spots within 1e-4 of the maximum
and slopes within 1e-3 px RMS; FP16 with per-subap scale slopes within
1e-3 px RMS.
"""
import ctypes as C

import numpy as np
import pytest

import capi
from oracle import centroid as oc
from oracle import wfs as ow
from shwfs import configs, system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [capi.METHOD_DFT_FP16, capi.METHOD_DFT_BF16X3])
def test_mfma_probe_matches_matmul(method):
    import torch
    lib = capi.load()
    rng = np.random.default_rng(method)
    A = rng.integers(-8, 9, size=(16, 32)).astype(np.float32)
    B = rng.integers(-8, 9, size=(32, 16)).astype(np.float32)
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(B).cuda()
    dD = torch.zeros((16, 16), dtype=torch.float32, device="cuda")
    assert lib.shwfs_mfma_probe(capi.ptr(dA), capi.ptr(dB), capi.ptr(dD), method, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dD.cpu().numpy(), A @ B)


# Shapes that take the other operand and binning variants of the MFMA
# kernels: 32-point windows (two 16-row operand blocks), windows smaller than
# the operand tile (masked), N = 64, one-pixel and 2x2 register binning, and
# the general binmap gather (block origin not on the rebin grid); b74: the
# 74x74 WFS of the uneven batch (8-point windows in a 16-point transform).
SHAPES = {
    "b74": dict(diam=8.0, nxsub=74, pdiam=8, npix=6, nfft=16, nrebin=2),
    "q2_n64": dict(diam=8.0, nxsub=8, pdiam=32, npix=16, nfft=64, nrebin=4),
    "q2_masked": dict(diam=8.0, nxsub=20, pdiam=24, npix=8, nfft=32, nrebin=4),
    "rebin1": dict(diam=4.0, nxsub=10, pdiam=16, npix=12, nfft=16, nrebin=1),
    "gather": dict(diam=8.0, nxsub=24, pdiam=16, npix=7, nfft=32, nrebin=4),
}


def _config(name):
    if name in SHAPES:
        return configs.sh_config(**SHAPES[name], r0=0.16, seeds=(3,), name=name)
    return configs.get(name)


def _run(name, method):
    c = _config(name)
    c["p_centroider"].type = "cog"
    s = system.ShSystem(c, method=method)
    phase = system.make_screens(c, 1)[0]
    s.set_phase(phase)
    s.comp_image(noise=False)
    s.do_centroids()
    t = s.tables()
    b, inten, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                                  t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                                  t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                                  t["nxsub"])
    w = c["p_wfs"]
    ref, _ = oc.cog(img, w._validsubsx, w._validsubsy, w.npix, s.centro.offset, s.centro.scale)
    spot = float(np.abs(s.bincube() - b).max() / np.abs(b).max())
    flux = float(np.abs(s.intensities() - inten).max() / np.abs(inten).max())
    rms_px = float(np.sqrt(np.mean((s.slopes() - ref) ** 2)) / w.pixsize)
    print("%s method %d: spot err %.2e, flux err %.2e, slope rms err %.2e px"
          % (name, method, spot, flux, rms_px), flush=True)
    return spot, flux, rms_px


@pytest.mark.parametrize("name", ["small", "large"])
def test_bf16x3_accuracy(name):
    spot, flux, rms = _run(name, capi.METHOD_DFT_BF16X3)
    assert spot < 1e-4 and flux < 1e-4 and rms < 1e-3


@pytest.mark.parametrize("name", ["small", "large"])
def test_fp16_accuracy(name):
    spot, flux, rms = _run(name, capi.METHOD_DFT_FP16)
    assert rms < 1e-3 and flux < 1e-2


@pytest.mark.parametrize("name", sorted(SHAPES))
@pytest.mark.parametrize("method", [capi.METHOD_DFT_BF16X3, capi.METHOD_DFT_FP16])
def test_mfma_variants_accuracy(name, method):
    spot, flux, rms = _run(name, method)
    if method == capi.METHOD_DFT_BF16X3:
        assert spot < 1e-4 and flux < 1e-4 and rms < 1e-3
    else:
        assert rms < 1e-3 and flux < 1e-2
