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


def _run(name, method):
    c = configs.get(name)
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
