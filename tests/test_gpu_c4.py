"""The uneven batch at full size against the float64 oracle: all eight SH WFS
(six 74x74 at pdiam 8 / Nfft 16 / 6x6 px, one 40x40 at Nfft 32 / 8x8 px, one
8x8 at pdiam 32 / Nfft 64 / 16x16 px, each on its own screen and r0) in one
launch per frame through sutra.WfsBatch, then a captured hipGraph of several
frames whose last frame is checked the same way.

Bars (per WFS, relative to the largest value, the reference tests' metric):
FFT FP32 spots / image within 1e-5 and CoG slopes within 1e-5; split-BF16x3
spots within 1e-4 and slopes within 1e-3 px RMS; FP16 slopes within 1e-3 px
RMS.  The per-WFS geometry of every batch member also has its own oracle
parity (tests/test_gpu_parity.py, tests/test_gpu_mfma.py: b74, q2_n64, large
shapes).
"""
import numpy as np
import pytest

from conftest import relative_array_error
from oracle import centroid as oc
from oracle import wfs as ow
from shwfs import configs, system

pytestmark = pytest.mark.gpu

FFT, BF16X3, FP16 = 0, 1, 2


def _oracle(s, phase):
    t = s.tables()
    b, _, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"],
                              t["jstart"], t["validsubsx"], t["validsubsy"], t["binmap"],
                              t["fluxPerSub"], t["nphot"], t["pdiam"], t["Nfft"], t["npix"],
                              t["nxsub"])
    w = s.p_wfs
    sl, _ = oc.cog(img, w._validsubsx, w._validsubsy, w.npix, s.centro.offset, s.centro.scale)
    return b, img, sl


def _check(bs, refs, methods, label):
    for i, s in enumerate(bs.systems):
        b, img, sl = refs[i]
        got_b, got_i, got_s = s.bincube(), s.binimg(), bs.slopes(i)
        eb = relative_array_error(got_b, b)
        ei = relative_array_error(got_i, img)
        es = relative_array_error(got_s, sl)
        rms = float(np.sqrt(np.mean((got_s - sl) ** 2)) / s.p_wfs.pixsize)
        m = methods[i]
        print("%s %s (nvalid %d, method %d): spot %.2e image %.2e slopes %.2e (rms %.2e px)"
              % (label, s.cfg["name"], s.p_wfs._nvalid, m, eb, ei, es, rms), flush=True)
        if m == FFT:
            assert eb < 1e-5 and ei < 1e-5 and es < 1e-5
        elif m == BF16X3:
            assert eb < 1e-4 and ei < 1e-4 and rms < 1e-3
        else:
            assert rms < 1e-3


@pytest.fixture(scope="module")
def c4():
    cfgs = configs.batch()
    bs = system.BatchSystem(cfgs)
    phases = [system.make_screens(c, 1)[0] for c in cfgs]
    bs.set_phases(phases)
    refs = [_oracle(s, ph) for s, ph in zip(bs.systems, phases)]
    assert [s.p_wfs._nvalid for s in bs.systems][:6] == [4268] * 6
    return bs, refs


@pytest.mark.parametrize("method", [FFT, BF16X3, FP16])
def test_eight_wfs_one_launch_match_oracle(c4, method):
    bs, refs = c4
    n = len(bs.systems)
    bs.batch.set_tiles([method] * n, [0] * n)
    bs.batch.launch()
    _check(bs, refs, [method] * n, "one launch")


def test_autotuned_graph_replay_matches_oracle(c4):
    """The FP32-accurate configuration the autotuner picks (FFT / BF16x3 per
    WFS, launch grid, XCD mapping), captured as a graph of 5 frames and
    replayed twice: the tenth frame equals the oracle (noise off, so every
    frame is the same image), and the frame counter advanced by 10."""
    bs, refs = c4
    us = bs.batch.autotune(method_mask=(1 << FFT) | (1 << BF16X3), reps=5)
    methods = bs.batch.methods
    print("autotuned %.2f us/frame: methods %s grid %d remap %d" % (us, methods, bs.batch.grid,
                                                                   bs.batch.remap), flush=True)
    for s in bs.systems:  # clear the outputs of the autotuner's trial frames
        s.wfs.set_bincube(np.zeros(s.bincube().shape, np.float32, order="F"),
                          s.bincube().size)
    bs.batch.set_frame(0)
    bs.batch.capture(5)
    bs.batch.replay(2)
    assert bs.batch.frame == 10
    _check(bs, refs, methods, "graph frame 10")
