"""Batched frames over uneven SH WFS (one launch per frame), hipGraph replay
of successive frames and the tile autotuner, checked against the per-WFS
path (itself checked against the float64 oracle in test_gpu_parity.py)."""
import numpy as np
import pytest

from conftest import relative_array_error
from shwfs import configs, system

pytestmark = pytest.mark.gpu

FFT, BF16X3, FP16 = 0, 1, 2


def _batch_cfgs(n74=2):
    cfgs = configs.batch()
    return cfgs[:n74] + cfgs[6:]


def _single(s, noise=False):
    """Standalone fused frame of one system: bincube, binimg, slopes."""
    import sutra
    sutra.comp_frame(s.wfs, s.centro, s.rtc.d_control[0].d_centroids, noise)
    return s.bincube(), s.binimg(), s.slopes()


@pytest.mark.parametrize("method", [FFT, BF16X3, FP16])
def test_batch_matches_per_wfs_frames(method):
    cfgs = _batch_cfgs()
    bs = system.BatchSystem(cfgs, method=method)
    bs.set_phases([system.make_screens(c, 1)[0] for c in cfgs])
    bs.batch.set_tiles([method] * len(cfgs), [0] * len(cfgs))
    bs.batch.launch()
    got = [(s.bincube(), s.binimg(), bs.slopes(i)) for i, s in enumerate(bs.systems)]
    for i, s in enumerate(bs.systems):
        ref = _single(s)
        errs = [relative_array_error(g, r) for g, r in zip(got[i], ref)]
        print("%s method %d: spot %.1e image %.1e slopes %.1e" % (cfgs[i]["name"], method, *errs),
              flush=True)
        if method == FFT:
            assert max(errs) == 0.0
        else:
            # the batch runs four sub-apertures per wave with one shared tail;
            # spots and images are the same computation, and so are the float
            # moments up to 64 pixels a spot; larger spots (16x16 px) sum their
            # float moments in another order: float rounding
            npix2 = cfgs[i]["p_wfs"].npix ** 2
            assert max(errs[:2]) <= 1e-6
            assert errs[2] <= (1e-6 if npix2 <= 64 else 5e-6)


@pytest.mark.parametrize("method", [BF16X3, FP16])
@pytest.mark.parametrize("centroider", ["cog", "tcog"])
def test_same_kind_batch_matches_per_wfs_frames(method, centroider):
    """Three WFS of one geometry (own screens and outputs) run the same-kind
    kernel (four sub-apertures per wave, one tail for four); the per-WFS path
    runs one sub-aperture per wave.  Both add the spot and the moments in the
    same order, so the results are equal bit for bit."""
    cfgs = [configs.get("noisy") for _ in range(3)]
    for k, c in enumerate(cfgs):
        c["p_wfs"].noise = -1.0
        c["seeds"] = [11 + k]
        c["p_centroider"].type = centroider
        c["p_centroider"].thresh = 0.05
    bs = system.BatchSystem(cfgs, method=method)
    bs.set_phases([system.make_screens(c, 1)[0] for c in cfgs])
    bs.batch.set_tiles([method] * len(cfgs), [0] * len(cfgs))
    bs.batch.launch()
    got = [(s.bincube(), s.binimg(), bs.slopes(i)) for i, s in enumerate(bs.systems)]
    for i, s in enumerate(bs.systems):
        ref = _single(s)
        errs = [relative_array_error(g, r) for g, r in zip(got[i], ref)]
        print("wfs %d method %d %s: spot %.1e image %.1e slopes %.1e"
              % (i, method, centroider, *errs), flush=True)
        assert errs == [0.0, 0.0, 0.0]


def test_batch_mixed_methods_and_persistent_grid():
    cfgs = _batch_cfgs()
    bs = system.BatchSystem(cfgs)
    bs.set_phases([system.make_screens(c, 1)[0] for c in cfgs])
    methods = [FFT, BF16X3, FP16, FFT][:len(cfgs)]
    splits = [1, 0, 0, 0][:len(cfgs)]
    for grid, remap in ((0, 0), (64, 1), (300, 0)):
        bs.batch.set_tiles(methods, splits, grid, remap)
        bs.batch.launch()
        got = [s.bincube() for s in bs.systems]
        for i, s in enumerate(bs.systems):
            s.wfs.set_method(methods[i])
            ref = _single(s)[0]
            err = relative_array_error(got[i], ref)
            # WFS 0 runs the FFT with one row transform per row (split 1); the
            # per-WFS path splits it in two (pdiam <= nfft / 2): float rounding only
            assert err <= 1e-6, (grid, remap, i, err)


def test_graph_replay_advances_the_noise_frame():
    cfgs = [configs.get("small"), configs.get("small")]
    cfgs[1]["seeds"] = [2]
    for c in cfgs:
        c["p_wfs"].noise = 0.5
        c["nphot"] = 500.0
    bs = system.BatchSystem(cfgs, noise=True)
    bs.set_phases([system.make_screens(c, 1)[0] for c in cfgs])
    for k, s in enumerate(bs.systems):
        s.wfs.set_noise(0.5, 77 + k)
    # recreate so the batch takes the new noise seeds (call counter 0)
    bs.batch = bs.sutra.WfsBatch(bs.systems[0].ctx, [s.wfs for s in bs.systems],
                                 [s.centro for s in bs.systems], True)
    bs.batch.set_frame(0)
    bs.batch.capture(3)
    bs.batch.replay(1)
    assert bs.batch.frame == 3
    got = [s.bincube() for s in bs.systems]
    for k, s in enumerate(bs.systems):
        s.wfs.set_noise(0.5, 77 + k)
        for _ in range(3):
            s.wfs.comp_image(True)  # third frame: Philox call 2
        err = relative_array_error(got[k], s.bincube())
        print("wfs %d frame 2: %.1e" % (k, err), flush=True)
        assert err == 0.0
    # a second replay continues with frames 3..5
    bs.batch.replay(1)
    assert bs.batch.frame == 6


def test_autotune_keeps_a_valid_fast_configuration():
    cfgs = _batch_cfgs()
    bs = system.BatchSystem(cfgs)
    bs.set_phases([system.make_screens(c, 1)[0] for c in cfgs])
    us = bs.batch.autotune(method_mask=(1 << FFT) | (1 << BF16X3), reps=10)
    print("autotuned: %.2f us/frame, methods %s, splits %s, grid %d, remap %d"
          % (us, bs.batch.methods, bs.batch.splits, bs.batch.grid, bs.batch.remap), flush=True)
    assert us > 0
    assert all(m in (FFT, BF16X3) for m in bs.batch.methods)
    bs.batch.launch()
    for i, s in enumerate(bs.systems):
        got = s.bincube()
        s.wfs.set_method(bs.batch.methods[i])
        ref = _single(s)[0]
        assert relative_array_error(got, ref) <= 1e-6
