"""Benchmark of the SH WFS hot path on one MI355X.

A step is one frame of the 80x80 Shack-Hartmann WFS (D = 16 m, 4964 valid
sub-apertures, 16x16 phase points zero-padded to a 32x32 transform, 4x4
binning to 8x8-pixel spots, noise off) followed by its slopes (thresholded CoG
with float64 moments, threshold 0.1 * max): phase screen -> spots -> detector
image -> slopes, inputs resident in HBM.  Frames are issued back to back by a
native loop; wall clock around K frames, synchronised on both sides.

Prints one JSON line: the metric and its value (frames/s of the best method,
one frame per launch), every method's rate and error against the float64
oracle, the roofline of the spot kernel (algorithmic bytes over the measured
kernel time, HBM traffic from the committed rocprofv3 summary), the matrix-core
use, the batched and uneven-batch forms, and the float64 CPU oracle's rate.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "compass-saxoplus_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "SH WFS frames/s (spots+slopes) at 80x80 subaps vs plain FP32; slope err vs FP64"
HBM_PEAK_GBS = 8000.0
METHOD_NAMES = {0: "fft_fp32", 1: "dft_bf16x3", 2: "dft_fp16"}


def algorithmic_bytes(w, nvalid):
    """Bytes one frame must move at minimum: phase + pupil windows, tip-tilt
    offsets, spots, detector image, per-subap flux/intensity/slopes/indices."""
    pd, npix, nx = w._pdiam, w.npix, w.nxsub
    return (nvalid * pd * pd * 8 + pd * pd * 4 + nvalid * npix * npix * 4
            + (npix * nx) ** 2 * 4 + nvalid * 32)


def profiled_traffic(method_name):
    """HBM bytes per launch of the spot kernel of this method on the 80x80
    frame, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary."""
    path = os.path.join(ROOT, "profiles", "spot_kernels_80x80_traffic.txt")
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 4 and f[0] == method_name:
                return int(f[3])
    except OSError:
        pass
    return None


MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 / fp16


def mfma_utilisation(w, results, batched):
    """Matrix-core use of the DFT methods on the 80x80 frame: useful flops
    8 (pdiam^2 N + N^2 pdiam) per sub-aperture (the two complex products of the
    pruned DFT) and issued flops (16x16x32 tiles: Q (N/8) Q for stage A,
    (N/16) (N/8) Q for stage B, x3 for split-BF16x3), against the dense peak."""
    n, p, nv = w._Nfft, w._pdiam, w._nvalid
    q = 1 if p <= 16 else 2
    tiles = q * (n // 8) * q + (n // 16) * (n // 8) * q
    useful = 8.0 * (p * p * n + n * n * p) * nv
    out = {}
    for m, name, mult in ((2, "dft_fp16", 1), (1, "dft_bf16x3", 3)):
        issued = tiles * 16384.0 * mult * nv
        row = {}
        if m in results:
            t = results[m]["kernel_us"] * 1e-6
            row["single_frame"] = dict(useful_tflops=round(useful / t / 1e12, 1),
                                       issued_tflops=round(issued / t / 1e12, 1))
        for k, per in batched.items():
            if name in per:
                t = per[name]["us_per_frame"] * 1e-6
                row["batched_x%d" % k] = dict(useful_tflops=round(useful / t / 1e12, 1),
                                             issued_tflops=round(issued / t / 1e12, 1))
        for v in row.values():
            v["frac_useful"] = round(v["useful_tflops"] / MFMA_PEAK_TFLOPS, 4)
            v["frac_issued"] = round(v["issued_tflops"] / MFMA_PEAK_TFLOPS, 4)
        if row:
            out[name] = row
    return dict(peak_tflops=MFMA_PEAK_TFLOPS, useful_flops_per_frame=useful,
                tiles_per_subap=tiles, methods=out) if out else None


def sync():
    import carma
    carma.deviceSync()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def measure(method, cfg_name, steps, warmup, phase, thr, wg=0):
    import sutra
    from shwfs import configs, system
    c = configs.get(cfg_name)
    s = system.ShSystem(c, method=method, fp64_moments=True, threshold=thr)
    if wg:
        s.wfs.set_wg_size(wg)
    s.set_phase(phase)
    slopes = s.rtc.d_control[0].d_centroids
    sutra.run_frames(s.wfs, s.centro, slopes, warmup, False, True)
    sync()
    t0 = time.perf_counter()
    gpu_ms = sutra.run_frames(s.wfs, s.centro, slopes, steps, False, True)
    sync()
    wall = time.perf_counter() - t0
    out = dict(frames_per_s=steps / wall, ms_per_frame=1e3 * wall / steps,
               kernel_ms=gpu_ms / steps, slopes=np.array(slopes), bincube=s.bincube())
    return out, s


def measure_batched(method, k, phase, thr, graphs=10, frames_per_graph=20):
    """k independent 80x80 frames (own screens and outputs) per launch through
    sutra.WfsBatch, replayed from a captured hipGraph."""
    import copy
    from shwfs import configs, system
    cfg0 = system.geometry(configs.get("large"))
    systems = []
    for i in range(k):
        s = system.ShSystem(copy.deepcopy(cfg0), method=method, fp64_moments=True, threshold=thr)
        s.set_phase(np.roll(phase, 37 * i, axis=0))
        systems.append(s)
    import sutra
    b = sutra.WfsBatch(systems[0].ctx, [s.wfs for s in systems], [s.centro for s in systems], False)
    b.set_tiles([method] * k, [0] * k)
    b.autotune(method_mask=1 << method, reps=10)  # launch grid, XCD mapping (and FFT split)
    b.capture(frames_per_graph)
    b.replay(2)
    sync()
    t0 = time.perf_counter()
    ms = b.replay(graphs)
    sync()
    wall = time.perf_counter() - t0
    frames = graphs * frames_per_graph * k
    return dict(frames_per_s=frames / wall, gpu_frames_per_s=frames / (ms * 1e-3),
                us_per_frame=1e3 * ms / frames, frames_per_launch=k, launches=graphs * frames_per_graph,
                grid=b.grid, remap=b.remap, splits=b.splits[:1])


def measure_uneven(method_mask=3, graphs=10, frames_per_graph=100):
    """The eight uneven SH WFS of the batching workload, one launch per frame,
    tiles from the autotuner over the methods of method_mask (3: the
    FP32-accurate FFT and BF16x3; 7: FP16 too), 1000 hipGraph frames."""
    from shwfs import configs, system
    cfgs = configs.batch()
    bs = system.BatchSystem(cfgs)
    bs.set_phases([system.make_screens(c, 1)[0] for c in cfgs])
    us = bs.batch.autotune(method_mask=method_mask, reps=20)
    bs.batch.capture(frames_per_graph)
    bs.batch.replay(1)
    sync()
    t0 = time.perf_counter()
    ms = bs.batch.replay(graphs)
    sync()
    wall = time.perf_counter() - t0
    frames = graphs * frames_per_graph
    return dict(frames_per_s=frames / wall, gpu_frames_per_s=frames / (ms * 1e-3),
                us_per_frame=1e3 * ms / frames, autotune_us=us, methods=bs.batch.methods,
                splits=bs.batch.splits, grid=bs.batch.grid, remap=bs.batch.remap,
                nvalid=[int(s.wfs.nvalid) for s in bs.systems])


def oracle_frame(s, phase, thr):
    from oracle import centroid as oc
    from oracle import wfs as ow
    t = s.tables()
    b, _, img = ow.comp_image(phase, t["pupil"], t["scale"], t["offsets"], t["istart"], t["jstart"],
                              t["validsubsx"], t["validsubsy"], t["binmap"], t["fluxPerSub"],
                              t["nphot"], t["pdiam"], t["Nfft"], t["npix"], t["nxsub"])
    w = s.p_wfs
    sl, _, _ = oc.tcog(img, w._validsubsx, w._validsubsy, w.npix, s.centro.offset,
                       s.centro.scale, thr)
    return b, sl


def cpu_baseline(s, phase, thr, budget_s=12.0):
    n, t0 = 0, time.perf_counter()
    while True:
        oracle_frame(s, phase, thr)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 50:
            break
    el = time.perf_counter() - t0
    return dict(value=n / el, unit="frames/s", cores=1, kind="port",
                sample="%d frames of the 80x80 workload by the float64 NumPy oracle "
                       "(fields, pocketfft, binning, normalisation, TCoG), %.1f s" % (n, el))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--methods", default="0,1,2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--wg", type=int, default=0, help="spot-kernel workgroup size (0 = default)")
    ap.add_argument("--batched", type=str, default="8,16",
                    help="comma-separated frames of the 80x80 WFS per launch in the batched "
                         "form (0 = skip)")
    ap.add_argument("--no-uneven", action="store_true", help="skip the eight-WFS batch")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="run a device copy of this many MiB first (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (the path has no multi-GPU stage); "
              "--gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: single-GPU build, WORLD_SIZE > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    from shwfs import configs, system
    import sutra  # noqa: F401

    if a.copy_calib_mib > 0:
        # a plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
        # counters give the scale of the spot kernel's counters
        import carma
        ctx = carma.context.get_instance()
        n = a.copy_calib_mib * (1 << 20) // 4
        src = carma.obj_float(ctx, np.ones(n, dtype=np.float32))
        dst = carma.obj_float(ctx, np.zeros(n, dtype=np.float32))
        for _ in range(3):
            dst.copy_from(src, n)
        sync()
        del src, dst

    cfg = system.geometry(configs.get("large"))
    w = cfg["p_wfs"]
    phase = system.make_screens(cfg, 1)[0]

    # threshold = 0.1 * max of the noise-free image, fixed for the run
    probe = system.ShSystem(configs.get("large"))
    probe.set_phase(phase)
    probe.comp_image(noise=False)
    thr = 0.1 * float(probe.binimg().max())
    ref_b, ref_s = oracle_frame(probe, phase, thr)
    del probe

    results = {}
    for m in [int(x) for x in a.methods.split(",") if x != ""]:
        try:
            r, s = measure(m, "large", a.steps, a.warmup, phase, thr, a.wg)
        except RuntimeError as e:
            print("method %s unavailable: %s" % (METHOD_NAMES[m], e), file=sys.stderr)
            continue
        err_px = float(np.sqrt(np.mean((r["slopes"] - ref_s) ** 2)) / w.pixsize)
        spot_err = float(np.abs(r["bincube"] - ref_b).max() / np.abs(ref_b).max())
        results[m] = dict(frames_per_s=r["frames_per_s"], ms_per_frame=r["ms_per_frame"],
                          kernel_us=1e3 * r["kernel_ms"], slope_rms_err_px=err_px,
                          spot_rel_err=spot_err, system=s)
        print("[bench] %s: %.0f frames/s, %.2f us/frame (gpu events), slope rms err %.2e px, "
              "spot err %.2e" % (METHOD_NAMES[m], r["frames_per_s"], 1e3 * r["kernel_ms"], err_px,
                                  spot_err), file=sys.stderr, flush=True)
    if not results:
        raise RuntimeError("no method ran")
    best = max(results, key=lambda k: results[k]["frames_per_s"])
    rb = results[best]
    nbytes = algorithmic_bytes(w, w._nvalid)
    achieved = nbytes / (rb["kernel_us"] * 1e-6) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=profiled_traffic(METHOD_NAMES[best]),
                traffic_source="profiles/spot_kernels_80x80_traffic.txt (rocprofv3 FETCH_SIZE x 2.0 "
                               "+ WRITE_SIZE x 1.0 per launch)",
                kernel="spot+TCoG fused kernel (%s)" % METHOD_NAMES[best],
                bytes_per_launch=nbytes, avg_launch_us=round(rb["kernel_us"], 3))
    batched = {}  # frames per launch -> method -> result
    for k in [int(x) for x in a.batched.split(",") if x.strip() and int(x) > 0]:
        batched[k] = {}
        for m in (0, 2, 1):
            try:
                r = measure_batched(m, k, phase, thr)
            except RuntimeError as e:
                print("batched %s unavailable: %s" % (METHOD_NAMES[m], e), file=sys.stderr)
                continue
            batched[k][METHOD_NAMES[m]] = {kk: (round(v, 3) if isinstance(v, float) else v)
                                           for kk, v in r.items()}
            print("[bench] batched x%d %s: %.0f frames/s (%.2f us/frame)"
                  % (k, METHOD_NAMES[m], r["frames_per_s"], r["us_per_frame"]),
                  file=sys.stderr, flush=True)
    uneven = None
    if not a.no_uneven:
        uneven = {}
        for mask, key in ((3, "fp32_accurate"), (7, "any_method")):
            r = measure_uneven(mask)
            uneven[key] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()}
            print("[bench] uneven 8-WFS batch (%s): %.0f frames/s (%.2f us/frame), methods %s"
                  % (key, r["frames_per_s"], r["us_per_frame"], r["methods"]),
                  file=sys.stderr, flush=True)
    mfma = mfma_utilisation(w, results, batched)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(rb["system"], phase, thr)
    plain = results.get(0)
    line = dict(
        metric=METRIC, value=round(rb["frames_per_s"], 1), unit="frames/s", n_gpus=1,
        steps=a.steps, warmup=a.warmup, ms_per_step=round(rb["ms_per_frame"], 6),
        higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
        data="synthetic",
        config=dict(workload="80x80 SH, D=16 m, 4964 valid subaps, 16x16 phase px, Nfft 32, "
                             "4x4 binning to 8x8 px, noise off, TCoG thr 0.1*max fp64 moments",
                    method=METHOD_NAMES[best], nvalid=w._nvalid, pdiam=w._pdiam, nfft=w._Nfft,
                    npix=w.npix, screen="von Karman r0=0.16 m L0=25 m seed 1"),
        slope_rms_err_px=rb["slope_rms_err_px"], spot_rel_err=rb["spot_rel_err"],
        methods={METHOD_NAMES[k]: dict(frames_per_s=round(v["frames_per_s"], 1),
                                       kernel_us=round(v["kernel_us"], 3),
                                       slope_rms_err_px=v["slope_rms_err_px"],
                                       spot_rel_err=v["spot_rel_err"])
                 for k, v in results.items()},
        speedup_vs_plain_fp32=(round(rb["frames_per_s"] / plain["frames_per_s"], 3)
                               if plain else None),
        roofline=roof, cpu_baseline=cpu, mfma=mfma,
        batched_80x80=dict(
            workload="K independent 80x80 frames (own screen and outputs) per launch, "
                     "hipGraph replay, for K in frames_per_launch",
            frames_per_launch={str(k): v for k, v in batched.items()},
            speedup_best_vs_plain_fp32_single=(
                round(max(x["frames_per_s"] for v in batched.values() for x in v.values())
                      / plain["frames_per_s"], 3)
                if batched and plain and any(batched.values()) else None)) if batched else None,
        uneven_batch=dict(workload="6x 74x74 (Nfft 16) + 40x40 (Nfft 32) + 8x8 (Nfft 64) SH, "
                                   "one launch per frame, autotuned tiles, 1000 hipGraph frames; "
                                   "fp32_accurate: methods FFT/BF16x3, any_method: FP16 too",
                          **uneven) if uneven else None,
    )
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
