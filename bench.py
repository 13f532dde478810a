"""Benchmark of the H.261 encoder core on one MI355X.

Workload (the batched configuration): 256 CIF + 256 QCIF streams, each of a
seed-drawn length uniform in 60..600 frames with a per-stream global motion
in [-15, 15] px/frame, rate control on with a per-stream target uniform in
64..384 kbit/s (VBV = target, initial quantiser 10).  Every frame of every
stream is generated into HBM before the timed region; the timed region codes
the whole job: K frame steps, each one pass of the engine over every active
stream (motion estimation, mode decision, rate control, the fused
macroblock coder, denoise update, exact frame size).

Also measured: the plain single stream (one CIF stream, 300 frames, fixed
QUANT=8) frames/s and per-frame latency p50/p99, the per-kernel device
times behind the roofline, and the oracle (the reference path restated in
C) on the host cores as the CPU baseline.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CIF frames/s on 1 MI355X, batched vs plain, bit-exact; p99 frame latency"
HBM_PEAK_GBS = 8000.0
SAD_PEAK = 157.3e12  # byte-SADs/s: 256 CU x 64 lanes x 2.4 GHz x 4 (v_sad_u8)

# algorithmic bytes per frame (cur + ref luma + 16 B/MB results for motion
# estimation; cur + ref YUV read, recon YUV written, int16 levels + 16 B/MB
# side information for the macroblock coder)
ME_BYTES = {396: 352 * 288 * 2 + 16 * 396, 99: 176 * 144 * 2 + 16 * 99}
MB_BYTES = {396: 3 * 152064 + 396 * (768 + 16), 99: 3 * 38016 + 99 * (768 + 16)}


def coarse_valid_candidates(mb_w, mb_h):
    """valid coarse candidates of a frame (depends on geometry only)"""
    grid = np.array([-14, -11, -8, -5, -2, 0, 2, 5, 8, 11, 14])
    n = 0
    for my in range(mb_h):
        for mx in range(mb_w):
            bx, by = 16 * mx, 16 * my
            lx0, lx1 = max(bx - 15, 0) - bx, min(bx + 15, (mb_w - 1) * 16) - bx
            ly0, ly1 = max(by - 15, 0) - by, min(by + 15, (mb_h - 1) * 16) - by
            n += int(((grid >= lx0) & (grid <= lx1)).sum() * ((grid >= ly0) & (grid <= ly1)).sum())
    return n


# byte-SADs per inter frame: valid coarse candidates + at most 49 refine
# candidates per MB, 256 byte-SADs each (an upper bound for the refine part)
SAD_OPS = {396: (coarse_valid_candidates(22, 18) + 49 * 396) * 256,
           99: (coarse_valid_candidates(11, 9) + 49 * 99) * 256}


def batched_streams(seed=2026):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(512):
        cif = i < 256
        specs.append(dict(size=(352, 288) if cif else (176, 144), length=int(rng.integers(60, 601)),
                          bitrate=int(rng.integers(64, 385)) * 1000, seed=int(rng.integers(1, 2 ** 62)),
                          motion=15))
    return specs


def open_streams(eng, specs, sigma=2):
    ids = []
    for s in specs:
        sid = eng.open(*s["size"], bitrate=s["bitrate"], vbv_size=s["bitrate"], init_quant=10,
                       fixed_quant=s.get("fixed", 0))
        eng.generate_input(sid, s["seed"], s["length"], sigma=sigma, motion=s["motion"])
        ids.append(sid)
    eng.sync()
    return ids


def run_job(eng, ids, specs, steps, compact_every=60):
    """K frame steps (graph replay), closing finished streams and re-packing
    the MB space every compact_every steps.  Returns (seconds, frames, mbs)."""
    lengths = np.array([s["length"] for s in specs])
    nmb = np.array([(s["size"][0] // 16) * (s["size"][1] // 16) for s in specs])
    done = 0
    open_mask = np.ones(len(ids), bool)
    eng.sync()
    t0 = time.perf_counter()
    while done < steps:
        chunk = min(compact_every, steps - done)
        eng.run(chunk)
        done += chunk
        fin = open_mask & (lengths <= done)
        if fin.any() and done < steps:
            eng.sync()
            for i in np.nonzero(fin)[0]:
                eng.close(ids[i])
            open_mask &= ~fin
            eng.compact()
    eng.sync()
    dt = time.perf_counter() - t0
    frames = np.minimum(lengths, steps)
    return dt, int(frames.sum()), int((frames * nmb).sum())


def plain_single_stream(Engine, frames=300, warmup=10):
    """one CIF stream, fixed QUANT 8: throughput (graph replay) and
    per-frame latency (one synchronised step per frame)"""
    eng = Engine(max_streams=1, mbs_per_wave=1)
    sid = eng.open(352, 288, fixed_quant=8)
    eng.generate_input(sid, 1, frames + warmup, sigma=2, motion=7)
    for _ in range(warmup):
        eng.step()
    eng.sync()
    t0 = time.perf_counter()
    eng.run(frames // 2)
    eng.sync()
    dt = time.perf_counter() - t0
    fps = (frames // 2) / dt
    lat = []
    for _ in range(frames - frames // 2):
        t = time.perf_counter()
        eng.step()
        eng.sync()
        lat.append((time.perf_counter() - t) * 1e3)
    eng.destroy()
    lat = np.array(lat)
    return fps, float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


def kernel_profile(Engine, specs, steps=12):
    """per-kernel device time (HIP events on the engine stream around each
    launch) with every stream active"""
    eng = Engine(max_streams=len(specs))
    ids = open_streams(eng, specs)
    eng.step()  # the intra frame (no search) is not representative
    eng.sync()
    eng.profile(True)
    for _ in range(steps):
        eng.step()
    eng.sync()
    t = eng.kernel_times()
    eng.destroy()
    return t


def copy_probe(nbytes):
    """one device-to-device copy of nbytes (the HIP runtime's copy kernel):
    the reference point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE"""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def cpu_baseline(frames=30, workers=None):
    """The oracle (C restatement of the reference path) on the host cores:
    one CIF stream of the plain configuration per worker process."""
    import multiprocessing as mp
    workers = workers or min(16, os.cpu_count() or 1)
    with mp.get_context("fork").Pool(workers) as pool:
        t0 = time.perf_counter()
        res = pool.starmap(_cpu_stream, [(i + 1, frames) for i in range(workers)])
        dt = time.perf_counter() - t0
    total = sum(r[0] for r in res)
    cpu_s = sum(r[1] for r in res)
    return total / dt, workers, dt, cpu_s


def _cpu_stream(seed, frames):
    from h261enc_cuda_amd import content_frame
    from oracle.binding import Oracle
    o = Oracle()
    s = o.stream(352, 288, fixed_quant=8)
    inputs = [content_frame(seed, 352, 288, t, sigma=2, motion=7) for t in range(frames)]
    t0 = time.process_time()
    for f in inputs:
        s.encode(*f)
    return frames, time.process_time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600, help="frame steps timed (600 = the whole job)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain single-stream line")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 256 MiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(streams would shard across GPUs with no exchange)"}))
        sys.exit(2)
    from h261enc_cuda_amd import Engine

    if args.copy_probe:
        copy_probe(256 << 20)

    specs = batched_streams()
    # warmup: a few steps of a small engine (module load, kernels, graph)
    w = Engine(max_streams=4)
    wids = open_streams(w, specs[:2] + specs[256:258])
    for _ in range(args.warmup):
        w.step()
    w.run(args.warmup)
    w.sync()
    w.destroy()

    eng = Engine(max_streams=len(specs))
    ids = open_streams(eng, specs)
    steps = min(args.steps, max(s["length"] for s in specs))
    dt, frames, mbs = run_job(eng, ids, specs, steps)
    eng.destroy()
    cif_equiv = mbs / 396.0
    value = cif_equiv / dt

    prof = kernel_profile(Engine, specs)
    n_cif, n_qcif = 256, 256
    me_ms = prof["me_ms"] / prof["steps"]
    mb_ms = prof["mb_ms"] / prof["steps"]
    me_bytes = n_cif * ME_BYTES[396] + n_qcif * ME_BYTES[99]
    mb_bytes = n_cif * MB_BYTES[396] + n_qcif * MB_BYTES[99]
    sad_ops = n_cif * SAD_OPS[396] + n_qcif * SAD_OPS[99]
    me_gbs = me_bytes / (me_ms * 1e-3) / 1e9
    mb_gbs = mb_bytes / (mb_ms * 1e-3) / 1e9
    sad_rate = sad_ops / (me_ms * 1e-3)
    dominant = "mb_kernel" if mb_ms >= me_ms else "me_kernel"
    if dominant == "mb_kernel":
        roof = {"kernel": "mb_kernel", "bound": "hbm", "achieved": round(mb_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(mb_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_ms": round(mb_ms, 4), "algorithmic_bytes_per_launch": mb_bytes}
    else:
        roof = {"kernel": "me_kernel", "bound": "hbm", "achieved": round(me_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(me_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_ms": round(me_ms, 4), "algorithmic_bytes_per_launch": me_bytes}
    roof["me_kernel"] = {"avg_launch_ms": round(me_ms, 4), "hbm_gbs": round(me_gbs, 1),
                         "sad_rate_tops": round(sad_rate / 1e12, 3), "sad_frac": round(sad_rate / SAD_PEAK, 4)}
    roof["mb_kernel"] = {"avg_launch_ms": round(mb_ms, 4), "hbm_gbs": round(mb_gbs, 1),
                         "hbm_frac": round(mb_gbs / HBM_PEAK_GBS, 4)}
    roof["other_ms_per_step"] = round(prof["other_ms"] / prof["steps"], 4)

    plain = None
    if not args.no_plain:
        fps, p50, p99 = plain_single_stream(Engine)
        plain = {"fps": round(fps, 1), "latency_ms_p50": round(p50, 3), "latency_ms_p99": round(p99, 3)}

    cpu = None
    if not args.no_cpu:
        cfps, cores, cdt, cpu_s = cpu_baseline()
        cpu = {"value": round(cfps, 2), "unit": "CIF frames/s", "cores": cores, "kind": "port",
               "sample": f"{cores} CIF streams x 30 frames (fixed QUANT 8, the plain configuration's content), "
                         f"one per process, {cdt:.1f} s wall, {cpu_s:.1f} s CPU"}

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "CIF frames/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3 / steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "batched: 256 CIF + 256 QCIF streams, lengths U{60..600}, global motion +-15 px, "
                               "rate control 64-384 kbit/s, inputs resident in HBM, whole job",
                   "frames": frames, "cif_equivalent_frames": round(cif_equiv, 1),
                   "frames_per_s": round(frames / dt, 1), "seconds": round(dt, 3),
                   "plain_single_cif": plain,
                   "batched_over_plain": round(value / plain["fps"], 2) if plain else None,
                   "parallelism": "1 GPU, all streams per frame step"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
