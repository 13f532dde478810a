"""Benchmark of the H.261 encoder core on one MI355X.

Workload (the batched configuration): 256 CIF + 256 QCIF streams, each of a
seed-drawn length uniform in 60..600 frames with a per-stream global motion
in [-15, 15] px/frame, rate control on with a per-stream target uniform in
64..384 kbit/s (VBV = target, initial quantiser 10).  Every frame of every
stream is generated into HBM before the timed region.  After W untimed
frame steps the timed region codes the rest of the job: K frame steps
(default 600 - W), each one pass of the engine over every active stream
(motion estimation, mode decision, rate control, the fused macroblock coder,
denoise update, exact frame size).  A second, untimed pass repeats the same
W + K steps with HIP events around each kernel launch: the per-kernel
average durations behind the roofline.

Also measured: the plain single stream (one CIF stream, 300 frames, fixed
QUANT=8) frames/s and per-frame latency p50/p99; the hard-motion batch
(1024 streams, per-MB local motion, scene cuts); the serving loop (Poisson
arrivals, continuous batching, per-frame latency); and the oracle (the
reference path restated in C) on the host cores as the CPU baseline.

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

from h261enc_cuda_amd.workloads import batched_streams, content_kwargs, hard_streams, open_streams, run_job  # noqa: E402

METRIC = "CIF frames/s on 1 MI355X, batched vs plain, bit-exact; p99 frame latency"
HBM_PEAK_GBS = 8000.0
SAD_PEAK = 157.3e12  # byte-SADs/s: 256 CU x 64 lanes x 2.4 GHz x 4 (v_sad_u8)

# algorithmic bytes per frame.  Motion estimation: read the current and the
# reference luma, write 44 B/MB (the ME result, 20 B, and the coding mode,
# 24 B).  Macroblock coder: read the current and reference YUV and the 24 B
# mode, write the reconstructed YUV, the int8 levels (384 B) and the 8 B side
# information per MB.
ME_BYTES = {396: 352 * 288 * 2 + 44 * 396, 99: 176 * 144 * 2 + 44 * 99}
MB_BYTES = {396: 3 * 152064 + 396 * (24 + 384 + 8), 99: 3 * 38016 + 99 * (24 + 384 + 8)}


# HBM bytes per macroblock actually moved by each kernel with the XCD remap
# of motion estimation and coder: FETCH_SIZE and WRITE_SIZE per dispatch from
# two rocprofv3 --pmc passes over `bench.py --steps 8 --no-cpu --no-plain
# --no-extra --copy-probe --shape 4,3,R,T` (the event pass's last 8 steps: all
# 512 streams active, 126,720 MBs per dispatch; 4 MBs per coder wave), each
# scaled by what the same counter reads for the 256 MiB device copy of
# --copy-probe, divided by the dispatch's MB count
# (profiles/batched_512_streams_hbm_traffic.md).  Motion estimation depends on
# its run and tile (R, T): chosen per launch size (0, 0; runs of 2 MBs in one
# tile at these launches) 80.3 MB read + 7.7 MB written, runs of 4 in tiles of
# 2: 74.2 + 9.3 MB.  The coder 131.8 + 148.6 MB (at 6 waves per SIMD, with
# more MBs in flight per XCD; 107.0 + 126.7 MB at 5), the bitstream writer
# (not remapped) 20.1 + 6.5 MB.  Other launch shapes: not measured.
TRAFFIC_PER_MB = {("me_kernel", True, (0, 0)): (80.30e6 + 7.66e6) / 126720,
                  ("me_kernel", True, (4, 2)): (74.22e6 + 9.28e6) / 126720,
                  ("mb_kernel", True, None): (131.84e6 + 148.56e6) / 126720,
                  ("bitstream_kernel", False, None): (20.10e6 + 6.47e6) / 126720}
REMAP_BIT = {"me_kernel": 1, "mb_kernel": 2, "bitstream_kernel": 4}


def traffic_per_launch(kernel, mbs_per_launch, tuned):
    """HBM traffic of an average launch: measured bytes per MB (for the
    kernel's XCD remap setting and, for motion estimation, its run and tile)
    times the launch's MBs"""
    shape = (tuned["me_mbs_per_wave"], tuned["me_tile"]) if kernel == "me_kernel" else None
    t = TRAFFIC_PER_MB.get((kernel, bool(tuned["xcd_remap"] & REMAP_BIT[kernel]), shape))
    return None if t is None else int(t * mbs_per_launch)


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
# candidates per MB, 256 byte-SADs each (an upper bound for the refine part;
# used only when the exact count below is missing)
SAD_OPS = {396: (coarse_valid_candidates(22, 18) + 49 * 396) * 256,
           99: (coarse_valid_candidates(11, 9) + 49 * 99) * 256}

# the exact candidates the reference's two-stage search evaluates in every
# frame step of the batched job (valid coarse + valid refine candidates,
# counted by the oracle over every stream; tests/golden/make_sad_counts.py)
SAD_COUNTS = os.path.join(ROOT, "tests", "golden", "batched_512_sad_candidates.json")


def exact_sad_per_step():
    """byte-SADs of each frame step of the batched job, or None"""
    try:
        with open(SAD_COUNTS) as f:
            d = json.load(f)
    except OSError:
        return None
    return [256 * c for c in d["candidates_per_step"]]


def algorithmic(specs, t0, t1, sad_steps=None):
    """algorithmic HBM bytes of the ME and MB kernels and byte-SADs of frame
    steps t0..t1-1 (intra frames: no reference read, no search); byte-SADs
    from sad_steps (exact, per step) when given, else the upper bound"""
    me = mb = sad = mbs = 0
    for s in specs:
        w, h = s["size"]
        nmb, fsz = w * h // 256, w * h * 3 // 2
        n = max(0, min(s["length"], t1) - t0)
        intra = 1 if (t0 == 0 and n > 0) else 0
        inter = n - intra
        me += inter * ME_BYTES[nmb] + intra * (w * h + 44 * nmb)
        mb += inter * MB_BYTES[nmb] + intra * (2 * fsz + nmb * (24 + 384 + 8))
        sad += inter * SAD_OPS[nmb]
        mbs += n * nmb
    if sad_steps is not None:
        sad = sum(sad_steps[t0:t1])
    return me, mb, sad, mbs


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


def hard_motion(Engine, steps=120, mbs_per_wave=0):
    """the hard-motion batch over its first `steps` frame steps (the first
    two untimed)"""
    specs = hard_streams()
    eng = Engine(max_streams=len(specs), mbs_per_wave=mbs_per_wave)
    ids = open_streams(eng, specs, frames=steps)
    eng.step()  # untimed: the intra frame, then one replay that captures the graph
    eng.run(1)
    dt, frames, mbs = run_job(eng, ids, specs, steps - 2, start=2, compact_every=steps)
    eng.destroy()
    return {"streams": len(specs), "steps": steps - 2, "frames": frames, "seconds": round(dt, 3),
            "frames_per_s": round(frames / dt, 1), "cif_frames_per_s": round(mbs / 396.0 / dt, 1)}


def host_fed(Engine, steps=20, warmup=3):
    """The batched streams with their input in host memory: each step uploads
    every stream's frame from one pinned host buffer (h261g_engine_put_frames:
    one copy across PCIe and one kernel that places the frames, nothing
    waits), codes it, and fetches every bitstream back to host memory
    (h261g_engine_fetch_bitstreams, collected one step later).  PCIe is
    inside the timed region.  Two frames per stream alternate (the search
    cost does not depend on the content); all 512 streams are active in every
    step."""
    import torch
    from h261enc_cuda_amd import content_frame
    specs = batched_streams()
    eng = Engine(max_streams=len(specs))
    ids = [eng.open(*s["size"], bitrate=s["bitrate"], vbv_size=s["bitrate"], init_quant=10) for s in specs]
    sets = []
    for t in range(2):
        buf, offs = Engine.pack_frames([content_frame(s["seed"], *s["size"], t, **content_kwargs(s)) for s in specs])
        pin = torch.empty(buf.size, dtype=torch.uint8, pin_memory=True)
        pbuf = pin.numpy()
        pbuf[...] = buf
        sets.append((pin, pbuf, offs))
    up_bytes = sets[0][1].size

    def one_step(t):
        _, pbuf, offs = sets[t % 2]
        eng.put_frames(ids, pbuf, offs)
        eng.run(1)
        eng.fetch_bitstreams(ids)  # waits for the previous step's fetch

    for t in range(warmup):
        one_step(t)
    eng.sync()
    eng.fetch_wait()
    t0 = time.perf_counter()
    for t in range(warmup, warmup + steps):
        one_step(t)
    down = eng.fetch_wait()
    dt = time.perf_counter() - t0
    eng.destroy()
    cif = steps * sum((s["size"][0] // 16) * (s["size"][1] // 16) for s in specs) / 396.0
    return {"streams": len(specs), "steps": steps, "seconds": round(dt, 3),
            "cif_frames_per_s": round(cif / dt, 1), "ms_per_step": round(dt * 1e3 / steps, 4),
            "upload_mb_per_step": round(up_bytes / 1e6, 2),
            "upload_gb_s": round(up_bytes * steps / dt / 1e9, 1),
            "last_step_bitstream_kb": round(down / 1e3, 1),
            "input": "one pinned host buffer per step (put_frames), every bitstream fetched to host memory"}


def serving(rate=2000.0, duration=0.5):
    """Poisson arrivals (60% CIF / 40% QCIF, lognormal lengths, 64-1024
    kbit/s, local-motion content), admission by HBM budget, continuous
    batching; every step's bitstreams delivered to host memory"""
    from h261enc_cuda_amd.serving import ServingLoop, poisson_requests
    reqs = poisson_requests(rate=rate, duration=duration)
    loop = ServingLoop(hbm_budget=96 << 30, max_streams=4096)
    st = loop.run(reqs, deliver=True)
    loop.close()
    return {"rate_req_s": rate, "arrival_window_s": duration, "requests": st.requests, "frames": st.frames,
            "latency_ends": "bitstream in host memory", "bitstream_mb_delivered": round(st.delivered_bytes / 1e6, 2),
            "seconds": round(st.seconds, 3), "frames_per_s": round(st.frames_per_s, 1),
            "cif_frames_per_s": round(st.cif_frames_per_s, 1),
            "frame_latency_ms_p50": round(st.frame_latency_ms_p50, 3),
            "frame_latency_ms_p99": round(st.frame_latency_ms_p99, 3),
            "request_latency_ms_p50": round(st.request_latency_ms_p50, 1),
            "request_latency_ms_p99": round(st.request_latency_ms_p99, 1),
            "input_wait_ms_p50": round(st.input_wait_ms_p50, 2), "input_wait_ms_p99": round(st.input_wait_ms_p99, 2),
            "max_concurrent_streams": st.max_concurrent, "steps": st.steps}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(frames=30, runs=3, max_workers=16):
    """The oracle (the C restatement of the reference path) on the host
    cores, in the plain CPU configuration: one CIF stream of `frames` frames
    (fixed QUANT 8, global motion +-7, noise sigma 2) per worker process,
    each pinned to its own core of this process's affinity set (at most
    max_workers: the CPU share of a one-GPU machine), distinct seeds, all started together
    behind a barrier after generating their input.  Median of `runs` runs of
    aggregate frames/s (total frames / span from the first start to the last
    end) and of per-core frames/s."""
    import multiprocessing as mp
    cores = sorted(os.sched_getaffinity(0))[:max_workers]
    ctx = mp.get_context("spawn")  # fresh interpreters: nothing of this process's GPU state
    agg, per_core, cpu_s = [], [], 0.0
    for r in range(runs):
        barrier = ctx.Barrier(len(cores))
        q = ctx.Queue()
        procs = [ctx.Process(target=_cpu_stream, args=(c, 1000 * r + i + 1, frames, barrier, q))
                 for i, c in enumerate(cores)]
        for p in procs:
            p.start()
        res = [q.get(timeout=600) for _ in procs]
        for p in procs:
            p.join()
        t0 = min(x[1] for x in res)
        t1 = max(x[2] for x in res)
        agg.append(sum(x[0] for x in res) / (t1 - t0))
        per_core.append(float(np.median([x[0] / (x[2] - x[1]) for x in res])))
        cpu_s += sum(x[3] for x in res)
    return {"value": round(float(np.median(agg)), 2), "per_core": round(float(np.median(per_core)), 2),
            "cores": len(cores), "runs": [round(a, 2) for a in agg], "cpu_s": round(cpu_s, 1)}


def _cpu_stream(core, seed, frames, barrier, q):
    os.sched_setaffinity(0, {core})
    from h261enc_cuda_amd import content_frame
    from oracle.binding import Oracle
    s = Oracle().stream(352, 288, fixed_quant=8)
    inputs = [content_frame(seed, 352, 288, t, sigma=2, motion=7) for t in range(frames)]
    barrier.wait()
    c0, t0 = time.process_time(), time.perf_counter()
    for f in inputs:
        s.encode(*f)
    t1 = time.perf_counter()
    q.put((frames, t0, t1, time.process_time() - c0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="frame steps timed after the warmup (default: the rest of the job, 600 - W)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain single-stream line")
    ap.add_argument("--no-extra", action="store_true", help="skip the hard-motion and serving lines")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 256 MiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--chain-ladder", type=int, default=0, choices=(0, 1, 2),
                    help="two-chain ladder (H261G_OPT_CHAIN_LADDER): 0 by size, 1 strict, 2 one step of slack")
    ap.add_argument("--fused-tails", type=int, default=3, choices=(0, 1, 2, 3),
                    help="rate-control begin (H261G_OPT_FUSED_TAILS): 1 in the tail of motion estimation, "
                         "2 as its own launch, 3 by size (default)")
    ap.add_argument("--compact-every", type=int, default=30,
                    help="steps between closing finished streams and re-packing the MB space")
    ap.add_argument("--shape", default=None,
                    help="MBS_PER_WAVE,XCD_REMAP[,ME_MBS_PER_WAVE[,ME_TILE[,ME_VARIANT]]]: a fixed launch shape instead of the "
                         "autotuner's (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus above 1 is not supported "
                                   "(streams would shard across GPUs with no exchange)"}))
        sys.exit(2)
    from h261enc_cuda_amd import Engine, device_info
    from h261enc_cuda_amd import _native as nat
    dev = device_info()

    if args.copy_probe:
        copy_probe(256 << 20)

    specs = batched_streams()
    job = max(s["length"] for s in specs)
    W = max(0, min(args.warmup, job - 1))
    K = job - W if args.steps is None else max(1, min(args.steps, job - W))

    # pass A (the timed region): W untimed steps of the job, the last one a
    # graph replay so capture happens before the clock starts, then K steps
    # launch shape (coder MBs per wave, XCD remap) tuned on a sample of this
    # workload before anything is timed
    from h261enc_cuda_amd import autotune
    if args.shape:
        v = [int(x) for x in args.shape.split(",")]
        tuned = {"mbs_per_wave": v[0], "xcd_remap": v[1], "me_mbs_per_wave": v[2] if len(v) > 2 else 0,
                 "me_tile": v[3] if len(v) > 3 else 0, "me_variant": v[4] if len(v) > 4 else 0, "fixed": True}
    else:
        tuned = autotune.tune([s["size"] for s in specs])

    eng = Engine(max_streams=len(specs))
    autotune.apply(eng, tuned)
    eng.config(nat.OPT_CHAIN_LADDER, args.chain_ladder)
    eng.config(nat.OPT_FUSED_TAILS, args.fused_tails)
    ids = open_streams(eng, specs)
    for _ in range(W - 1):
        eng.step()
    if W:
        eng.run(1)
    eng.sync()
    dt, frames, mbs = run_job(eng, ids, specs, K, start=W, compact_every=args.compact_every)
    eng.destroy()
    cif_equiv = mbs / 396.0
    value = cif_equiv / dt
    shortest = min(s["length"] for s in specs)
    timed_window = (f"frame steps {W}..{W + K - 1} of the job's {job}"
                    + (f" (all {len(specs)} streams active: the shortest has {shortest} frames)"
                       if W + K <= shortest else " (streams finish and are re-packed during it)"))

    # pass B: the same W + K steps again, each kernel bracketed by HIP events
    # on the engine stream; the averages are over the same launches as pass A.
    # Each kernel runs alone here: the bitstream writer at the end of its own
    # step rather than beside the next step's motion estimation, so no event
    # span covers two kernels
    eng = Engine(max_streams=len(specs))
    autotune.apply(eng, tuned)
    eng.config(nat.OPT_CHAIN_LADDER, args.chain_ladder)
    eng.config(nat.OPT_FUSED_TAILS, args.fused_tails)
    eng.config(nat.OPT_OVERLAP_WRITER, 0)
    ids = open_streams(eng, specs)
    for _ in range(W):
        eng.step()
    eng.sync()
    eng.profile(True)
    first = {}

    def after_chunk(done):
        if not first:
            first.update(eng.kernel_times(), done=done)

    run_job(eng, ids, specs, K, start=W, compact_every=args.compact_every, after_chunk=after_chunk)
    prof = eng.kernel_times()
    eng.destroy()

    # pass C: the same job without device bitstreams -- levels, side
    # information, exact frame sizes and reconstructions ready (the device
    # layer's boundary; the bitstream is then written on the host when read)
    eng = Engine(max_streams=len(specs))
    eng.config(nat.OPT_DEVICE_BITSTREAM, 0)
    autotune.apply(eng, tuned)
    eng.config(nat.OPT_CHAIN_LADDER, args.chain_ladder)
    eng.config(nat.OPT_FUSED_TAILS, args.fused_tails)
    ids = open_streams(eng, specs)
    for _ in range(W - 1):
        eng.step()
    if W:
        eng.run(1)
    eng.sync()
    dt_l, _, mbs_l = run_job(eng, ids, specs, K, start=W, compact_every=args.compact_every)
    eng.destroy()
    levels_fps = mbs_l / 396.0 / dt_l

    sad_steps = exact_sad_per_step()

    def kernel_figures(t, t0, t1):
        n = t["steps"]
        me_b, mb_b, sad, mbs = algorithmic(specs, t0, t1, sad_steps)
        me_ms, mb_ms = t["me_ms"] / n, t["mb_ms"] / n
        return dict(launches=n, me_ms=me_ms, mb_ms=mb_ms, me_bytes=me_b / n, mb_bytes=mb_b / n, mbs=mbs / n,
                    me_gbs=me_b / n / (me_ms * 1e-3) / 1e9, mb_gbs=mb_b / n / (mb_ms * 1e-3) / 1e9,
                    sad=sad / n, sad_rate=sad / n / (me_ms * 1e-3), other_ms=t["other_ms"] / n,
                    bits_ms=t["bits_ms"] / n)

    job_k = kernel_figures(prof, W, W + K)
    full_k = kernel_figures(first, W, first["done"])
    dominant = "mb_kernel" if job_k["mb_ms"] >= job_k["me_ms"] else "me_kernel"
    key = "mb" if dominant == "mb_kernel" else "me"
    hbm = {"achieved": round(job_k[key + "_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(job_k[key + "_gbs"] / HBM_PEAK_GBS, 4)}
    # the binding roof: T_roof = max(bytes / HBM bandwidth, byte-SADs / v_sad_u8
    # rate) for motion estimation, bytes / HBM bandwidth for the coder
    t_hbm = job_k[key + "_bytes"] / (HBM_PEAK_GBS * 1e9)
    t_sad = job_k["sad"] / SAD_PEAK if dominant == "me_kernel" else 0.0
    if t_sad > t_hbm:
        achieved = job_k["sad_rate"] / 1e12
        roof = {"kernel": dominant, "bound": "sad", "achieved": round(achieved, 3), "peak": SAD_PEAK / 1e12,
                "unit": "T byte-SAD/s", "frac": round(achieved * 1e12 / SAD_PEAK, 4)}
    else:
        roof = {"kernel": dominant, "bound": "hbm", **hbm}
    roof.update({"traffic": traffic_per_launch(dominant, job_k["mbs"], tuned),
                 "avg_launch_ms": round(job_k[key + "_ms"], 4),
                 "algorithmic_bytes_per_launch": int(job_k[key + "_bytes"]),
                 "algorithmic_byte_sads_per_launch": int(job_k["sad"]) if dominant == "me_kernel" else None,
                 "sad_count": ("exact: the reference's valid candidates per step (tests/golden/"
                               "batched_512_sad_candidates.json)" if sad_steps else "upper bound (49 refine per MB)"),
                 "t_roof_ms": round(max(t_hbm, t_sad) * 1e3, 4), "hbm": hbm, "launches": job_k["launches"]})
    for name, f in (("job_average", job_k), ("full_load", full_k)):
        roof[name] = {"launches": f["launches"],
                      "mbs_per_launch": round(f["mbs"], 1),
                      "me_kernel": {"avg_launch_ms": round(f["me_ms"], 4), "hbm_gbs": round(f["me_gbs"], 1),
                                    "traffic": traffic_per_launch("me_kernel", f["mbs"], tuned),
                                    "sad_rate_tops": round(f["sad_rate"] / 1e12, 3),
                                    "sad_frac": round(f["sad_rate"] / SAD_PEAK, 4)},
                      "mb_kernel": {"avg_launch_ms": round(f["mb_ms"], 4), "hbm_gbs": round(f["mb_gbs"], 1),
                                    "traffic": traffic_per_launch("mb_kernel", f["mbs"], tuned),
                                    "hbm_frac": round(f["mb_gbs"] / HBM_PEAK_GBS, 4)},
                      "frame_tail_ms_per_step": round(f["other_ms"], 4),
                      "bitstream_kernel_ms_per_step": round(f["bits_ms"], 4)}

    plain = None
    if not args.no_plain:
        fps, p50, p99 = plain_single_stream(Engine)
        plain = {"fps": round(fps, 1), "latency_ms_p50": round(p50, 3), "latency_ms_p99": round(p99, 3)}

    hard = serve = fed = None
    if not args.no_extra:
        hard = hard_motion(Engine)
        serve = serving()
        fed = host_fed(Engine)

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline()
        cpu = {"value": c["value"], "unit": "CIF frames/s", "cores": c["cores"], "kind": "port",
               "per_core": c["per_core"], "cpu_model": cpu_model(), "runs": c["runs"],
               "sample": f"the plain CPU configuration: one CIF stream of 30 frames (fixed QUANT 8, global motion "
                         f"+-7, sigma 2, distinct seeds) per process, each pinned to its own core, {c['cores']} "
                         f"cores; median of 3 runs ({c['cpu_s']} s of CPU in all)"}

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "CIF frames/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": round(dt * 1e3 / K, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "batched: 256 CIF + 256 QCIF streams, lengths U{60..600}, global motion +-15 px, "
                               "rate control 64-384 kbit/s, inputs resident in HBM, every frame's H.261 bitstream written in HBM "
                               f"by the step; timed: {timed_window}",
                   "frames": frames, "cif_equivalent_frames": round(cif_equiv, 1),
                   "frames_per_s": round(frames / dt, 1), "seconds": round(dt, 3),
                   "levels_ready_cif_frames_per_s": round(levels_fps, 1),
                   "plain_single_cif": plain,
                   "batched_over_plain": round(value / plain["fps"], 2) if plain else None,
                   "hard_motion_batch": hard, "serving": serve,
                   "host_fed_batch": fed,
                   "parallelism": "1 GPU, all streams per frame step",
                   "autotune": tuned,
                   "device": f"{dev['name']} {dev['arch']}, {dev['compute_units']} CUs, {dev['clock_mhz']} MHz"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
