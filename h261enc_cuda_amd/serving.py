"""Serving: encode requests admitted as they arrive, coded together.

A request is a whole synthetic sequence (QCIF or CIF, its length and
bitrate target).  The loop admits every arrived request whose device
footprint (stream state + its frames in HBM) fits the HBM budget, then runs
one frame step over all admitted streams -- requests that arrive while
others are coding join the next step (continuous batching); finished
streams are closed and their memory and macroblock range reused.

Per-frame latency = time from the frame becoming eligible (its input
available and its previous frame done) to the end of the step that coded it
(levels, side information, reconstruction and the H.261 bitstream on the
device) -- with `deliver`, to the bitstream's arrival in host memory (every
step's bitstreams are gathered into pinned host memory by one queued
operation).  A request's input is a synthetic sequence generated on the
device; the time from its arrival until that input is available is reported
separately as the input wait, and is part of the request latency (arrival to
last frame).
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Dict, List, Optional

import numpy as np

from . import _native as nat
from .encoder import Engine


@dataclasses.dataclass
class Request:
    arrival: float          # seconds after the start
    width: int
    height: int
    frames: int
    bitrate: int
    seed: int
    sigma: int = 4
    motion: int = 0
    local: bool = True
    scene_cut: int = 50


def poisson_requests(rate: float, duration: float, seed: int = 4, cif_share: float = 0.6,
                     median_frames: int = 120, sigma: float = 0.8, cap: int = 1800,
                     rate_lo: int = 64000, rate_hi: int = 1024000) -> List[Request]:
    """Poisson arrivals; 60% CIF / 40% QCIF; lognormal length (median 120,
    sigma 0.8, cap 1800); bitrate uniform in [64, 1024] kbit/s; content with
    per-macroblock local motion and scene cuts every 50 frames."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        cif = rng.random() < cif_share
        n = int(min(cap, max(1, round(median_frames * np.exp(sigma * rng.standard_normal())))))
        out.append(Request(arrival=t, width=352 if cif else 176, height=288 if cif else 144, frames=n,
                           bitrate=int(rng.integers(rate_lo, rate_hi + 1)), seed=int(rng.integers(1, 2 ** 62))))


@dataclasses.dataclass
class ServingStats:
    requests: int
    frames: int
    cif_equivalent_frames: float
    seconds: float
    frame_latency_ms_p50: float
    frame_latency_ms_p99: float
    request_latency_ms_p50: float
    request_latency_ms_p99: float
    input_wait_ms_p50: float
    input_wait_ms_p99: float
    max_concurrent: int
    steps: int
    delivered_bytes: int = 0

    @property
    def frames_per_s(self) -> float:
        return self.frames / self.seconds

    @property
    def cif_frames_per_s(self) -> float:
        return self.cif_equivalent_frames / self.seconds


class ServingLoop:
    def __init__(self, hbm_budget: int = 64 << 30, max_streams: int = 4096, mbs_per_wave: int = 0,
                 async_input: bool = True):
        self.eng = Engine(max_streams=max_streams, hbm_budget=hbm_budget, mbs_per_wave=mbs_per_wave)
        # a request's frames are generated beside the frame steps; it joins
        # the first step after they are ready
        self.eng.config(nat.OPT_ASYNC_INPUT, int(async_input))

    def run(self, requests: List[Request], collect: bool = False, realtime: bool = True, deliver: bool = False):
        """Serves the requests.  deliver=True brings every step's bitstreams
        to host memory before the step counts as done; collect=True (implies
        deliver) also returns every request's per-frame bitstreams."""
        deliver = deliver or collect
        eng = self.eng
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting: collections.deque = collections.deque()
        index = {id(r): i for i, r in enumerate(requests)}
        outputs: Dict[int, List[bytes]] = {}
        # active streams as parallel arrays (per-step bookkeeping is vectorised)
        sids = np.zeros(0, np.int64)
        req_idx = np.zeros(0, np.int64)
        eligible = np.zeros(0)
        arrival = np.zeros(0)
        left = np.zeros(0, np.int64)
        nmb = np.zeros(0, np.int64)
        started = np.zeros(0, bool)  # input ready: the stream codes a frame per step
        frame_lat: List[np.ndarray] = []
        req_lat: List[float] = []
        input_wait: List[float] = []
        t_start = time.perf_counter()
        frames = mbs = steps = max_conc = delivered = 0
        while pending or waiting or len(sids):
            now = time.perf_counter() - t_start
            while pending and (pending[0].arrival <= now or not realtime):
                waiting.append(pending.popleft())
            new = []
            while waiting:
                r = waiting[0]
                try:
                    sid = eng.open(r.width, r.height, denoise=5, init_quant=10, bitrate=r.bitrate,
                                   vbv_size=r.bitrate)
                except RuntimeError:
                    break  # no slot / budget: admit when memory frees up
                try:
                    eng.generate_input(sid, r.seed, r.frames, sigma=r.sigma, motion=r.motion, local=r.local,
                                       scene_cut=r.scene_cut)
                except RuntimeError:
                    eng.close(sid)
                    break
                waiting.popleft()
                new.append((sid, r))
                if collect:
                    outputs[index[id(r)]] = []
            if new:
                t_adm = time.perf_counter() - t_start
                sids = np.concatenate([sids, [sid for sid, _ in new]])
                req_idx = np.concatenate([req_idx, [index[id(r)] for _, r in new]])
                eligible = np.concatenate([eligible, np.full(len(new), t_adm)])
                arrival = np.concatenate([arrival, [r.arrival if realtime else 0.0 for _, r in new]])
                left = np.concatenate([left, [r.frames for _, r in new]])
                nmb = np.concatenate([nmb, [(r.width // 16) * (r.height // 16) for _, r in new]])
                started = np.concatenate([started, np.zeros(len(new), bool)])
            if not len(sids):
                if pending and realtime:
                    time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t_start), 0.01)))
                continue
            max_conc = max(max_conc, len(sids))
            t_call = time.perf_counter() - t_start
            eng.step()
            # streams activated by this step (their input became ready) code
            # their first frame in it: eligible from this step's start
            for k in np.nonzero(~started)[0].tolist():
                if not eng.input_pending(int(sids[k])):
                    started[k] = True
                    eligible[k] = t_call
                    input_wait.append((t_call - arrival[k]) * 1e3)
            run = started
            if deliver:
                eng.fetch_bitstreams(sids[run])
            eng.sync()
            if deliver:
                if collect:
                    for ri, bits in zip(req_idx[run].tolist(), eng.fetched()):
                        outputs[ri].append(bits)
                else:
                    delivered += eng.fetch_wait()
            steps += 1
            t_end = time.perf_counter() - t_start
            frame_lat.append((t_end - eligible[run]) * 1e3)
            eligible[run] = t_end
            left[run] -= 1
            frames += int(run.sum())
            mbs += int(nmb[run].sum())
            done = left <= 0
            if done.any():
                for sid in sids[done].tolist():
                    eng.close(sid)
                req_lat.extend(((t_end - arrival[done]) * 1e3).tolist())
                keep = ~done
                sids, req_idx, eligible, arrival, left, nmb, started = (
                    sids[keep], req_idx[keep], eligible[keep], arrival[keep], left[keep], nmb[keep], started[keep])
        seconds = time.perf_counter() - t_start
        fl = np.concatenate(frame_lat) if frame_lat else np.zeros(1)
        rl = np.array(req_lat) if req_lat else np.zeros(1)
        iw = np.array(input_wait) if input_wait else np.zeros(1)
        stats = ServingStats(requests=len(requests), frames=frames, cif_equivalent_frames=mbs / 396.0,
                             seconds=seconds, frame_latency_ms_p50=float(np.percentile(fl, 50)),
                             frame_latency_ms_p99=float(np.percentile(fl, 99)),
                             request_latency_ms_p50=float(np.percentile(rl, 50)),
                             request_latency_ms_p99=float(np.percentile(rl, 99)),
                             input_wait_ms_p50=float(np.percentile(iw, 50)),
                             input_wait_ms_p99=float(np.percentile(iw, 99)), max_concurrent=max_conc,
                             steps=steps, delivered_bytes=delivered)
        return (stats, outputs) if collect else stats

    def close(self):
        self.eng.destroy()
