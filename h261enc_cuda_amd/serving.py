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
operation, and the loop queues the next step before it collects them, so
the device never waits for the host's bookkeeping).  A request's input is a
synthetic sequence generated on the
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
        deliver) also returns every request's per-frame bitstreams.

        With deliver, one step runs ahead: step k+1 is queued before step k's
        bitstreams are collected, so the host's bookkeeping (admissions,
        latencies, closes) overlaps the device's next step instead of leaving
        it idle.  Which streams code a frame in a step is known when it is
        queued (input ready, frames left), so the per-request state is
        updated then; a request is closed once its last frame's bitstream
        has arrived."""
        deliver = deliver or collect
        eng = self.eng
        R = len(requests)
        pending = collections.deque(sorted(range(R), key=lambda i: requests[i].arrival))
        waiting: collections.deque = collections.deque()
        outputs: Dict[int, List[bytes]] = {}
        # per-request state, indexed by the request's position in `requests`
        sid_of = np.full(R, -1, np.int64)
        eligible = np.zeros(R)
        arrival = np.array([r.arrival if realtime else 0.0 for r in requests]) if R else np.zeros(0)
        left = np.array([r.frames for r in requests], np.int64) if R else np.zeros(0, np.int64)
        nmb = np.array([(r.width // 16) * (r.height // 16) for r in requests], np.int64) if R else np.zeros(0, np.int64)
        started = np.zeros(R, bool)  # input ready: the stream codes a frame per step
        active = np.zeros(R, bool)   # open on the engine
        frame_lat: List[np.ndarray] = []
        req_lat: List[float] = []
        input_wait: List[float] = []
        t_start = time.perf_counter()
        frames = mbs = steps = max_conc = delivered = 0
        # the queued step whose results are not collected yet: the requests
        # that code a frame in it, and those for which it is the last frame
        inflight = None
        while pending or waiting or active.any() or inflight is not None:
            now = time.perf_counter() - t_start
            while pending and (requests[pending[0]].arrival <= now or not realtime):
                waiting.append(pending.popleft())
            while waiting:
                ri = waiting[0]
                r = requests[ri]
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
                sid_of[ri] = sid
                active[ri] = True
                if collect:
                    outputs[ri] = []
            act = np.nonzero(active)[0]
            queued = None
            if len(act):
                max_conc = max(max_conc, len(act))
                t_call = time.perf_counter() - t_start
                eng.step()
                steps += 1
                # streams activated by this step (their input became ready)
                # code their first frame in it: eligible from this step's start
                for ri in act[~started[act]].tolist():
                    if not eng.input_pending(int(sid_of[ri])):
                        started[ri] = True
                        eligible[ri] = t_call
                        input_wait.append((t_call - arrival[ri]) * 1e3)
                run = act[started[act] & (left[act] > 0)]
                if len(run):
                    left[run] -= 1
                    queued = (run, run[left[run] == 0])
            if deliver and inflight is not None:
                # step k's bitstreams (the fetch waits for step k only; step
                # k+1, queued above, keeps the device busy meanwhile)
                ran, last = inflight
                if collect:
                    for ri, bits in zip(ran.tolist(), eng.fetched()):
                        outputs[ri].append(bits)
                else:
                    delivered += eng.fetch_wait()
                self._step_done(ran, last, eng, sid_of, eligible, arrival, active, frame_lat, req_lat, t_start)
                frames += len(ran)
                mbs += int(nmb[ran].sum())
                inflight = None
            if queued is not None:
                if deliver:
                    eng.fetch_bitstreams(sid_of[queued[0]])
                    inflight = queued
                else:
                    eng.sync()
                    self._step_done(*queued, eng, sid_of, eligible, arrival, active, frame_lat, req_lat, t_start)
                    frames += len(queued[0])
                    mbs += int(nmb[queued[0]].sum())
            elif len(act):
                # a step that coded nothing (inputs still being generated):
                # wait for it rather than queue empty steps
                eng.sync()
            elif inflight is None and pending and realtime:
                time.sleep(max(0.0, min(requests[pending[0]].arrival - (time.perf_counter() - t_start), 0.01)))
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

    @staticmethod
    def _step_done(ran, last, eng, sid_of, eligible, arrival, active, frame_lat, req_lat, t_start):
        """The frames of a step are done (on the device, or delivered): their
        latencies, each stream's next frame eligible from now, and the
        requests whose last frame this was (`last`) closed."""
        t_end = time.perf_counter() - t_start
        frame_lat.append((t_end - eligible[ran]) * 1e3)
        eligible[ran] = t_end
        for ri in last.tolist():
            eng.close(int(sid_of[ri]))
            active[ri] = False
            req_lat.append((t_end - arrival[ri]) * 1e3)

    def close(self):
        self.eng.destroy()
