"""The benchmark workloads as stream specifications, and the job driver that
runs them through an Engine.

batched_streams  256 CIF + 256 QCIF streams, lengths U{60..600}, global motion
                 in [-15, 15] px/frame, rate control 64..384 kbit/s
hard_streams     1024 mixed QCIF/CIF streams, per-MB local motion (random
                 affine warp + 10% random vectors), scene cuts every 50
                 frames, noise sigma 4, rate control on
plain_stream     one CIF stream, fixed QUANT 8, global motion in [-7, 7]

A spec is a dict: size (w, h), length (frames), bitrate (bit/s; VBV = bitrate,
initial quantiser 10), seed (content), motion, and optionally local,
scene_cut, sigma, fixed (quantiser override).
"""
from __future__ import annotations

import time
from typing import Callable, List, Optional, Sequence

import numpy as np

CIF = (352, 288)
QCIF = (176, 144)


def batched_streams(seed: int = 2026) -> List[dict]:
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(512):
        cif = i < 256
        specs.append(dict(size=CIF if cif else QCIF, length=int(rng.integers(60, 601)),
                          bitrate=int(rng.integers(64, 385)) * 1000, seed=int(rng.integers(1, 2 ** 62)),
                          motion=15))
    return specs


def hard_streams(seed: int = 2027) -> List[dict]:
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(1024):
        cif = i < 512
        specs.append(dict(size=CIF if cif else QCIF, length=int(rng.integers(60, 601)),
                          bitrate=int(rng.integers(64, 385)) * 1000, seed=int(rng.integers(1, 2 ** 62)),
                          motion=0, local=True, scene_cut=50, sigma=4))
    return specs


def plain_stream(frames: int = 300) -> dict:
    return dict(size=CIF, length=frames, bitrate=512000, seed=1, motion=7, sigma=2, fixed=8)


def content_kwargs(spec: dict, sigma: int = 2) -> dict:
    """keyword arguments of content_frame / Engine.generate_input for a spec"""
    return dict(sigma=spec.get("sigma", sigma), motion=spec["motion"], local=spec.get("local", False),
                scene_cut=spec.get("scene_cut", 0))


def open_streams(eng, specs: Sequence[dict], sigma: int = 2, frames: Optional[int] = None) -> List[int]:
    """opens every spec's stream with its device-generated input (the first
    `frames` frames, default all) and waits until the inputs are in HBM"""
    ids = []
    for s in specs:
        sid = eng.open(*s["size"], bitrate=s["bitrate"], vbv_size=s["bitrate"], init_quant=10,
                       fixed_quant=s.get("fixed", 0))
        n = s["length"] if frames is None else min(frames, s["length"])
        eng.generate_input(sid, s["seed"], n, **content_kwargs(s, sigma))
        ids.append(sid)
    eng.sync()
    return ids


def run_job(eng, ids: Sequence[int], specs: Sequence[dict], steps: int, start: int = 0, compact_every: int = 30,
            after_chunk: Optional[Callable[[int], None]] = None, chunk_steps: Optional[int] = None):
    """frame steps start..start+steps-1 (graph replay; eager with events when
    the engine profiles), closing finished streams and re-packing the MB
    space every compact_every steps (closing and re-packing are queued
    behind the steps: nothing waits).  chunk_steps: replay that many steps
    per run call (default compact_every); after_chunk(done) runs after each.
    Returns (seconds, frames, mbs)."""
    lengths = np.array([s["length"] for s in specs])
    nmb = np.array([(s["size"][0] // 16) * (s["size"][1] // 16) for s in specs])
    done = start
    end = start + steps
    open_mask = lengths > start
    per_run = chunk_steps or compact_every
    eng.sync()
    t0 = time.perf_counter()
    while done < end:
        chunk = min(compact_every, end - done)
        left = chunk
        while left > 0:
            k = min(per_run, left)
            eng.run(k)
            done += k
            left -= k
            if after_chunk is not None:
                after_chunk(done)
        fin = open_mask & (lengths <= done)
        if fin.any() and done < end:
            for i in np.nonzero(fin)[0]:
                eng.close(ids[i])
            open_mask &= ~fin
            eng.compact()
    eng.sync()
    dt = time.perf_counter() - t0
    frames = np.minimum(lengths, end) - np.minimum(lengths, start)
    return dt, int(frames.sum()), int((frames * nmb).sum())
