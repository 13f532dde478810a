"""Autotuner for the engine's launch shape.

The macroblock coder codes `mbs_per_wave` macroblocks in sequence per wave
(statistics for the denoise update are reduced in LDS and flushed to HBM
once per wave, so more MBs per wave means fewer global atomics but fewer,
longer waves).  The tuner times each candidate on a sample of the target
workload -- the same mix of stream sizes, all streams active, inter frames
-- with HIP events around the kernel launches, and returns the fastest.
"""
from __future__ import annotations

from typing import Dict, Iterable, List, Tuple

from .encoder import Engine

_CACHE: Dict[Tuple, int] = {}


def tune_mbs_per_wave(n_cif: int, n_qcif: int, candidates: Iterable[int] = (1, 2, 4, 8), steps: int = 6,
                      seed: int = 99) -> Tuple[int, List[dict]]:
    """Returns (best mbs_per_wave, per-candidate timings)."""
    key = (n_cif, n_qcif, tuple(candidates))
    results = []
    for mpw in candidates:
        eng = Engine(max_streams=n_cif + n_qcif, mbs_per_wave=mpw)
        for i in range(n_cif + n_qcif):
            size = (352, 288) if i < n_cif else (176, 144)
            sid = eng.open(*size)
            eng.generate_input(sid, seed + i, steps + 2, sigma=2, motion=15)
        eng.step()  # intra frame
        eng.sync()
        eng.profile(True)
        for _ in range(steps):
            eng.step()
        eng.sync()
        t = eng.kernel_times()
        eng.destroy()
        results.append({"mbs_per_wave": mpw, "mb_ms": t["mb_ms"] / max(1, t["steps"]),
                        "me_ms": t["me_ms"] / max(1, t["steps"])})
    best = min(results, key=lambda r: r["mb_ms"])["mbs_per_wave"]
    _CACHE[key] = best
    return best, results


def cached(n_cif: int, n_qcif: int, candidates: Iterable[int] = (1, 2, 4, 8)):
    return _CACHE.get((n_cif, n_qcif, tuple(candidates)))
