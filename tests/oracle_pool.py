"""The oracle over many streams at once, in CPU worker processes.

Test infrastructure: the long-sequence and full-batch parity tests need the
oracle's output for hundreds of streams and tens to hundreds of frames.  Each
stream is a sequential chain (reference frame, rate control, denoise state),
so streams are spread over a pool of spawned processes, one stream per task.
The workers load only liboracle.so and the host content generator
(libh261content.so): never the HIP runtime, so they do not count as GPU
processes and may run while the parent drives the GPU.

Per frame a worker returns a 16-byte digest of the oracle's bitstream, its
length, the picture quantiser, and on request a digest of the reconstruction
and the bitstream bytes themselves.
"""
from __future__ import annotations

import hashlib
import multiprocessing as mp
import os
from typing import Dict, List, NamedTuple, Optional, Sequence

import numpy as np


def digest(data: bytes) -> bytes:
    return hashlib.blake2b(data, digest_size=16).digest()


def recon_digest(planes) -> bytes:
    return digest(b"".join(np.ascontiguousarray(p).tobytes() for p in planes))


class Frame(NamedTuple):
    bits: bytes          # digest of the bitstream
    nbytes: int
    quant: int
    recon: Optional[bytes]   # digest of the reconstruction (on request)
    raw: Optional[bytes]     # the bitstream itself (on request)
    sad: int             # byte-SAD candidates the search evaluated


def _stream_job(args) -> List[Frame]:
    spec, n_frames, want_recon, want_raw = args
    from h261enc_cuda_amd.encoder import content_frame
    from h261enc_cuda_amd.workloads import content_kwargs
    from oracle.binding import Oracle
    w, h = spec["size"]
    s = Oracle().stream(w, h, init_quant=10, bitrate=spec["bitrate"], vbv=spec["bitrate"],
                        fixed_quant=spec.get("fixed", 0))
    kw = content_kwargs(spec)
    out = []
    for t in range(n_frames):
        f = content_frame(spec["seed"], w, h, t, **kw)
        bits, rec, _, _, q = s.encode(*f)
        out.append(Frame(digest(bits), len(bits), q, recon_digest(rec) if want_recon else None,
                         bits if want_raw else None, s.sad_candidates()))
    s.close()
    return out


def _ping() -> int:
    from oracle.binding import Oracle  # noqa: F401
    from h261enc_cuda_amd.encoder import content_frame  # noqa: F401
    return os.getpid()


def pool_workers() -> int:
    """at most 16 processes (the CPU share of a one-GPU machine)"""
    return max(1, min(16, os.cpu_count() or 1))


class OracleRuns:
    """Starts the oracle over `specs` (frames 0 .. min(length, n_frames) - 1
    each) in the background; `result()` waits and returns, per spec, its list
    of Frame records.  recon_ids / raw_ids: spec indices whose reconstruction
    digests / raw bitstreams are kept."""

    def __init__(self, specs: Sequence[dict], n_frames: int, recon_ids=(), raw_ids=(),
                 workers: Optional[int] = None):
        self.n = len(specs)
        jobs = [(dict(s), min(s["length"], n_frames), i in set(recon_ids), i in set(raw_ids))
                for i, s in enumerate(specs)]
        # the longest chains first, so the pool drains evenly
        self.order = sorted(range(self.n), key=lambda i: -jobs[i][1] * specs[i]["size"][0] * specs[i]["size"][1])
        self.pool = mp.get_context("spawn").Pool(workers or pool_workers())
        # a worker that cannot start is replaced by the pool forever: fail
        # fast instead
        try:
            self.pool.apply_async(_ping).get(120)
        except BaseException:
            self.pool.terminate()
            self.pool.join()
            raise RuntimeError("oracle worker processes failed to start")
        self.async_res = self.pool.map_async(_stream_job, [jobs[i] for i in self.order], chunksize=1)
        self._res: Optional[List[List[Frame]]] = None

    def result(self, timeout: Optional[float] = None) -> List[List[Frame]]:
        if self._res is None:
            got = self.async_res.get(timeout)
            res: List[Optional[List[Frame]]] = [None] * self.n
            for k, i in enumerate(self.order):
                res[i] = got[k]
            self._res = res  # type: ignore[assignment]
            self.pool.close()
            self.pool.join()
        return self._res

    def close(self) -> None:
        if self._res is None:
            self.pool.terminate()
            self.pool.join()


def first_mismatch(got: Dict[int, List[tuple]], want: List[List[Frame]]) -> Optional[str]:
    """got[i] = [(frame, bits digest, nbytes, quant), ...] for spec i; the
    first disagreement with the oracle, or None"""
    for i, frames in got.items():
        for t, d, nb, q in frames:
            w = want[i][t]
            if (d, nb, q) != (w.bits, w.nbytes, w.quant):
                return (f"stream {i} frame {t}: device {nb} bytes quant {q}, "
                        f"oracle {w.nbytes} bytes quant {w.quant}, bitstreams {'equal' if d == w.bits else 'differ'}")
    return None
