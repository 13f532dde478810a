"""Parity with the oracle at the lengths and batch shapes the benchmark runs.

* plain: one CIF stream, fixed QUANT 8, 300 frames (past the temporal
  reference's wrap at 32, src/h261_picture.c:11), through the engine in the
  benchmark's plain shape and through the reference's own frame loop
  (src/h261_main.c:186-249) over the drop-in ABI;
* batched: the benchmark's 512 streams (256 CIF + 256 QCIF, rate control
  64-384 kbit/s, global motion +-15) in its launch shape -- coder 4 MBs per
  wave, XCD remap of motion estimation and coder, two chains, rate-control
  begin as its own launch, graphs replayed in chunks, finished streams closed
  and the MB space re-packed every 30 steps -- for 70 frame steps (the first
  re-packing closes two streams at step 65);
* hard motion: the 1024-stream local-motion batch (noise sigma 4, rate
  control on) across its scene cut at frame 50.

Bar: bit-exact bitstreams and quantisers for every checked frame, and
reconstructions of a sample of streams.  The oracle runs in CPU worker
processes (tests/oracle_pool.py) while the GPU codes.
"""
import numpy as np
import pytest

from conftest import announce
from oracle_pool import OracleRuns, digest, first_mismatch, recon_digest
from h261enc_cuda_amd import Engine, content_frame
from h261enc_cuda_amd import _native as nat
from h261enc_cuda_amd.workloads import (batched_streams, content_kwargs, hard_streams, open_streams, plain_stream,
                                        run_job)

pytestmark = pytest.mark.gpu

# the launch shape the autotuner picks for the batched workload (coder MBs
# per wave, XCD remap bits) and the benchmark's job parameters
BENCH_SHAPE = (4, 3)
WARMUP = 5
COMPACT_EVERY = 30


class Tracker:
    """Which streams are open and which frame each coded last, following
    run_job's rule: after the chunk ending at step `done` (< end), streams
    whose length is <= done are closed."""

    def __init__(self, specs, n_frames, start, end):
        self.len_gen = np.array([min(s["length"], n_frames) for s in specs])
        self.length = np.array([s["length"] for s in specs])
        self.start, self.end = start, end
        self.boundary = -1

    def open_ids(self):
        return [i for i in range(len(self.length)) if self.length[i] > self.boundary]

    def after(self, done):
        if (done - self.start) % COMPACT_EVERY == 0 and done < self.end:
            self.boundary = done


def fetch(eng, ids, which):
    eng.fetch_bitstreams([ids[i] for i in which])
    return eng.fetched()


# ---------------------------------------------------------------------------
# plain: one CIF stream, 300 frames


@pytest.fixture(scope="module")
def plain_oracle(oracle):
    spec = plain_stream(310)
    w, h = spec["size"]
    s = oracle.stream(w, h, fixed_quant=8)
    frames = []
    for t in range(spec["length"]):
        f = content_frame(spec["seed"], w, h, t, **content_kwargs(spec))
        bits, rec, _, _, q = s.encode(*f)
        frames.append((bits, recon_digest(rec), q))
    return spec, frames


@pytest.mark.timeout(600)
def test_plain_cif_300_frames_matches_oracle(plain_oracle):
    """The benchmark's plain stream (one wave per MB, 310 generated frames):
    every frame stepped and read back, then the benchmark's own pattern
    (10 steps, a 150-step graph replay, 150 synchronised steps)."""
    spec, want = plain_oracle
    w, h = spec["size"]
    n = spec["length"]
    announce("plain CIF 310 frames, per step")
    eng = Engine(max_streams=1, mbs_per_wave=1)
    sid = eng.open(w, h, fixed_quant=8)
    eng.generate_input(sid, spec["seed"], n, **content_kwargs(spec))
    for t in range(n):
        eng.step()
        eng.sync()
        info = eng.info(sid)
        bits = eng.bitstream(sid)
        ob, orec, oq = want[t]
        assert info.frame_num == t
        assert info.quant == oq == 8, f"frame {t}: quantiser {info.quant}"
        assert bits == ob, f"frame {t}: bitstream differs ({len(bits)} vs {len(ob)} bytes)"
        assert recon_digest(eng.recon(sid)) == orec, f"frame {t}: reconstruction differs"
    eng.destroy()

    announce("plain CIF 310 frames, benchmark pattern")
    eng = Engine(max_streams=1, mbs_per_wave=1)
    sid = eng.open(w, h, fixed_quant=8)
    eng.generate_input(sid, spec["seed"], n, **content_kwargs(spec))
    for _ in range(10):
        eng.step()
    eng.sync()
    eng.run(150)
    eng.sync()
    assert eng.info(sid).frame_num == 159
    assert eng.bitstream(sid) == want[159][0], "frame 159 after the graph replay"
    assert recon_digest(eng.recon(sid)) == want[159][1]
    for t in range(160, n):
        eng.step()
        eng.sync()
        assert eng.bitstream(sid) == want[t][0], f"frame {t}: bitstream differs"
    assert recon_digest(eng.recon(sid)) == want[n - 1][1]
    eng.destroy()


@pytest.mark.timeout(600)
def test_dropin_reference_loop_300_frames_matches_oracle(plain_oracle):
    """The reference's frame loop, mode decision, rate control and writer
    (built from its sources) over libh261gpu.so for the 300-frame plain
    stream: every frame's bitstream and reconstruction equal the oracle's."""
    import os
    from oracle.binding import REF_DROPIN_SO, RefDropin
    if not os.path.exists(REF_DROPIN_SO):
        pytest.skip("oracle/_ref drop-in build not present")
    spec, want = plain_oracle
    w, h = spec["size"]
    n = 300
    flat = np.stack([np.concatenate([p.ravel() for p in content_frame(spec["seed"], w, h, t,
                                                                      **content_kwargs(spec))])
                     for t in range(n)])
    announce("drop-in reference loop, plain CIF 300 frames")
    bits, sizes, recon, quants, _ = RefDropin().encode(flat, w, h, fixed_quant=8)
    off = 0
    ysz, csz = w * h, w * h // 4
    for t in range(n):
        ob, orec, oq = want[t]
        assert quants[t] == oq, f"frame {t}: quantiser"
        assert bits[off:off + sizes[t]] == ob, f"frame {t}: bitstream differs"
        r = recon[t]
        assert recon_digest((r[:ysz], r[ysz:ysz + csz], r[ysz + csz:])) == orec, f"frame {t}: recon differs"
        off += sizes[t]
    assert off == len(bits)


# ---------------------------------------------------------------------------
# batched: the benchmark's 512 streams in its launch shape

BATCH_FRAMES = 70


def _batched_sample(specs):
    """4 CIF + 4 QCIF streams: the two that close at the first re-packing and
    others spread over the MB space"""
    lengths = np.array([s["length"] for s in specs])
    closers = [int(i) for i in np.nonzero(lengths <= WARMUP + 2 * COMPACT_EVERY)[0]]
    cif = [i for i in closers if i < 256][:1] + [0, 131, 255]
    qcif = [i for i in closers if i >= 256][:1] + [256, 389, 511]
    return sorted(set(cif[:4] + qcif[:4]))


def _bench_engine(specs, n_frames, mbs_per_wave, remap, max_streams=None):
    eng = Engine(max_streams=max_streams or len(specs))
    eng.config(nat.OPT_MBS_PER_WAVE, mbs_per_wave)
    eng.config(nat.OPT_XCD_REMAP, remap)
    ids = open_streams(eng, specs, frames=n_frames)
    return eng, ids


@pytest.mark.timeout(900)
def test_batched_bench_shape_matches_oracle():
    specs = batched_streams()
    T = BATCH_FRAMES
    sample = _batched_sample(specs)
    assert any(specs[i]["length"] <= WARMUP + 2 * COMPACT_EVERY for i in sample), "no stream closes in the window"
    runs = OracleRuns(specs, T, recon_ids=sample, raw_ids=sample)
    try:
        # (1) the benchmark's exact pattern: warmup (4 steps + one replay that
        # captures), then replays of 30 steps with closes and re-packing;
        # every open stream's last frame checked at each chunk end
        announce("batched 512 streams, benchmark pattern (30-step replays)")
        eng, ids = _bench_engine(specs, T, *BENCH_SHAPE)
        for _ in range(WARMUP - 1):
            eng.step()
        eng.run(1)
        tr = Tracker(specs, T, WARMUP, T)
        got_a = {}
        recon_a = {}

        def chunk_end(done):
            open_ids = tr.open_ids()
            for i, b in zip(open_ids, fetch(eng, ids, open_ids)):
                t = min(tr.len_gen[i], done) - 1
                got_a.setdefault(i, []).append((t, digest(b), len(b), eng.info(ids[i]).quant))
            for i in sample:
                if i in open_ids:
                    recon_a[(i, min(tr.len_gen[i], done) - 1)] = recon_digest(eng.recon(ids[i]))
            tr.after(done)

        run_job(eng, ids, specs, T - WARMUP, start=WARMUP, compact_every=COMPACT_EVERY, after_chunk=chunk_end)
        eng.destroy()

        # (2) the same shape stepped one graph replay at a time: every frame
        # of every stream checked
        announce("batched 512 streams, every frame")
        eng, ids = _bench_engine(specs, T, *BENCH_SHAPE)
        got_b = {}
        recon_b = {}
        raw_b = {i: [] for i in sample}

        def record(done, which):
            for i, b in zip(which, fetch(eng, ids, which)):
                got_b.setdefault(i, []).append((done - 1, digest(b), len(b), eng.info(ids[i]).quant))
                if i in raw_b:
                    raw_b[i].append(b)
                    recon_b[(i, done - 1)] = recon_digest(eng.recon(ids[i]))

        for k in range(WARMUP):
            if k < WARMUP - 1:
                eng.step()
            else:
                eng.run(1)
            record(k + 1, list(range(len(specs))))
        tr = Tracker(specs, T, WARMUP, T)

        def step_end(done):
            record(done, [i for i in tr.open_ids() if tr.len_gen[i] >= done])
            tr.after(done)

        run_job(eng, ids, specs, T - WARMUP, start=WARMUP, compact_every=COMPACT_EVERY, after_chunk=step_end,
                chunk_steps=1)
        eng.destroy()

        want = runs.result(timeout=900)
    finally:
        runs.close()
    assert sum(len(v) for v in got_b.values()) == sum(min(s["length"], T) for s in specs)
    bad = first_mismatch(got_a, want)
    assert bad is None, f"benchmark pattern: {bad}"
    bad = first_mismatch(got_b, want)
    assert bad is None, f"every frame: {bad}"
    for (i, t), d in list(recon_a.items()) + list(recon_b.items()):
        assert d == want[i][t].recon, f"stream {i} frame {t}: reconstruction differs"
    for i in sample:
        assert [digest(b) for b in raw_b[i]] == [f.bits for f in want[i]]


# ---------------------------------------------------------------------------
# hard motion: 1024 streams across the scene cut at frame 50

HARD_FRAMES = 56


@pytest.mark.timeout(900)
def test_hard_motion_batch_across_scene_cut_matches_oracle():
    specs = hard_streams()
    T = HARD_FRAMES
    sample = list(range(0, 512, 16)) + list(range(512, 1024, 16))  # 32 CIF + 32 QCIF
    runs = OracleRuns([specs[i] for i in sample], T, recon_ids=range(0, len(sample), 8))
    try:
        announce("hard motion 1024 streams, 56 frames")
        eng = Engine(max_streams=len(specs))
        ids = open_streams(eng, specs, frames=T)
        got = {}
        recon = {}

        def record(done):
            for k, b in enumerate(fetch(eng, ids, sample)):
                got.setdefault(k, []).append((done - 1, digest(b), len(b), eng.info(ids[sample[k]]).quant))
                if k % 8 == 0:
                    recon[(k, done - 1)] = recon_digest(eng.recon(ids[sample[k]]))

        eng.step()  # as the benchmark: the intra frame, then one replay that captures
        record(1)
        eng.run(1)
        record(2)
        run_job(eng, ids, specs, T - 2, start=2, compact_every=T, after_chunk=record, chunk_steps=1)
        eng.destroy()
        want = runs.result(timeout=900)
    finally:
        runs.close()
    bad = first_mismatch(got, want)
    assert bad is None, bad
    for (k, t), d in recon.items():
        assert d == want[k][t].recon, f"stream {sample[k]} frame {t}: reconstruction differs"
    # the scene cut is in the window: frame 50 (a new texture) costs far more
    # than frame 49 in nearly every stream, under rate control
    assert np.mean([want[k][50].nbytes > 1.3 * want[k][49].nbytes for k in range(len(sample))]) > 0.9


# ---------------------------------------------------------------------------
# engine state across graph replays, re-packing and input activation


def _small_specs():
    return [dict(size=(352, 288), length=40, bitrate=256000, seed=701, motion=15),
            dict(size=(176, 144), length=40, bitrate=64000, seed=702, motion=15),
            dict(size=(176, 144), length=40, bitrate=128000, seed=703, motion=15),
            dict(size=(352, 288), length=40, bitrate=384000, seed=704, motion=15)]


def _oracle_frames(oracle, spec, n):
    w, h = spec["size"]
    s = oracle.stream(w, h, bitrate=spec["bitrate"], vbv=spec["bitrate"])
    return [s.encode(*content_frame(spec["seed"], w, h, t, **content_kwargs(spec)))[0] for t in range(n)]


def test_overlapped_writer_across_repeated_graph_replays(oracle):
    """With the writer beside the next step's motion estimation, a second run
    of the same captured graph (no close or re-packing between) still ends
    with the last step's bitstream written: run(k); sync(); run(k); sync()."""
    specs = _small_specs()
    engines = []
    for overlap in (1, 0):
        eng = Engine(max_streams=8)
        eng.config(nat.OPT_OVERLAP_WRITER, overlap)
        ids = open_streams(eng, specs, frames=30)
        engines.append((eng, ids))
    announce("overlapped writer, repeated graph replays")
    want = [_oracle_frames(oracle, s, 13) for s in specs]
    done = 0
    for k in (3, 3, 1, 1, 5):
        for eng, _ in engines:
            eng.run(k)
            eng.sync()
        done += k
        for i in range(len(specs)):
            a = engines[0][0].bitstream(engines[0][1][i])
            b = engines[1][0].bitstream(engines[1][1][i])
            assert a == b == want[i][done - 1], f"after {done} steps: stream {i}"
    for eng, _ in engines:
        eng.destroy()


def test_compact_while_async_input_generates(oracle):
    """Re-packing while a stream's asynchronous input is still being
    generated: the stream must join at its new MB base (its activation writes
    the whole descriptor)."""
    specs = _small_specs()
    eng = Engine(max_streams=8)
    eng.config(nat.OPT_ASYNC_INPUT, 1)
    announce("compact with asynchronous input pending")
    ids = [None] * len(specs)
    for i in (0, 1, 2):
        ids[i] = eng.open(*specs[i]["size"], bitrate=specs[i]["bitrate"], vbv_size=specs[i]["bitrate"])
        eng.generate_input(ids[i], specs[i]["seed"], 20, **content_kwargs(specs[i]))
    while any(eng.input_pending(ids[i]) for i in (0, 1, 2)):
        eng.step()
        eng.sync()
    for _ in range(3):
        eng.step()
    eng.sync()
    done = {i: eng.frames_done(ids[i]) for i in (0, 1, 2)}
    eng.close(ids[0])  # the lowest MB range: the others move down
    i = 3
    ids[i] = eng.open(*specs[i]["size"], bitrate=specs[i]["bitrate"], vbv_size=specs[i]["bitrate"])
    eng.generate_input(ids[i], specs[i]["seed"], 600, **content_kwargs(specs[i]))  # long: still generating
    # activation happens only in a later step call, so the input is pending
    # at the re-packing whatever the device has done meanwhile
    assert eng.input_pending(ids[i])
    eng.compact()
    coded = {1: [], 2: [], 3: []}
    for _ in range(12):
        eng.step()
        eng.sync()
        for j in coded:
            if j == 3 and eng.input_pending(ids[3]):
                continue
            if j == 3 and eng.frames_done(ids[3]) == len(coded[3]):
                continue
            coded[j].append(eng.bitstream(ids[j]))
    eng.destroy()
    for j in (1, 2):
        want = _oracle_frames(oracle, specs[j], done[j] + 12)
        assert coded[j] == want[done[j]:], f"stream {j}"
    assert coded[3], "the new stream never joined"
    want = _oracle_frames(oracle, specs[3], len(coded[3]))
    assert coded[3] == want, "stream opened during the re-packing"


def test_generated_input_after_host_frames(oracle):
    """A stream fed host frames, then switched to device-generated input
    while its last step is still queued: generation must continue from the
    frame the device has reached."""
    spec = _small_specs()[0]
    w, h = spec["size"]
    eng = Engine(max_streams=2)
    sid = eng.open(w, h, bitrate=spec["bitrate"], vbv_size=spec["bitrate"])
    announce("host frames, then generated input")
    got = []
    for t in range(4):
        eng.put_frame(sid, *content_frame(spec["seed"], w, h, t, **content_kwargs(spec)))
        eng.step()
        if t < 3:
            eng.sync()
            got.append(eng.bitstream(sid))
    # no sync: the fourth step may still be queued
    eng.generate_input(sid, spec["seed"], 10, **content_kwargs(spec))
    eng.sync()
    got.append(eng.bitstream(sid))
    for _ in range(6):
        eng.step()
        eng.sync()
        got.append(eng.bitstream(sid))
    eng.destroy()
    assert got == _oracle_frames(oracle, spec, 10)
