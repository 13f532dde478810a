"""Parity of the HIP device layer with the oracle, through the C ABI.

Bar: bit-exact.  Motion vectors, the three SATD costs, macroblock types,
flags, CBPs, quantised levels, reconstructions and bitstream bytes must
equal the oracle's (a CPU restatement of the reference path) on the same
seeded input.
"""
import numpy as np
import pytest

from conftest import announce
from h261enc_cuda_amd import Engine, LegacyDevice, content_frame
from h261enc_cuda_amd.encoder import CIF, QCIF

pytestmark = pytest.mark.gpu


def seq(seed, size, n, **kw):
    w, h = size
    return [content_frame(seed, w, h, t, **kw) for t in range(n)]


def zeros(size):
    w, h = size
    return (np.zeros((h, w), np.uint8), np.zeros((h // 2, w // 2), np.uint8), np.zeros((h // 2, w // 2), np.uint8))


def special_frames(size, rng):
    """Content that stresses ties and limits: flat, uniform noise, extremes."""
    w, h = size
    flat = tuple(np.full(s, 128, np.uint8) for s in ((h, w), (h // 2, w // 2), (h // 2, w // 2)))
    noise = tuple(rng.integers(0, 256, s, dtype=np.uint8) for s in ((h, w), (h // 2, w // 2), (h // 2, w // 2)))
    yy, xx = np.mgrid[0:h, 0:w]
    checker = (((yy // 3 + xx // 5) & 1) * 255).astype(np.uint8)
    ext = (checker, checker[::2, ::2].copy(), (255 - checker[::2, ::2]).copy())
    return [flat, noise, ext]


def assert_me_equal(got, want, what):
    for f in ("intra_cost", "inter_cost", "filter_cost", "mvx", "mvy"):
        bad = np.nonzero(got[f] != want[f])[0]
        assert bad.size == 0, f"{what}: {f} differs at MBs {bad[:10]}: got {got[f][bad[:5]]} want {want[f][bad[:5]]}"


@pytest.mark.parametrize("variant", ["0", "1"], ids=["two-stage", "surface"])
@pytest.mark.parametrize("size", [CIF, QCIF], ids=["cif", "qcif"])
def test_legacy_me_matches_oracle(oracle, size, variant, monkeypatch):
    monkeypatch.setenv("H261GPU_ME_VARIANT", variant)
    rng = np.random.default_rng(7)
    frames = seq(11, size, 4, sigma=2, motion=7)
    cases = [(frames[0], zeros(size), "frame0 vs zero ref")]
    cases += [(frames[t], frames[t - 1], f"frame{t}") for t in range(1, 4)]
    sp = special_frames(size, rng)
    cases += [(sp[0], sp[0], "flat/flat"), (sp[1], sp[2], "noise/checker"), (sp[2], sp[1], "checker/noise"),
              (frames[1], sp[0], "texture/flat")]
    dev = LegacyDevice(*size)
    try:
        for cur, ref, name in cases:
            announce(f"legacy me {size} {name}")
            dev.set_frames(cur, ref)
            got = dev.me()
            want = oracle.me_frame(cur[0], ref[0])
            assert_me_equal(got, want, f"{size} {name}")
    finally:
        dev.close()


@pytest.mark.parametrize("size,fixed,full_swing", [(CIF, 8, False), (QCIF, 0, False), (CIF, 1, True)],
                         ids=["cif-q8", "qcif-rc", "cif-q1-full-swing"])
def test_legacy_encode_matches_oracle(oracle, size, fixed, full_swing):
    """encode/decode through the drop-in entry points with the oracle's mode
    decision; macroblock structs and reconstructions compared per frame."""
    w, h = size
    frames = extreme_frames(size, np.random.default_rng(98)) if full_swing else seq(21, size, 6, sigma=2, motion=7)
    os_ = oracle.stream(w, h, fixed_quant=fixed)
    dev = LegacyDevice(w, h)
    ref = zeros(size)
    try:
        for t, f in enumerate(frames):
            announce(f"legacy encode {size} frame {t}")
            bits, rec, me, mbs, q = os_.encode(*f, want_details=True)
            dev.set_frames(f, ref)
            got_me = dev.me()
            assert_me_equal(got_me, me, f"frame {t}")
            md = oracle.mode_decision(me, 0 if t == 0 else 1)
            got_mbs = dev.encode(md, q)
            got_rec = dev.decode()
            for fld in ("type", "flags", "quant", "mv", "cbp", "tcoeff", "mb_x", "mb_y", "address"):
                bad = np.nonzero((got_mbs[fld] != mbs[fld]).reshape(len(mbs), -1).any(axis=1))[0]
                assert bad.size == 0, f"frame {t}: {fld} differs at MBs {bad[:10]}"
            for p in range(3):
                assert np.array_equal(got_rec[p], rec[p]), f"frame {t}: recon plane {p} differs"
            ref = rec
    finally:
        dev.close()


@pytest.mark.parametrize("size,fixed,n", [(CIF, 8, 30), (QCIF, 0, 30), (CIF, 0, 12)],
                         ids=["cif-q8-config0", "qcif-rc", "cif-rc"])
def test_dropin_reference_host_matches_oracle(oracle, size, fixed, n):
    """The reference's own frame loop, mode decision, rate control and writer
    (compiled from its sources) over libh261gpu.so: bitstream and
    reconstruction byte-identical to the oracle."""
    from oracle.binding import REF_DROPIN_SO, RefDropin
    import os
    if not os.path.exists(REF_DROPIN_SO):
        pytest.skip("oracle/_ref drop-in build not present")
    w, h = size
    frames = seq(31, size, n, sigma=2, motion=7)
    flat = np.stack([np.concatenate([p.ravel() for p in f]) for f in frames])
    announce(f"drop-in encode {size} fixed={fixed} n={n}")
    bits, sizes, recon, quants, _ = RefDropin().encode(flat, w, h, fixed_quant=fixed)
    os_ = oracle.stream(w, h, fixed_quant=fixed)
    off = 0
    for t, f in enumerate(frames):
        ob, orec, _, _, oq = os_.encode(*f)
        assert quants[t] == oq, f"frame {t}: quantiser {quants[t]} vs {oq}"
        assert bits[off:off + sizes[t]] == ob, f"frame {t}: bitstream differs ({sizes[t]} vs {len(ob)} bytes)"
        assert np.array_equal(recon[t], np.concatenate([p.ravel() for p in orec])), f"frame {t}: recon differs"
        off += sizes[t]
    assert off == len(bits)


def run_engine_host_fed(eng, sid, frames):
    out = []
    for t, f in enumerate(frames):
        eng.put_frame(sid, *f)
        eng.step()
        eng.sync()
        info = eng.info(sid)
        bs = eng.bitstream(sid)
        out.append((bs, eng.recon(sid), info.quant, info.bytes, eng.macroblocks(sid)))
    return out


@pytest.mark.parametrize("variant", [0, 1], ids=["two-stage", "surface"])
@pytest.mark.parametrize("size,fixed", [(CIF, 8), (QCIF, 0), (CIF, 0)], ids=["cif-q8", "qcif-rc", "cif-rc"])
def test_engine_single_stream_matches_oracle(oracle, size, fixed, variant):
    w, h = size
    n = 10
    frames = seq(41, size, n, sigma=2, motion=7)
    eng = Engine(max_streams=2, me_variant=variant)
    sid = eng.open(w, h, fixed_quant=fixed)
    announce(f"engine single {size} fixed={fixed}")
    res = run_engine_host_fed(eng, sid, frames)
    os_ = oracle.stream(w, h, fixed_quant=fixed)
    for t, f in enumerate(frames):
        ob, orec, _, ombs, oq = os_.encode(*f, want_details=True)
        bs, rec, q, nbytes, mbs = res[t]
        assert q == oq, f"frame {t}: quantiser {q} vs {oq}"
        assert nbytes == len(ob), f"frame {t}: device frame size {nbytes} vs {len(ob)}"
        assert bs == ob, f"frame {t}: bitstream differs"
        for p in range(3):
            assert np.array_equal(rec[p], orec[p]), f"frame {t}: recon plane {p} differs"
    eng.destroy()


def extreme_frames(size, rng):
    """Full-swing content: flat 0 and 255 frames, 0/255 checkerboards and
    stripes and their inversions, uniform noise.  Residuals reach +-255; at
    quantiser 1 about 11,000 AC levels per sequence sit at the +-127 clamp,
    SATD costs reach about 110,000 and CIF frames reach 320 KB."""
    w, h = size
    shapes = ((h, w), (h // 2, w // 2), (h // 2, w // 2))
    yy, xx = np.mgrid[0:h, 0:w]
    checker = (((yy >> 3) + (xx >> 3)) & 1).astype(np.uint8) * 255
    stripes = ((xx & 1) * 255).astype(np.uint8)

    def planes(y):
        return (y, y[::2, ::2].copy(), (255 - y[::2, ::2]).copy())

    flat0 = tuple(np.zeros(s, np.uint8) for s in shapes)
    flat255 = tuple(np.full(s, 255, np.uint8) for s in shapes)
    noise = tuple(rng.integers(0, 256, s, dtype=np.uint8) for s in shapes)
    shifted = tuple(np.roll(p, (5, -7) if i == 0 else (2, -3), axis=(0, 1)) for i, p in enumerate(noise))
    return [flat0, flat255, planes(checker), planes(255 - checker), noise, shifted, planes(stripes), flat0]


@pytest.mark.parametrize("variant", [0, 1], ids=["two-stage", "surface"])
@pytest.mark.parametrize("size,fixed", [(CIF, 1), (QCIF, 31), (CIF, 0)], ids=["cif-q1", "qcif-q31", "cif-rc"])
def test_engine_extreme_content_matches_oracle(oracle, size, fixed, variant):
    w, h = size
    frames = extreme_frames(size, np.random.default_rng(97))
    eng = Engine(max_streams=2, me_variant=variant)
    sid = eng.open(w, h, fixed_quant=fixed)
    announce(f"engine extreme content {size} fixed={fixed} variant={variant}")
    res = run_engine_host_fed(eng, sid, frames)
    os_ = oracle.stream(w, h, fixed_quant=fixed)
    for t, f in enumerate(frames):
        ob, orec, _, _, oq = os_.encode(*f, want_details=True)
        bs, rec, q, nbytes, _ = res[t]
        assert q == oq, f"frame {t}: quantiser {q} vs {oq}"
        assert nbytes == len(ob), f"frame {t}: device frame size {nbytes} vs {len(ob)}"
        assert bs == ob, f"frame {t}: bitstream differs"
        for p in range(3):
            assert np.array_equal(rec[p], orec[p]), f"frame {t}: recon plane {p} differs"
    eng.destroy()


def test_engine_batched_matches_oracle_and_alone(oracle):
    """Mixed CIF/QCIF streams with device-generated input, rate control on,
    coded together: every stream equals the oracle, and equals itself coded
    alone."""
    specs = [(CIF, 101, 0, 256000), (QCIF, 102, 0, 64000), (CIF, 103, 8, 0), (QCIF, 104, 0, 384000),
             (QCIF, 105, 12, 0), (CIF, 106, 0, 1024000)]
    n = 8
    eng = Engine(max_streams=8, me_variant=1)
    sids = []
    for size, seed, fixed, rate in specs:
        sid = eng.open(*size, fixed_quant=fixed, bitrate=rate or 512000, vbv_size=rate or 512000)
        eng.generate_input(sid, seed, n, sigma=2, motion=15)
        sids.append(sid)
    announce("engine batched 6 streams")
    got = {sid: [] for sid in sids}
    inputs = {sid: [] for sid in sids}
    for t in range(n):
        for sid in sids:
            inputs[sid].append(eng.input(sid))
        eng.step()
        eng.sync()
        for sid in sids:
            got[sid].append((eng.bitstream(sid), eng.recon(sid), eng.info(sid)))
    for (size, seed, fixed, rate), sid in zip(specs, sids):
        w, h = size
        os_ = oracle.stream(w, h, fixed_quant=fixed, bitrate=rate or 512000, vbv=rate or 512000)
        for t in range(n):
            f = content_frame(seed, w, h, t, sigma=2, motion=15)
            for p in range(3):
                assert np.array_equal(inputs[sid][t][p], f[p]), f"stream {sid} frame {t}: device content differs"
            ob, orec, _, _, oq = os_.encode(*f)
            bs, rec, info = got[sid][t]
            assert info.quant == oq and info.bytes == len(ob), f"stream {sid} frame {t}"
            assert bs == ob, f"stream {sid} frame {t}: bitstream differs"
            for p in range(3):
                assert np.array_equal(rec[p], orec[p]), f"stream {sid} frame {t}: recon differs"
    eng.destroy()
    # the same stream alone
    size, seed, fixed, rate = specs[1]
    alone = Engine(max_streams=1)
    sid = alone.open(*size, fixed_quant=fixed, bitrate=rate, vbv_size=rate)
    alone.generate_input(sid, seed, n, sigma=2, motion=15)
    announce("engine alone")
    for t in range(n):
        alone.step()
        alone.sync()
        assert alone.bitstream(sid) == got[sids[1]][t][0], f"alone vs batched, frame {t}"
    alone.destroy()


def test_engine_first_stream_closed_keeps_others_exact(oracle):
    """The stream holding MB 0 closes while the others go on (its range stays
    unmapped, no re-packing): the others' bitstreams and reconstructions
    still equal the oracle's, in steps and in graph replays.  Covers the
    kernels' handling of lanes and slots with nothing to do, which must not
    read through an unmapped MB's stream."""
    specs = [(QCIF, 421), (CIF, 422), (QCIF, 423)]
    n = 7
    eng = Engine(max_streams=len(specs))
    ids = []
    for size, seed in specs:
        sid = eng.open(*size, bitrate=160000, vbv_size=160000)
        eng.generate_input(sid, seed, n, sigma=3, motion=15)
        ids.append(sid)
    got = {k: [] for k in range(1, len(specs))}
    announce("first stream closed: 2 steps, close, 2 steps, graph replay of 3")
    for t in range(n):
        if t == 2:
            eng.close(ids[0])
        if t < 4:
            eng.step()
        elif t == 4:
            eng.run(3)
        eng.sync()
        if t < 4 or t == n - 1:
            for k in got:
                got[k].append((t, eng.bitstream(ids[k])))
    for k in got:
        size, seed = specs[k]
        s = oracle.stream(*size, bitrate=160000, vbv=160000)
        want = [s.encode(*content_frame(seed, *size, t, sigma=3, motion=15))[0] for t in range(n)]
        for t, bits in got[k]:
            assert bits == want[t], f"stream {k} frame {t}"
    eng.destroy()


def test_engine_graph_run_matches_stepwise():
    """h261g_engine_run (HIP graph replay) codes the same frames as step()."""
    n = 6
    a = Engine(max_streams=3)
    b = Engine(max_streams=3)
    ids_a, ids_b = [], []
    for size, seed in ((CIF, 201), (QCIF, 202), (QCIF, 203)):
        for eng, ids in ((a, ids_a), (b, ids_b)):
            sid = eng.open(*size)
            eng.generate_input(sid, seed, n + 2)
            ids.append(sid)
    announce("engine graph vs steps")
    for _ in range(n):
        a.step()
    a.sync()
    b.run(n)
    b.sync()
    for sa, sb in zip(ids_a, ids_b):
        assert a.bitstream(sa) == b.bitstream(sb)
        ra, rb = a.recon(sa), b.recon(sb)
        for p in range(3):
            assert np.array_equal(ra[p], rb[p])
        assert a.frames_done(sa) == b.frames_done(sb) == n
    a.destroy()
    b.destroy()


def test_put_frames_matches_put_frame(oracle):
    """Host-fed input for many streams in one call (h261g_engine_put_frames:
    the batch packed in one host buffer, one copy and one kernel, nothing
    waits) codes exactly what one h261g_stream_put_frame per stream codes,
    from pageable and from pinned host memory, with frames in any order and
    some streams skipping a step; one stream also against the oracle.  An
    invalid call queues nothing."""
    import torch
    specs = [(CIF, 501, 0, 256000), (QCIF, 502, 0, 96000), (QCIF, 503, 8, 0), (CIF, 504, 8, 0)]
    n = 6
    extra = 3
    frames = [seq(seed, size, n + extra, sigma=2, motion=9) for size, seed, _, _ in specs]
    engines = []
    for _ in range(3):
        eng = Engine(max_streams=8)
        sids = [eng.open(*size, fixed_quant=fixed, bitrate=rate or 512000, vbv_size=rate or 512000)
                for size, _, fixed, rate in specs]
        engines.append((eng, sids))
    (a, ia), (b, ib), (c, ic) = engines
    announce("put_frames: per-stream, packed pageable, packed pinned")

    size0, _, fixed0, rate0 = specs[0]
    os_ = oracle.stream(*size0, fixed_quant=fixed0, bitrate=rate0, vbv=rate0)
    # stream 1 skips step 2, stream 2 skips step 3 (no input: not coded)
    skip = {(1, 2), (2, 3)}
    done = [0] * len(specs)
    for t in range(n):
        fed = [k for k in range(len(specs)) if (k, t) not in skip]
        for k in fed:
            a.put_frame(ia[k], *frames[k][done[k]])
        # packed in reverse order, so offsets do not follow the ids
        order = fed[::-1]
        buf, offs = Engine.pack_frames([frames[k][done[k]] for k in order])
        b.put_frames([ib[k] for k in order], buf, offs)
        pin = torch.empty(buf.size, dtype=torch.uint8, pin_memory=True)
        pbuf = pin.numpy()
        pbuf[...] = buf
        c.put_frames([ic[k] for k in order], pbuf, offs)
        for eng, _ in engines:
            eng.step()
            eng.sync()
        del pin, pbuf  # the pinned buffer may go once the step has finished
        for k in fed:
            want = a.bitstream(ia[k])
            assert b.bitstream(ib[k]) == want, f"packed pageable, step {t} stream {k}"
            assert c.bitstream(ic[k]) == want, f"packed pinned, step {t} stream {k}"
            assert a.info(ia[k]).quant == b.info(ib[k]).quant == c.info(ic[k]).quant
            ra, rb, rc = a.recon(ia[k]), b.recon(ib[k]), c.recon(ic[k])
            for p in range(3):
                assert np.array_equal(ra[p], rb[p]) and np.array_equal(ra[p], rc[p]), f"step {t} stream {k} recon"
            done[k] += 1
        ob = os_.encode(*frames[0][t])[0]
        assert c.bitstream(ic[0]) == ob, f"packed pinned vs oracle, frame {t}"
    for k, (eng, sids) in enumerate(engines):
        assert [eng.frames_done(s) for s in sids] == done, f"engine {k}"

    # steps queued back to back: each call's copy runs beside the steps
    # queued before it, into the staging buffer the call before last used;
    # nothing is synchronised until the end (every pinned buffer kept)
    keep = []
    for t in range(extra):
        for k in range(len(specs)):
            a.put_frame(ia[k], *frames[k][done[k] + t])
        a.step()
        a.sync()
        buf, offs = Engine.pack_frames([frames[k][done[k] + t] for k in range(len(specs))])
        pin = torch.empty(buf.size, dtype=torch.uint8, pin_memory=True)
        pin.numpy()[...] = buf
        keep.append(pin)
        c.put_frames(ic, pin.numpy(), offs)
        c.step()
    c.sync()
    del keep
    done = [d + extra for d in done]
    for k in range(len(specs)):
        assert c.bitstream(ic[k]) == a.bitstream(ia[k]), f"queued steps: stream {k}"
        ra, rc = a.recon(ia[k]), c.recon(ic[k])
        for p in range(3):
            assert np.array_equal(ra[p], rc[p]), f"queued steps: stream {k} recon"
    assert [c.frames_done(s) for s in ic] == done

    # invalid calls queue nothing: a repeated id, a stream with device input,
    # an id that is not open, an offset that is not a multiple of 16, a
    # frame past the buffer's end
    buf, offs = Engine.pack_frames([frames[0][0], frames[0][1]])  # two CIF frames
    with pytest.raises(RuntimeError):
        b.put_frames([ib[0], ib[0]], buf, offs)
    gen = b.open(*QCIF)
    b.generate_input(gen, 7, 2)
    with pytest.raises(RuntimeError):
        b.put_frames([ib[1], gen], buf, offs)
    with pytest.raises(RuntimeError):
        b.put_frames([ib[1], 7], buf, offs)
    with pytest.raises(RuntimeError):
        b.put_frames([ib[1]], buf, [8])
    with pytest.raises(RuntimeError):
        b.put_frames([ib[0], ib[3]], buf, [offs[1], offs[1] + 16])
    b.step()
    b.sync()
    assert [b.frames_done(s) for s in ib] == [d - extra for d in done], "a rejected call activated a stream"
    assert b.frames_done(gen) == 1
    for eng, _ in engines:
        eng.destroy()


def test_put_frames_more_streams_than_one_placement_launch():
    """One h261g_engine_put_frames call for 300 QCIF streams (more than the
    256 a placement kernel takes, so two launches share the staged copy),
    with two streams left out of the call, codes what one put_frame per
    stream codes, over three steps."""
    ns = 300
    frames = [seq(900 + k, QCIF, 3, sigma=2, motion=5) for k in range(ns)]
    a, b = Engine(max_streams=ns), Engine(max_streams=ns)
    ia = [a.open(*QCIF, fixed_quant=8) for _ in range(ns)]
    ib = [b.open(*QCIF, fixed_quant=8) for _ in range(ns)]
    announce(f"put_frames: {ns} QCIF streams in one call")
    left_out = {7, 263}
    for t in range(3):
        fed = [k for k in range(ns) if k not in left_out or t == 0]
        for k in fed:
            a.put_frame(ia[k], *frames[k][t])
        order = fed[::-1]
        buf, offs = Engine.pack_frames([frames[k][t] for k in order])
        b.put_frames([ib[k] for k in order], buf, offs)
        a.step()
        b.step()
        a.sync()
        b.sync()
        for k in fed:
            assert b.bitstream(ib[k]) == a.bitstream(ia[k]), f"step {t} stream {k}"
    assert [b.frames_done(s) for s in ib] == [a.frames_done(s) for s in ia]
    assert b.frames_done(ib[7]) == 1 and b.frames_done(ib[8]) == 3
    a.destroy()
    b.destroy()


def test_put_frames_at_benchmark_scale_matches_generated_input():
    """512 host-fed streams (256 CIF + 256 QCIF: two chains, full-size
    launches) fed by one h261g_engine_put_frames call per step code exactly
    what the same streams code from device-generated input, every stream's
    bitstream compared at every step."""
    ns, steps = 512, 3
    sizes = [CIF] * 256 + [QCIF] * 256
    a, b = Engine(max_streams=ns), Engine(max_streams=ns)
    ia, ib = [], []
    for k, size in enumerate(sizes):
        rate = 64000 + 1000 * (k % 320)
        ia.append(a.open(*size, bitrate=rate, vbv_size=rate))
        ib.append(b.open(*size, bitrate=rate, vbv_size=rate))
        a.generate_input(ia[-1], 700 + k, steps, sigma=2, motion=15)
    announce(f"put_frames at scale: {ns} streams, {steps} steps")
    for t in range(steps):
        buf, offs = Engine.pack_frames([content_frame(700 + k, *sizes[k], t, sigma=2, motion=15)
                                        for k in range(ns)])
        b.put_frames(ib, buf, offs)
        a.step()
        b.step()
        a.fetch_bitstreams(ia)
        got_a = a.fetched()
        b.fetch_bitstreams(ib)
        got_b = b.fetched()
        for k in range(ns):
            assert got_b[k] == got_a[k], f"step {t} stream {k}"
    assert [b.frames_done(s) for s in ib] == [steps] * ns
    a.destroy()
    b.destroy()


@pytest.mark.parametrize("local", [False, True], ids=["global-motion", "local-motion-scene-cuts"])
def test_generated_input_equals_host_generator(local):
    """Frames that h261g_stream_generate_input writes on the device equal the
    host generator's (content_frame) plane for plane, frame after frame: CIF
    and QCIF, global motion, or local motion across scene cuts."""
    kw = dict(sigma=4, motion=15, local=True, scene_cut=3) if local else dict(sigma=2, motion=15)
    specs = [(CIF, 931), (QCIF, 932)]
    n = 7
    eng = Engine(max_streams=len(specs))
    sids = [eng.open(*size, fixed_quant=8) for size, _ in specs]
    for sid, (size, seed) in zip(sids, specs):
        eng.generate_input(sid, seed, n, **kw)
    announce(f"generated input vs host generator, local={local}")
    for t in range(n):
        for sid, (size, seed) in zip(sids, specs):
            got = eng.input(sid)
            want = content_frame(seed, *size, t, **kw)
            for p in range(3):
                assert np.array_equal(got[p], want[p]), f"frame {t} {size} plane {p}"
        eng.step()
        eng.sync()
    eng.destroy()


def test_engine_api_edges(oracle):
    """The engine's edges: an engine with nothing open steps, replays and
    syncs; invalid sizes, quantisers, options, ids and input calls are
    refused and change nothing; a full engine refuses another stream; a
    host-fed stream without a new frame and a generated stream past its last
    frame are not coded; a closed id is refused and its slot is reused."""
    from h261enc_cuda_amd import _native as nat
    eng = Engine(max_streams=2)
    announce("engine API edges")
    assert eng.step() == 0
    eng.run(0)
    eng.run(3)
    eng.sync()
    for w, h in ((320, 240), (352, 144), (0, 0), (176, 288)):
        with pytest.raises(RuntimeError):
            eng.open(w, h)
    with pytest.raises(RuntimeError):
        eng.open(*QCIF, init_quant=0)
    with pytest.raises(RuntimeError):
        eng.open(*QCIF, fixed_quant=32)
    with pytest.raises(RuntimeError):
        eng.config(999, 0)
    with pytest.raises(RuntimeError):
        eng.config(nat.OPT_XCD_REMAP, 8)
    with pytest.raises(RuntimeError):
        eng.info(0)  # nothing open
    a = eng.open(*QCIF, fixed_quant=8)
    b = eng.open(*CIF, fixed_quant=8)
    assert (a, b) == (0, 1)
    with pytest.raises(RuntimeError):
        eng.open(*QCIF)  # both slots taken
    with pytest.raises(RuntimeError):
        eng.generate_input(a, 1, 0)
    eng.generate_input(b, 5, 2, sigma=2, motion=7)
    with pytest.raises(RuntimeError):
        eng.put_frame(b, *zeros(CIF))  # a device-generated stream takes no host frames
    with pytest.raises(RuntimeError):
        eng.generate_input(b, 6, 2)  # a stream gets one generated sequence
    fa = seq(77, QCIF, 1, sigma=2, motion=7)[0]
    eng.put_frame(a, *fa)
    eng.step()
    eng.sync()
    ob = oracle.stream(*QCIF, fixed_quant=8).encode(*fa)[0]
    assert eng.bitstream(a) == ob
    first_a = eng.bitstream(a)
    # a has no new frame: only b is coded; then b's sequence has ended
    for _ in range(3):
        eng.step()
        eng.sync()
    assert (eng.frames_done(a), eng.frames_done(b)) == (1, 2)
    assert eng.bitstream(a) == first_a
    ob_b = oracle.stream(*CIF, fixed_quant=8)
    for t in range(2):
        want_b = ob_b.encode(*content_frame(5, *CIF, t, sigma=2, motion=7))[0]
    assert eng.bitstream(b) == want_b
    eng.close(a)
    with pytest.raises(RuntimeError):
        eng.close(a)
    with pytest.raises(RuntimeError):
        eng.bitstream(a)
    with pytest.raises(RuntimeError):
        eng.put_frame(a, *fa)
    c = eng.open(*QCIF, fixed_quant=8)
    assert c == a
    eng.put_frame(c, *fa)
    eng.step()
    eng.sync()
    assert eng.bitstream(c) == ob and eng.frames_done(c) == 1
    eng.destroy()


def test_engine_two_chains_match_one_chain(oracle):
    """Two chains (the MB space split at a stream boundary, each half stepped
    on its own HIP stream), tied by the strict ladder or by the ladder with
    one step of slack, code every stream as one chain does, stepwise and
    under graph replay, across a close and a re-packing; one stream is also
    checked against the oracle."""
    from h261enc_cuda_amd._native import OPT_CHAIN_LADDER, OPT_CHAINS, OPT_FUSED_TAILS
    specs = [(CIF, 301, 0, 256000), (QCIF, 302, 0, 64000), (QCIF, 303, 6, 0), (CIF, 304, 0, 384000),
             (QCIF, 305, 0, 128000)]
    n = 10
    engines = []
    # one chain; two chains with the strict ladder and with one step of slack
    for chains, ladder in ((1, 0), (2, 1), (2, 2)):
        eng = Engine(max_streams=8)
        eng.config(OPT_CHAINS, chains)
        eng.config(OPT_CHAIN_LADDER, ladder)
        if chains == 2:  # each half's rate-control begin as its own launch, as in full-size steps
            eng.config(OPT_FUSED_TAILS, 2)
        sids = []
        for size, seed, fixed, rate in specs:
            sid = eng.open(*size, fixed_quant=fixed, bitrate=rate or 512000, vbv_size=rate or 512000)
            eng.generate_input(sid, seed, 3 * n, sigma=2, motion=15)
            sids.append(sid)
        engines.append((eng, sids))
    announce("engine two chains vs one chain")

    def compare(what):
        a, ia = engines[0]
        for j, (b, ib) in enumerate(engines[1:], 1):
            for k, (sa, sb) in enumerate(zip(ia, ib)):
                if sa is None:
                    continue
                assert a.bitstream(sa) == b.bitstream(sb), f"{what}: engine {j} stream {k} bitstream"
                assert a.info(sa).quant == b.info(sb).quant, f"{what}: engine {j} stream {k} quantiser"
                ra, rb = a.recon(sa), b.recon(sb)
                for p in range(3):
                    assert np.array_equal(ra[p], rb[p]), f"{what}: engine {j} stream {k} recon"

    size, seed, fixed, rate = specs[3]
    os_ = oracle.stream(*size, fixed_quant=fixed, bitrate=rate, vbv=rate)
    for t in range(n):
        for eng, _ in engines:
            eng.step()
            eng.sync()
        compare(f"step {t}")
        ob, _, _, _, _ = os_.encode(*content_frame(seed, *size, t, sigma=2, motion=15))
        for eng, sids in engines[1:]:
            assert eng.bitstream(sids[3]) == ob, f"two chains vs oracle, frame {t}"
    for eng, _ in engines:
        eng.run(5)
        eng.sync()
    compare("graph replay")
    for eng, sids in engines:
        eng.close(sids[1])
        sids[1] = None
        eng.compact()
        eng.run(4)
        eng.step()
        eng.sync()
    compare("after close and re-packing")
    for eng, _ in engines:
        eng.destroy()


@pytest.mark.skip(reason="two-chain steps with each chain's writer on a side stream beside its next search "
                  "are switched off: their first graph capture ended the process with a host segfault in "
                  "h261g_engine_run (the second chain's writer stream forked from the second chain's stream); "
                  "cause not found")
def test_engine_two_chains_overlapped_writers_match_one_chain():
    """Two chains whose bitstream writers run beside each chain's next search
    (OPT_OVERLAP_WRITER 1) code every stream as one chain does, stepwise and
    under graph replay."""
    from h261enc_cuda_amd._native import OPT_CHAIN_LADDER, OPT_CHAINS, OPT_FUSED_TAILS, OPT_OVERLAP_WRITER
    specs = [(CIF, 301, 0, 256000), (QCIF, 302, 0, 64000), (QCIF, 303, 6, 0), (CIF, 304, 0, 384000),
             (QCIF, 305, 0, 128000)]
    engines = []
    for chains, ladder, overlap in ((1, 0, 2), (2, 1, 1), (2, 2, 1)):
        eng = Engine(max_streams=8)
        eng.config(OPT_CHAINS, chains)
        eng.config(OPT_CHAIN_LADDER, ladder)
        eng.config(OPT_OVERLAP_WRITER, overlap)
        if chains == 2:
            eng.config(OPT_FUSED_TAILS, 2)
        sids = []
        for size, seed, fixed, rate in specs:
            sid = eng.open(*size, fixed_quant=fixed, bitrate=rate or 512000, vbv_size=rate or 512000)
            eng.generate_input(sid, seed, 20, sigma=2, motion=15)
            sids.append(sid)
        engines.append((eng, sids))
    announce("engine two chains, overlapped writers, vs one chain")
    for what, steps in (("stepwise", 4), ("graph replay", 5)):
        for eng, _ in engines:
            if what == "stepwise":
                for _ in range(steps):
                    eng.step()
            else:
                eng.run(steps)
            eng.sync()
        a, ia = engines[0]
        for j, (b, ib) in enumerate(engines[1:], 1):
            for k, (sa, sb) in enumerate(zip(ia, ib)):
                assert a.bitstream(sa) == b.bitstream(sb), f"{what}: engine {j} stream {k} bitstream"
    for eng, _ in engines:
        eng.destroy()


@pytest.mark.parametrize("fused", [1, 2], ids=["rc-begin-in-me", "rc-begin-launch"])
def test_engine_me_runs_agree(oracle, fused):
    """Motion estimation with 1, 2, 4 or 8 consecutive MBs per wave (runs
    that cross stream boundaries, one cost atomic per stream segment, and in
    the fused form the wave completing a stream's count starting its
    picture) and search-window tiles of 1, 2 or 4 MBs (tiles cut at row and
    stream ends, at the right frame edge, at the top and bottom rows) codes
    the same bytes as the oracle."""
    from h261enc_cuda_amd._native import OPT_FUSED_TAILS, OPT_ME_MBS_PER_WAVE, OPT_ME_TILE
    specs = [(QCIF, 311), (CIF, 312), (QCIF, 313), (CIF, 314), (QCIF, 315)]
    n = 4
    results = []
    for runs, tile in ((1, 1), (2, 1), (2, 2), (4, 1), (4, 2), (4, 4), (8, 2), (8, 4)):
        announce(f"motion estimation {runs} MBs per wave, tile {tile}, fused={fused}")
        eng = Engine(max_streams=len(specs))
        eng.config(OPT_FUSED_TAILS, fused)
        eng.config(OPT_ME_MBS_PER_WAVE, runs)
        eng.config(OPT_ME_TILE, tile)
        ids = []
        for size, seed in specs:
            sid = eng.open(*size, bitrate=96000, vbv_size=96000)
            eng.generate_input(sid, seed, n, sigma=3, motion=15)
            ids.append(sid)
        out = []
        for _ in range(n):
            eng.step()
            eng.sync()
            out.append([(eng.bitstream(sid), eng.info(sid).quant) for sid in ids])
        eng.destroy()
        results.append(out)
    for r in results[1:]:
        assert r == results[0]
    for k, (size, seed) in enumerate(specs):
        s = oracle.stream(*size, bitrate=96000, vbv=96000)
        for t in range(n):
            ob, _, _, _, oq = s.encode(*content_frame(seed, *size, t, sigma=3, motion=15))
            assert results[0][t][k] == (ob, oq), f"stream {k} frame {t}"


def test_engine_launch_shapes_agree(oracle):
    """Fused and unfused frame tails, rate-control begin in the ME tail or as
    its own launch, 1/3/8 MBs per wave (workgroups that span several
    streams) and both search kernels code identical bytes."""
    from h261enc_cuda_amd._native import OPT_FUSED_TAILS, OPT_MBS_PER_WAVE
    specs = [(QCIF, 301), (CIF, 302), (QCIF, 303), (QCIF, 304), (CIF, 305)]
    n = 5
    shapes = [(1, 1, 0), (0, 3, 0), (1, 8, 1), (0, 1, 1), (2, 4, 0), (2, 1, 1)]
    results = []
    for fused, mpw, variant in shapes:
        announce(f"launch shape fused={fused} mbs_per_wave={mpw} variant={variant}")
        eng = Engine(max_streams=len(specs), me_variant=variant)
        eng.config(OPT_FUSED_TAILS, fused)
        eng.config(OPT_MBS_PER_WAVE, mpw)
        ids = []
        for size, seed in specs:
            sid = eng.open(*size, bitrate=128000, vbv_size=128000)
            eng.generate_input(sid, seed, n, sigma=3, motion=15)
            ids.append(sid)
        out = []
        for _ in range(n):
            eng.step()
            eng.sync()
            out.append([eng.bitstream(sid) for sid in ids])
        eng.destroy()
        results.append(out)
    for r in results[1:]:
        assert r == results[0]
    w, h = specs[1][0]
    s = oracle.stream(w, h, bitrate=128000, vbv=128000)
    for t in range(n):
        f = content_frame(specs[1][1], w, h, t, sigma=3, motion=15)
        assert s.encode(*f)[0] == results[0][t][1], f"frame {t}"
