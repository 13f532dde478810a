"""The bitstream written on the device (one wave per coded macroblock, placed
by the frame tail) against the host writer on the same macroblocks, whose
bytes the golden fixtures pin to the reference writer; and the project
decoder's closed loop on it: decode(device bitstream) == device
reconstruction.  Bar: bit-exact."""
import numpy as np
import pytest

from conftest import announce
from h261enc_cuda_amd import Engine, content_frame, write_picture
from h261enc_cuda_amd import _native as nat
from h261enc_cuda_amd.decoder import Decoder
from h261enc_cuda_amd.encoder import CIF, QCIF

pytestmark = pytest.mark.gpu


def host_bits(eng, sid):
    w, h = eng.sizes[sid]
    info = eng.info(sid)
    return write_picture(eng.macroblocks(sid), w, h, info.frame_num, info.quant)


def check_stream(eng, sid, what):
    info = eng.info(sid)
    dev = eng.bitstream(sid)
    ref = host_bits(eng, sid)
    assert len(dev) == info.bytes, what
    if dev != ref:
        i = next(k for k in range(min(len(dev), len(ref))) if dev[k] != ref[k]) if dev[:len(ref)] != ref else len(ref)
        raise AssertionError(f"{what}: {len(dev)} vs {len(ref)} bytes, first difference at byte {i}")
    return dev


def test_device_bitstream_batched_mix():
    """32 streams (CIF and QCIF, rate control, local motion and scene cuts),
    stepped one at a time and by graph replay"""
    eng = Engine(max_streams=32)
    try:
        rng = np.random.default_rng(5)
        sids = []
        for i in range(32):
            size = CIF if i % 2 else QCIF
            sid = eng.open(*size, init_quant=int(rng.integers(4, 20)), bitrate=int(rng.integers(64, 1024)) * 1000,
                           vbv_size=int(rng.integers(64, 1024)) * 1000)
            eng.generate_input(sid, seed=100 + i, n_frames=40, sigma=4, motion=15, local=bool(i % 3 == 0),
                               scene_cut=7)
            sids.append(sid)
        decoders = {sid: Decoder(*eng.sizes[sid]) for sid in sids}
        for t in range(4):
            announce(f"device bitstream batched step {t}")
            eng.step()
            eng.sync()
            for sid in sids:
                bits = check_stream(eng, sid, f"stream {sid} frame {t}")
                dec, _, tr = decoders[sid].decode(bits)
                assert tr == t & 31
                rec = eng.recon(sid)
                for p in range(3):
                    assert np.array_equal(dec[p], rec[p]), (sid, t, p)
        announce("device bitstream batched run(5)")
        eng.run(5)
        eng.sync()
        for sid in sids:
            assert eng.info(sid).frame_num == 8
            check_stream(eng, sid, f"stream {sid} after run")
    finally:
        eng.destroy()


@pytest.mark.parametrize("fused", [1, 0], ids=["fused", "separate"])
def test_device_bitstream_large_frames(fused):
    """quantiser 1 (escape codes, 63-coefficient blocks): with worst-case
    buffers the device writes every frame; equal to the host writer and to
    the bytes with no device bitstream"""
    sizes = [CIF, QCIF, CIF]
    out = {}
    for mode in (2, 0):
        eng = Engine(max_streams=4)
        try:
            eng.config(nat.OPT_DEVICE_BITSTREAM, mode)
            eng.config(nat.OPT_FUSED_TAILS, fused)
            sids = [eng.open(*s, fixed_quant=1 + 7 * k) for k, s in enumerate(sizes)]
            got = []
            for t in range(3):
                announce(f"device bitstream large frames mode {mode} fused {fused} frame {t}")
                for k, sid in enumerate(sids):
                    w, h = sizes[k]
                    eng.put_frame(sid, *content_frame(40 + k, w, h, t, sigma=30, motion=15))
                eng.step()
                eng.sync()
                for sid in sids:
                    got.append(check_stream(eng, sid, f"mode {mode} stream {sid} frame {t}"))
            out[mode] = got
        finally:
            eng.destroy()
    assert out[2] == out[0]
    assert max(len(b) for b in out[2]) > 128 * 1024  # escape-heavy frames past the default buffer


def test_device_bitstream_overflow_falls_back_to_host_writer():
    """a QCIF intra frame at quantiser 1 outgrows the default 32 KiB buffer:
    its bitstream comes from the host writer, the next (small) frames from
    the device again; all equal the host writer's bytes"""
    eng = Engine(max_streams=2)
    try:
        sid = eng.open(*QCIF, fixed_quant=1)
        frame = content_frame(77, *QCIF, 0, sigma=16, motion=3)
        sizes = []
        for t in range(3):  # the same picture again: small inter frames
            announce(f"device bitstream overflow frame {t}")
            eng.put_frame(sid, *frame)
            eng.step()
            eng.sync()
            sizes.append(len(check_stream(eng, sid, f"overflow frame {t}")))
        assert sizes[0] > 32 * 1024 and sizes[1] < 32 * 1024, sizes
    finally:
        eng.destroy()


def test_device_bitstream_option_needs_no_open_stream():
    eng = Engine(max_streams=2)
    try:
        eng.open(*QCIF)
        with pytest.raises(RuntimeError):
            eng.config(nat.OPT_DEVICE_BITSTREAM, 0)
    finally:
        eng.destroy()


def test_bulk_fetch_matches_per_stream_reads():
    """h261g_engine_fetch_bitstreams delivers every listed stream's last
    frame in one queued operation, the same bytes as the per-stream read,
    including a frame that outgrew its device buffer (host-written)"""
    eng = Engine(max_streams=16)
    try:
        sids = []
        for i in range(12):
            size = CIF if i % 3 else QCIF
            sid = eng.open(*size, bitrate=int(96000 + 40000 * i), vbv_size=int(96000 + 40000 * i))
            eng.generate_input(sid, 900 + i, 6, sigma=4, motion=15, local=bool(i % 2))
            sids.append(sid)
        big = eng.open(*QCIF, fixed_quant=1)
        big_frame = content_frame(77, *QCIF, 0, sigma=16, motion=3)
        for t in range(3):
            announce(f"bulk fetch step {t}")
            eng.put_frame(big, *big_frame)
            eng.step()
            order = sids[::-1] + [big] + sids[:3]
            eng.fetch_bitstreams(order)
            got = eng.fetched()
            eng.sync()
            want = [eng.bitstream(sid) for sid in order]
            assert got == want, t
            assert len(got[len(sids)]) > 32 * 1024 or t > 0
        eng.fetch_bitstreams([])
        assert eng.fetched() == []
    finally:
        eng.destroy()


def test_overlapped_writer_equals_in_step_writer():
    """The writer beside the next step's motion estimation (default) and at
    the end of its own step give the same bytes, across graph replays,
    closes, re-packing and opens between steps"""
    def job(overlap):
        eng = Engine(max_streams=12)
        out = []
        try:
            eng.config(nat.OPT_OVERLAP_WRITER, overlap)
            lengths = [3, 9, 5, 9, 2, 7, 9, 4]
            sids = []
            for i, n in enumerate(lengths):
                sid = eng.open(*(CIF if i % 2 else QCIF), bitrate=160000, vbv_size=160000)
                eng.generate_input(sid, 300 + i, n, sigma=4, motion=15, local=bool(i % 3 == 0), scene_cut=4)
                sids.append(sid)
            live = dict(zip(sids, lengths))
            done = 0
            for chunk in (1, 2, 3, 3):
                announce(f"overlapped writer {overlap} chunk {chunk} at {done}")
                eng.run(chunk)
                done += chunk
                eng.sync()
                out.append([(sid, eng.info(sid).frame_num, eng.bitstream(sid)) for sid in sorted(live)])
                for sid in [s for s, n in live.items() if n <= done]:
                    eng.close(sid)
                    del live[sid]
                eng.compact()
                if chunk == 2:  # a new stream joins after re-packing
                    sid = eng.open(*QCIF, bitrate=96000, vbv_size=96000)
                    eng.generate_input(sid, 777, 6, sigma=4, motion=15)
                    live[sid] = done + 6
            return out
        finally:
            eng.destroy()
    assert job(1) == job(0)
