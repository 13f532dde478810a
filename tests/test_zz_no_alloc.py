"""h261g_engine_fetch_bitstreams_into, the allocation-free bulk bitstream
fetch, through the C ABI: the same sizes, places and bytes as
h261g_engine_fetch_bitstreams on two calls in a row (the second with fewer
streams), buffers in device memory and in pinned host memory; ids of no open
stream give size -1; oversized and pageable arguments are rejected.

The test allocates its buffers itself (hipMalloc / hipHostMalloc from the
HIP runtime libh261gpu.so links); the entry point allocates nothing."""
import ctypes

import numpy as np
import pytest

from conftest import announce
from h261enc_cuda_amd import Engine
from h261enc_cuda_amd import _native as nat
from h261enc_cuda_amd.workloads import open_streams

pytestmark = pytest.mark.gpu

ERR_ARG = -1  # H261G_ERR_ARG

SPECS = [dict(size=(352, 288), length=12, bitrate=256000, seed=901, motion=15),
         dict(size=(176, 144), length=12, bitrate=64000, seed=902, motion=15),
         dict(size=(176, 144), length=12, bitrate=384000, seed=903, motion=15),
         dict(size=(352, 288), length=12, bitrate=128000, seed=904, motion=15),
         dict(size=(176, 144), length=12, bitrate=1024000, seed=905, motion=15, fixed=2)]


class Hip:
    """the HIP runtime libh261gpu.so is linked against (found in this
    process's mappings), for the test's own buffers"""

    def __init__(self):
        nat.load()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    path = line.split()[-1]
                    break
        assert path, "libamdhip64 not mapped"
        self.lib = ctypes.CDLL(path)
        self.bufs = []

    def device(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        self.bufs.append((p.value, False))
        return p.value

    def pinned(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipHostMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes), ctypes.c_uint(0)) == 0
        self.bufs.append((p.value, True))
        return p.value

    def to_host(self, ptr, nbytes):
        if any(q == ptr and pinned for q, pinned in self.bufs):
            return ctypes.string_at(ptr, nbytes)
        out = np.zeros(max(1, nbytes), np.uint8)
        if nbytes:
            assert self.lib.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr),
                                      ctypes.c_size_t(nbytes), 2) == 0  # device to host
        return out[:nbytes].tobytes()

    def to_device(self, ptr, data: bytes):
        assert self.lib.hipMemcpy(ctypes.c_void_p(ptr), ctypes.c_char_p(data), ctypes.c_size_t(len(data)), 1) == 0

    def free(self):
        for p, pinned in self.bufs:
            (self.lib.hipHostFree if pinned else self.lib.hipFree)(ctypes.c_void_p(p))
        self.bufs = []


def reference_fetch(eng, sids):
    """h261g_engine_fetch_bitstreams + h261g_engine_fetched: sizes, offsets,
    bytes"""
    lib = eng.lib
    ids = np.ascontiguousarray(np.asarray(sids, np.int32))
    assert lib.h261g_engine_fetch_bitstreams(eng.h, nat.ptr(ids), len(ids)) == 0
    data, sizes, offs = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    n = lib.h261g_engine_fetched(eng.h, ctypes.byref(data), ctypes.byref(sizes), ctypes.byref(offs))
    assert n == len(sids)
    sz = np.ctypeslib.as_array(ctypes.cast(sizes, ctypes.POINTER(ctypes.c_int32)), (n,)).copy()
    of = np.ctypeslib.as_array(ctypes.cast(offs, ctypes.POINTER(ctypes.c_int32)), (n + 1,)).copy()
    buf = ctypes.string_at(data, int(of[n]))
    return sz, of, buf


def fetch_into(hip, eng, sids, ids_ptr, out_ptr, cap, meta_ptr):
    n = len(sids)
    ids = np.asarray(sids, np.int32).tobytes()
    if ids_ptr_is_pinned(hip, ids_ptr):
        ctypes.memmove(ids_ptr, ids, len(ids))
    else:
        hip.to_device(ids_ptr, ids)
    rc = eng.lib.h261g_engine_fetch_bitstreams_into(eng.h, ctypes.c_void_p(ids_ptr), n, ctypes.c_void_p(out_ptr),
                                                     ctypes.c_int64(cap), ctypes.c_void_p(meta_ptr))
    assert rc == 0, rc
    eng.sync()
    meta = np.frombuffer(hip.to_host(meta_ptr, 4 * (2 * n + 1)), np.int32)
    sz, of = meta[:n], meta[n:]
    used = min(int(of[n]), cap)
    return sz.copy(), of.copy(), hip.to_host(out_ptr, used)


def ids_ptr_is_pinned(hip, p):
    return any(q == p and pinned for q, pinned in hip.bufs)


def check_equal(sids, got, want, cap=None):
    gs, go, gb = got
    ws, wo, wb = want
    assert list(go) == list(wo), "offsets differ"
    for i in range(len(sids)):
        if cap is not None and int(wo[i]) + ((int(ws[i]) + 3) & ~3) > cap:
            assert gs[i] == -1, f"stream {sids[i]}: past the capacity but size {gs[i]}"
            continue
        assert gs[i] == ws[i], f"stream {sids[i]}: size {gs[i]} vs {ws[i]}"
        if ws[i] > 0:
            assert gb[go[i]:go[i] + gs[i]] == wb[wo[i]:wo[i] + ws[i]], f"stream {sids[i]}: bytes differ"


def test_fetch_into_matches_fetch_bitstreams():
    hip = Hip()
    eng = Engine(max_streams=8)
    try:
        ids = open_streams(eng, SPECS, frames=12)
        announce("fetch_bitstreams_into: 5 streams, device buffers")
        for _ in range(3):
            eng.step()
        eng.sync()
        cap = 1 << 20
        ids_dev, out_dev, meta_dev = hip.device(4 * 64), hip.device(cap), hip.device(4 * 129)
        want = reference_fetch(eng, ids)
        assert all(want[0] > 0)
        got = fetch_into(hip, eng, ids, ids_dev, out_dev, cap, meta_dev)
        check_equal(ids, got, want)

        # second call: a later frame, fewer streams in another order, pinned
        # host buffers
        announce("fetch_bitstreams_into: 2 streams, pinned host buffers")
        eng.step()
        eng.sync()
        sub = [ids[3], ids[1]]
        ids_pin, out_pin, meta_pin = hip.pinned(4 * 64), hip.pinned(cap), hip.pinned(4 * 129)
        want2 = reference_fetch(eng, sub)
        got2 = fetch_into(hip, eng, sub, ids_pin, out_pin, cap, meta_pin)
        check_equal(sub, got2, want2)

        # a capacity that holds only the first frames: the rest get -1 at the
        # same places
        announce("fetch_bitstreams_into: small capacity, unknown ids")
        want3 = reference_fetch(eng, ids)
        small = int(want3[1][2])  # room for the first two frames
        got3 = fetch_into(hip, eng, ids, ids_dev, out_dev, small, meta_dev)
        check_equal(ids, got3, want3, cap=small)
        assert list(got3[0][2:]) == [-1] * (len(ids) - 2)

        # ids of no open stream: a closed one, a never-opened slot, out of range
        eng.close(ids[2])
        bad = [ids[2], 7, 12345, -3, ids[0]]
        gs, go, gb = fetch_into(hip, eng, bad, ids_dev, out_dev, cap, meta_dev)
        assert list(gs[:4]) == [-1, -1, -1, -1]
        w0 = reference_fetch(eng, [ids[0]])
        assert gs[4] == w0[0][0] and gb[go[4]:go[4] + gs[4]] == w0[2][:w0[0][0]]

        # rejected before anything is queued
        big = nat.FETCH_INTO_MAX_STREAMS + 1
        assert eng.lib.h261g_engine_fetch_bitstreams_into(eng.h, ctypes.c_void_p(ids_dev), big,
                                                          ctypes.c_void_p(out_dev), ctypes.c_int64(cap),
                                                          ctypes.c_void_p(meta_dev)) == ERR_ARG
        pageable = np.asarray(ids[:2], np.int32)
        assert eng.lib.h261g_engine_fetch_bitstreams_into(eng.h, ctypes.c_void_p(pageable.ctypes.data), 2,
                                                          ctypes.c_void_p(out_dev), ctypes.c_int64(cap),
                                                          ctypes.c_void_p(meta_dev)) == ERR_ARG
    finally:
        eng.destroy()
        hip.free()
