"""ctypes binding of the oracle (liboracle.so) and of the reference host
build (_ref/libh261ref.so, _ref/libh261ref_dropin.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libh261ref.so")
REF_DROPIN_SO = os.path.join(HERE, "_ref", "libh261ref_dropin.so")

ME_MB_DTYPE = np.dtype(
    [("intra_cost", "<i4"), ("zero_cost", "<i4"), ("inter_cost", "<i4"), ("filter_cost", "<i4"),
     ("mvx", "i1"), ("mvy", "i1"), ("pad", "<i2")]
)
MACROBLOCK_DTYPE = np.dtype(
    [("address", "<i4"), ("type", "<i4"), ("flags", "<i4"), ("quant", "<i4"), ("mv", "<i4", (2,)),
     ("cbp", "<i4"), ("tcoeff", "<i4", (6, 64)), ("mb_x", "<i4"), ("mb_y", "<i4")]
)
MODE_DTYPE = np.dtype(
    [("cost", "<i4"), ("quant_adjust", "<i4"), ("type", "<i4"), ("flags", "<i4"), ("mv_x", "<i4"), ("mv_y", "<i4")]
)
DENOISE_DTYPE = np.dtype(
    [("denoise", "<i4"), ("intra_count", "<i4"), ("intra_noise", "<i4", (64,)), ("intra_offset", "<i4", (64,)),
     ("inter_count", "<i4"), ("inter_noise", "<i4", (64,)), ("inter_offset", "<i4", (64,))]
)
RC_DTYPE = np.dtype(
    [("mode", "<i4"), ("current_quant", "<i4"), ("picture_cost", "<i4"), ("prediction_cost", "<i4", (2,)),
     ("prediction_size", "<i4", (2,)), ("bitrate", "<i4"), ("vbv_size", "<i4"), ("vbv_fill", "<i4")]
)

P = ctypes.c_void_p
I = ctypes.c_int


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class Oracle:
    """The C restatement of the reference path (h261_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        """path: the oracle library (another build of h261_oracle.c, e.g. a
        deliberately mutated one in a test, may be given)"""
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        lib = ctypes.CDLL(path)
        self.lib = lib
        sig = {
            "ora_mv_limits": (None, [I, I, P]),
            "ora_sad16": (I, [P, P, I, I, I, I, I]),
            "ora_satd4x4": (I, [P]),
            "ora_select_quadtree": (I, [P, I]),
            "ora_me_frame": (None, [P, P, I, I, P]),
            "ora_fdct8x8": (None, [P, P]),
            "ora_idct8x8": (None, [P, P]),
            "ora_quant_block": (I, [P, I, I]),
            "ora_dequant_block": (None, [P, I, I, I]),
            "ora_predict8x8": (None, [P, I, I, I, I, P]),
            "ora_stream_create": (P, [I, I, I, I, I, I, I]),
            "ora_stream_destroy": (None, [P]),
            "ora_stream_encode_frame": (I, [P, P, P, P, P, I, P, P, P, P, P, P]),
            "ora_stream_last_sad_candidates": (ctypes.c_long, [P]),
            "ora_stream_last_picture_cost": (I, [P]),
            "ora_mode_decision": (None, [P, I, I, P]),
            "ora_rc_init": (None, [P, I, I, I, I]),
            "ora_rc_picture_cost": (I, [P, I]),
            "ora_rc_init_picture": (I, [P, I]),
            "ora_rc_update_picture": (None, [P, I, I, I]),
            "ora_encode_macroblocks": (None, [P, P, I, I, P, I, P, P, P]),
            "ora_write_picture": (I, [P, I, I, I, I, P, I]),
            "ora_layout": (I, [ctypes.POINTER(ctypes.c_long), I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args

    # -- unit functions --------------------------------------------------
    def me_frame(self, cur_y: np.ndarray, ref_y: np.ndarray) -> np.ndarray:
        h, w = cur_y.shape
        out = np.zeros((h // 16) * (w // 16), ME_MB_DTYPE)
        self.lib.ora_me_frame(_p(np.ascontiguousarray(cur_y)), _p(np.ascontiguousarray(ref_y)), w, h, _p(out))
        return out

    def select_quadtree(self, cost: np.ndarray, dim: int) -> int:
        c = np.ascontiguousarray(cost, dtype=np.uint16)
        return self.lib.ora_select_quadtree(_p(c), dim)

    def satd4x4(self, d: np.ndarray) -> int:
        a = np.ascontiguousarray(d, dtype=np.int16).reshape(16)
        return self.lib.ora_satd4x4(_p(a))

    def fdct(self, blk: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(blk, dtype=np.int32).reshape(64)
        o = np.zeros(64, np.int32)
        self.lib.ora_fdct8x8(_p(a), _p(o))
        return o.reshape(8, 8)

    def idct(self, blk: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(blk, dtype=np.int32).reshape(64)
        o = np.zeros(64, np.int32)
        self.lib.ora_idct8x8(_p(a), _p(o))
        return o.reshape(8, 8)

    def quant(self, blk: np.ndarray, q: int, intra: bool):
        a = np.ascontiguousarray(blk, dtype=np.int32).reshape(64).copy()
        coded = self.lib.ora_quant_block(_p(a), q, int(intra))
        return a.reshape(8, 8), bool(coded)

    def dequant(self, blk: np.ndarray, q: int, intra: bool, coded: bool) -> np.ndarray:
        a = np.ascontiguousarray(blk, dtype=np.int32).reshape(64).copy()
        self.lib.ora_dequant_block(_p(a), q, int(intra), int(coded))
        return a.reshape(8, 8)

    def predict(self, ref: np.ndarray, px: int, py: int, filt: bool) -> np.ndarray:
        r = np.ascontiguousarray(ref, dtype=np.uint8)
        o = np.zeros(64, np.int32)
        self.lib.ora_predict8x8(_p(r), r.shape[1], px, py, int(filt), _p(o))
        return o.reshape(8, 8)

    def mode_decision(self, me: np.ndarray, frame_type: int) -> np.ndarray:
        md = np.zeros(len(me), MODE_DTYPE)
        self.lib.ora_mode_decision(_p(me), len(me), frame_type, _p(md))
        return md

    def encode_macroblocks(self, cur, ref, md: np.ndarray, pic_quant: int, dn: np.ndarray):
        """The MB coder of one frame (ora_encode_macroblocks) for given mode
        decisions and denoise state.  cur, ref: (Y, Cb, Cr).  dn: one
        DENOISE_DTYPE record, updated in place (the next frame's offsets).
        Returns (macroblocks, reconstruction planes)."""
        h, w = cur[0].shape
        n_mb = (w // 16) * (h // 16)
        assert len(md) == n_mb and dn.dtype == DENOISE_DTYPE and dn.shape == (1,)
        cur = [np.ascontiguousarray(p, dtype=np.uint8) for p in cur]
        ref = [np.ascontiguousarray(p, dtype=np.uint8) for p in ref]
        rec = [np.zeros_like(p) for p in cur]
        mbs = np.zeros(n_mb, MACROBLOCK_DTYPE)
        cp = (ctypes.c_void_p * 3)(*[_p(p) for p in cur])
        rp = (ctypes.c_void_p * 3)(*[_p(p) for p in ref])
        op = (ctypes.c_void_p * 3)(*[_p(p) for p in rec])
        m = np.ascontiguousarray(md, dtype=MODE_DTYPE)
        self.lib.ora_encode_macroblocks(ctypes.addressof(cp), ctypes.addressof(rp), w, h, _p(m), pic_quant, _p(dn),
                                        _p(mbs), ctypes.addressof(op))
        return mbs, tuple(rec)

    def write_picture(self, mbs: np.ndarray, mb_w: int, source_format: int, tr: int, quant: int,
                      cap: int = 1 << 20) -> bytes:
        out = np.zeros(cap, np.uint8)
        n = self.lib.ora_write_picture(_p(mbs), mb_w, source_format, tr, quant, _p(out), cap)
        if n < 0:
            raise RuntimeError("oracle writer: invalid macroblock data or overflow")
        return out[:n].tobytes()

    def layout(self) -> list:
        buf = (ctypes.c_long * 64)()
        n = self.lib.ora_layout(buf, 64)
        return list(buf[:n])

    # -- whole stream ------------------------------------------------------
    def stream(self, width: int, height: int, denoise: int = 5, init_quant: int = 10, bitrate: int = 512000,
               vbv: int = 512000, fixed_quant: int = 0) -> "OracleStream":
        return OracleStream(self, width, height, denoise, init_quant, bitrate, vbv, fixed_quant)


class OracleStream:
    def __init__(self, o: Oracle, w, h, denoise, init_quant, bitrate, vbv, fixed_quant):
        self.o, self.w, self.h = o, w, h
        self.n_mb = (w // 16) * (h // 16)
        self.handle = o.lib.ora_stream_create(w, h, denoise, init_quant, bitrate, vbv, fixed_quant)
        if not self.handle:
            raise ValueError("bad stream size")

    def encode(self, y: np.ndarray, cb: np.ndarray, cr: np.ndarray, want_details: bool = False):
        """Returns (bitstream bytes, (recon y, cb, cr), me, mbs, quant)."""
        cap = 1 << 20
        bits = np.zeros(cap, np.uint8)
        ry = np.zeros_like(y)
        rcb = np.zeros_like(cb)
        rcr = np.zeros_like(cr)
        me = np.zeros(self.n_mb, ME_MB_DTYPE) if want_details else None
        mbs = np.zeros(self.n_mb, MACROBLOCK_DTYPE) if want_details else None
        q = ctypes.c_int(0)
        n = self.o.lib.ora_stream_encode_frame(
            self.handle, _p(np.ascontiguousarray(y)), _p(np.ascontiguousarray(cb)), _p(np.ascontiguousarray(cr)),
            _p(bits), cap, _p(ry), _p(rcb), _p(rcr), _p(me) if want_details else None,
            _p(mbs) if want_details else None, ctypes.addressof(q))
        if n < 0:
            raise RuntimeError("oracle encode failed")
        return bits[:n].tobytes(), (ry, rcb, rcr), me, mbs, q.value

    def sad_candidates(self) -> int:
        return self.o.lib.ora_stream_last_sad_candidates(self.handle)

    def picture_cost(self) -> int:
        return self.o.lib.ora_stream_last_picture_cost(self.handle)

    def close(self):
        if self.handle:
            self.o.lib.ora_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Reference:
    """The reference's own host code (writer, mode decision, rate control)
    compiled from its sources by oracle/Makefile."""

    def __init__(self):
        if not os.path.exists(REF_SO):
            raise RuntimeError(f"{REF_SO} missing: run make -C oracle ref")
        lib = ctypes.CDLL(REF_SO)
        self.lib = lib
        lib.refdrv_layout.restype = I
        lib.refdrv_layout.argtypes = [ctypes.POINTER(ctypes.c_long), I]
        lib.refdrv_write_picture.restype = I
        lib.refdrv_write_picture.argtypes = [P, I, I, I, I, I, I, P, I]
        lib.refdrv_mode_decision.restype = None
        lib.refdrv_mode_decision.argtypes = [P, I, I, I, P]
        lib.refdrv_rc_init.restype = I
        lib.refdrv_rc_init.argtypes = [I, I, I, I]
        lib.refdrv_rc_begin.restype = I
        lib.refdrv_rc_begin.argtypes = [I, P, I, I, I, I, I]
        lib.refdrv_rc_end.restype = None
        lib.refdrv_rc_end.argtypes = [I, I]
        lib.refdrv_rc_state.restype = None
        lib.refdrv_rc_state.argtypes = [I, P]

    def layout(self) -> list:
        buf = (ctypes.c_long * 64)()
        n = self.lib.refdrv_layout(buf, 64)
        return list(buf[:n])

    def write_picture(self, mbs: np.ndarray, mb_w: int, mb_h: int, source_format: int, frame_num: int,
                      quant: int, frame_type: int = 1) -> bytes:
        cap = 1 << 20
        out = np.zeros(cap, np.uint8)
        n = self.lib.refdrv_write_picture(_p(mbs), mb_w, mb_h, source_format, frame_num, quant, frame_type,
                                          _p(out), cap)
        return out[:n].tobytes()

    def mode_decision(self, me: np.ndarray, mb_w: int, mb_h: int, frame_type: int) -> np.ndarray:
        md = np.zeros(mb_w * mb_h, MODE_DTYPE)
        self.lib.refdrv_mode_decision(_p(me), mb_w, mb_h, frame_type, _p(md))
        return md


class RefDropin:
    """The reference frame loop (h261_main.c) over libh261gpu.so's drop-in ABI."""

    def __init__(self):
        if not os.path.exists(REF_DROPIN_SO):
            raise RuntimeError(f"{REF_DROPIN_SO} missing: run make -C oracle ref")
        lib = ctypes.CDLL(REF_DROPIN_SO)
        lib.refdrv_dropin_encode.restype = ctypes.c_long
        lib.refdrv_dropin_encode.argtypes = [I, I, I, I, P, I, I, I, I, I, P, ctypes.c_long, P, P, P, P]
        self.lib = lib

    def encode(self, frames: np.ndarray, width: int, height: int, denoise=5, init_quant=10, bitrate=512000,
               vbv=512000, fixed_quant=0):
        n = frames.shape[0]
        fsz = width * height * 3 // 2
        cap = n * (1 << 18) + (1 << 20)
        bits = np.zeros(cap, np.uint8)
        sizes = np.zeros(n, np.int32)
        recon = np.zeros((n, fsz), np.uint8)
        quants = np.zeros(n, np.int32)
        last = np.zeros((width // 16) * (height // 16), MACROBLOCK_DTYPE)
        sf = 1 if width == 352 else 0
        total = self.lib.refdrv_dropin_encode(width, height, sf, n, _p(np.ascontiguousarray(frames)), denoise,
                                              init_quant, bitrate, vbv, fixed_quant, _p(bits), cap, _p(sizes),
                                              _p(recon), _p(quants), _p(last))
        if total < 0:
            raise RuntimeError("drop-in encode failed")
        return bits[:total].tobytes(), sizes, recon, quants, last
