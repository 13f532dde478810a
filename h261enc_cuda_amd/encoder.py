"""Host side of the encoder core over the C ABI of libh261gpu.so.

Engine        the multi-stream engine (include/h261gpu.h part 2)
LegacyDevice  the 16 drop-in entry points of the reference's src/gpu_device.h,
              driven the way src/h261_main.c drives them
content_frame the synthetic content generator (same pels as on the device)
write_picture the H.261 writer over a frame's macroblock structs
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Tuple

import numpy as np

from . import _native as nat
from ._native import MACROBLOCK_DTYPE, ME_MB_DTYPE, MODE_DTYPE, check, ptr

CIF = (352, 288)
QCIF = (176, 144)


def content_frame(seed: int, width: int, height: int, frame: int, sigma: int = 2, motion: int = 7,
                  local: bool = False, scene_cut: int = 0) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Frame `frame` of the synthetic sequence (Y, Cb, Cr as uint8 arrays),
    made on the host (libh261content.so; the device generator makes the same
    pels)."""
    lib = nat.load_content()
    cp = nat.ContentParams(seed=seed, sigma=sigma, motion=motion, local=int(local), scene_cut=scene_cut)
    y = np.zeros((height, width), np.uint8)
    cb = np.zeros((height // 2, width // 2), np.uint8)
    cr = np.zeros_like(cb)
    check(lib.h261g_content_frame_host(ctypes.byref(cp), width, height, frame, ptr(y), ptr(cb), ptr(cr)),
          "content_frame")
    return y, cb, cr


def write_picture(mbs: np.ndarray, width: int, height: int, frame_num: int, quant: int,
                  cap: int = 1 << 20) -> bytes:
    """H.261 bitstream of one picture from its macroblock structs."""
    lib = nat.load()
    m = np.ascontiguousarray(mbs, dtype=MACROBLOCK_DTYPE)
    out = np.zeros(cap, np.uint8)
    n = check(lib.h261g_write_picture(ptr(m), width, height, frame_num, quant, ptr(out), cap), "write_picture")
    return out[:n].tobytes()


def device_info() -> dict:
    """the current HIP device: compute units, clock, L2, memory, name, arch"""
    lib = nat.load()
    d = nat.DeviceInfo()
    check(lib.h261g_device_info(ctypes.byref(d)), "device_info")
    return {"compute_units": d.compute_units, "clock_mhz": d.clock_khz // 1000, "l2_bytes": d.l2_bytes,
            "memory_gib": round(d.memory_bytes / 2 ** 30, 1), "name": d.name.decode(errors="replace"),
            "arch": d.arch.decode(errors="replace")}


class Engine:
    """Many independent QCIF/CIF streams, one frame step at a time."""

    def __init__(self, max_streams: int = 16, hbm_budget: int = 0, mbs_per_wave: int = 0,
                 me_variant: Optional[int] = None):
        self.lib = nat.load()
        h = ctypes.c_void_p()
        check(self.lib.h261g_engine_create(ctypes.byref(h), max_streams, hbm_budget, mbs_per_wave), "engine_create")
        self.h = h
        self.sizes = {}
        if me_variant is not None:
            self.config(nat.OPT_ME_VARIANT, me_variant)

    def config(self, option: int, value: int) -> None:
        check(self.lib.h261g_engine_config(self.h, option, value), "engine_config")

    def open(self, width: int, height: int, denoise: int = 5, init_quant: int = 10, bitrate: int = 512000,
             vbv_size: int = 512000, fixed_quant: int = 0) -> int:
        p = nat.StreamParams(width=width, height=height, denoise=denoise, init_quant=init_quant, bitrate=bitrate,
                             vbv_size=vbv_size, fixed_quant=fixed_quant, reserved=0)
        sid = check(self.lib.h261g_stream_open(self.h, ctypes.byref(p)), "stream_open")
        self.sizes[sid] = (width, height)
        return sid

    def close(self, sid: int) -> None:
        check(self.lib.h261g_stream_close(self.h, sid), "stream_close")
        self.sizes.pop(sid, None)

    def put_frame(self, sid: int, y: np.ndarray, cb: np.ndarray, cr: np.ndarray) -> None:
        check(self.lib.h261g_stream_put_frame(self.h, sid, ptr(np.ascontiguousarray(y)),
                                              ptr(np.ascontiguousarray(cb)), ptr(np.ascontiguousarray(cr))),
              "put_frame")

    def put_frames(self, sids, buf: np.ndarray, offsets) -> None:
        """Input of the next step for several host-fed streams at once
        (h261g_engine_put_frames): stream sids[i]'s I420 frame (Y, Cb, Cr
        back to back) starts at byte offsets[i] of the uint8 buffer buf.  The
        copy is queued, not waited for: a buffer in pinned memory must stay
        unchanged until the next step has finished."""
        ids = np.ascontiguousarray(np.asarray(sids, dtype=np.int32))
        offs = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        if buf.dtype != np.uint8 or buf.ndim != 1 or not buf.flags["C_CONTIGUOUS"]:
            raise ValueError("put_frames: a contiguous one-dimensional uint8 buffer")
        if len(ids) != len(offs):
            raise ValueError("put_frames: one offset per stream id")
        check(self.lib.h261g_engine_put_frames(self.h, len(ids), ptr(ids), ptr(buf), buf.size, ptr(offs)),
              "put_frames")

    @staticmethod
    def pack_frames(frames, align: int = 16):
        """Packs (y, cb, cr) frames into one uint8 buffer (each frame's planes
        back to back, frames at multiples of `align`); returns (buf, offsets)."""
        offs, pos = [], 0
        for y, cb, cr in frames:
            offs.append(pos)
            pos += -(-(y.size + cb.size + cr.size) // align) * align
        buf = np.zeros(pos, np.uint8)
        for (y, cb, cr), o in zip(frames, offs):
            n0, n1 = y.size, y.size + cb.size
            buf[o:o + n0] = y.reshape(-1)
            buf[o + n0:o + n1] = cb.reshape(-1)
            buf[o + n1:o + n1 + cr.size] = cr.reshape(-1)
        return buf, offs

    def generate_input(self, sid: int, seed: int, n_frames: int, sigma: int = 2, motion: int = 7,
                       local: bool = False, scene_cut: int = 0) -> None:
        cp = nat.ContentParams(seed=seed, sigma=sigma, motion=motion, local=int(local), scene_cut=scene_cut)
        check(self.lib.h261g_stream_generate_input(self.h, sid, ctypes.byref(cp), n_frames), "generate_input")

    def step(self) -> int:
        return check(self.lib.h261g_engine_step(self.h), "step")

    def run(self, n_steps: int) -> None:
        check(self.lib.h261g_engine_run(self.h, n_steps), "run")

    def sync(self) -> None:
        check(self.lib.h261g_engine_sync(self.h), "sync")

    def compact(self) -> None:
        check(self.lib.h261g_engine_compact(self.h), "compact")

    def input_pending(self, sid: int) -> bool:
        return bool(check(self.lib.h261g_stream_input_pending(self.h, sid), "input_pending"))

    def frames_done(self, sid: int) -> int:
        return check(self.lib.h261g_stream_frames_done(self.h, sid), "frames_done")

    def info(self, sid: int) -> nat.FrameInfo:
        fi = nat.FrameInfo()
        check(self.lib.h261g_stream_info(self.h, sid, ctypes.byref(fi)), "info")
        return fi

    def macroblocks(self, sid: int) -> np.ndarray:
        w, h = self.sizes[sid]
        out = np.zeros((w // 16) * (h // 16), MACROBLOCK_DTYPE)
        check(self.lib.h261g_stream_read_macroblocks(self.h, sid, ptr(out)), "read_macroblocks")
        return out

    def me(self, sid: int) -> np.ndarray:
        w, h = self.sizes[sid]
        out = np.zeros((w // 16) * (h // 16), ME_MB_DTYPE)
        check(self.lib.h261g_stream_read_me(self.h, sid, ptr(out)), "read_me")
        return out

    def recon(self, sid: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        w, h = self.sizes[sid]
        y = np.zeros((h, w), np.uint8)
        cb = np.zeros((h // 2, w // 2), np.uint8)
        cr = np.zeros_like(cb)
        check(self.lib.h261g_stream_read_recon(self.h, sid, ptr(y), ptr(cb), ptr(cr)), "read_recon")
        return y, cb, cr

    def input(self, sid: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        w, h = self.sizes[sid]
        y = np.zeros((h, w), np.uint8)
        cb = np.zeros((h // 2, w // 2), np.uint8)
        cr = np.zeros_like(cb)
        check(self.lib.h261g_stream_read_input(self.h, sid, ptr(y), ptr(cb), ptr(cr)), "read_input")
        return y, cb, cr

    def bitstream(self, sid: int, cap: int = 1 << 20) -> bytes:
        out = np.zeros(cap, np.uint8)
        n = check(self.lib.h261g_stream_read_bitstream(self.h, sid, ptr(out), cap), "read_bitstream")
        return out[:n].tobytes()

    def fetch_bitstreams(self, sids) -> None:
        """Queues delivery of the listed streams' last-frame bitstreams to
        host memory (returns at once; `fetched` collects them)."""
        ids = np.ascontiguousarray(np.asarray(sids, dtype=np.int32))
        check(self.lib.h261g_engine_fetch_bitstreams(self.h, ptr(ids), len(ids)), "fetch_bitstreams")
        self._fetch_ids = ids.copy()

    def fetch_wait(self) -> int:
        """Waits for the last fetch_bitstreams; returns the bytes delivered
        (the bitstreams stay in the engine's pinned buffer)."""
        data, sizes, offs = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        n = check(self.lib.h261g_engine_fetched(self.h, ctypes.byref(data), ctypes.byref(sizes), ctypes.byref(offs)),
                  "fetched")
        return int(ctypes.cast(offs, ctypes.POINTER(ctypes.c_int32))[n]) if n else 0

    def fetched(self) -> List[bytes]:
        """Waits for the last fetch_bitstreams; one bytes object per listed
        stream, in order (a frame that outgrew its device buffer is written
        on the host here)."""
        data, sizes, offs = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        n = check(self.lib.h261g_engine_fetched(self.h, ctypes.byref(data), ctypes.byref(sizes), ctypes.byref(offs)),
                  "fetched")
        if n == 0:
            return []
        sz = np.ctypeslib.as_array(ctypes.cast(sizes, ctypes.POINTER(ctypes.c_int32)), (n,)).copy()
        of = np.ctypeslib.as_array(ctypes.cast(offs, ctypes.POINTER(ctypes.c_int32)), (n + 1,)).copy()
        buf = ctypes.string_at(data, int(of[n])) if of[n] else b""
        out = []
        for i in range(n):
            if sz[i] >= 0:
                out.append(buf[of[i]:of[i] + sz[i]])
            else:
                out.append(self.bitstream(int(self._fetch_ids[i])))
        return out

    def profile(self, enable: bool) -> None:
        check(self.lib.h261g_engine_profile(self.h, int(enable)), "profile")

    def kernel_times(self):
        a, b, c, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int64()
        check(self.lib.h261g_engine_kernel_times(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                 ctypes.byref(d), ctypes.byref(n)), "kernel_times")
        return {"me_ms": a.value, "mb_ms": b.value, "other_ms": c.value, "bits_ms": d.value, "steps": n.value}

    def destroy(self) -> None:
        if self.h:
            self.lib.h261g_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# offsets into the reference structs (include/h261gpu_abi.h, x86-64)
_ME_RESULT_OFF = 24          # me_gpu_t.prgs_macroblocks_result
_MBS_MACROBLOCKS_OFF = 16    # h261_macroblocks_t.ps_macroblocks
_MBS_RECON_Y_OFF = 136       # .pui8_reconstructed_Y_host
_MBS_RECON_CR_OFF = 152      # .pui8_reconstructed_Cr_host
_MBS_RECON_CB_OFF = 168      # .pui8_reconstructed_Cb_host
_CTX_SIZE = 248              # h261_context_t
_CTX_QUANT_OFF = 88 + 72     # .s_picture.i_quantiser
_CTX_MODE_OFF = 192          # .ps_mode_decision
_CTX_MBS_OFF = 240           # .ps_macroblocks


def _read_ptr(addr: int) -> int:
    return ctypes.c_void_p.from_address(addr).value


class LegacyDevice:
    """The drop-in device layer (src/gpu_device.h) as src/h261_main.c uses it:
    init, per frame set_current/set_reference, me, encode, decode."""

    def __init__(self, width: int, height: int, denoise: int = 5):
        self.lib = nat.load()
        self.w, self.h = width, height
        self.n_mb = (width // 16) * (height // 16)
        self.lib.h261_gpu_init()
        self.lib.h261_gpu_device_init_textures(width, height)
        self.mbs = ctypes.c_void_p()
        self.lib.h261_init_macroblocks(ctypes.byref(self.mbs), width, height, denoise)
        self.me_gpu = ctypes.c_void_p()
        self.lib.h261_gpu_device_init_me(ctypes.byref(self.me_gpu), width, height)
        err = self.lib.h261gpu_last_error()
        if err or not self.mbs.value or not self.me_gpu.value:
            raise RuntimeError(f"legacy init failed: {nat.ERRORS.get(err, err)}")
        self._md = np.zeros(self.n_mb, MODE_DTYPE)
        self._md_struct = (ctypes.c_int32 * 4)()  # h261_mode_decision_t: w, h, prgs_mb
        self._ctx = ctypes.create_string_buffer(_CTX_SIZE)

    def set_frames(self, cur, ref) -> None:
        w, h = self.w, self.h
        c = [np.ascontiguousarray(p) for p in cur]
        r = [np.ascontiguousarray(p) for p in ref]
        self.lib.h261_gpu_device_set_current(ptr(c[0]), w, h)
        self.lib.h261_gpu_device_set_current_chroma_cb(ptr(c[1]), w // 2, h // 2)
        self.lib.h261_gpu_device_set_current_chroma_cr(ptr(c[2]), w // 2, h // 2)
        self.lib.h261_gpu_device_set_reference(ptr(r[0]), w, h)
        self.lib.h261_gpu_device_set_reference_chroma_cb(ptr(r[1]), w // 2, h // 2)
        self.lib.h261_gpu_device_set_reference_chroma_cr(ptr(r[2]), w // 2, h // 2)
        self._raise_if_error("set_frames")

    def me(self) -> np.ndarray:
        self.lib.h261_gpu_device_me(self.me_gpu)
        self._raise_if_error("me")
        res = _read_ptr(self.me_gpu.value + _ME_RESULT_OFF)
        return np.frombuffer((ctypes.c_char * (20 * self.n_mb)).from_address(res), ME_MB_DTYPE).copy()

    def encode(self, md: np.ndarray, pic_quant: int) -> np.ndarray:
        self._md[:] = md
        mdp = self._md.ctypes.data
        ctypes.memmove(ctypes.addressof(self._md_struct), (ctypes.c_int32 * 2)(self.w // 16, self.h // 16), 8)
        ctypes.c_void_p.from_address(ctypes.addressof(self._md_struct) + 8).value = mdp
        base = ctypes.addressof(self._ctx)
        ctypes.c_int32.from_address(base + _CTX_QUANT_OFF).value = pic_quant
        ctypes.c_void_p.from_address(base + _CTX_MODE_OFF).value = ctypes.addressof(self._md_struct)
        ctypes.c_void_p.from_address(base + _CTX_MBS_OFF).value = self.mbs.value
        self.lib.h261_gpu_device_encode_macroblocks(base)
        self._raise_if_error("encode")
        arr = _read_ptr(self.mbs.value + _MBS_MACROBLOCKS_OFF)
        return np.frombuffer((ctypes.c_char * (1572 * self.n_mb)).from_address(arr), MACROBLOCK_DTYPE).copy()

    def decode(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        self.lib.h261_gpu_device_decode_macroblocks(ctypes.addressof(self._ctx))
        self._raise_if_error("decode")
        w, h = self.w, self.h
        planes = []
        for off, (pw, ph) in ((_MBS_RECON_Y_OFF, (w, h)), (_MBS_RECON_CB_OFF, (w // 2, h // 2)),
                              (_MBS_RECON_CR_OFF, (w // 2, h // 2))):
            a = _read_ptr(self.mbs.value + off)
            planes.append(np.frombuffer((ctypes.c_char * (pw * ph)).from_address(a), np.uint8).reshape(ph, pw).copy())
        return tuple(planes)

    def _raise_if_error(self, what: str) -> None:
        err = self.lib.h261gpu_last_error()
        if err:
            raise RuntimeError(f"{what}: {nat.ERRORS.get(err, err)}")

    def close(self) -> None:
        if self.me_gpu.value:
            self.lib.h261_gpu_device_deinit_me(ctypes.byref(self.me_gpu))
        if self.mbs.value:
            self.lib.h261_deinit_macroblocks(self.mbs)
            self.mbs = ctypes.c_void_p()
        self.lib.h261_gpu_device_deinit_textures()


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    """PSNR of one plane as src/h261_main.c:38-91 computes it: the integer
    sum of squared differences, 100.0 when it is zero, else
    10 log10(255.0 * 255.0 * pels / ssd) in double, in that order"""
    ssd = int(((a.astype(np.int64) - b.astype(np.int64)) ** 2).sum())
    if ssd == 0:
        return 100.0
    return 10 * math.log10(255.0 * 255.0 * (a.shape[0] * a.shape[1]) / float(ssd))
