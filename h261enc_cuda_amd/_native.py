"""ctypes binding of libh261gpu.so (include/h261gpu.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  There is
no fallback -- importing this module without the built library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libh261gpu.so")
# the host content generator alone: no HIP runtime, so CPU-only processes
# (checker pools, file writers) can make input frames without touching the GPU
CONTENT_LIB_PATH = os.path.join(_HERE, "libh261content.so")

# layouts of the ABI structs (include/h261gpu_abi.h), little endian
ME_MB_DTYPE = np.dtype(
    [("intra_cost", "<i4"), ("zero_cost", "<i4"), ("inter_cost", "<i4"), ("filter_cost", "<i4"),
     ("mvx", "i1"), ("mvy", "i1"), ("pad", "<i2")]
)
assert ME_MB_DTYPE.itemsize == 20

MACROBLOCK_DTYPE = np.dtype(
    [("address", "<i4"), ("type", "<i4"), ("flags", "<i4"), ("quant", "<i4"), ("mv", "<i4", (2,)),
     ("cbp", "<i4"), ("tcoeff", "<i4", (6, 64)), ("mb_x", "<i4"), ("mb_y", "<i4")]
)
assert MACROBLOCK_DTYPE.itemsize == 1572

MODE_DTYPE = np.dtype(
    [("cost", "<i4"), ("quant_adjust", "<i4"), ("type", "<i4"), ("flags", "<i4"), ("mv_x", "<i4"), ("mv_y", "<i4")]
)

ERRORS = {
    0: "ok", -1: "invalid argument", -2: "HIP runtime failure", -3: "allocation failed",
    -4: "no free stream slot or HBM budget exhausted", -5: "call out of order",
}

# every extern "C" symbol declared in include/h261gpu.h
EXPORTS = [
    "h261_gpu_init", "h261_gpu_device_init_textures", "h261_gpu_device_deinit_textures",
    "h261_gpu_device_set_current", "h261_gpu_device_set_current_chroma_cb",
    "h261_gpu_device_set_current_chroma_cr", "h261_gpu_device_set_reference",
    "h261_gpu_device_set_reference_chroma_cb", "h261_gpu_device_set_reference_chroma_cr",
    "h261_gpu_device_init_me", "h261_gpu_device_deinit_me", "h261_gpu_device_me",
    "h261_init_macroblocks", "h261_deinit_macroblocks", "h261_gpu_device_encode_macroblocks",
    "h261_gpu_device_decode_macroblocks", "h261gpu_last_error", "h261gpu_error_string",
    "h261g_engine_create", "h261g_engine_destroy", "h261g_stream_footprint", "h261g_stream_open",
    "h261g_stream_close", "h261g_stream_put_frame", "h261g_stream_generate_input",
    "h261g_stream_frames_done", "h261g_engine_compact", "h261g_engine_step", "h261g_engine_run",
    "h261g_engine_sync", "h261g_stream_info", "h261g_stream_read_macroblocks", "h261g_stream_read_recon",
    "h261g_stream_read_me", "h261g_stream_read_bitstream", "h261g_stream_read_input",
    "h261g_engine_profile", "h261g_engine_kernel_times", "h261g_content_frame_host", "h261g_write_picture",
    "h261g_engine_config", "h261g_device_info", "h261g_stream_input_pending", "h261g_engine_fetch_bitstreams",
    "h261g_engine_fetched", "h261g_engine_fetch_bitstreams_into", "h261g_engine_put_frames",
]
FETCH_INTO_MAX_STREAMS = 4096  # H261G_FETCH_INTO_MAX_STREAMS
OPT_ME_VARIANT = 1
OPT_MBS_PER_WAVE = 2
OPT_FUSED_TAILS = 3
OPT_XCD_REMAP = 4
OPT_ASYNC_INPUT = 5
OPT_DEVICE_BITSTREAM = 6
OPT_OVERLAP_WRITER = 7
OPT_CHAINS = 8
OPT_ME_MBS_PER_WAVE = 9
OPT_ME_TILE = 10
OPT_CHAIN_LADDER = 11


class StreamParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("denoise", ctypes.c_int32),
                ("init_quant", ctypes.c_int32), ("bitrate", ctypes.c_int32), ("vbv_size", ctypes.c_int32),
                ("fixed_quant", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ContentParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("sigma", ctypes.c_int32), ("motion", ctypes.c_int32),
                ("local", ctypes.c_int32), ("scene_cut", ctypes.c_int32)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("compute_units", ctypes.c_int32), ("clock_khz", ctypes.c_int32), ("l2_bytes", ctypes.c_int64),
                ("memory_bytes", ctypes.c_int64), ("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 64)]


class FrameInfo(ctypes.Structure):
    _fields_ = [("frame_num", ctypes.c_int32), ("frame_type", ctypes.c_int32), ("quant", ctypes.c_int32),
                ("bytes", ctypes.c_int32), ("cost", ctypes.c_int32)]


_lib = None


def load() -> ctypes.CDLL:
    """Loads libh261gpu.so (RTLD_GLOBAL, so the reference host built against it
    resolves the same copy); raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() (or make -C h261enc_cuda_amd)")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    P = ctypes.c_void_p
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    sig = {
        "h261gpu_last_error": (ctypes.c_int, []),
        "h261gpu_error_string": (ctypes.c_char_p, [ctypes.c_int]),
        "h261g_engine_create": (ctypes.c_int, [ctypes.POINTER(P), i32, i64, i32]),
        "h261g_engine_destroy": (ctypes.c_int, [P]),
        "h261g_engine_config": (ctypes.c_int, [P, i32, i32]),
        "h261g_device_info": (ctypes.c_int, [ctypes.POINTER(DeviceInfo)]),
        "h261g_stream_footprint": (i64, [i32, i32]),
        "h261g_stream_open": (ctypes.c_int, [P, ctypes.POINTER(StreamParams)]),
        "h261g_stream_close": (ctypes.c_int, [P, i32]),
        "h261g_stream_put_frame": (ctypes.c_int, [P, i32, P, P, P]),
        "h261g_engine_put_frames": (ctypes.c_int, [P, i32, P, P, i64, P]),
        "h261g_stream_generate_input": (ctypes.c_int, [P, i32, ctypes.POINTER(ContentParams), i32]),
        "h261g_stream_frames_done": (ctypes.c_int, [P, i32]),
        "h261g_stream_input_pending": (ctypes.c_int, [P, i32]),
        "h261g_engine_compact": (ctypes.c_int, [P]),
        "h261g_engine_step": (ctypes.c_int, [P]),
        "h261g_engine_run": (ctypes.c_int, [P, i32]),
        "h261g_engine_sync": (ctypes.c_int, [P]),
        "h261g_stream_info": (ctypes.c_int, [P, i32, ctypes.POINTER(FrameInfo)]),
        "h261g_stream_read_macroblocks": (ctypes.c_int, [P, i32, P]),
        "h261g_stream_read_recon": (ctypes.c_int, [P, i32, P, P, P]),
        "h261g_stream_read_me": (ctypes.c_int, [P, i32, P]),
        "h261g_stream_read_bitstream": (ctypes.c_int, [P, i32, P, i32]),
        "h261g_engine_fetch_bitstreams": (ctypes.c_int, [P, P, i32]),
        "h261g_engine_fetched": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P)]),
        "h261g_engine_fetch_bitstreams_into": (ctypes.c_int, [P, P, i32, P, i64, P]),
        "h261g_stream_read_input": (ctypes.c_int, [P, i32, P, P, P]),
        "h261g_engine_profile": (ctypes.c_int, [P, i32]),
        "h261g_engine_kernel_times": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_double),
                                                      ctypes.POINTER(ctypes.c_double),
                                                      ctypes.POINTER(ctypes.c_double),
                                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]),
        "h261g_content_frame_host": (ctypes.c_int, [ctypes.POINTER(ContentParams), i32, i32, i32, P, P, P]),
        "h261g_write_picture": (ctypes.c_int, [P, i32, i32, i32, i32, P, i32]),
        # legacy drop-in (src/gpu_device.h); Void returns
        "h261_gpu_init": (None, []),
        "h261_gpu_device_init_textures": (None, [i32, i32]),
        "h261_gpu_device_deinit_textures": (None, []),
        "h261_gpu_device_set_current": (None, [P, i32, i32]),
        "h261_gpu_device_set_current_chroma_cb": (None, [P, i32, i32]),
        "h261_gpu_device_set_current_chroma_cr": (None, [P, i32, i32]),
        "h261_gpu_device_set_reference": (None, [P, i32, i32]),
        "h261_gpu_device_set_reference_chroma_cb": (None, [P, i32, i32]),
        "h261_gpu_device_set_reference_chroma_cr": (None, [P, i32, i32]),
        "h261_gpu_device_init_me": (None, [ctypes.POINTER(P), i32, i32]),
        "h261_gpu_device_deinit_me": (None, [ctypes.POINTER(P)]),
        "h261_gpu_device_me": (None, [P]),
        "h261_init_macroblocks": (None, [ctypes.POINTER(P), i32, i32, i32]),
        "h261_deinit_macroblocks": (None, [P]),
        "h261_gpu_device_encode_macroblocks": (None, [P]),
        "h261_gpu_device_decode_macroblocks": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


_content_lib = None


def load_content() -> ctypes.CDLL:
    """Loads libh261content.so (h261g_content_frame_host only); raises if it
    has not been built."""
    global _content_lib
    if _content_lib is not None:
        return _content_lib
    if not os.path.exists(CONTENT_LIB_PATH):
        raise RuntimeError(
            f"{CONTENT_LIB_PATH} is missing: build it with __graft_entry__.build() (or make -C h261enc_cuda_amd)")
    lib = ctypes.CDLL(CONTENT_LIB_PATH)
    lib.h261g_content_frame_host.restype = ctypes.c_int
    lib.h261g_content_frame_host.argtypes = [ctypes.POINTER(ContentParams), ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    _content_lib = lib
    return lib


def ptr(a: np.ndarray) -> int:
    """Address of a C-contiguous numpy array."""
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise RuntimeError(f"{what}: {ERRORS.get(rc, rc)}")
    return rc
