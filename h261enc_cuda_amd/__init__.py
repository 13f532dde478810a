"""h261enc_cuda_amd -- MI355X (gfx950) H.261 encoder core.

The product is libh261gpu.so (HIP kernels + C ABI, include/h261gpu.h); this
package is its Python host side: the ctypes binding (_native), the
multi-stream engine wrapper and the synthetic content generator (encoder).
"""
from . import _native  # noqa: F401
from .encoder import Engine, LegacyDevice, content_frame, device_info, write_picture  # noqa: F401

__all__ = ["Engine", "LegacyDevice", "content_frame", "device_info", "write_picture"]
