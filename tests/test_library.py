"""The C-ABI library loads and exports every symbol include/h261gpu.h declares
(no compute calls: these run without a GPU)."""
import os
import re
import subprocess

from h261enc_cuda_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "h261gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:H261G_API\s+)?(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n.startswith("h261")))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    names = declared_functions()
    assert len(names) >= 16 + 20, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_native.EXPORTS) == names


def test_legacy_prototypes_are_the_reference_sixteen():
    legacy = [n for n in declared_functions() if not n.startswith(("h261g_", "h261gpu_"))]
    assert sorted(legacy) == sorted([
        "h261_gpu_init", "h261_gpu_device_init_textures", "h261_gpu_device_deinit_textures",
        "h261_gpu_device_set_current", "h261_gpu_device_set_current_chroma_cb",
        "h261_gpu_device_set_current_chroma_cr", "h261_gpu_device_set_reference",
        "h261_gpu_device_set_reference_chroma_cb", "h261_gpu_device_set_reference_chroma_cr",
        "h261_gpu_device_init_me", "h261_gpu_device_deinit_me", "h261_gpu_device_me",
        "h261_init_macroblocks", "h261_deinit_macroblocks", "h261_gpu_device_encode_macroblocks",
        "h261_gpu_device_decode_macroblocks"])


def test_library_is_built_for_gfx950():
    """The shared object carries gfx950 device code objects."""
    blob = open(_native.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_product_does_not_reference_the_oracle():
    """No product source names the oracle or links it."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "h261enc_cuda_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")) or f == "Makefile":
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), f"{f} mentions the oracle"


def test_xcd_remap_is_a_bijection():
    """The XCD-aware workgroup order (h261gpu_internal.h xcd_logical) covers
    every logical index exactly once for any launch size."""
    def grid(n):
        return ((n + 7) // 8) * 8

    def logical(i, n):
        per = (n + 7) // 8
        return (i % 8) * per + i // 8

    for n in list(range(1, 200)) + [396, 99, 126720, 63360 + 5]:
        seen = [logical(i, n) for i in range(grid(n))]
        inside = sorted(g for g in seen if g < n)
        assert inside == list(range(n)), n
