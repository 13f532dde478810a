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


def test_option_numbers_match_the_header():
    """Every H261G_OPT_* of include/h261gpu.h has its OPT_* twin in the
    Python binding with the same number, and no other."""
    text = open(os.path.join(ROOT, "include", "h261gpu.h")).read()
    header = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define H261G_OPT_([A-Z0-9_]+)\s+(\d+)", text)}
    py = {k[4:]: v for k, v in vars(_native).items() if k.startswith("OPT_") and isinstance(v, int)}
    assert header and header == py


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


def test_xcd_remap_is_a_bijection(tmp_path):
    """The XCD-aware workgroup order (launch_grid / launch_index of
    h261gpu_internal.h, compiled here by the host compiler) covers every
    logical index exactly once for any launch size, padding aside; the
    chunks of consecutive indexes alternate between the 8 XCDs."""
    import subprocess
    src = tmp_path / "xcd.cpp"
    src.write_text(
        '#include <cstdio>\n#include "h261gpu_internal.h"\n'
        "int main() {\n"
        "  const int sizes[] = {1, 7, 8, 9, 63, 64, 65, 99, 396, 511, 512, 513, 4095, 31680, 63365, 126720};\n"
        "  for (int n : sizes) {\n"
        "    const int g = h261g::launch_grid(n, true);\n"
        "    std::printf(\"%d %d\", n, g);\n"
        "    for (int i = 0; i < g; i++) std::printf(\" %d\", h261g::launch_index(i, n, true));\n"
        "    std::printf(\"\\n\");\n"
        "  }\n"
        "  std::printf(\"chunk %d\\n\", h261g::kXcdChunk);\n"
        "}\n")
    exe = tmp_path / "xcd"
    csrc = os.path.join(ROOT, "h261enc_cuda_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", csrc, str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    chunk = int(lines[-2].split()[1])
    for line in lines[:-2]:
        v = [int(x) for x in line.split()]
        n, g, idx = v[0], v[1], v[2:]
        assert len(idx) == g and g % 8 == 0 and g - n < 8 * chunk, n
        assert sorted(idx) == list(range(g)), n  # a permutation of the padded grid
        # workgroup i runs on XCD i % 8; its logical index lies in a chunk
        # dealt to that XCD
        for i, l in enumerate(idx):
            assert (l // chunk) % 8 == i % 8, (n, i)


def test_host_content_library_needs_no_hip_runtime():
    """libh261content.so (the host content generator for CPU-only processes)
    exports h261g_content_frame_host and links no HIP library; it makes the
    same frames as libh261gpu.so's copy."""
    import ctypes
    import numpy as np
    path = _native.CONTENT_LIB_PATH
    deps = subprocess.run(["ldd", path], check=True, capture_output=True, text=True).stdout
    assert "amdhip" not in deps and "hsa" not in deps, deps
    a = _native.load_content()
    b = _native.load()
    cp = _native.ContentParams(seed=77, sigma=3, motion=9, local=0, scene_cut=0)
    for lib in (a, b):
        lib.h261g_content_frame_host.restype = ctypes.c_int
    out = []
    for lib in (a, b):
        y = np.zeros((144, 176), np.uint8)
        c = np.zeros((72, 88), np.uint8)
        r = np.zeros((72, 88), np.uint8)
        assert lib.h261g_content_frame_host(ctypes.byref(cp), 176, 144, 5, _native.ptr(y), _native.ptr(c),
                                            _native.ptr(r)) == 0
        out.append((y, c, r))
    for p in range(3):
        assert np.array_equal(out[0][p], out[1][p])
