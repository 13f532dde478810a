"""The oracle's DCT denoise (src/mbgpu_kernels.cuh:478-636, launch order
src/gpu_device.cu:691-705) pinned by closed forms, and the pin shown to
catch plausible mistakes.

Construction: every 8x8 block of the frame is flat, so its DCT is exactly
DC = 8 x (block value, or residual against a flat reference) and AC = 0.
The reference's denoise then has a closed form per block type t (intra 0,
inter 1) and coefficient position i, with O the incoming offsets staged as
UInt16 (mbgpu:513, 525-528) and d the strength:
    c' = sign(c) * max(|c| - O[t][i], 0)                 (this frame's levels)
    count[t] = blocks of type t, noise[t][i] = sum |c| before the subtraction
    O'[t][i] = (count[t] * d + noise[t][i] / 2) / (noise[t][i] + 1)
O' is the next frame's offset (macroblocks_update_dct_denoise runs after
macroblocks_dct_denoise), and counts and noise sums restart at zero.  Levels
follow from c' by the quantiser: intra DC clamp(c'/8, 1, 254) with 128 ->
255 (mbgpu:689-702), inter trunc(c' * (float)(1/(2q))) clamped to +-127.

test_denoise_checks_catch_mutants compiles copies of oracle/h261_oracle.c
with one planted mistake each and requires the same checks to fail."""
import os
import subprocess

import numpy as np
import pytest

from oracle.binding import DENOISE_DTYPE, MODE_DTYPE, Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SRC = os.path.join(ROOT, "oracle", "h261_oracle.c")

W, H = 176, 144
MB_W, MB_H = W // 16, H // 16
N_MB = MB_W * MB_H
INTRA, INTER = 0, 2
REF_LEVEL = 100   # the flat reference frame
STRENGTH = 5      # the reference application's denoise strength
QUANT = 2  # inter levels of a residual DC of 8 (level 2) show an offset of 2 or more


def construction(seed):
    """flat blocks: intra MBs (every third) with block values 0..255, inter
    MBs (mv 0, no loop filter) with residual +-1 in one block in five and 0
    elsewhere against the flat reference.  The inter noise sums stay small
    enough that the outgoing inter offset (about count * d / sum) is neither
    0 nor count * d, so it depends on what the sums include, on when they
    are taken and on the rounding term."""
    rng = np.random.default_rng(seed)
    md = np.zeros(N_MB, MODE_DTYPE)
    md["type"] = np.where(np.arange(N_MB) % 3 == 0, INTRA, INTER)
    y = np.zeros((H, W), np.int32)
    c = [np.zeros((H // 2, W // 2), np.int32) for _ in range(2)]
    vals = np.zeros((N_MB, 6), np.int64)
    for m in range(N_MB):
        mx, my = m % MB_W, m // MB_W
        for b in range(6):
            if md["type"][m] == INTRA:
                v = int(rng.integers(0, 256))
            else:
                v = REF_LEVEL + (int(rng.choice([-1, 1])) if rng.integers(0, 5) == 0 else 0)
            vals[m, b] = v
            if b < 4:
                y[16 * my + 8 * (b >> 1):16 * my + 8 * (b >> 1) + 8, 16 * mx + 8 * (b & 1):16 * mx + 8 * (b & 1) + 8] = v
            else:
                c[b - 4][8 * my:8 * my + 8, 8 * mx:8 * mx + 8] = v
    cur = (y.astype(np.uint8), c[0].astype(np.uint8), c[1].astype(np.uint8))
    ref = tuple(np.full_like(p, REF_LEVEL) for p in cur)
    return md, cur, ref, vals


def closed_form(md, vals, off_in):
    """levels [mb][6][64] and the outgoing offsets (2, 64) for incoming
    offsets off_in (2, 64) as stored (Int32; staged as UInt16)"""
    o = np.asarray(off_in, np.int64) & 0xFFFF
    levels = np.zeros((N_MB, 6, 64), np.int64)
    count = [0, 0]
    noise = np.zeros((2, 64), np.int64)
    inv = np.float32(1.0 / float(np.float32(2 * QUANT)))
    for m in range(N_MB):
        t = 0 if md["type"][m] == INTRA else 1
        for b in range(6):
            dc = 8 * int(vals[m, b] - (0 if t == 0 else REF_LEVEL))
            count[t] += 1
            noise[t][0] += abs(dc)
            mag = max(abs(dc) - int(o[t][0]), 0)
            cd = -mag if dc < 0 else mag
            if t == 0:
                lv = min(max(cd // 8, 1), 254)
                levels[m, b, 0] = 255 if lv == 128 else lv
            else:
                levels[m, b, 0] = min(127, max(-127, int(np.float32(cd) * inv)))
    out = np.zeros((2, 64), np.int64)
    for t in range(2):
        for i in range(64):
            out[t][i] = (count[t] * STRENGTH + noise[t][i] // 2) // (noise[t][i] + 1)
    return levels, out


def check_denoise(o: Oracle, seed=5):
    """Two frames of the construction through the oracle's MB coder with one
    denoise state: levels and outgoing offsets equal the closed form.  The
    first frame's incoming offsets are chosen (one above 65535); the second
    frame's are the first frame's outgoing ones."""
    md, cur, ref, vals = construction(seed)
    dn = np.zeros(1, DENOISE_DTYPE)
    dn["denoise"] = STRENGTH
    off = np.zeros((2, 64), np.int64)
    off[0, 0], off[1, 0] = 65536 + 40, 24
    off[0, 1:], off[1, 1:] = 3, 7
    dn["intra_offset"][0] = off[0]
    dn["inter_offset"][0] = off[1]
    for frame in range(2):
        mbs, _ = o.encode_macroblocks(cur, ref, md, QUANT, dn)
        want_levels, want_off = closed_form(md, vals, off)
        got = mbs["tcoeff"].reshape(N_MB, 6, 64)
        bad = np.argwhere(got != want_levels)
        assert bad.size == 0, f"frame {frame}: levels differ at (mb, block, pos) {bad[:3].tolist()}"
        assert np.array_equal(dn["intra_offset"][0], want_off[0]), f"frame {frame}: intra offsets"
        assert np.array_equal(dn["inter_offset"][0], want_off[1]), f"frame {frame}: inter offsets"
        assert dn["intra_count"][0] == dn["inter_count"][0] == 0, f"frame {frame}: counts not restarted"
        assert not dn["intra_noise"].any() and not dn["inter_noise"].any(), f"frame {frame}: sums not restarted"
        off = want_off


def test_flat_blocks_have_an_exact_dct(oracle):
    for v in (-255, -3, 0, 1, 100, 255):
        c = oracle.fdct(np.full((8, 8), v, np.int32)).ravel()
        assert c[0] == 8 * v and not c[1:].any(), v


@pytest.mark.parametrize("seed", [5, 6])
def test_denoise_closed_form(oracle, seed):
    check_denoise(oracle, seed)


def test_closed_form_is_not_trivial():
    """the construction makes the checks bite: both block types present,
    offsets that change levels, and outgoing offsets where the rounding term
    of the update and the per-block count both matter"""
    md, _, _, vals = construction(5)
    off = np.zeros((2, 64), np.int64)
    off[0, 0], off[1, 0] = 65536 + 40, 24
    lv_with, out = closed_form(md, vals, off)
    lv_without, _ = closed_form(md, vals, np.zeros((2, 64), np.int64))
    intra = md["type"] == INTRA
    assert ((lv_with != lv_without)[intra]).sum() > 100 and ((lv_with != lv_without)[~intra]).sum() > 50
    # outgoing offsets: positions 1..63 are count * d; position 0 depends on
    # the noise sums (inter: between 1 and count * d, and changed by the
    # rounding term)
    n_intra = 6 * int(intra.sum())
    n_inter = 6 * N_MB - n_intra
    assert out[0, 1] == n_intra * STRENGTH and out[1, 1] == n_inter * STRENGTH
    noise = 8 * int(np.abs(vals[~intra] - REF_LEVEL).sum())
    assert 0 < out[1, 0] < n_inter * STRENGTH and out[1, 0] != (n_inter * STRENGTH) // (noise + 1)


# each mutation: (name, exact text in h261_oracle.c, replacement)
MUTANTS = [
    ("offsets applied to the frame that measured them",
     "                a = imax(a - (int)off[t][i], 0);\n                coef[m][b][i] = c < 0 ? -a : a;\n",
     ""),
    ("statistics after the subtraction",
     "                noise[t][i] += a;\n                a = imax(a - (int)off[t][i], 0);\n",
     "                a = imax(a - (int)off[t][i], 0);\n                noise[t][i] += a;\n"),
    ("intra and inter offset tables swapped",
     "        off[0][i] = (uint16_t)dn->rgi_intra_signal_offset[i];\n"
     "        off[1][i] = (uint16_t)dn->rgi_inter_signal_offset[i];\n",
     "        off[0][i] = (uint16_t)dn->rgi_inter_signal_offset[i];\n"
     "        off[1][i] = (uint16_t)dn->rgi_intra_signal_offset[i];\n"),
    ("intra and inter statistics swapped",
     "            int t = intra ? 0 : 1;\n            count[t]++;\n",
     "            int t = intra ? 0 : 1;\n            count[1 - t]++;\n"),
    ("offsets not staged as UInt16",
     "        off[0][i] = (uint16_t)dn->rgi_intra_signal_offset[i];\n",
     "        off[0][i] = (uint16_t)(dn->rgi_intra_signal_offset[i] > 65535 ? 65535 : dn->rgi_intra_signal_offset[i]);\n"),
    ("count per macroblock instead of per block",
     "            count[t]++;\n", "            if (b == 0) count[t]++;\n"),
    ("rounding term of the update dropped",
     "((int)count[1] * dn->i_denoise + ni / 2) / (ni + 1)", "((int)count[1] * dn->i_denoise) / (ni + 1)"),
    ("chroma blocks left out of the statistics",
     "                noise[t][i] += a;\n", "                if (b < 4) noise[t][i] += a;\n"),
    ("DC position left out of the denoise",
     "                a = imax(a - (int)off[t][i], 0);\n", "                if (i) a = imax(a - (int)off[t][i], 0);\n"),
]


def _mutant_offsets_applied_late(src):
    """the first mutation removes the subtraction; it is re-inserted after
    the update with the new offsets, i.e. the offsets act on the frame whose
    statistics produced them"""
    anchor = "    dn->i_inter_count = 0;\n    dn->i_intra_count = 0;\n"
    late = (anchor +
            "    for (int m = 0; m < n_mb; m++)\n"
            "        for (int b = 0; b < 6; b++) {\n"
            "            int t = md[m].i_mb_type == H261_MB_TYPE_INTRA ? 0 : 1;\n"
            "            for (int i = 0; i < 64; i++) {\n"
            "                int c = coef[m][b][i], a = abs(c);\n"
            "                int o = (uint16_t)(t ? dn->rgi_inter_signal_offset[i] : dn->rgi_intra_signal_offset[i]);\n"
            "                a = imax(a - o, 0);\n"
            "                coef[m][b][i] = c < 0 ? -a : a;\n"
            "            }\n"
            "        }\n")
    assert src.count(anchor) == 1
    return src.replace(anchor, late)


@pytest.mark.parametrize("k", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_denoise_checks_catch_mutants(tmp_path, k):
    name, old, new = MUTANTS[k]
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, f"mutation site of '{name}' not found once in h261_oracle.c"
    src = src.replace(old, new)
    if k == 0:
        src = _mutant_offsets_applied_late(src)
    (tmp_path / "m.c").write_text(src)
    so = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-w", "-I", os.path.join(ROOT, "oracle"),
                    "-I", os.path.join(ROOT, "include"), str(tmp_path / "m.c"), "-o", str(so), "-lm"], check=True)
    with pytest.raises(AssertionError):
        check_denoise(Oracle(str(so)))
