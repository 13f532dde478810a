"""Deterministic inputs of the golden fixtures, shared by make_golden.py
(which records the expected outputs) and tests/test_golden.py (which checks
them).  Every case is a function of its seed only."""
import hashlib

import numpy as np

from oracle.binding import MACROBLOCK_DTYPE, ME_MB_DTYPE, MODE_DTYPE


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(a if isinstance(a, (bytes, bytearray)) else np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def random_macroblocks(rng, mb_w, mb_h, pic_q, intra_frame):
    """Valid macroblock structs covering every MTYPE: intra (+MQUANT), inter
    skipped/MC/MC+filter with and without coefficients and MQUANT, levels up
    to +-127 (escapes), vectors over the whole [-15, 15] range."""
    n = mb_w * mb_h
    mbs = np.zeros(n, MACROBLOCK_DTYPE)
    for m in range(n):
        mb = mbs[m]
        mb["mb_x"], mb["mb_y"] = 16 * (m % mb_w), 16 * (m // mb_w)
        intra = intra_frame or rng.random() < 0.15
        q = pic_q if rng.random() < 0.8 else int(rng.integers(1, 32))
        mb["quant"] = q
        lv = np.zeros((6, 64), np.int32)
        for b in range(6):
            k = rng.integers(0, 12)
            pos = rng.integers(0, 64, k)
            big = rng.random(k) < 0.1
            lv[b, pos] = np.where(big, rng.integers(-127, 128, k), rng.integers(-3, 4, k))
        if intra:
            mb["type"] = 0
            mb["flags"] = 4
            mb["cbp"] = 63
            lv[:, 0] = rng.integers(1, 255, 6)
            lv[rng.random(6) < 0.1, 0] = 255
        else:
            mb["type"] = 2
            flags = int(rng.choice([0, 1, 3, 4, 5, 7]))
            if flags & 1 or flags == 0:
                mb["mv"] = rng.integers(-15, 16, 2) if flags & 1 else (0, 0)
            cbp = 0
            if flags & 4:
                cbp = int(rng.integers(1, 64))
                for b in range(6):
                    if cbp & (32 >> b) and not lv[b].any():
                        lv[b, rng.integers(0, 64)] = int(rng.choice([-2, -1, 1, 2]))
            mb["flags"] = flags
            mb["cbp"] = cbp
        mb["tcoeff"] = lv
    return mbs


def writer_cases():
    """(name, mbs, mb_w, mb_h, source_format, frame_num, quant, frame_type)"""
    out = []
    for (w, h) in ((352, 288), (176, 144)):
        mb_w, mb_h = w // 16, h // 16
        sf = 1 if w == 352 else 0
        rng = np.random.default_rng(1000 + w)
        for trial in range(6):
            q = int(rng.integers(1, 32))
            intra = trial % 3 == 0
            mbs = random_macroblocks(rng, mb_w, mb_h, q, intra)
            out.append((f"{w}x{h}-{trial}", mbs, mb_w, mb_h, sf, trial, q, 0 if intra else 1))
    return out


def mode_decision_cases():
    """(name, me results, frame_type) on CIF"""
    rng = np.random.default_rng(2002)
    out = []
    for trial in range(12):
        n = 396
        me = np.zeros(n, ME_MB_DTYPE)
        scale = int(rng.integers(10, 5000))
        for f in ("intra_cost", "inter_cost", "filter_cost"):
            me[f] = rng.integers(0, scale, n)
        ties = rng.random(n) < 0.2
        me["filter_cost"][ties] = me["inter_cost"][ties]
        ties = rng.random(n) < 0.2
        me["intra_cost"][ties] = me["inter_cost"][ties]
        me["mvx"] = rng.integers(-15, 16, n)
        me["mvy"] = rng.integers(-15, 16, n)
        zero = rng.random(n) < 0.3
        me["mvx"][zero] = 0
        me["mvy"][zero] = 0
        out.append((f"md-{trial}", me, 0 if trial % 4 == 0 else 1))
    return out


def rate_control_cases():
    """(name, quant0, rate, vbv, fixed, frames): frames = [(frame_type,
    per-MB costs, frame bytes)]"""
    rng = np.random.default_rng(3003)
    out = []
    for trial in range(8):
        quant0 = int(rng.integers(1, 32))
        rate = int(rng.integers(8000, 2000000))
        vbv = int(rate * rng.uniform(0.5, 2.0))
        n_mb = 99 if trial % 2 else 396
        fixed = 8 if trial == 5 else 0
        frames = []
        for t in range(40):
            ft = 0 if t == 0 or (trial % 3 == 0 and t % 25 == 0) else 1
            md = np.zeros(n_mb, MODE_DTYPE)
            md["cost"] = rng.integers(0, int(rng.choice([50, 500, 5000])), n_mb)
            frames.append((ft, md, int(rng.integers(10, 60000))))
        out.append((f"rc-{trial}", quant0, rate, vbv, fixed, n_mb, frames))
    return out


def stream_cases():
    """(name, width, height, stream params, content params, frames): the plain
    configuration (CIF, QUANT 8, 30 frames) and a rate-controlled QCIF one"""
    return [
        ("cif-q8", 352, 288, dict(fixed_quant=8), dict(seed=1, sigma=2, motion=7), 30),
        ("qcif-rc", 176, 144, dict(bitrate=64000, vbv=64000, init_quant=10), dict(seed=9, sigma=4, motion=15), 30),
    ]


def numeric_blocks(seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, (n, 8, 8)).astype(np.int32)
