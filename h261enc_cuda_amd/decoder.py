"""An H.261 decoder (ITU-T H.261 clause 4) for checking the encoder in a
closed loop: decode(bitstream) must reproduce the encoder's reconstruction
byte for byte.

It parses the picture / GOB / MB / block layers with the VLC tables of the
standard, reconstructs with the standard's dequantiser, motion
compensation and loop filter, and inverts the transform with the same AAN
float/double operation sequence the encoder's reconstruction uses
(src/mbgpu_kernels.cuh:827-931 of the reference), vectorised over blocks.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np

# ---- VLC tables (ITU-T H.261 tables 1-5) as (code, length) -> value ----

_MBA = [(1, 1), (3, 3), (2, 3), (3, 4), (2, 4), (3, 5), (2, 5), (7, 7), (6, 7), (11, 8), (10, 8), (9, 8), (8, 8),
        (7, 8), (6, 8), (23, 10), (22, 10), (21, 10), (20, 10), (19, 10), (18, 10), (35, 11), (34, 11), (33, 11),
        (32, 11), (31, 11), (30, 11), (29, 11), (28, 11), (27, 11), (26, 11), (25, 11), (24, 11)]  # MBA 1..33
_MBA_STUFFING = (15, 11)
# MTYPE: (intra, mquant, mvd, coeff, filter)
_MTYPE = {(1, 4): (1, 0, 0, 1, 0), (1, 7): (1, 1, 0, 1, 0), (1, 1): (0, 0, 0, 1, 0), (1, 5): (0, 1, 0, 1, 0),
          (1, 9): (0, 0, 1, 0, 0), (1, 8): (0, 0, 1, 1, 0), (1, 10): (0, 1, 1, 1, 0), (1, 3): (0, 0, 1, 0, 1),
          (1, 2): (0, 0, 1, 1, 1), (1, 6): (0, 1, 1, 1, 1)}
_CBP = {}
for _cbp, (_c, _l) in enumerate([
        (0, 0), (11, 5), (9, 5), (13, 6), (13, 4), (23, 7), (19, 7), (31, 8), (12, 4), (22, 7), (18, 7), (30, 8),
        (19, 5), (27, 8), (23, 8), (19, 8), (11, 4), (21, 7), (17, 7), (29, 8), (17, 5), (25, 8), (21, 8), (17, 8),
        (15, 6), (15, 8), (13, 8), (3, 9), (15, 5), (11, 8), (7, 8), (7, 9), (10, 4), (20, 7), (16, 7), (28, 8),
        (14, 6), (14, 8), (12, 8), (2, 9), (16, 5), (24, 8), (20, 8), (16, 8), (14, 5), (10, 8), (6, 8), (6, 9),
        (18, 5), (26, 8), (22, 8), (18, 8), (13, 5), (9, 8), (5, 8), (5, 9), (12, 5), (8, 8), (4, 8), (4, 9),
        (7, 3), (10, 5), (8, 5), (12, 6)]):
    if _cbp:
        _CBP[(_c, _l)] = _cbp
# MVD: value d in -16..15 (and d+-32)
_MVD = {}
for _d, (_c, _l) in zip(range(-16, 16), [
        (25, 11), (27, 11), (29, 11), (31, 11), (33, 11), (35, 11), (19, 10), (21, 10), (23, 10), (7, 8), (9, 8),
        (11, 8), (7, 7), (3, 5), (3, 4), (3, 3), (1, 1), (2, 3), (2, 4), (2, 5), (6, 7), (10, 8), (8, 8), (6, 8),
        (22, 10), (20, 10), (18, 10), (34, 11), (32, 11), (30, 11), (28, 11), (26, 11)]):
    _MVD[(_c, _l)] = _d
# TCOEFF (run, level) without sign; EOB = (2, 2); escape = (1, 6)
_TCOEFF = {
    (3, 2): (0, 1), (4, 4): (0, 2), (5, 5): (0, 3), (6, 7): (0, 4), (38, 8): (0, 5), (33, 8): (0, 6),
    (10, 10): (0, 7), (29, 12): (0, 8), (24, 12): (0, 9), (19, 12): (0, 10), (16, 12): (0, 11), (26, 13): (0, 12),
    (25, 13): (0, 13), (24, 13): (0, 14), (23, 13): (0, 15), (3, 3): (1, 1), (6, 6): (1, 2), (37, 8): (1, 3),
    (12, 10): (1, 4), (27, 12): (1, 5), (22, 13): (1, 6), (21, 13): (1, 7), (5, 4): (2, 1), (4, 7): (2, 2),
    (11, 10): (2, 3), (20, 12): (2, 4), (20, 13): (2, 5), (7, 5): (3, 1), (36, 8): (3, 2), (28, 12): (3, 3),
    (19, 13): (3, 4), (6, 5): (4, 1), (15, 10): (4, 2), (18, 12): (4, 3), (7, 6): (5, 1), (9, 10): (5, 2),
    (18, 13): (5, 3), (5, 6): (6, 1), (30, 12): (6, 2), (4, 6): (7, 1), (21, 12): (7, 2), (7, 7): (8, 1),
    (17, 12): (8, 2), (5, 7): (9, 1), (17, 13): (9, 2), (39, 8): (10, 1), (16, 13): (10, 2), (35, 8): (11, 1),
    (34, 8): (12, 1), (32, 8): (13, 1), (14, 10): (14, 1), (13, 10): (15, 1), (8, 10): (16, 1),
    (31, 12): (17, 1), (26, 12): (18, 1), (25, 12): (19, 1), (23, 12): (20, 1), (22, 12): (21, 1),
    (31, 13): (22, 1), (30, 13): (23, 1), (29, 13): (24, 1), (28, 13): (25, 1), (27, 13): (26, 1),
}
_ZIGZAG = np.array([0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27,
                    20, 13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58,
                    59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63])


class BitstreamError(ValueError):
    pass


class _Bits:
    def __init__(self, data: bytes):
        self.v = int.from_bytes(data, "big") if data else 0
        self.n = 8 * len(data)
        self.pos = 0

    def peek(self, k: int) -> int:
        if self.pos + k <= self.n:
            return (self.v >> (self.n - self.pos - k)) & ((1 << k) - 1)
        # past the end: pad with zeros
        have = self.n - self.pos
        if have <= 0:
            return 0
        return ((self.v & ((1 << have) - 1)) << (k - have))

    def read(self, k: int) -> int:
        if self.pos + k > self.n:
            raise BitstreamError("read past the end")
        x = self.peek(k)
        self.pos += k
        return x

    def vlc(self, table: Dict[Tuple[int, int], object], maxlen: int = 13):
        for ln in range(1, maxlen + 1):
            key = (self.peek(ln), ln)
            if key in table:
                self.pos += ln
                return table[key]
        raise BitstreamError(f"invalid VLC at bit {self.pos}")


_MBA_TAB = {k: i + 1 for i, k in enumerate(_MBA)}
_MBA_TAB[_MBA_STUFFING] = "stuffing"


def _read_block(b: _Bits, intra: bool) -> np.ndarray:
    lv = np.zeros(64, np.int32)
    i = 0
    if intra:
        lv[0] = b.read(8)
        i = 1
    else:
        if b.peek(1) == 1:  # first coefficient '1s' = run 0, level +-1
            b.read(1)
            lv[0] = -1 if b.read(1) else 1
            i = 1
    while True:
        if b.peek(2) == 2:  # EOB
            b.read(2)
            return lv
        if b.peek(6) == 1:  # escape
            b.read(6)
            run = b.read(6)
            level = b.read(8)
            level = level - 256 if level >= 128 else level
        else:
            run, level = b.vlc(_TCOEFF)
            if b.read(1):
                level = -level
        i += run
        if i >= 64:
            raise BitstreamError("coefficient index beyond 63")
        lv[_ZIGZAG[i]] = level
        i += 1


# ---- reconstruction arithmetic ----

_INV_PRE = np.array([1.0, 1.387040, 1.306563, 1.175876, 1.0, 0.785695, 0.541196, 0.275899], np.float64).astype(
    np.float32)


def _f(x):
    return np.asarray(x, np.float32)


def _dmul(x, c):
    """float * double literal, rounded to float"""
    return (np.asarray(x, np.float32).astype(np.float64) * c).astype(np.float32)


def _idct_1d(t):
    """AAN inverse on axis -1 of float32 t (shape (..., 8))"""
    t0, t1, t2, t3, t4, t5, t6, t7 = (t[..., k] for k in range(8))
    tmp13 = _f(t2 + t6)
    tmp12 = _f(t2 - t6)
    tmp12 = _dmul(tmp12, 1.414214)
    tmp12 = _f(tmp12 - tmp13)
    tmp10 = _f(t0 + t4)
    tmp11 = _f(t0 - t4)
    a0, a1, a2, a3 = _f(tmp10 + tmp13), _f(tmp11 + tmp12), _f(tmp11 - tmp12), _f(tmp10 - tmp13)
    z11, z12, z13, z10 = _f(t1 + t7), _f(t1 - t7), _f(t5 + t3), _f(t5 - t3)
    tmp13 = _f(z11 + z13)
    tmp11 = _dmul(_f(z11 - z13), 1.414214)
    z5 = _dmul(_f(z12 + z10), 1.84776)
    tmp10 = (z12.astype(np.float64) * 1.082392 - z5.astype(np.float64)).astype(np.float32)
    tmp12 = (z10.astype(np.float64) * -2.613126 + z5.astype(np.float64)).astype(np.float32)
    tmp12 = _f(tmp12 - tmp13)
    tmp11 = _f(tmp11 - tmp12)
    tmp10 = _f(tmp10 + tmp11)
    return np.stack([_f(a0 + tmp13), _f(a1 + tmp12), _f(a2 + tmp11), _f(a3 - tmp10), _f(a3 + tmp10),
                     _f(a2 - tmp11), _f(a1 - tmp12), _f(a0 - tmp13)], axis=-1)


def idct_blocks(coef: np.ndarray) -> np.ndarray:
    """coef (N, 8, 8) int -> residual (N, 8, 8) int in [-256, 255]."""
    c = np.asarray(coef, np.int64)
    cs = (_INV_PRE.astype(np.float64) * 0.125).astype(np.float32)  # per row
    t = _f(_f(c.astype(np.float32) * _INV_PRE[None, None, :]) * cs[None, :, None])
    t = np.swapaxes(_idct_1d(np.swapaxes(t, 1, 2)), 1, 2)  # columns first
    o = _idct_1d(t)                                          # then rows
    r = np.floor(o.astype(np.float64) + 0.49999).astype(np.int64)
    return np.clip(r, -256, 255)


def dequant(lv: np.ndarray, q: int, intra: bool) -> np.ndarray:
    """ITU-T H.261 4.2.4 (intra DC: 255 -> 1024, else 8x)."""
    x = lv.astype(np.int64)
    s = np.sign(x)
    r = q * (2 * x + s) - (s if q % 2 == 0 else 0)
    r = np.clip(r, -2048, 2047)
    if intra:
        r[0] = 1024 if lv[0] == 255 else 8 * lv[0]
    return r


def _loop_filter(a: np.ndarray) -> np.ndarray:
    a = a.astype(np.int64)
    h = a.copy() * 4
    h[:, 1:7] = a[:, 0:6] + 2 * a[:, 1:7] + a[:, 2:8]
    v = h.copy() * 4
    v[1:7] = h[0:6] + 2 * h[1:7] + h[2:8]
    return (v + 8) >> 4


class Decoder:
    """Decodes a sequence of H.261 pictures (one bytes object per picture)."""

    def __init__(self, width: int, height: int):
        if (width, height) not in ((352, 288), (176, 144)):
            raise ValueError("H.261 pictures are CIF or QCIF")
        self.w, self.h = width, height
        self.ref = [np.zeros((height, width), np.uint8), np.zeros((height // 2, width // 2), np.uint8),
                    np.zeros((height // 2, width // 2), np.uint8)]

    def decode(self, data: bytes):
        """Returns the reconstruction (Y, Cb, Cr) and the per-MB syntax."""
        b = _Bits(data)
        if b.read(20) != 16:
            raise BitstreamError("no picture start code")
        tr = b.read(5)
        ptype = b.read(6)
        cif = (ptype >> 2) & 1
        if (self.w == 352) != bool(cif):
            raise BitstreamError("source format differs from the decoder's")
        while b.read(1):  # PEI / PSPARE
            b.read(8)
        mb_w = self.w // 16
        n_gob = 12 if cif else 3
        mbs: List[Optional[dict]] = [None] * (mb_w * (self.h // 16))
        for g in range(n_gob):
            if b.read(16) != 1:
                raise BitstreamError(f"no GOB start code for GOB {g}")
            gn = b.read(4)
            want = g + 1 if cif else 2 * g + 1
            if gn != want:
                raise BitstreamError(f"GOB number {gn}, expected {want}")
            quant = b.read(5)
            while b.read(1):
                b.read(8)
            gx = ((gn - 1) % 2) * 11 if cif else 0
            gy = ((gn - 1) // 2) * 3
            addr = 0
            prev_mv = None
            while True:
                rem = b.n - b.pos
                # next GOB start code, or only the zero padding of the last byte
                if rem <= 0 or (rem >= 16 and b.peek(16) == 1) or (rem < 16 and b.peek(rem) == 0):
                    break
                d = b.vlc(_MBA_TAB, 11)
                if d == "stuffing":
                    continue
                if addr + d > 33:
                    raise BitstreamError("MBA beyond the GOB")
                if d != 1:
                    prev_mv = None
                addr += d
                i = addr - 1
                if i in (0, 11, 22):
                    prev_mv = None
                intra, mq, mvd, coeff, fil = b.vlc(_MTYPE, 10)
                if mq:
                    quant = b.read(5)
                mv = (0, 0)
                if mvd:
                    dx, dy = b.vlc(_MVD, 11), b.vlc(_MVD, 11)
                    px, py = prev_mv if prev_mv is not None else (0, 0)
                    vx, vy = px + dx, py + dy
                    vx += 32 if vx < -15 else (-32 if vx > 15 else 0)
                    vy += 32 if vy < -15 else (-32 if vy > 15 else 0)
                    mv = (vx, vy)
                    prev_mv = mv
                else:
                    prev_mv = None
                cbp = 63 if intra else (b.vlc(_CBP, 9) if coeff else 0)
                levels = np.zeros((6, 64), np.int32)
                for k in range(6):
                    if cbp & (32 >> k):
                        levels[k] = _read_block(b, bool(intra))
                mx, my = gx + i % 11, gy + i // 11
                mbs[my * mb_w + mx] = dict(intra=bool(intra), quant=quant, mv=mv, filter=bool(fil), cbp=cbp,
                                           levels=levels, mc=bool(mvd), coeff=bool(coeff))
        return self._reconstruct(mbs), mbs, tr

    def _reconstruct(self, mbs):
        w, h = self.w, self.h
        out = [p.copy() for p in self.ref]
        mb_w = w // 16
        coded, where = [], []
        preds = {}
        for m, mb in enumerate(mbs):
            bx, by = 16 * (m % mb_w), 16 * (m // mb_w)
            if mb is None:  # not transmitted: copy of the previous picture
                continue
            for k in range(6):
                p = 0 if k < 4 else k - 3
                px = bx + 8 * (k & 1) if k < 4 else bx // 2
                py = by + 8 * (k >> 1) if k < 4 else by // 2
                if mb["intra"]:
                    pred = np.zeros((8, 8), np.int64)
                else:
                    vx, vy = mb["mv"]
                    if k >= 4:  # chroma vector: luma / 2 truncated toward zero
                        vx, vy = int(vx / 2), int(vy / 2)
                    src = self.ref[p][py + vy:py + vy + 8, px + vx:px + vx + 8]
                    if src.shape != (8, 8):
                        raise BitstreamError("motion vector leaves the picture")
                    pred = _loop_filter(src) if mb["filter"] else src.astype(np.int64)
                preds[(m, k)] = (p, px, py, pred)
                if mb["cbp"] & (32 >> k):
                    coded.append(dequant(mb["levels"][k], mb["quant"], mb["intra"]).reshape(8, 8))
                    where.append((m, k))
        res = idct_blocks(np.stack(coded)) if coded else np.zeros((0, 8, 8), np.int64)
        resid = {mk: res[i] for i, mk in enumerate(where)}
        for (m, k), (p, px, py, pred) in preds.items():
            r = resid.get((m, k), 0)
            out[p][py:py + 8, px:px + 8] = np.clip(pred + r, 0, 255).astype(np.uint8)
        self.ref = out
        return tuple(out)
