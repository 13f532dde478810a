"""Writes tests/golden/golden.json: expected outputs of the fixture cases
(tests/golden/cases.py).

* writer, mode decision, rate control: the reference's own host code
  (src/h261_picture.c, h261_gob.c, h261_macroblock.c, h261_bitstream.c,
  h261_tables.c, h261_moded.c, h261_ratectrl.c), built from its sources by
  `make -C oracle ref` into oracle/_ref/libh261ref.so;
* whole streams, device arithmetic (DCT, IDCT, quantiser, dequantiser, loop
  filter, SATD, motion search) and the content generator: the oracle, whose
  writer/MD/RC are the ones checked against the reference above.

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from h261enc_cuda_amd import content_frame  # noqa: E402
from oracle.binding import RC_DTYPE, Oracle, Reference  # noqa: E402
from tests.golden.cases import (mode_decision_cases, numeric_blocks, rate_control_cases, sha,  # noqa: E402
                                stream_cases, writer_cases)


def main():
    ref = Reference()
    ora = Oracle()
    g = {"source": "tests/golden/make_golden.py"}

    g["writer"] = {}
    for name, mbs, mb_w, mb_h, sf, fn, q, ft in writer_cases():
        bits = ref.write_picture(mbs, mb_w, mb_h, sf, fn, q, ft)
        g["writer"][name] = {"bytes": len(bits), "sha256": sha(bits), "head": bits[:48].hex()}

    g["mode_decision"] = {}
    for name, me, ft in mode_decision_cases():
        md = ref.mode_decision(me, 22, 18, ft)
        g["mode_decision"][name] = {"sha256": sha(md["type"], md["flags"], md["mv_x"], md["mv_y"], md["cost"]),
                                    "types": md["type"][:24].tolist(), "flags": md["flags"][:24].tolist()}

    g["rate_control"] = {}
    for slot, (name, quant0, rate, vbv, fixed, n_mb, frames) in enumerate(rate_control_cases()):
        ref.lib.refdrv_rc_init(slot, quant0, rate, vbv)
        mb_w = 11 if n_mb == 99 else 22
        quants, states = [], []
        for t, (ft, md, nbytes) in enumerate(frames):
            quants.append(int(ref.lib.refdrv_rc_begin(slot, md.ctypes.data, mb_w, n_mb // mb_w, t, ft, fixed)))
            ref.lib.refdrv_rc_end(slot, nbytes)
            st = (ctypes.c_int * 9)()
            ref.lib.refdrv_rc_state(slot, st)
            states.append(list(st))
        g["rate_control"][name] = {"quants": quants, "final_state": states[-1], "states_sha256": sha(np.array(states))}

    g["streams"] = {}
    for name, w, h, sp, cp, n in stream_cases():
        s = ora.stream(w, h, **sp)
        frames = []
        for t in range(n):
            y, cb, cr = content_frame(cp["seed"], w, h, t, sigma=cp["sigma"], motion=cp["motion"])
            bits, (ry, rcb, rcr), _, _, q = s.encode(y, cb, cr, want_details=True)
            frames.append({"bytes": len(bits), "bits_sha256": sha(bits), "recon_sha256": sha(ry, rcb, rcr), "quant": q})
        s.close()
        g["streams"][name] = frames

    g["content"] = {}
    for name, w, h, sp, cp, n in stream_cases():
        g["content"][name] = [sha(*content_frame(cp["seed"], w, h, t, sigma=cp["sigma"], motion=cp["motion"]))
                              for t in range(3)]

    num = {}
    blocks = numeric_blocks(11, 4000, -255, 255)
    num["fdct_residual"] = sha(np.stack([ora.fdct(b) for b in blocks]))
    blocks = numeric_blocks(12, 4000, 0, 255)
    num["fdct_intra"] = sha(np.stack([ora.fdct(b) for b in blocks]))
    blocks = numeric_blocks(13, 4000, -2048, 2047)
    num["idct"] = sha(np.stack([ora.idct(b) for b in blocks]))
    coef = np.arange(-4096, 4096, dtype=np.int32).reshape(-1, 8, 8)
    qt = []
    for q in range(1, 32):
        for intra in (False, True):
            for b in coef:
                lv, coded = ora.quant(b, q, intra)
                qt.append(lv)
                qt.append(np.array([coded], np.int32))
    num["quant_table"] = sha(*qt)
    lv = numeric_blocks(14, 2000, -127, 127)
    dq = []
    for i, b in enumerate(lv):
        q = 1 + i % 31
        intra = i % 5 == 0
        if intra:
            b = b.copy()
            b[0, 0] = 1 + (i * 37) % 255
        dq.append(ora.dequant(b, q, intra, True))
    num["dequant"] = sha(np.stack(dq))
    rng = np.random.default_rng(15)
    ref_plane = rng.integers(0, 256, (64, 64)).astype(np.uint8)
    pr = [ora.predict(ref_plane, int(px), int(py), bool(f)) for px, py, f in
          zip(rng.integers(0, 56, 500), rng.integers(0, 56, 500), rng.integers(0, 2, 500))]
    num["predict"] = sha(np.stack(pr))
    d = rng.integers(-255, 256, (5000, 16)).astype(np.int16)
    num["satd"] = sha(np.array([ora.satd4x4(x) for x in d], np.int32))
    rng = np.random.default_rng(16)
    cur = rng.integers(0, 256, (144, 176)).astype(np.uint8)
    refy = np.roll(cur, (3, -5), axis=(0, 1)).copy()
    refy[rng.random(refy.shape) < 0.05] = 0
    me = ora.me_frame(cur, refy)
    num["me_qcif"] = sha(me)
    g["numerics"] = num

    path = os.path.join(ROOT, "tests", "golden", "golden.json")
    with open(path, "w") as f:
        json.dump(g, f, indent=1, sort_keys=True)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
