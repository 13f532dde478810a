"""The oracle's writer, mode decision and rate control against the
reference's own host code, compiled from its sources into oracle/_ref
(src/h261_picture.c, h261_gob.c, h261_macroblock.c, h261_bitstream.c,
h261_tables.c, h261_moded.c, h261_ratectrl.c), and the product's host writer
against both."""
import ctypes

import numpy as np
import pytest

from h261enc_cuda_amd import content_frame, write_picture
from oracle.binding import MACROBLOCK_DTYPE, ME_MB_DTYPE, MODE_DTYPE, RC_DTYPE


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


@pytest.mark.parametrize("size", [(352, 288), (176, 144)], ids=["cif", "qcif"])
def test_writer_matches_reference_on_random_macroblocks(oracle, reference, size):
    w, h = size
    mb_w, mb_h = w // 16, h // 16
    sf = 1 if w == 352 else 0
    rng = np.random.default_rng(w)
    for trial in range(40):
        q = int(rng.integers(1, 32))
        intra = trial % 5 == 0
        mbs = random_macroblocks(rng, mb_w, mb_h, q, intra)
        ref = reference.write_picture(mbs, mb_w, mb_h, sf, trial, q, 0 if intra else 1)
        ora = oracle.write_picture(mbs, mb_w, sf, trial & 31, q)
        prod = write_picture(mbs, w, h, trial, q)
        assert ora == ref, f"trial {trial}: oracle writer differs from the reference"
        assert prod == ref, f"trial {trial}: product writer differs from the reference"


def test_writer_matches_reference_on_encoded_frames(oracle, reference):
    w, h = 352, 288
    s = oracle.stream(w, h)
    for t in range(4):
        f = content_frame(77, w, h, t, sigma=4, motion=15)
        bits, _, _, mbs, q = s.encode(*f, want_details=True)
        assert reference.write_picture(mbs, 22, 18, 1, t, q, 0 if t == 0 else 1) == bits
        assert write_picture(mbs, w, h, t, q) == bits


def test_mode_decision_matches_reference(oracle, reference):
    rng = np.random.default_rng(2)
    for trial in range(50):
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
        ft = 0 if trial % 4 == 0 else 1
        assert (oracle.mode_decision(me, ft) == reference.mode_decision(me, 22, 18, ft)).all(), trial


def test_rate_control_matches_reference(oracle, reference):
    lib = oracle.lib
    rng = np.random.default_rng(4)
    for trial in range(12):
        quant0 = int(rng.integers(1, 32))
        rate = int(rng.integers(8000, 2000000))
        vbv = int(rate * rng.uniform(0.5, 2.0))
        slot = trial % 16
        reference.lib.refdrv_rc_init(slot, quant0, rate, vbv)
        rc = np.zeros(1, RC_DTYPE)
        lib.ora_rc_init(rc.ctypes.data, quant0, rate, vbv, 2)
        n_mb = 99 if trial % 2 else 396
        mb_w = 11 if n_mb == 99 else 22
        for t in range(60):
            ft = 0 if t == 0 or (trial % 3 == 0 and t % 25 == 0) else 1
            md = np.zeros(n_mb, MODE_DTYPE)
            md["cost"] = rng.integers(0, int(rng.choice([50, 500, 5000])), n_mb)
            fixed = 8 if trial == 5 else 0
            qr = reference.lib.refdrv_rc_begin(slot, md.ctypes.data, mb_w, n_mb // mb_w, t, ft, fixed)
            rc["picture_cost"] = lib.ora_rc_picture_cost(md.ctypes.data, n_mb)
            qo = lib.ora_rc_init_picture(rc.ctypes.data, ft)
            if fixed:
                qo = fixed
            qo = min(31, max(1, qo))
            assert qo == qr, (trial, t)
            nbytes = int(rng.integers(10, 60000))
            reference.lib.refdrv_rc_end(slot, nbytes)
            lib.ora_rc_update_picture(rc.ctypes.data, ft, qo, nbytes)
            st = (ctypes.c_int * 9)()
            reference.lib.refdrv_rc_state(slot, st)
            r = rc[0]
            mine = [r["mode"], r["current_quant"], r["picture_cost"], *r["prediction_cost"], *r["prediction_size"],
                    r["bitrate"], r["vbv_fill"]]
            assert [int(v) for v in mine] == list(st), (trial, t)


def test_abi_layout_matches_reference_headers(oracle, reference):
    """include/h261gpu_abi.h lays out every struct exactly as the reference's
    src/h261_types.h (sizes and the offsets the device layer uses)."""
    assert oracle.layout() == reference.layout()
