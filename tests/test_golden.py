"""The committed golden fixtures (tests/golden/golden.json, written by
tests/golden/make_golden.py from the reference's own host code and from the
oracle) against the oracle, the product's host writer and content generator
(CPU), and against the engine on the GPU.  Runs without the reference."""
import json
import os

import numpy as np
import pytest

from h261enc_cuda_amd import content_frame, write_picture
from oracle.binding import RC_DTYPE
from tests.conftest import announce
from tests.golden.cases import (mode_decision_cases, numeric_blocks, rate_control_cases, sha, stream_cases,
                                writer_cases)

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def test_writer_golden(oracle):
    """oracle and product writers reproduce the reference writer's bytes"""
    for name, mbs, mb_w, mb_h, sf, fn, q, ft in writer_cases():
        g = GOLDEN["writer"][name]
        ora = oracle.write_picture(mbs, mb_w, sf, fn & 31, q)
        prod = write_picture(mbs, mb_w * 16, mb_h * 16, fn, q)
        assert (len(ora), sha(ora)) == (g["bytes"], g["sha256"]), name
        assert (len(prod), sha(prod)) == (g["bytes"], g["sha256"]), name


def test_mode_decision_golden(oracle):
    for name, me, ft in mode_decision_cases():
        md = oracle.mode_decision(me, ft)
        g = GOLDEN["mode_decision"][name]
        assert md["type"][:24].tolist() == g["types"], name
        assert sha(md["type"], md["flags"], md["mv_x"], md["mv_y"], md["cost"]) == g["sha256"], name


def test_rate_control_golden(oracle):
    lib = oracle.lib
    for name, quant0, rate, vbv, fixed, n_mb, frames in rate_control_cases():
        g = GOLDEN["rate_control"][name]
        rc = np.zeros(1, RC_DTYPE)
        lib.ora_rc_init(rc.ctypes.data, quant0, rate, vbv, 2)
        quants = []
        for ft, md, nbytes in frames:
            rc["picture_cost"] = lib.ora_rc_picture_cost(md.ctypes.data, n_mb)
            q = lib.ora_rc_init_picture(rc.ctypes.data, ft)
            q = min(31, max(1, fixed if fixed else q))
            quants.append(q)
            lib.ora_rc_update_picture(rc.ctypes.data, ft, q, nbytes)
        r = rc[0]
        state = [r["mode"], r["current_quant"], r["picture_cost"], *r["prediction_cost"], *r["prediction_size"],
                 r["bitrate"], r["vbv_fill"]]
        assert quants == g["quants"], name
        assert [int(v) for v in state] == g["final_state"], name


@pytest.mark.parametrize("case", stream_cases(), ids=[c[0] for c in stream_cases()])
def test_oracle_streams_golden(oracle, case):
    name, w, h, sp, cp, n = case
    s = oracle.stream(w, h, **sp)
    for t, g in enumerate(GOLDEN["streams"][name]):
        f = content_frame(cp["seed"], w, h, t, sigma=cp["sigma"], motion=cp["motion"])
        bits, rec, _, _, q = s.encode(*f, want_details=True)
        assert (len(bits), sha(bits), sha(*rec), q) == (g["bytes"], g["bits_sha256"], g["recon_sha256"],
                                                        g["quant"]), (name, t)
    s.close()


def test_content_golden():
    for name, w, h, sp, cp, n in stream_cases():
        for t, g in enumerate(GOLDEN["content"][name]):
            assert sha(*content_frame(cp["seed"], w, h, t, sigma=cp["sigma"], motion=cp["motion"])) == g, (name, t)


def test_numerics_golden(oracle):
    num = GOLDEN["numerics"]
    assert sha(np.stack([oracle.fdct(b) for b in numeric_blocks(11, 4000, -255, 255)])) == num["fdct_residual"]
    assert sha(np.stack([oracle.fdct(b) for b in numeric_blocks(12, 4000, 0, 255)])) == num["fdct_intra"]
    assert sha(np.stack([oracle.idct(b) for b in numeric_blocks(13, 4000, -2048, 2047)])) == num["idct"]
    rng = np.random.default_rng(16)
    cur = rng.integers(0, 256, (144, 176)).astype(np.uint8)
    refy = np.roll(cur, (3, -5), axis=(0, 1)).copy()
    refy[rng.random(refy.shape) < 0.05] = 0
    assert sha(oracle.me_frame(cur, refy)) == num["me_qcif"]


def test_golden_detects_a_planted_mistake(oracle):
    """a one-off change of the writer's input (one level) changes the
    recorded digest: the fixtures are sensitive at this granularity"""
    name, mbs, mb_w, mb_h, sf, fn, q, ft = writer_cases()[1]
    mbs = mbs.copy()
    m = int(np.nonzero(mbs["cbp"])[0][0])
    mbs[m]["tcoeff"][0][5] += 1 if mbs[m]["tcoeff"][0][5] < 100 else -1
    assert sha(oracle.write_picture(mbs, mb_w, sf, fn & 31, q)) != GOLDEN["writer"][name]["sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", stream_cases(), ids=[c[0] for c in stream_cases()])
def test_engine_streams_golden(native, case):
    """the engine's bitstream, reconstruction and quantiser per frame equal
    the committed digests"""
    from h261enc_cuda_amd import Engine
    name, w, h, sp, cp, n = case
    announce(f"test_engine_streams_golden {name}")
    eng = Engine(max_streams=1)
    sid = eng.open(w, h, denoise=5, init_quant=sp.get("init_quant", 10), bitrate=sp.get("bitrate", 512000),
                   vbv_size=sp.get("vbv", 512000), fixed_quant=sp.get("fixed_quant", 0))
    try:
        for t, g in enumerate(GOLDEN["streams"][name]):
            eng.put_frame(sid, *content_frame(cp["seed"], w, h, t, sigma=cp["sigma"], motion=cp["motion"]))
            eng.step()
            eng.sync()
            bits = eng.bitstream(sid)
            info = eng.info(sid)
            assert (len(bits), sha(bits), sha(*eng.recon(sid)), info.quant) == (
                g["bytes"], g["bits_sha256"], g["recon_sha256"], g["quant"]), (name, t)
            assert info.bytes == g["bytes"], (name, t)
    finally:
        eng.destroy()
