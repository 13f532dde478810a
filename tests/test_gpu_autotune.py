"""The launch-shape autotuner: it picks one of the candidates it measured,
and the shape it picks changes timing only -- the bitstreams are the same
bytes under any shape."""
import pytest

from conftest import announce
from h261enc_cuda_amd import Engine, autotune
from h261enc_cuda_amd.encoder import CIF, QCIF

pytestmark = pytest.mark.gpu


def encode(sizes, mpw, remap, me_mpw=0, me_tile=0, n=4):
    eng = Engine(max_streams=len(sizes))
    try:
        autotune.apply(eng, {"mbs_per_wave": mpw, "xcd_remap": remap, "me_mbs_per_wave": me_mpw, "me_tile": me_tile})
        sids = []
        for i, size in enumerate(sizes):
            sid = eng.open(*size, bitrate=128000, vbv_size=128000)
            eng.generate_input(sid, 500 + i, n, sigma=3, motion=15)
            sids.append(sid)
        out = []
        for _ in range(n):
            eng.step()
            eng.sync()
            out.append([eng.bitstream(sid) for sid in sids])
        return out
    finally:
        eng.destroy()


def test_autotune_picks_a_measured_candidate_and_keeps_results():
    sizes = [CIF] * 3 + [QCIF] * 3
    announce("autotune 6 streams, 2 + 4 + 2 candidates")
    r = autotune.tune(sizes, mbs_per_wave=(1, 4), xcd_remap=(0, 2), me_mbs_per_wave=(1, 2), me_tile=(1, 2), steps=2)
    assert r["mbs_per_wave"] in (1, 4) and r["xcd_remap"] in (0, 2)
    assert (r["me_mbs_per_wave"], r["me_tile"]) in ((0, 0), (1, 1), (2, 1), (2, 2))
    assert len(r["trials"]) == 2 * 2 + 2 * 4 + 2 and all(t["step_ms"] > 0 for t in r["trials"])
    key = ("mbs_per_wave", "xcd_remap", "me_mbs_per_wave", "me_tile")
    pick = tuple(r[k] for k in key)
    assert r["step_ms"] == min(t["step_ms"] for t in r["trials"] if tuple(t[k] for k in key) == pick)
    other = (1 if r["mbs_per_wave"] != 1 else 4, 2 if r["xcd_remap"] != 2 else 0,
             *((2, 2) if (r["me_mbs_per_wave"], r["me_tile"]) != (2, 2) else (1, 1)))
    announce(f"autotune shapes {pick} vs {other}")
    assert encode(sizes, *pick) == encode(sizes, *other)
