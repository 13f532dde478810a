"""Serving: requests admitted while others code produce, per request, the
same bitstream as the request coded alone (bit-exact, the stated
tolerance), and the first ones equal the oracle."""
import numpy as np
import pytest

from conftest import announce
from h261enc_cuda_amd import Engine, content_frame
from h261enc_cuda_amd.serving import Request, ServingLoop, poisson_requests

pytestmark = pytest.mark.gpu


def alone(r: Request):
    eng = Engine(max_streams=1)
    sid = eng.open(r.width, r.height, bitrate=r.bitrate, vbv_size=r.bitrate)
    eng.generate_input(sid, r.seed, r.frames, sigma=r.sigma, motion=r.motion, local=r.local,
                       scene_cut=r.scene_cut)
    out = []
    for _ in range(r.frames):
        eng.step()
        eng.sync()
        out.append(eng.bitstream(sid))
    eng.destroy()
    return out


@pytest.mark.parametrize("async_input,chains", [(True, 0), (False, 0), (True, 2)],
                         ids=["async-input", "sync-input", "async-input-two-chains"])
def test_serving_mix_equals_alone(oracle, async_input, chains):
    reqs = poisson_requests(rate=400.0, duration=0.05, seed=11, median_frames=6, sigma=0.5, cap=12)
    assert len(reqs) >= 8
    announce(f"serving {len(reqs)} requests, async input {async_input}, chains {chains}")
    # a tight budget forces admission to wait for memory
    loop = ServingLoop(hbm_budget=6 * (1 << 20) * 3, max_streams=16, async_input=async_input)
    if chains:
        # the halves of the MB space as separate chains, with rate-control
        # begin launched, as in steps that fill the device
        from h261enc_cuda_amd import _native as nat
        loop.eng.config(nat.OPT_CHAINS, chains)
        loop.eng.config(nat.OPT_FUSED_TAILS, 2)
    stats, outs = loop.run(reqs, collect=True, realtime=False)
    loop.close()
    assert stats.frames == sum(r.frames for r in reqs)
    assert stats.max_concurrent >= 2
    for i, r in enumerate(reqs):
        announce(f"alone request {i}")
        assert outs[i] == alone(r), f"request {i}: mix differs from alone"
    # the first two requests also against the oracle
    for i in range(2):
        r = reqs[i]
        s = oracle.stream(r.width, r.height, bitrate=r.bitrate, vbv=r.bitrate)
        for t in range(r.frames):
            f = content_frame(r.seed, r.width, r.height, t, sigma=r.sigma, motion=r.motion, local=r.local,
                              scene_cut=r.scene_cut)
            bits = s.encode(*f)[0]
            assert bits == outs[i][t], f"request {i} frame {t}: differs from the oracle"
