"""Closed loop: the project's H.261 decoder (parsing with the standard's VLC
tables, the standard's dequantiser, MC and loop filter) reproduces the
oracle's reconstruction from the oracle's bitstream, byte for byte.  This
pins the writer's syntax and the reconstruction path of the oracle against
an implementation that shares no code with it."""
import numpy as np
import pytest

from h261enc_cuda_amd import content_frame
from h261enc_cuda_amd.decoder import BitstreamError, Decoder, idct_blocks


@pytest.mark.parametrize("size,fixed,motion,sigma,local", [
    ((352, 288), 8, 7, 2, False), ((176, 144), 0, 15, 4, False), ((352, 288), 0, 15, 4, False),
    ((176, 144), 0, 0, 4, True)], ids=["cif-q8", "qcif-rc", "cif-rc", "qcif-local"])
def test_decode_reproduces_oracle_reconstruction(oracle, size, fixed, motion, sigma, local):
    w, h = size
    s = oracle.stream(w, h, fixed_quant=fixed)
    d = Decoder(w, h)
    for t in range(6):
        f = content_frame(9, w, h, t, sigma=sigma, motion=motion, local=local, scene_cut=4 if local else 0)
        bits, rec, _, mbs, q = s.encode(*f, want_details=True)
        dec, dmbs, tr = d.decode(bits)
        assert tr == t & 31
        for p in range(3):
            assert np.array_equal(dec[p], rec[p]), (t, p)
        for m, mb in enumerate(dmbs):
            if mb is None:
                assert mbs[m]["flags"] == 0
                continue
            assert mb["cbp"] == (mbs[m]["cbp"] if mbs[m]["type"] == 2 else 63)
            if mb["mc"]:
                assert tuple(mb["mv"]) == tuple(mbs[m]["mv"])


def test_idct_matches_oracle(oracle):
    rng = np.random.default_rng(12)
    coef = np.zeros((400, 8, 8), np.int64)
    for i in range(400):
        k = rng.integers(1, 20)
        coef[i].ravel()[rng.integers(0, 64, k)] = rng.integers(-2048, 2048, k)
    got = idct_blocks(coef)
    for i in range(400):
        assert (got[i] == oracle.idct(coef[i])).all()


def test_decoder_rejects_garbage():
    d = Decoder(176, 144)
    with pytest.raises(BitstreamError):
        d.decode(b"\xff\xff\xff\xff")
    with pytest.raises(BitstreamError):
        d.decode(b"")
