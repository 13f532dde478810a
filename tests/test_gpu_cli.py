"""`python -m h261enc_cuda_amd.cli encode` (the counterpart of
src/h261_main.c) on a synthesized I420 file: the bitstream file is the
oracle's frames back to back, the reconstruction file the oracle's
reconstructions (h261_main.c:253-255), and each printed line carries the
oracle's quantiser, picture cost and size and the PSNR of the input against
that reconstruction (h261_main.c:257-264), then the mean PSNR and kbit/s
summary (h261_main.c:274-279)."""
import re

import numpy as np
import pytest

from conftest import announce
from h261enc_cuda_amd import cli
from h261enc_cuda_amd.encoder import psnr

pytestmark = pytest.mark.gpu

LINE = re.compile(r"frame (\d{4}), quant=(\d+), cost=(\d+), size=(\d+), ypsnr=([\d.]+), cbpsnr=([\d.]+), "
                  r"crpsnr=([\d.]+), time=[\d.]+ms")


@pytest.mark.parametrize("qcif,quant,n", [(True, 0, 12), (False, 8, 6)], ids=["qcif-rc", "cif-q8"])
def test_cli_encode_matches_oracle(tmp_path, capsys, oracle, qcif, quant, n):
    w, h = (176, 144) if qcif else (352, 288)
    src, out, rec = tmp_path / "in.yuv", tmp_path / "out.h261", tmp_path / "rec.yuv"
    size = ["--qcif"] if qcif else []
    assert cli.main(["synth", str(src), "--frames", str(n), "--seed", "13", "--sigma", "2", "--motion", "7"] + size) == 0
    announce(f"cli encode {w}x{h} quant={quant}")
    args = ["encode", str(src), str(out), str(rec)] + size + (["--quant", str(quant)] if quant else [])
    assert cli.main(args) == 0
    lines = [ln for ln in capsys.readouterr().out.splitlines() if not ln.startswith("[case]")]
    s = oracle.stream(w, h, fixed_quant=quant)
    bits_all, rec_all, ps_sum = b"", b"", np.zeros(3)
    frames = list(cli.read_frames(str(src), w, h))
    assert len(frames) == n and len(lines) == n + 2
    for t, f in enumerate(frames):
        bits, orec, _, _, q = s.encode(*f)
        bits_all += bits
        rec_all += b"".join(p.tobytes() for p in orec)
        m = LINE.fullmatch(lines[t])
        assert m, lines[t]
        ps = [psnr(a, b) for a, b in zip(f, orec)]
        ps_sum += ps
        assert (int(m.group(1)), int(m.group(2)), int(m.group(3)), int(m.group(4))) == \
               (t, q, s.picture_cost(), len(bits)), lines[t]
        assert m.group(5, 6, 7) == tuple(f"{p:f}" for p in ps), lines[t]
    assert out.read_bytes() == bits_all
    assert rec.read_bytes() == rec_all
    assert lines[n] == "PSNR Mean: %f, %f, %f" % tuple(ps_sum / n)
    kbit = len(bits_all) * 8 / (n / 0.02997)
    assert lines[n + 1].startswith("%f kBit/s, " % kbit)
