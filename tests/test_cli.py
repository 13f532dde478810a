"""The command line counterpart of src/h261_main.c, on the host: I420 file
round trip of the content generator (h261_main.c:178-184 reads planes back
to back) and the PSNR of h261_main.c:38-91 on cases with a known squared
error."""
import math

import numpy as np

from h261enc_cuda_amd import cli, content_frame
from h261enc_cuda_amd.encoder import psnr


def test_synth_writes_i420_frames_read_back(tmp_path):
    path = tmp_path / "s.yuv"
    assert cli.main(["synth", str(path), "--frames", "3", "--qcif", "--seed", "4", "--sigma", "1",
                     "--motion", "5"]) == 0
    assert path.stat().st_size == 3 * 176 * 144 * 3 // 2
    frames = list(cli.read_frames(str(path), 176, 144))
    assert len(frames) == 3
    for t, f in enumerate(frames):
        want = content_frame(4, 176, 144, t, sigma=1, motion=5)
        for p in range(3):
            assert np.array_equal(f[p], want[p])
    assert len(list(cli.read_frames(str(path), 176, 144, limit=2))) == 2


def test_psnr_is_the_reference_formula():
    a = np.full((144, 176), 90, np.uint8)
    assert psnr(a, a) == 100.0
    b = a.copy()
    b[3, 7] = 91  # one pel off by one: ssd 1
    assert psnr(a, b) == 10 * math.log10(255.0 * 255.0 * 176 * 144 / 1.0)
    c = a.copy()
    c[:, :] = 80  # every pel off by 10 (unsigned wrap-around in the reference squares the same)
    assert abs(psnr(a, c) - 10 * math.log10(255.0 ** 2 / 100.0)) < 1e-12
    d = np.zeros((72, 88), np.uint8)
    e = np.full((72, 88), 255, np.uint8)
    assert psnr(d, e) == 0.0 and psnr(e, d) == 0.0
