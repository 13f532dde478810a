"""The synthetic content generator: ranges, determinism and the motion it
claims (a frame is the previous one shifted by the drawn vector)."""
import numpy as np

from h261enc_cuda_amd import content_frame
from h261enc_cuda_amd.serving import poisson_requests


def test_ranges_and_determinism():
    y, cb, cr = content_frame(5, 352, 288, 0, sigma=0)
    assert 16 <= y.min() and y.max() <= 235
    assert y.std() > 10  # a texture, not a flat field
    y2, _, _ = content_frame(5, 352, 288, 0, sigma=0)
    assert np.array_equal(y, y2)
    y3, _, _ = content_frame(6, 352, 288, 0, sigma=0)
    assert not np.array_equal(y, y3)


def test_global_motion_is_an_exact_shift():
    """Without noise, frame t+1 = frame t displaced by a vector in [-M, M]^2."""
    for seed, m in ((3, 7), (4, 15)):
        prev = content_frame(seed, 176, 144, 0, sigma=0, motion=m)[0].astype(int)
        for t in range(1, 6):
            cur = content_frame(seed, 176, 144, t, sigma=0, motion=m)[0].astype(int)
            found = None
            for dy in range(-m, m + 1):
                for dx in range(-m, m + 1):
                    a = cur[20:124, 20:156]
                    b = prev[20 + dy:124 + dy, 20 + dx:156 + dx]
                    if np.array_equal(a, b):
                        found = (dx, dy)
                        break
                if found:
                    break
            assert found is not None, (seed, t)
            prev = cur


def test_noise_statistics():
    a = content_frame(8, 352, 288, 3, sigma=0)[0].astype(int)
    b = content_frame(8, 352, 288, 3, sigma=4)[0].astype(int)
    d = (b - a)[(a > 30) & (a < 220)]
    assert abs(d.mean()) < 0.1
    assert 3.6 < d.std() < 4.4


def test_chroma_follows_luma_texture():
    y, cb, cr = content_frame(9, 352, 288, 0, sigma=0)
    s = y[0::2, 0::2].astype(int) + y[1::2, 0::2] + y[0::2, 1::2] + y[1::2, 1::2]
    assert np.array_equal(cb.astype(int), np.clip(128 + np.trunc((s - 504) / 4).astype(int), 0, 255))
    assert np.array_equal(cr.astype(int), np.clip(128 - np.trunc((s - 504) / 8).astype(int), 0, 255))


def test_local_motion_scene_cuts():
    a = content_frame(10, 176, 144, 49, sigma=0, local=True, scene_cut=50)[0]
    b = content_frame(10, 176, 144, 50, sigma=0, local=True, scene_cut=50)[0]
    c = content_frame(10, 176, 144, 51, sigma=0, local=True, scene_cut=50)[0]
    # a new scene starts at frame 50: unrelated to 49
    assert np.mean(a == b) < 0.2
    # frame 50 is the unshifted texture, so every MB of frame 51 is frame 50
    # displaced by that MB's vector in [-15, 15]^2
    for (mx, my) in ((2, 2), (5, 4), (8, 6)):
        blk = c[16 * my:16 * my + 16, 16 * mx:16 * mx + 16]
        hit = any(np.array_equal(blk, b[16 * my + dy:16 * my + dy + 16, 16 * mx + dx:16 * mx + dx + 16])
                  for dy in range(-15, 16) for dx in range(-15, 16)
                  if 0 <= 16 * my + dy <= 128 and 0 <= 16 * mx + dx <= 160)
        assert hit, (mx, my)


def test_poisson_request_mix():
    reqs = poisson_requests(rate=2000.0, duration=1.0, seed=1)
    assert 1800 < len(reqs) < 2200
    cif = np.mean([r.width == 352 for r in reqs])
    assert 0.55 < cif < 0.65
    lens = np.array([r.frames for r in reqs])
    assert 100 < np.median(lens) < 140 and lens.max() <= 1800
    rates = np.array([r.bitrate for r in reqs])
    assert rates.min() >= 64000 and rates.max() <= 1024000


def test_pack_frames_layout():
    """Engine.pack_frames (the host buffer of h261g_engine_put_frames): each
    frame's Y, Cb, Cr back to back, frames at multiples of 16, nothing
    overlapping."""
    from h261enc_cuda_amd import Engine
    frames = [content_frame(s, w, h, 0) for s, (w, h) in ((1, (352, 288)), (2, (176, 144)), (3, (176, 144)))]
    buf, offs = Engine.pack_frames(frames)
    assert buf.dtype == np.uint8 and buf.ndim == 1
    assert all(o % 16 == 0 for o in offs)
    ends = [o + sum(p.size for p in f) for o, f in zip(offs, frames)]
    assert all(e <= o2 for e, o2 in zip(ends, offs[1:])) and ends[-1] <= buf.size
    for (y, cb, cr), o in zip(frames, offs):
        assert np.array_equal(buf[o:o + y.size].reshape(y.shape), y)
        assert np.array_equal(buf[o + y.size:o + y.size + cb.size].reshape(cb.shape), cb)
        assert np.array_equal(buf[o + y.size + cb.size:o + y.size + cb.size + cr.size].reshape(cr.shape), cr)
