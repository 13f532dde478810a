"""Command line encoder: the counterpart of the reference's src/h261_main.c.

  python -m h261enc_cuda_amd.cli encode IN.yuv OUT.h261 RECON.yuv [--qcif]
         [--quant Q | --bitrate BPS] [--denoise 5] [--frames N]
  python -m h261enc_cuda_amd.cli synth OUT.yuv --frames N [--qcif] [--seed S]
         [--sigma 2] [--motion 7]

encode reads I420 frames (Y, Cb, Cr planes back to back, as h261_main.c:
178-184), codes them on the GPU, writes the bitstream and the
reconstruction, and prints per-frame quantiser, picture cost, size, PSNR and
time (h261_main.c:257-264) and the
summary line (mean PSNR, kbit/s at 29.97 frames/s, frames/s) of
h261_main.c:257-279.  Rate control defaults to the reference's settings
(initial quantiser 10, 512 kbit/s, VBV 512 kbit, h261_main.c:165); --quant
fixes the picture quantiser instead.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from .encoder import Engine, content_frame, psnr


def read_frames(path, w, h, limit=None):
    fsz = w * h * 3 // 2
    with open(path, "rb") as f:
        n = 0
        while limit is None or n < limit:
            buf = f.read(fsz)
            if len(buf) < fsz:
                return
            a = np.frombuffer(buf, np.uint8)
            yield (a[:w * h].reshape(h, w), a[w * h:w * h * 5 // 4].reshape(h // 2, w // 2),
                   a[w * h * 5 // 4:].reshape(h // 2, w // 2))
            n += 1


def encode(args) -> int:
    w, h = (176, 144) if args.qcif else (352, 288)
    eng = Engine(max_streams=1, mbs_per_wave=1)
    sid = eng.open(w, h, denoise=args.denoise, init_quant=10, bitrate=args.bitrate, vbv_size=args.bitrate,
                   fixed_quant=args.quant)
    psnr_sum = np.zeros(3)
    total_bytes = 0
    n = 0
    t_start = time.perf_counter()
    with open(args.output, "wb") as fout, open(args.recon, "wb") as frec:
        for f in read_frames(args.input, w, h, args.frames):
            t0 = time.perf_counter()
            eng.put_frame(sid, *f)
            eng.step()
            eng.sync()
            bits = eng.bitstream(sid)
            info = eng.info(sid)
            rec = eng.recon(sid)
            dt = (time.perf_counter() - t0) * 1e3
            fout.write(bits)
            for p in rec:
                frec.write(p.tobytes())
            ps = [psnr(a, b) for a, b in zip(f, rec)]
            psnr_sum += ps
            total_bytes += len(bits)
            print(f"frame {n:04d}, quant={info.quant}, cost={info.cost}, size={len(bits)}, ypsnr={ps[0]:f}, "
                  f"cbpsnr={ps[1]:f}, crpsnr={ps[2]:f}, time={dt:.3f}ms")
            n += 1
    elapsed = time.perf_counter() - t_start
    eng.destroy()
    if n == 0:
        print("no frames read", file=sys.stderr)
        return 1
    print("PSNR Mean: %f, %f, %f" % tuple(psnr_sum / n))
    print("%f kBit/s, %f fps" % (total_bytes * 8 / (n / 0.02997), n / elapsed))
    return 0


def synth(args) -> int:
    w, h = (176, 144) if args.qcif else (352, 288)
    with open(args.output, "wb") as f:
        for t in range(args.frames):
            for p in content_frame(args.seed, w, h, t, sigma=args.sigma, motion=args.motion):
                f.write(p.tobytes())
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="h261enc_cuda_amd.cli")
    sub = ap.add_subparsers(dest="cmd", required=True)
    e = sub.add_parser("encode")
    e.add_argument("input")
    e.add_argument("output")
    e.add_argument("recon")
    e.add_argument("--qcif", action="store_true")
    e.add_argument("--quant", type=int, default=0)
    e.add_argument("--bitrate", type=int, default=512000)
    e.add_argument("--denoise", type=int, default=5)
    e.add_argument("--frames", type=int, default=None)
    s = sub.add_parser("synth")
    s.add_argument("output")
    s.add_argument("--frames", type=int, default=30)
    s.add_argument("--qcif", action="store_true")
    s.add_argument("--seed", type=int, default=1)
    s.add_argument("--sigma", type=int, default=2)
    s.add_argument("--motion", type=int, default=7)
    args = ap.parse_args(argv)
    return encode(args) if args.cmd == "encode" else synth(args)


if __name__ == "__main__":
    sys.exit(main())
