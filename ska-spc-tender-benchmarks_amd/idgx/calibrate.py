"""HBM-counter calibration and copy ceiling: a plain device-to-device copy of known size.

Run under `rocprofv3 --pmc FETCH_SIZE` (and separately `--pmc WRITE_SIZE`): the counters of the copy kernel divided
by the bytes it must move give the scale factor applied to the gridder/degridder counters to turn them into HBM
bytes per launch.  With --time it also reports the copy's own rate (read + write bytes per second, HIP events over
`repeat` copies after one warm-up copy): the practical ceiling of a kernel that streams each byte once in and once
out, against which the HBM-bound kernels are read.

    python -m idgx.calibrate [--mib 1024] [--repeat 3] [--time]
"""
import argparse
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--time", action="store_true", help="report the copy rate")
    args = ap.parse_args()
    import torch

    n = args.mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    if args.time:
        dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
    for _ in range(args.repeat):
        dst.copy_(src)
    if args.time:
        e1.record()
    torch.cuda.synchronize()
    print(f"calibrate: copied {args.repeat} x {args.mib} MiB (read {args.mib} MiB + write {args.mib} MiB per copy)")
    if args.time:
        ms = e0.elapsed_time(e1) / args.repeat
        rate = 2 * args.mib * 1024 * 1024 / (ms * 1e-3) / 1e9
        print(f"calibrate: {ms:.3f} ms per copy, {rate:.0f} GB/s read + write ({rate / 8000:.2f} of 8 TB/s)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
