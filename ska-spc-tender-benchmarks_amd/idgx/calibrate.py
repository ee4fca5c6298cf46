"""HBM-counter calibration: a plain device-to-device copy of known size.

Run under `rocprofv3 --pmc FETCH_SIZE` (and separately `--pmc WRITE_SIZE`): the counters of the copy kernel divided
by the bytes it must move give the scale factor applied to the gridder/degridder counters to turn them into HBM
bytes per launch.

    python -m idgx.calibrate [--mib 1024] [--repeat 3]
"""
import argparse
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--repeat", type=int, default=3)
    args = ap.parse_args()
    import torch

    n = args.mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.repeat):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"calibrate: copied {args.repeat} x {args.mib} MiB (read {args.mib} MiB + write {args.mib} MiB per copy)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
