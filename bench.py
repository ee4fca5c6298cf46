"""Benchmark: IDG gridder + degridder round trip on one MI355X.

A step is one gridder pass plus one degridder pass over the whole synthetic observation (default: the
SKA-Low-like configuration: 128 stations, 256 x 1 s, 64 channels, S=32, G=4096, full w-term), with every input
already resident in HBM.  Prints ONE JSON line (rank 0) with the round-trip throughput in Mvis/s, the per-direction
rates, the roofline of the dominant kernel measured with HIP events, and the CPU baseline (the oracle's faithful
restatement of the reference OpenMP CPU path) timed on a deterministic subset of the same workload.

    python bench.py [--steps K] [--warmup W] [--config ska_low] [--precision fp32]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ska-spc-tender-benchmarks_amd"))

import numpy as np  # noqa: E402

FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 (vector and matrix), published
HBM_PEAK_GBS = 8000.0


def flop_model(md, ch, S, C_all):
    """Algorithmic work of one gridder (or degridder) pass: per subgrid P*(5*T + 34*V) + 134*P flops and P*V
    sin/cos pairs (P = S^2 pixels, T timesteps, V = T * channels visibilities)."""
    P = S * S
    T = md["nr_timesteps"].astype(np.int64)
    nc = ch["nr_channels"].astype(np.int64) if ch is not None else np.full(T.shape, C_all, np.int64)
    V = T * nc
    flops = int((P * (5 * T + 34 * V) + 134 * P).sum())
    sincos = int((P * V).sum())
    return flops, sincos, int(V.sum()), int(T.sum())


def algorithmic_bytes(md, ch, S, C_all, n_slots, n_stations, direction):
    """HBM bytes one pass must move: visibilities 32 B, uvw 12 B per timestep, metadata 36 B and channel window
    8 B per subgrid, wavenumbers, taper, the A-terms once, and the subgrids (32 B per pixel)."""
    P = S * S
    T = md["nr_timesteps"].astype(np.int64)
    nc = ch["nr_channels"].astype(np.int64) if ch is not None else np.full(T.shape, C_all, np.int64)
    n = md.size
    b = 32 * int((T * nc).sum()) + 12 * int(T.sum()) + 44 * n + 4 * C_all + 4 * P + 32 * P * n_slots * n_stations
    b += 32 * P * n
    return b


def cpu_baseline(w, every, nthreads):
    """Oracle FP32 restatement (same loop nest / OpenMP structure as the reference CPU path) on every k-th subgrid."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from idgx import observation
    from helpers import oracle_degrid, oracle_grid

    idx = np.arange(0, w["nr_subgrids"], every)
    sw = observation.subset(w, idx)
    t0 = time.perf_counter()
    sub = oracle_grid(sw, nthreads=nthreads)
    t1 = time.perf_counter()
    oracle_degrid(sw, sub, nthreads=nthreads)
    t2 = time.perf_counter()
    nvis = sw["nr_visibilities"]
    return dict(grid_s=t1 - t0, degrid_s=t2 - t1, nvis=nvis, nsub=idx.size)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="ska_low")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--cpu-every", type=int, default=0, help="CPU baseline on every k-th subgrid (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the hot path is per-GPU; see the sharding design); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    import idgx
    from idgx import observation
    from idgx.device import DeviceWorkload

    dev = torch.device("cuda:0")
    t_setup = time.perf_counter()
    w = observation.make_observation(args.config, visibilities=False)
    C = int(w["wavenumbers"].shape[0])
    n_rows = int(w["nr_rows"])
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(w["config"]["seed"]))
    # Synthetic complex64 visibilities, real/imag ~ N(0,1), generated in place in HBM.
    d_vis = torch.randn(n_rows * C * 8, generator=gen, device=dev, dtype=torch.float32)
    dw = DeviceWorkload(w, device=dev, visibilities_on_device=d_vis)
    plan = dw.plan(precision=args.precision)
    stream = torch.cuda.current_stream()
    print(f"bench: {args.config}: {w['nr_subgrids']} subgrids, {w['nr_visibilities']} visibilities, "
          f"setup {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        dw.grid(plan, stream=stream)
        dw.degrid(plan, stream=stream)
    torch.cuda.synchronize()

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        dw.grid(plan, stream=stream)
        ev[k][1].record(stream)
        dw.degrid(plan, stream=stream)
        ev[k][2].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_grid = sum(e[0].elapsed_time(e[1]) for e in ev) / args.steps / 1e3
    t_degrid = sum(e[1].elapsed_time(e[2]) for e in ev) / args.steps / 1e3

    md, ch, S = w["metadata"], w["channels"], int(w["subgrid_size"])
    flops, sincos, nvis, _ = flop_model(md, ch, S, C)
    step_s = wall / args.steps
    value = nvis / (t_grid + t_degrid) / 1e6
    dominant, t_dom = ("degridder", t_degrid) if t_degrid > t_grid else ("gridder", t_grid)
    achieved = flops / t_dom / 1e12
    bytes_dom = algorithmic_bytes(md, ch, S, C, int(w["nr_aterm_slots"]), int(w["nr_stations"]), dominant)

    cpu = None
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        every = args.cpu_every or max(1, w["nr_subgrids"] // 256)
        host_vis = d_vis.view(torch.uint8).cpu().numpy().view(np.complex64).reshape(n_rows, C, 4)
        w["visibilities"] = host_vis
        r = cpu_baseline(w, every, nthreads)
        cpu = {
            "value": r["nvis"] / (r["grid_s"] + r["degrid_s"]) / 1e6,
            "unit": "Mvis/s",
            "cores": nthreads,
            "kind": "port",
            "sample": f"every {every}th subgrid of the plan ({r['nsub']} subgrids, {r['nvis']} visibilities); "
                      f"gridder {r['nvis'] / r['grid_s'] / 1e6:.3f} Mvis/s, degridder "
                      f"{r['nvis'] / r['degrid_s'] / 1e6:.3f} Mvis/s; oracle FP32 restatement of the reference "
                      f"OpenMP CPU path, -O3 -ffp-contract=off",
        }

    out = {
        "metric": "Mvis/s gridder+degridder, 1 MI355X: FP32 vs MFMA mixed at rel-err<=5e-5",
        "value": value,
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else args.precision,
        "data": "synthetic (seeded layout/uvw/plan; N(0,1) complex64 visibilities generated in HBM)",
        "config": {
            "workload": f"{args.config}: {w['config']['n_stations']} stations, {w['config']['n_times']} x 1 s, "
                        f"{C} channels, S={S}, G={w['grid_size']}, image {float(w['image_size'])} rad",
            "subgrids": int(w["nr_subgrids"]),
            "visibilities": nvis,
            "precision": args.precision,
        },
        "gridder_mvis_s": nvis / t_grid / 1e6,
        "degridder_mvis_s": nvis / t_degrid / 1e6,
        "gridder_ms": t_grid * 1e3,
        "degridder_ms": t_degrid * 1e3,
        "gflops_per_pass": flops / 1e9,
        "gsincos_per_pass": sincos / 1e9,
        "roofline": {
            "bound": "mfma",
            "kernel": dominant,
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "algorithmic_bytes": bytes_dom,
            "hbm_gbs_algorithmic": bytes_dom / t_dom / 1e9,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
