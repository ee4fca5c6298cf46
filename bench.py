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
# The CPUs this process may use, taken before any OpenMP runtime starts: with OMP_PROC_BIND set, the runtime pins
# the main thread to one of them, and a child process would inherit that single CPU.
HOST_CPUS = sorted(os.sched_getaffinity(0))
# CPU baseline threads stay on neighbouring cores.
os.environ.setdefault("OMP_PROC_BIND", "close")

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


def flop_model_selected(md, vis_offset, S):
    """flop_model with V_s = the selected visibilities of subgrid s (vis_offset[s+1] - vis_offset[s])."""
    P = S * S
    T = md["nr_timesteps"].astype(np.int64)
    V = np.diff(np.asarray(vis_offset, np.int64))
    return int((P * (5 * T + 34 * V) + 134 * P).sum()), int((P * V).sum()), int(V.sum())


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


def cpu_model():
    """Host CPU model name and the number of logical CPUs this process may use."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, len(HOST_CPUS)


def cpu_baseline(w, every, nthreads, repeats=3):
    """Oracle FP32 restatement (same loop nest / OpenMP structure as the reference CPU path) on every k-th subgrid;
    median of `repeats` runs per direction."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from idgx import observation
    from helpers import oracle_degrid, oracle_grid

    idx = np.arange(0, w["nr_subgrids"], every)
    sw = observation.subset(w, idx)
    tg, td = [], []
    for _ in range(repeats):
        t0 = time.perf_counter()
        sub = oracle_grid(sw, nthreads=nthreads)
        t1 = time.perf_counter()
        oracle_degrid(sw, sub, nthreads=nthreads)
        t2 = time.perf_counter()
        tg.append(t1 - t0)
        td.append(t2 - t1)
    nvis = sw["nr_visibilities"]
    return dict(grid_s=float(np.median(tg)), degrid_s=float(np.median(td)), nvis=nvis, nsub=idx.size)


def reference_as_shipped(nthreads):
    """The reference CPU launchers themselves, compiled from its sources as shipped (-g, i.e. -O0; oracle/_ref),
    on the reference generator's workload of the small configuration's shape (10 stations, 1 time slot, 128
    timesteps, 16 channels, S=32, G=1024: 92,160 visibilities).  None when the binary was not built."""
    import subprocess

    exe = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(exe):
        return None
    env = dict(os.environ, OMP_NUM_THREADS=str(nthreads), OMP_PROC_BIND="close")
    out = subprocess.run([exe, "-", "10", "1", "128", "16", "32", "1024"], env=env, capture_output=True, text=True,
                         timeout=300, check=True, preexec_fn=lambda: os.sched_setaffinity(0, HOST_CPUS)
                         ).stdout.strip().splitlines()[-1]
    t = json.loads(out)
    nvis = 45 * 128 * 16
    tg, td = t["gridder_us"] / 1e6, t["degridder_us"] / 1e6
    return {"value": nvis / (tg + td) / 1e6, "unit": "Mvis/s", "cores": nthreads, "kind": "reference",
            "sample": f"reference gridding/src/{{gridding,degridding}}_benchmark.cpp CPU launchers as shipped (-g), "
                      f"reference generator srand(0), 10 stations x 1 slot x 128 timesteps x 16 channels, S=32, "
                      f"G=1024 ({nvis} visibilities); gridder {nvis / tg / 1e6:.4f} Mvis/s, degridder "
                      f"{nvis / td / 1e6:.4f} Mvis/s"}


# v_sin / v_cos issue at half rate on gfx950 (8 cycles per wave64, measured): 256 CUs x 4 SIMDs x 8 lanes per cycle.
TRANS_PEAK_T_PER_S = 256 * 4 * 8 * 2.4e9 / 1e12


def trans_rate(sincos_pairs, seconds):
    """Transcendental rate of a pass: one sine and one cosine per (pixel, visibility) pair, against the half-rate
    v_sin / v_cos peak."""
    rate = 2.0 * sincos_pairs / seconds / 1e12
    return {"achieved": rate, "peak": TRANS_PEAK_T_PER_S, "unit": "T/s", "frac": rate / TRANS_PEAK_T_PER_S}


def kernel_key(precision, path, direction, shape=None, plain_shape=None):
    """Name of the kernel instantiation that runs a pass, as rocprofv3 reports it (the rows of the PMC summaries).
    plain_shape: (gridder pixels, degridder channels) per thread of the plain one-shot kernels, from the library's
    own rule (idgx_plain_kernel_shape)."""
    if precision == "fp32":
        if path == "one-shot":
            return f"{direction}_plain<{plain_shape[0] if direction == 'gridder' else plain_shape[1]}, 1>"
        if shape == "tiles per class":
            return "gridder_plain_queue_tiles<1>"
        if isinstance(shape, list):
            if len(set(shape)) != 1:
                return None
            shape = shape[0]
        return f"{direction}_plain_queue<{shape[0] if direction == 'gridder' else shape[1]}, 1>"
    dt = 0 if precision.startswith("bf16") else 1
    split = "true" if precision.endswith("x3") else "false"
    if path == "one-shot":
        return f"{direction}_mfma<{dt}, {split}, {8 if direction == 'gridder' else 4}, 256, false>"
    if shape == "tiles per class":  # one launch, every item with its class's tiles per wave
        return f"gridder_mfma_queue_tiles<{dt}, {split}>"
    if isinstance(shape, list):  # one shape per class: one kernel only when they agree
        if len(set(shape)) != 1:
            return None
        shape = shape[0]
    tiles, batch, pm = shape
    return f"{direction}_mfma_queue<{dt}, {split}, {tiles}, {batch}, {'true' if pm else 'false'}>"


def pipe_profile(config, key):
    """VALU and matrix-core utilisation of kernel `key` from the committed rocprofv3 --pmc summary for this workload
    (profiles/pipe_utilisation_<config>.json), or None when it was not profiled."""
    path = os.path.join(ROOT, "profiles", f"pipe_utilisation_{config}.json")
    if key is None or not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f).get(key)
    return None if k is None else dict(k, source=os.path.relpath(path, ROOT))


MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 / BF16 matrix peak, published


def roofline_record(config, precision, key, dominant, flops, sincos, t_dom):
    """The roofline of the dominant kernel.  The gridder and degridder are bound by VALU issue (phases, sin/cos and,
    for the mixed forms, the hi/lo split), far from HBM and from the matrix-core peak, so `frac` is the
    FP32-equivalent algorithmic rate against the FP32 peak, with the committed counters beside it: VALU busy and
    VALU instructions per (pixel, visibility) pair, the matrix cores' utilisation (counter and algorithmic MFMA
    rate against the FP16/BF16 peak) and the sin/cos rate against the half-rate v_sin / v_cos peak."""
    achieved = flops / t_dom / 1e12
    traffic = measured_traffic(config, key)
    pipes = pipe_profile(config, key)
    split = 1.0 if precision.endswith("x3") else 0.5
    mfma_flops = 0 if precision == "fp32" else 128 * sincos * split
    valu = None
    if pipes is not None:
        valu = {"busy_pct": pipes.get("VALUBusy"), "insts_per_pair": pipes["SQ_INSTS_VALU"] * 64 / sincos,
                "mfma_util_pct": pipes.get("MfmaUtil"), "source": pipes["source"],
                "note": "counters of the committed profile of this kernel on this workload (mean per launch); "
                        "insts_per_pair = SQ_INSTS_VALU x 64 lanes / (pixel, visibility) pairs of the pass"}
    return {
        "bound": "valu",
        "kernel": key,
        "pass": dominant,
        "achieved": achieved,
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP32_PEAK_TFLOPS,
        "traffic": None if traffic is None else traffic["bytes"],
        "traffic_detail": traffic,
        "flop_model": "P*(5*T + 34*V) + 134*P per subgrid (FP32-equivalent algorithmic flops), sincos separate",
        "valu": valu,
        "mfma": None if not mfma_flops else {
            "achieved_tflops": mfma_flops / t_dom / 1e12, "peak_tflops": MFMA_PEAK_TFLOPS,
            "frac": mfma_flops / t_dom / 1e12 / MFMA_PEAK_TFLOPS,
            "note": "algorithmic MFMA flops: 128 per pair for the 3-product split (2 instructions), 64 single"},
        "transcendental": trans_rate(sincos, t_dom),
    }


def form_rates(precision, tg, td, flops, sincos, bytes_by_direction=None):
    """Per-direction rates of one form: FP32-equivalent rate against the FP32 peak, algorithmic MFMA rate against
    the FP16/BF16 matrix peak (128 flops per pair with the 3-product split, 64 single) and, when the algorithmic
    bytes are known, HBM GB/s at those bytes."""
    split = 1.0 if precision.endswith("x3") else 0.5
    mfma_flops = 0 if precision == "fp32" else 128 * sincos * split
    out = {}
    for d, t in (("gridder", tg), ("degridder", td)):
        r = {"fp32_frac": flops / t / 1e12 / FP32_PEAK_TFLOPS,
             "mfma_frac": mfma_flops / t / 1e12 / MFMA_PEAK_TFLOPS}
        if bytes_by_direction is not None:
            r["hbm_gbs_algorithmic"] = bytes_by_direction[d] / t / 1e9
        out[d] = r
    return out


def measured_traffic(config, key):
    """HBM bytes per launch of kernel `key` from the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration) for this workload, or None when it was not profiled."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{config}.json")
    if key is None or not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f).get("kernels", {}).get(key)
    return None if k is None else {"bytes": k["traffic_bytes"], "read_bytes": k["read_bytes"],
                                   "write_bytes": k["write_bytes"], "kernel": key,
                                   "source": os.path.relpath(path, ROOT)}


class OneShot:
    """One launch per direction over the whole plan (idgx_grid / idgx_degrid)."""

    def __init__(self, dw, precision):
        self.dw, self.precision, self.path = dw, precision, "one-shot"
        self.plan = dw.plan(precision=precision)
        self.shapes = {"gridder": None, "degridder": None}

    def grid(self, stream):
        self.dw.grid(self.plan, stream=stream)

    def degrid(self, stream):
        self.dw.degrid(self.plan, stream=stream)

    def buffers(self):
        return self.dw.d_subgrids, self.dw.d_vis_out


class Batched:
    """Persistent work-queue kernels over the whole plan (idgx_batch_*), shapes chosen by the autotuner.  A plan of
    several subgrid classes runs in the batch's default mode: one persistent launch per direction over the merged
    queue of every class (gridder items with their class's shape).  `db`: an existing DeviceBatch whose HBM buffers
    the batch reuses (its previous batch is released first)."""

    def __init__(self, classes, d_vis, precision, stream, db=None):
        from idgx.device import DeviceBatch

        self.precision, self.path = precision, "batched"
        if db is None:
            self.db = DeviceBatch(classes, precision=precision, visibilities_on_device=d_vis)
        else:
            db.rebatch(precision)
            self.db = db
        t0 = time.perf_counter()
        print(f"bench: autotuning the {precision} batch", file=sys.stderr, flush=True)
        tg, td = self.db.batch.autotune(self.db.data, repeats=2, stream=stream)
        print(f"bench: {precision} batch autotuned in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        self.describe(tg, td)

    def describe(self, tg, td):
        """Shapes, launch mode and queue sizes of the batch as it now runs."""
        classes = [c["w"] for c in self.db.classes]
        table = self.db.batch.shape_table()
        plain = self.precision == "fp32"
        per_class = [self.db.batch.class_shape(k) for k in range(len(classes))]
        single = self.db.batch.single_launch()
        self.shapes = {}
        for i, direction in enumerate(("gridder", "degridder")):
            if single[i] == -1:
                self.shapes[direction] = [table[pc[i]] for pc in per_class]
            elif single[i] < 0:  # one launch, each class with its own shape
                self.shapes[direction] = "tiles per class"
            else:
                self.shapes[direction] = table[single[i]]

        def shape_fields(g, d):
            if plain:
                return {"gridder": {"pixels_per_thread": table[g][0]},
                        "degridder": {"visibilities_per_thread": table[d][1]}}
            return {"gridder": {"tiles_per_wave": table[g][0], "lds_visibilities": table[g][1],
                                "phase_mfma": table[g][2]},
                    "degridder": {"tiles_per_wave": table[d][0], "lds_pixels": table[d][1],
                                  "phase_mfma": table[d][2]}}

        items = self.db.batch.items()
        self.info = {"launches": {d: ("one per class" if single[i] == -1 else "one over all classes")
                                  for i, d in enumerate(("gridder", "degridder"))},
                     "single_launch_shapes": {d: (self.shapes[d] if single[i] != -1 else None)
                                              for i, d in enumerate(("gridder", "degridder"))},
                     "shapes_per_class": [dict({"subgrid_size": int(c["subgrid_size"])}, **shape_fields(g, d))
                                          for c, (g, d) in zip(classes, per_class)],
                     "queue_items": {"gridder": items[0], "degridder": items[1]},
                     "persistent_workgroups": {"gridder": items[2], "degridder": items[3]},
                     "autotune_ms": {"gridder": tg, "degridder": td}}

    def grid(self, stream):
        self.db.grid(stream)

    def degrid(self, stream):
        self.db.degrid(stream)

    def buffers(self):
        return self.db.classes[0]["d_subgrids"], self.db.d_vis_out


def time_form(runner, steps, warmup, stream):
    """Average gridder and degridder kernel time (s) over `steps` round trips, HIP events on `stream`."""
    import torch

    for _ in range(warmup):
        runner.grid(stream)
        runner.degrid(stream)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        runner.grid(stream)
        ev[k][1].record(stream)
        runner.degrid(stream)
        ev[k][2].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_grid = sum(e[0].elapsed_time(e[1]) for e in ev) / steps / 1e3
    t_degrid = sum(e[1].elapsed_time(e[2]) for e in ev) / steps / 1e3
    return t_grid, t_degrid, wall / steps


def time_cycle(runner, dw, steps, warmup, stream):
    """Average time (ms) of each stage of an imaging cycle on resident buffers: clear the master grid, grid, subgrid
    FFT, add to the grid, master-grid FFT to the image and back (deconvolution would sit between the two), split
    from the grid, inverse subgrid FFT, degrid (HIP events on `stream`).  A plan with w-layers adds every subgrid to
    the grid of its layer, transforms every layer and combines the layers into one image with the w-screens (and
    back), all on resident [layers][4][G][G] grids, in one call each way (idgx_grid_to_image, idgx_image_to_grid:
    the screens and the sum over layers run inside the row pass).  The master-grid transforms are restricted to the
    subgrids' bounding box (the uv grids are zero outside it on the way to the image, and only its rows are needed
    on the way back), and the clear zeroes only what the transform to the image reads (idgx_grid_clear_box).  Timed
    beside the cycle: the full transforms and, with w-layers, the separate calls (idgx_grid_fft_box and
    idgx_wstack_to_image, and back)."""
    import torch

    from idgx import FFT_BACKWARD, FFT_FORWARD, grid_box, grid_clear_box, grid_fft_box, wstack_from_image, \
        wstack_to_image

    plan = dw.plan()
    box = grid_box(plan, dw.w["metadata"])
    d_sub, _ = runner.buffers()
    layered = float(dw.w.get("w_step") or 0.0) > 0.0
    z_min, nl = dw.layers() if layered else (0, 1)
    d_grid = dw.layer_buffer(nl) if layered else dw.grid_buffer()
    d_image = dw.grid_buffer() if layered else None
    stages = [("grid_clear", lambda: grid_clear_box(plan, d_grid, nl, box, stream)),
              ("gridder", lambda: runner.grid(stream)),
              ("subgrid_fft", lambda: dw.fft(plan, FFT_FORWARD, d_sub, stream)),
              ("adder", (lambda: dw.add_layered(plan, d_grid, d_sub, stream)) if layered else
               (lambda: dw.add(plan, d_grid, d_sub, stream)))]
    if layered:
        stages += [("grid_to_image", lambda: dw.grid_to_image(plan, d_grid, d_image, box, stream)),
                   ("image_to_grid", lambda: dw.image_to_grid(plan, d_image, d_grid, box, stream))]
    else:
        stages += [("grid_fft_to_image", lambda: grid_fft_box(plan, d_grid, nl, FFT_BACKWARD, box, stream)),
                   ("grid_fft_to_uv", lambda: grid_fft_box(plan, d_grid, nl, FFT_FORWARD, box, stream))]
    stages += [("splitter", (lambda: dw.split_layered(plan, d_grid, d_sub, stream)) if layered else
                (lambda: dw.split(plan, d_grid, d_sub, stream))),
               ("subgrid_ifft", lambda: dw.fft(plan, FFT_BACKWARD, d_sub, stream)),
               ("degridder", lambda: runner.degrid(stream))]
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            for _, f in stages:
                f()
        torch.cuda.synchronize()
        acc = [0.0] * len(stages)
        for _ in range(steps):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(stages) + 1)]
            ev[0].record(stream)
            for i, (_, f) in enumerate(stages):
                f()
                ev[i + 1].record(stream)
            torch.cuda.synchronize()
            for i in range(len(stages)):
                acc[i] += ev[i].elapsed_time(ev[i + 1])
        # The full transforms of the same grids, for comparison (not part of the cycle).
        full = {"to_image": lambda: dw.grid_fft(plan, d_grid, nl, FFT_BACKWARD, stream),
                "to_uv": lambda: dw.grid_fft(plan, d_grid, nl, FFT_FORWARD, stream)}
        if layered:
            def separate_to_image():
                grid_fft_box(plan, d_grid, nl, FFT_BACKWARD, box, stream)
                wstack_to_image(plan, d_grid, z_min, nl, d_image, stream)

            def separate_to_uv():
                wstack_from_image(plan, d_image, z_min, nl, d_grid, stream)
                grid_fft_box(plan, d_grid, nl, FFT_FORWARD, box, stream)

            full.update(separate_to_image=separate_to_image, separate_to_uv=separate_to_uv)
        full_ms = {}
        for name, f in full.items():
            f()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(steps):
                f()
            e1.record(stream)
            torch.cuda.synchronize()
            full_ms[name] = e0.elapsed_time(e1) / steps
    del d_grid, d_image
    return {name: acc[i] / steps for i, (name, _) in enumerate(stages)}, box, full_ms, nl


def time_pcie(nvis_bytes, sub_bytes, repeats=3):
    """Host<->device copy times (s) from pinned host memory for one gridder pass (visibilities in, subgrids out)
    and one degridder pass (subgrids in, visibilities out), as a caller handing over host buffers would pay."""
    import torch

    out = {}
    for name, (nin, nout) in {"gridder": (nvis_bytes, sub_bytes), "degridder": (sub_bytes, nvis_bytes)}.items():
        h_in = torch.empty(nin, dtype=torch.uint8, pin_memory=True)
        h_out = torch.empty(nout, dtype=torch.uint8, pin_memory=True)
        d_in = torch.empty(nin, dtype=torch.uint8, device="cuda")
        d_out = torch.empty(nout, dtype=torch.uint8, device="cuda")
        best_in = best_out = float("inf")
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            d_in.copy_(h_in, non_blocking=True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            h_out.copy_(d_out, non_blocking=True)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            best_in, best_out = min(best_in, t1 - t0), min(best_out, t2 - t1)
        out[name] = {"in_s": best_in, "out_s": best_out, "in_gbs": nin / best_in / 1e9, "out_gbs": nout / best_out / 1e9}
        del h_in, h_out, d_in, d_out
    return out


def accuracy(runner, w, host_vis, idx):
    """Relative L2 error of the gridder output and of the degridder output (degridding the same subgrids) on the
    sample of subgrids idx, against the FP64 oracle and against the oracle's FP32 restatement of the reference CPU
    path (the FP32 path) on the same inputs; (whole sample, worst subgrid) per direction and reference."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from idgx import observation
    from helpers import oracle_degrid, oracle_grid, per_subgrid_rel_l2, rel_l2

    import torch

    sw = observation.subset(dict(w, visibilities=host_vis), idx)
    S, C = int(w["subgrid_size"]), int(w["wavenumbers"].shape[0])
    d_sub, d_vis_out = runner.buffers()
    dev = d_sub.device
    subs = d_sub.view(-1, 4 * S * S * 8)[torch.from_numpy(idx).to(dev)].cpu().numpy()
    subs = subs.view(np.complex64).reshape(idx.size, 4, S, S)
    rows = torch.from_numpy(sw["source_rows"]).to(dev)
    vis = d_vis_out.view(-1, C * 4 * 8)[rows].cpu().numpy().view(np.complex64).reshape(-1, C, 4)
    md, ch = sw["metadata"], sw["channels"]
    windows = []
    mask = np.zeros(vis.shape[:2], bool)
    for s in range(md.size):
        r0, nt = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
        c0, nc = int(ch["channel_begin"][s]), int(ch["nr_channels"][s])
        windows.append((slice(r0, r0 + nt), slice(c0, c0 + nc)))
        mask[r0:r0 + nt, c0:c0 + nc] = True
    out = {}
    for name, fp64 in (("vs_fp64", True), ("vs_fp32_path", False)):
        g = oracle_grid(sw, fp64=fp64)
        v = oracle_degrid(sw, subs, fp64=fp64)
        out[name] = {
            "gridder": rel_l2(subs, g), "gridder_worst_subgrid": float(per_subgrid_rel_l2(subs, g).max()),
            "degridder": rel_l2(vis[mask], v[mask]),
            "degridder_worst_subgrid": max(rel_l2(vis[r, c], v[r, c]) for r, c in windows),
        }
    return out


def class_sample(bo, k, idx, d_vis, d_flags):
    """Host copy of subgrids idx of class k of a batched workload: the subset workload (rows gathered), its
    visibilities with the flagged samples zeroed, and the selection mask of its windows."""
    import torch

    from idgx import observation

    c = bo["classes"][k]
    C = int(bo["wavenumbers"].shape[0])
    sw = observation.subset(dict(c, visibilities=None), idx)
    rows = torch.from_numpy(sw["source_rows"]).to(d_vis.device)
    vis = d_vis.view(-1, C * 8)[rows].cpu().numpy().view(np.complex64).reshape(-1, C, 4)
    fl = d_flags[rows].cpu().numpy()
    sel = np.zeros(fl.shape, bool)
    md, ch = sw["metadata"], sw["channels"]
    for s in range(md.size):
        r0, nt = int(md["time_offset"][s]), int(md["nr_timesteps"][s])
        c0, nc = int(ch["channel_begin"][s]), int(ch["nr_channels"][s])
        sel[r0:r0 + nt, c0:c0 + nc] = True
    sel &= ~fl
    sw["visibilities"] = np.where(fl[..., None], 0, vis).astype(np.complex64)
    return sw, sel, rows


def run_batched_config(args):
    """Batched uneven workload: the whole plan as one persistent launch per direction over the merged queue of
    every subgrid class (gridder items with their class's autotuned tiles per wave), then, beside it, one launch
    per class and direction; flags as selection lists, all built in HBM; throughput counts the selected
    (unflagged) visibilities."""
    import torch

    from idgx import observation
    from idgx.device import device_flags, device_selection_lists

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from helpers import oracle_degrid, oracle_grid, per_subgrid_rel_l2, rel_l2

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    t_setup = time.perf_counter()
    bo = observation.make_batched_observation(args.config, visibilities=False, flags=False)
    cfg = bo["config"]
    C, n_rows = int(bo["wavenumbers"].shape[0]), int(bo["nr_rows"])
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(cfg["seed"]))
    d_vis = torch.randn(n_rows * C * 8, generator=gen, device=dev, dtype=torch.float32)
    n_bl = cfg["n_stations"] * (cfg["n_stations"] - 1) // 2
    d_flags = device_flags(n_bl, cfg["n_times"], C, cfg["flag_max"], cfg["seed"], device=dev)
    flops = sincos = nvis = 0
    for c in bo["classes"]:
        d_vi, d_off, h_off = device_selection_lists(c, d_flags)
        c.update(d_vis_index=d_vi, d_vis_offset=d_off, vis_offset=h_off, nr_visibilities=int(h_off[-1]))
        f, sc, v = flop_model_selected(c["metadata"], h_off, int(c["subgrid_size"]))
        flops, sincos, nvis = flops + f, sincos + sc, nvis + v
    torch.cuda.synchronize()
    # The plan again on the device (idgx_plan_segments): bit-identical to the host planner, and its time.
    from idgx import plan_segments_device
    planner = {"host_s": 0.0, "device_s": 0.0, "bit_identical": True}
    for c in bo["classes"]:
        t0 = time.perf_counter()
        md_d, ch_d = plan_segments_device(bo["uvw"], c["frequencies64"], *c["stations"], *c["segments"],
                                          subgrid_size=int(c["subgrid_size"]), grid_size=cfg["grid_size"],
                                          image_size=cfg["image_size"], nr_times=cfg["n_times"])
        planner["device_s"] += time.perf_counter() - t0
        planner["host_s"] += c["host_plan_s"]
        planner["bit_identical"] &= md_d.tobytes() == c["metadata"].tobytes() and \
            ch_d.tobytes() == c["channels"].tobytes()
    n_sub = sum(int(c["nr_subgrids"]) for c in bo["classes"])
    window = n_bl * cfg["n_times"] * C
    print(f"bench: {args.config}: {n_sub} subgrids in {len(bo['classes'])} classes "
          f"(S = {', '.join(str(c['subgrid_size']) for c in bo['classes'])}), {nvis} selected visibilities of "
          f"{window} ({1 - nvis / window:.1%} flagged or off-grid), setup {time.perf_counter() - t_setup:.1f} s",
          file=sys.stderr, flush=True)

    def fp64_errors(db):
        """Relative L2 error against the FP64 oracle on about 60 subgrids per class (selected samples only)."""
        acc = {}
        for k, c in enumerate(bo["classes"]):
            every = args.check_every or max(1, int(c["nr_subgrids"]) // 60)
            idx = np.arange(every // 2, int(c["nr_subgrids"]), every)
            sw, sel, rows = class_sample(bo, k, idx, d_vis, d_flags)
            S = int(c["subgrid_size"])
            d_sub = db.classes[k]["d_subgrids"]
            subs = d_sub.view(-1, 4 * S * S * 8)[torch.from_numpy(idx).to(dev)].cpu().numpy()
            subs = subs.view(np.complex64).reshape(idx.size, 4, S, S)
            vis_out = db.d_vis_out.view(-1, C * 32)[rows].cpu().numpy().view(np.complex64)
            vis_out = vis_out.reshape(-1, C, 4)
            g64 = oracle_grid(sw, fp64=True)
            v64 = oracle_degrid(sw, subs, fp64=True)
            acc[f"S{S}"] = {"subgrids": int(idx.size), "gridder": rel_l2(subs, g64),
                            "gridder_worst_subgrid": float(per_subgrid_rel_l2(subs, g64).max()),
                            "degridder": rel_l2(vis_out[sel], v64[sel])}
        return acc

    def cpu_errors(db):
        """Relative L2 error of the plain form against the CPU restatement of the reference path (FP32, flagged
        samples zeroed) on the same sample as fp64_errors."""
        acc = {}
        for k, c in enumerate(bo["classes"]):
            every = args.check_every or max(1, int(c["nr_subgrids"]) // 60)
            idx = np.arange(every // 2, int(c["nr_subgrids"]), every)
            sw, sel, rows = class_sample(bo, k, idx, d_vis, d_flags)
            S = int(c["subgrid_size"])
            subs = db.classes[k]["d_subgrids"].view(-1, 4 * S * S * 8)[torch.from_numpy(idx).to(dev)].cpu().numpy()
            subs = subs.view(np.complex64).reshape(idx.size, 4, S, S)
            vis_out = db.d_vis_out.view(-1, C * 32)[rows].cpu().numpy().view(np.complex64).reshape(-1, C, 4)
            g32 = oracle_grid(sw)
            v32 = oracle_degrid(sw, subs)
            acc[f"S{S}"] = {"subgrids": int(idx.size), "gridder": rel_l2(subs, g32),
                            "gridder_worst_subgrid": float(per_subgrid_rel_l2(subs, g32).max()),
                            "degridder": rel_l2(vis_out[sel], v32[sel])}
        return acc

    # The mixed form first, then plain FP32 on the same HBM buffers (the mixed batch released first: its prepared
    # pixels are as large as the subgrids).  Both run the persistent kernels over the same selection lists, so both
    # skip the flagged samples and the ratio compares equal work.
    precisions = ([] if args.precision == "fp32" else [args.precision]) + ([] if args.no_plain else ["fp32"])
    forms = []
    runners = []
    db = None
    torch.cuda.empty_cache()  # selection-list temporaries: give their cached memory back before the batch allocates
    for prec in precisions:
        runner = Batched(bo["classes"], d_vis, prec, stream, db=db)
        db = runner.db
        free, total = torch.cuda.mem_get_info()
        print(f"bench: {prec} batch resident, HBM in use {(total - free) / 1e9:.1f} of {total / 1e9:.1f} GB",
              file=sys.stderr, flush=True)
        tg, td, step = time_form(runner, args.steps, args.warmup, stream)
        rec = {"precision": prec, "path": runner.path, "gridder_ms": tg * 1e3, "degridder_ms": td * 1e3,
               "step_ms": step * 1e3, "round_trip_mvis_s": nvis / (tg + td) / 1e6,
               "gridder_mvis_s": nvis / tg / 1e6, "degridder_mvis_s": nvis / td / 1e6,
               "gridder_tflops": flops / tg / 1e12, "degridder_tflops": flops / td / 1e12,
               "rates": form_rates(prec, tg, td, flops, sincos)}
        rec.update(runner.info)
        launches = runner.db.batch.last_launches()
        rec["persistent_launches_per_pass"] = {"gridder": launches[0], "degridder": launches[1]}
        print(f"bench: {prec} batched (gridder {runner.info['launches']['gridder']}, {runner.shapes['gridder']}; "
              f"degridder {runner.info['launches']['degridder']}, {runner.shapes['degridder']}): gridder "
              f"{tg * 1e3:.2f} ms, degridder {td * 1e3:.2f} ms, {rec['round_trip_mvis_s']:.1f} Mvis/s",
              file=sys.stderr, flush=True)
        if not args.no_check:
            rec["rel_l2_vs_fp64"] = fp64_errors(runner.db)
            if prec == "fp32":
                rec["rel_l2_vs_cpu_restatement"] = cpu_errors(runner.db)
        if len(bo["classes"]) > 1:
            # Beside the single launch: one persistent launch per class and direction, with the per-class shapes
            # the same autotuner run chose.
            single = runner.db.batch.single_launch()
            runner.db.batch.set_single_launch(-1, -1)
            p_tg, p_td, _ = time_form(runner, args.steps, args.warmup, stream)
            rec["per_class_launches"] = {"gridder_ms": p_tg * 1e3, "degridder_ms": p_td * 1e3,
                                         "round_trip_mvis_s": nvis / (p_tg + p_td) / 1e6}
            print(f"bench: {prec} batched, one launch per class: gridder {p_tg * 1e3:.2f} ms, "
                  f"degridder {p_td * 1e3:.2f} ms, {rec['per_class_launches']['round_trip_mvis_s']:.1f} Mvis/s",
                  file=sys.stderr, flush=True)
            runner.db.batch.set_single_launch(*single)
        forms.append(rec)
        runners.append(runner)
    head, runner = forms[0], runners[0]
    plain = forms[-1] if len(forms) > 1 and forms[-1]["precision"] == "fp32" else None
    tg, td = head["gridder_ms"] / 1e3, head["degridder_ms"] / 1e3
    dominant, t_dom = ("degridder", td) if td > tg else ("gridder", tg)
    dom_key = kernel_key(head["precision"], "batched", dominant, runner.shapes[dominant])
    roofline = roofline_record(args.config, head["precision"], dom_key, dominant, flops, sincos, t_dom)
    if dom_key is None:  # one launch per class with different shapes
        roofline["kernel"] = " + ".join(kernel_key(head["precision"], "batched", dominant, sh)
                                        for sh in runner.shapes[dominant])
    roofline["flop_model"] = ("P*(5*T + 34*V) + 134*P per subgrid, V = selected visibilities (FP32-equivalent "
                              "algorithmic flops), sincos separate")

    cpu = None
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t_cpu = v_cpu = 0.0
        n_cpu = 0
        for k, c in enumerate(bo["classes"]):
            every = args.cpu_every or max(1, int(c["nr_subgrids"]) // 400)
            idx = np.arange(0, int(c["nr_subgrids"]), every)
            sw, sel, _ = class_sample(bo, k, idx, d_vis, d_flags)
            t0 = time.perf_counter()
            sub = oracle_grid(sw, nthreads=nthreads)
            oracle_degrid(sw, sub, nthreads=nthreads)
            t_cpu += time.perf_counter() - t0
            v_cpu += int(sel.sum())
            n_cpu += idx.size
        cpu = {"value": v_cpu / t_cpu / 1e6, "unit": "Mvis/s", "cores": nthreads, "kind": "port",
               "sample": f"about 400 subgrids per class ({n_cpu} subgrids, {int(v_cpu)} selected visibilities): "
                         f"oracle FP32 restatement of the reference OpenMP CPU path, gridder + degridder, flagged "
                         f"samples zeroed (the CPU path has no selection lists, so it processes every sample of "
                         f"each window); value counts the selected visibilities"}

    out = {
        "metric": "Mvis/s gridder+degridder, 1 MI355X: FP32 vs MFMA mixed at rel-err<=5e-5",
        "value": head["round_trip_mvis_s"],
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head["precision"] == "fp32" else
                 {"bf16x3": "bf16x3-mfma+f32acc", "fp16x3": "fp16x3-mfma+f32acc"}.get(head["precision"],
                                                                                 head["precision"] + "-mfma"),
        "data": "synthetic (seeded layout/uvw/plan; N(0,1) complex64 visibilities and per-baseline flags generated "
                "in HBM)",
        "config": {
            "workload": f"{args.config}: {cfg['n_stations']} stations, {cfg['n_times']} x 1 s, {C} channels "
                        f"{cfg['f_lo'] / 1e6:.0f}-{cfg['f_hi'] / 1e6:.0f} MHz, S = "
                        f"{'/'.join(str(c['subgrid_size']) for c in bo['classes'])} by baseline-length tercile, "
                        f"time chunks log-uniform in {list(cfg['t_range'])}, per-baseline flag fraction "
                        f"U(0, {cfg['flag_max']}), G={cfg['grid_size']}, image {cfg['image_size']} rad",
            "subgrids": n_sub,
            "visibilities": nvis,
            "samples_in_windows": window,
            "precision": head["precision"],
            "path": "batched",
            "classes": [{"subgrid_size": int(c["subgrid_size"]), "subgrids": int(c["nr_subgrids"]),
                         "visibilities": int(c["nr_visibilities"])} for c in bo["classes"]],
        },
        "speedup_vs_plain_fp32": None if plain is None else head["round_trip_mvis_s"] / plain["round_trip_mvis_s"],
        "speedup_note": "both forms batched over the same selection lists (flagged samples skipped by both): equal "
                        "work" if plain is not None else "plain form not timed (--no-plain)",
        "forms": forms,
        "planner": dict(planner, note="device time includes uploading the segments and downloading the plan"),
        "gflops_per_pass": flops / 1e9,
        "gsincos_per_pass": sincos / 1e9,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="ska_low",
                    help="ska_low (default), small, full_cycle, or a batched configuration: batched_uneven, "
                         "batched_uneven_small")
    ap.add_argument("--precision", default="fp16x3", help="headline mixed form (the plain fp32 form is timed beside it)")
    ap.add_argument("--no-batch", action="store_true", help="headline on the one-shot kernels instead of the batch")
    ap.add_argument("--cpu-every", type=int, default=0, help="CPU baseline on every k-th subgrid (0 = auto)")
    ap.add_argument("--check-every", type=int, default=0, help="FP64 accuracy sample: every k-th subgrid (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--cycle", action="store_true", help="also time the imaging-cycle stages (default for full_cycle)")
    ap.add_argument("--pcie", action="store_true", help="also time the host transfers a host-buffer caller pays")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched configurations: time only the mixed form (counter runs of its kernels)")
    ap.add_argument("--all-forms", action="store_true",
                    help="also time the other mixed forms (bf16, fp16, bf16x3 one-shot and bf16x3 batched) before "
                         "the headline form")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the hot path is per-GPU; see the sharding design); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    from idgx import observation

    if args.config in observation.BATCHED_CONFIGS:
        return run_batched_config(args)

    import torch

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
    stream = torch.cuda.current_stream()
    md, ch, S = w["metadata"], w["channels"], int(w["subgrid_size"])
    flops, sincos, nvis, _ = flop_model(md, ch, S, C)
    print(f"bench: {args.config}: {w['nr_subgrids']} subgrids, {nvis} visibilities, "
          f"setup {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    need_host = not (args.no_cpu and args.no_check)
    host_vis = d_vis.view(torch.uint8).cpu().numpy().view(np.complex64).reshape(n_rows, C, 4) if need_host else None
    check_every = args.check_every or max(1, w["nr_subgrids"] // 160)
    check_idx = np.arange(check_every // 2, w["nr_subgrids"], check_every)

    runners = [OneShot(dw, "fp32")]
    if not args.no_batch:
        runners.append(Batched([w], d_vis, "fp32", stream))
    if args.all_forms:
        others = [p for p in ("bf16", "fp16", "bf16x3", "fp16x3") if p != args.precision]
        runners += [OneShot(dw, p) for p in others]
        if not args.no_batch:
            runners += [Batched([w], d_vis, p, stream) for p in ("bf16x3",) if p != args.precision]
    if args.precision != "fp32":
        runners.append(OneShot(dw, args.precision))
        if not args.no_batch:
            runners.append(Batched([w], d_vis, args.precision, stream))
    forms = []
    bytes_by_direction = {d: algorithmic_bytes(md, ch, S, C, int(w["nr_aterm_slots"]), int(w["nr_stations"]), d)
                          for d in ("gridder", "degridder")}
    for runner in runners:
        prec = runner.precision
        tg, td, step = time_form(runner, args.steps, args.warmup, stream)
        rec = {"precision": prec, "path": runner.path, "gridder_ms": tg * 1e3, "degridder_ms": td * 1e3,
               "step_ms": step * 1e3,
               "round_trip_mvis_s": nvis / (tg + td) / 1e6, "gridder_mvis_s": nvis / tg / 1e6,
               "degridder_mvis_s": nvis / td / 1e6,
               "gridder_tflops": flops / tg / 1e12, "degridder_tflops": flops / td / 1e12}
        rec["rates"] = form_rates(prec, tg, td, flops, sincos, bytes_by_direction)
        if hasattr(runner, "info"):
            rec.update(runner.info)
        if not args.no_check:
            acc = accuracy(runner, w, host_vis, check_idx)
            rec["rel_l2_vs_fp64"], rec["rel_l2_vs_fp32_path"] = acc["vs_fp64"], acc["vs_fp32_path"]
        print(f"bench: {prec} {runner.path}: gridder {tg * 1e3:.2f} ms, degridder {td * 1e3:.2f} ms, "
              f"{rec['round_trip_mvis_s']:.1f} Mvis/s", file=sys.stderr, flush=True)
        forms.append(rec)
    head = forms[-1]
    plain_forms = [f for f in forms if f["precision"] == "fp32"]
    plain = max(plain_forms, key=lambda f: f["round_trip_mvis_s"])  # the faster plain FP32 form
    cycle = None
    if args.cycle or args.config == "full_cycle":
        cycle, box, full_ms, nl = time_cycle(runners[-1], dw, args.steps, args.warmup, stream)
        total = sum(cycle.values())
        cycle = {"stages_ms": cycle, "total_ms": total, "cycle_mvis_s": nvis / (total / 1e3) / 1e6,
                 "precision": head["precision"], "path": head["path"], "grid_fft_box": list(box),
                 "w_layers": nl, "beside_the_cycle_ms": full_ms}
        print(f"bench: imaging cycle ({head['precision']} {head['path']}): " +
              ", ".join(f"{k} {v:.2f} ms" for k, v in cycle["stages_ms"].items()) +
              f"; total {total:.2f} ms; master-grid transforms on the box {list(box)} (beside the cycle: " +
              ", ".join(f"{k} {v:.2f} ms" for k, v in full_ms.items()) + ")", file=sys.stderr, flush=True)
    tg, td = head["gridder_ms"] / 1e3, head["degridder_ms"] / 1e3
    dominant, t_dom = ("degridder", td) if td > tg else ("gridder", tg)
    bytes_dom = bytes_by_direction[dominant]
    head_runner = runners[-1]
    from idgx import plain_kernel_shape

    dom_key = kernel_key(head["precision"], head["path"], dominant, head_runner.shapes[dominant],
                         plain_shape=plain_kernel_shape(dw.plan(), ch is not None))
    roofline = roofline_record(args.config, head["precision"], dom_key, dominant, flops, sincos, t_dom)
    roofline.update(algorithmic_bytes=bytes_dom, hbm_gbs_algorithmic=bytes_dom / t_dom / 1e9)

    cpu = None
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        every = args.cpu_every or max(1, w["nr_subgrids"] // 850)
        r = cpu_baseline(dict(w, visibilities=host_vis), every, nthreads)
        model, ncpu = cpu_model()
        cpu = {
            "value": r["nvis"] / (r["grid_s"] + r["degrid_s"]) / 1e6,
            "unit": "Mvis/s",
            "cores": nthreads,
            "kind": "port",
            "sample": f"every {every}th subgrid of the plan ({r['nsub']} subgrids, {r['nvis']} visibilities), median "
                      f"of 3 runs: gridder {r['nvis'] / r['grid_s'] / 1e6:.3f} Mvis/s, degridder "
                      f"{r['nvis'] / r['degrid_s'] / 1e6:.3f} Mvis/s; oracle FP32 restatement of the reference "
                      f"OpenMP CPU path (same loop nest and OpenMP structure), -O3 -ffp-contract=off, "
                      f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}",
            "host_cpu": f"{model}, {ncpu} logical CPUs available",
        }
        ref = reference_as_shipped(nthreads)
        if ref is not None:
            cpu["reference_as_shipped"] = ref

    out = {
        "metric": "Mvis/s gridder+degridder, 1 MI355X: FP32 vs MFMA mixed at rel-err<=5e-5",
        "value": head["round_trip_mvis_s"],
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head["precision"] == "fp32" else
                 {"bf16x3": "bf16x3-mfma+f32acc", "fp16x3": "fp16x3-mfma+f32acc"}.get(head["precision"],
                                                                                 head["precision"] + "-mfma"),
        "data": "synthetic (seeded layout/uvw/plan; N(0,1) complex64 visibilities generated in HBM)",
        "config": {
            "workload": f"{args.config}: {w['config']['n_stations']} stations, {w['config']['n_times']} x 1 s, "
                        f"{C} channels {w['config']['f_lo'] / 1e6:.0f}-{w['config']['f_hi'] / 1e6:.0f} MHz, "
                        f"S={S}, {w['config']['timesteps_per_subgrid']} timesteps/subgrid, G={w['grid_size']}, "
                        f"image {float(w['image_size']):.3g} rad",
            "subgrids": int(w["nr_subgrids"]),
            "visibilities": nvis,
            "precision": head["precision"],
            "path": head["path"],
        },
        "speedup_vs_plain_fp32": head["round_trip_mvis_s"] / plain["round_trip_mvis_s"],
        "speedup_note": f"against the faster plain FP32 form of this run ({plain['path']}); every form processes "
                        f"the same visibilities",
        "forms": forms,
        "gflops_per_pass": flops / 1e9,
        "gsincos_per_pass": sincos / 1e9,
        "accuracy_sample": f"every {check_every}th subgrid ({check_idx.size} subgrids) vs the FP64 oracle and vs "
                           f"the FP32 path (the oracle's FP32 restatement of the reference CPU gridder and degridder)",
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cycle is not None:
        out["imaging_cycle"] = cycle
    if args.pcie:
        pc = time_pcie(nvis * 32, int(w["nr_subgrids"]) * 4 * S * S * 8)
        t_pc = tg + td + sum(v["in_s"] + v["out_s"] for v in pc.values())
        out["pcie_inclusive"] = {"copies": pc, "round_trip_mvis_s": nvis / t_pc / 1e6,
                                 "note": "host buffers handed over per pass from pinned memory (not the value)"}
        print(f"bench: PCIe-inclusive round trip {nvis / t_pc / 1e6:.1f} Mvis/s", file=sys.stderr, flush=True)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
