"""Write the golden fixtures for the gridding hot path from the compiled reference CPU sources.

Run in a container that holds the reference at /root/reference:

    python tests/golden/make_golden.py

It builds oracle/_ref/ref_driver with oracle/Makefile (the reference's gridding_benchmark.cpp,
degridding_benchmark.cpp, gridding_initialization.cpp and benchmarks_common.cpp compiled in place, as shipped:
-g, OpenMP backend), runs it for each shape below, and stores the dumped arrays as compressed .npz files next to
this script plus MANIFEST.json (shape, seed, SHA-256 of every array).  The fixtures are data only: the inputs the
reference's own generator produced (srand(0), gridding_initialization.cpp:7-206) and the outputs of its CPU
gridder (gridding_benchmark.cpp:21-182) and degridder (degridding_benchmark.cpp:23-199).

The degridder's input subgrids come from the reference's initialize_subgrids ramp, because launch_degridding
leaves them uninitialised (degridding_benchmark.cpp:221-230).  Shapes keep S <= 32: the reference CPU degridder
overruns its scratch buffer for larger subgrids (degridding_benchmark.cpp:62,118).
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))

# name: (nstations, ntimeslots, ntimesteps_per_subgrid, nchannels, subgrid_size, grid_size)
SHAPES = {
    "ref_s32": (3, 1, 32, 16, 32, 1024),
    "ref_s24": (4, 1, 20, 5, 24, 512),
    # gridding/examples/example.yaml: 2 stations, 2 timeslots, 128 timesteps, 16 channels, S=32, G=1024
    "ref_example": (2, 2, 128, 16, 32, 1024),
}

METADATA_DTYPE = np.dtype(
    [
        ("baseline_offset", "<i4"),
        ("time_offset", "<i4"),
        ("nr_timesteps", "<i4"),
        ("aterm_index", "<i4"),
        ("station1", "<u4"),
        ("station2", "<u4"),
        ("x", "<i4"),
        ("y", "<i4"),
        ("z", "<i4"),
    ]
)


def load(d, name, dtype, shape):
    a = np.fromfile(os.path.join(d, name), dtype=dtype)
    return a.reshape(shape)


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"])
    driver = os.path.join(REPO, "oracle", "_ref", "ref_driver")
    manifest = {}
    for name, shape in SHAPES.items():
        nst, nts, T, C, S, G = shape
        nbl = nst * (nst - 1) // 2
        nsub = nbl * nts
        with tempfile.TemporaryDirectory() as d:
            subprocess.check_call([driver, d] + [str(v) for v in shape], env=dict(os.environ, OMP_NUM_THREADS="4"))
            arrays = {
                "uvw": load(d, "uvw.bin", "<f4", (nsub, T, 3)),
                "frequencies": load(d, "frequencies.bin", "<f4", (C,)),
                "wavenumbers": load(d, "wavenumbers.bin", "<f4", (C,)),
                "visibilities": load(d, "visibilities.bin", "<c8", (nsub, T, C, 4)),
                "baselines": load(d, "baselines.bin", "<u4", (nbl, 2)),
                "spheroidal": load(d, "spheroidal.bin", "<f4", (S, S)),
                "aterms": load(d, "aterms.bin", "<c8", (nts, nst, S, S, 4)),
                "metadata": load(d, "metadata.bin", METADATA_DTYPE, (nsub,)),
                "grid_subgrids": load(d, "grid_subgrids.bin", "<c8", (nsub, 4, S, S)),
                "degrid_subgrids": load(d, "degrid_subgrids.bin", "<c8", (nsub, 4, S, S)),
                "degrid_visibilities": load(d, "degrid_visibilities.bin", "<c8", (nsub, T, C, 4)),
            }
        out = os.path.join(HERE, name + ".npz")
        np.savez_compressed(out, **arrays)
        manifest[name] = {
            "shape": dict(zip(["nstations", "ntimeslots", "ntimesteps_per_subgrid", "nchannels", "subgrid_size",
                               "grid_size"], shape)),
            "image_size": 0.1,
            "w_step_in_lambda": 0,
            "seed": "srand(0)",
            "build": "g++ -g -std=gnu++17 -fopenmp -DENABLE_OMP -DENABLE_CPU=1 (reference as shipped)",
            "sha256": {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() for k, v in arrays.items()},
        }
        print(f"{name}: {nsub} subgrids -> {os.path.relpath(out, REPO)} ({os.path.getsize(out)} bytes)")
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
