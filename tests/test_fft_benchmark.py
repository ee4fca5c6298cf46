"""The fft-benchmarks driver (C++, drop-in for the reference's FFT benchmark executable) and the plain batched 2-D
DFT behind it (idgx_fft2d): YAML handling and error behaviour on the CPU; on the GPU the transform against numpy's
FP64 FFT (this library's three-pass kernels and the rocFFT path, square and rectangular, in and out of place, FP32
and FP64), and a driver run with the reference's round-trip check and CSV layout."""
import os
import subprocess

import numpy as np
import pytest

from helpers import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ska-spc-tender-benchmarks_amd")
DRIVER = os.path.join(PKG, "bin", "fft-benchmarks")

BASE = """memorysize: 100000000
niterations: 2
dimensions: [[1024, 1024], [2048, 1024], [300, 200]]
float_types: [float, double]
transform_types: [forward, inverse]
hardware_types: [gpu]
"""


def _run(tmp_path, text, out="result.csv"):
    cfg = tmp_path / "c.yaml"
    cfg.write_text(text)
    return subprocess.run([DRIVER, str(cfg), out], cwd=tmp_path, capture_output=True, text=True, timeout=300)


def test_usage_message():
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "fft-benchmarks <configuration file> <output file>" in r.stderr


@pytest.mark.parametrize("edit, message", [
    (("dimensions: [[1024, 1024], [2048, 1024], [300, 200]]", "dimensions: []"), "dimensions section empty."),
    (("float_types: [float, double]", "float_types: [half]"), "Invalid float_type token half"),
    (("transform_types: [forward, inverse]", "transform_types: [backward]"), "Invalid transform_type token backward"),
    (("hardware_types: [gpu]", "hardware_types: [cpu]"), "CPU support is disabled"),
    (("hardware_types: [gpu]", "hardware_types: []"), "hardware_types section empty."),
    (("niterations: 2\n", ""), "Invalid niterations section"),
    (("memorysize: 100000000\n", ""), "Invalid memorysize section"),
])
def test_configuration_errors(tmp_path, edit, message):
    r = _run(tmp_path, BASE.replace(*edit))
    assert r.returncode != 0 and message in r.stderr, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("nx, ny, fp64, in_place", [
    (1024, 1024, False, False), (2048, 1024, False, True), (1024, 4096, False, False),
    (300, 200, False, False), (512, 256, True, False), (4096, 4096, False, True)])
def test_fft2d_matches_numpy_fp64(nx, ny, fp64, in_place):
    """Planes [batch][ny][nx]: sign -1 is numpy's fft2, sign +1 its unnormalised ifft2."""
    import torch
    from idgx import FFT_FP32, FFT_FP64, fft2d

    batch = 3
    rng = np.random.default_rng(nx + ny)
    x = rng.standard_normal((batch, ny, nx)) + 1j * rng.standard_normal((batch, ny, nx))
    x = x.astype(np.complex128 if fp64 else np.complex64)
    for sign in (-1, 1):
        d_in = torch.from_numpy(x.copy()).to("cuda")
        d_out = d_in if in_place else torch.empty_like(d_in)
        fft2d(d_in, d_out, nx, ny, batch, sign, FFT_FP64 if fp64 else FFT_FP32)
        torch.cuda.synchronize()
        got = d_out.cpu().numpy()
        xd = x.astype(np.complex128)
        ref = np.fft.fft2(xd) if sign < 0 else np.fft.ifft2(xd) * (nx * ny)
        err = max(rel_l2(got[b], ref[b]) for b in range(batch))
        print(f"fft2d {nx}x{ny} {'fp64' if fp64 else 'fp32'} sign {sign:+d} {'in place' if in_place else ''}: "
              f"{err:.2e}")
        assert err < (1e-13 if fp64 else 1e-6)


@pytest.mark.gpu
def test_driver_runs_and_writes_csv(tmp_path):
    r = _run(tmp_path, BASE)
    assert r.returncode == 0, r.stderr + r.stdout
    rows = [l.rstrip("\n").split(";")[:-1] for l in open(tmp_path / "result.csv")]
    assert len(rows) == 1 + 12 and len(rows[0]) == 14
    assert rows[0][2].strip() == "Status" and rows[0][9].strip() == "Mean compute time (us)"
    for row in rows[1:]:
        fp = "double" if int(row[0]) >= 6 else "float"  # float types expand before transform types and dimensions
        print(row[1], fp, row[2], row[3], row[5], row[9], row[12], row[13])
        assert row[2] == "OK" and int(row[5]) >= 1 and float(row[9]) > 0
        # the reference's check: the inverse of the output against the input pattern k mod 128, scaled 1/(nx ny)
        assert float(row[3]) < 2e-3
