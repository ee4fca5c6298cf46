"""The fft-benchmarks driver (C++, drop-in for the reference's FFT benchmark executable) and the plain batched 2-D
DFT behind it (idgx_fft2d): YAML handling, error behaviour and the kernel routing on the CPU; on the GPU the transform
against numpy's FP64 FFT on every path (LDS tiles for sides 16..512 in FP32 and FP64, the three-pass kernels, the
radix-3 decimation for sides 3 x 2^k, rocFFT for the rest; square and rectangular, in and out of place), the
24576^2 plane through size-independent properties, and a driver run with the reference's round-trip check and CSV
layout."""
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


# The reference's fft/examples/gpu.yaml:3 sizes (single precision).
GPU_YAML_SIDES = (64, 128, 512, 1024, 2048, 4096, 8192, 16384, 24576)


def test_every_reference_gpu_size_runs_own_kernels():
    from idgx import FFT_FP32, FFT_FP64, fft2d_own_kernels

    for n in GPU_YAML_SIDES:
        assert fft2d_own_kernels(n, n, FFT_FP32), n
    # FP64: every power-of-two plane with rows of up to 8192 points (LDS tiles, four-step columns)
    for nx, ny in ((16, 16), (64, 16), (512, 512), (256, 32), (1024, 1024), (8192, 8192), (512, 16384)):
        assert fft2d_own_kernels(nx, ny, FFT_FP64), (nx, ny)
    # FP32 power-of-two shapes outside the three-pass sizes (one side below 1024)
    for nx, ny in ((512, 4096), (64, 8192), (16384, 64), (16384, 16)):
        assert fft2d_own_kernels(nx, ny, FFT_FP32), (nx, ny)
    # rocFFT: other shapes, FP64 rows of 16384, column tiles wider than the plane
    for nx, ny, prec in ((300, 200, FFT_FP32), (3072, 1024, FFT_FP32), (8, 8, FFT_FP32), (16384, 16384, FFT_FP64),
                         (16, 2048, FFT_FP32)):
        assert not fft2d_own_kernels(nx, ny, prec), (nx, ny)
    from idgx import IdgxError
    with pytest.raises(IdgxError):
        fft2d_own_kernels(0, 64, FFT_FP32)


@pytest.mark.gpu
@pytest.mark.parametrize("nx, ny, fp64, in_place", [
    (16, 16, False, False), (64, 64, False, False), (128, 128, False, True), (512, 512, False, False),
    (16, 32, False, True), (64, 16, True, False), (256, 512, True, True), (128, 128, True, False),
    (128, 64, True, True), (512, 16, False, False), (16, 512, False, True), (256, 32, True, False),
    (1024, 1024, False, False), (2048, 1024, False, True), (1024, 4096, False, False),
    (300, 200, False, False), (512, 256, True, False), (4096, 4096, False, True), (1024, 1024, True, False),
    (3072, 3072, False, False), (3072, 3072, False, True), (6144, 6144, False, False), (12288, 12288, False, True),
    (2048, 1024, True, True), (4096, 4096, True, False), (8192, 2048, True, False), (8192, 8192, True, True),
    (512, 4096, False, False), (64, 8192, False, True), (16384, 64, False, False), (16384, 16, True, False)])
def test_fft2d_matches_numpy_fp64(nx, ny, fp64, in_place):
    """Planes [batch][ny][nx]: sign -1 is numpy's fft2, sign +1 its unnormalised ifft2."""
    import torch
    from idgx import FFT_FP32, FFT_FP64, fft2d, fft2d_own_kernels

    batch = 3 if nx * ny <= (1 << 24) else 1
    print(f"fft2d {nx}x{ny}: {'own kernels' if fft2d_own_kernels(nx, ny, FFT_FP64 if fp64 else FFT_FP32) else 'rocFFT'}",
          flush=True)
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


@pytest.mark.gpu
@pytest.mark.parametrize("fp64", [False, True])
def test_fft2d_every_power_of_two_shape_up_to_2048(fp64):
    """Every power-of-two shape 16 .. 2048 per side (64 shapes, batch 2, sign -1 out of place) against numpy's FP64
    FFT, whichever path runs it (LDS one- or two-pass, four-step columns, three-pass kernels, rocFFT)."""
    import torch
    from idgx import FFT_FP32, FFT_FP64, fft2d, fft2d_own_kernels

    prec = FFT_FP64 if fp64 else FFT_FP32
    sides = [16 << k for k in range(8)]
    worst, own = 0.0, 0
    for ny in sides:
        for nx in sides:
            rng = np.random.default_rng(nx * 7 + ny)
            x = rng.standard_normal((2, ny, nx)) + 1j * rng.standard_normal((2, ny, nx))
            x = x.astype(np.complex128 if fp64 else np.complex64)
            d_in = torch.from_numpy(x).to("cuda")
            d_out = torch.empty_like(d_in)
            fft2d(d_in, d_out, nx, ny, 2, -1, prec)
            torch.cuda.synchronize()
            ref = np.fft.fft2(x.astype(np.complex128))
            got = d_out.cpu().numpy()
            err = max(rel_l2(got[b], ref[b]) for b in range(2))
            assert err < (1e-13 if fp64 else 1e-6), (nx, ny, err)
            worst = max(worst, err)
            own += fft2d_own_kernels(nx, ny, prec)
    print(f"fft2d {'fp64' if fp64 else 'fp32'}: 64 shapes 16..2048, {own} on own kernels, worst {worst:.2e}",
          flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("n, batch", [(64, 9000), (16, 140003)])
def test_fft2d_small_planes_beyond_one_grid(n, batch):
    """More tiles than workgroups in a launch (8192), so that every workgroup loops over several tiles; the last
    tile of 16 x 16 planes is partial (16 planes per tile)."""
    import torch
    from idgx import FFT_FP32, fft2d

    rng = np.random.default_rng(n)
    x = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))).astype(np.complex64)
    d = torch.from_numpy(x).to("cuda")
    fft2d(d, d, n, n, batch, -1, FFT_FP32)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    ref = np.fft.fft2(x.astype(np.complex128))
    err = float(np.sqrt(np.sum(np.abs(got - ref) ** 2) / np.sum(np.abs(ref) ** 2)))
    print(f"fft2d {n}x{n} x {batch}: {err:.2e}", flush=True)
    assert err < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("nx, ny, fp64", [(512, 256, False), (256, 256, True)])
def test_fft2d_rows_then_columns_in_groups(nx, ny, fp64):
    """The row-then-column path runs 64 MB of planes at a time: 150 planes of 1 MB take groups of 64, 64 and 22
    planes; every plane against numpy's FP64 transform, out of place and both signs."""
    import torch
    from idgx import FFT_FP32, FFT_FP64, fft2d

    batch = 150
    rng = np.random.default_rng(nx * ny + batch)
    x = rng.standard_normal((batch, ny, nx)) + 1j * rng.standard_normal((batch, ny, nx))
    x = x.astype(np.complex128 if fp64 else np.complex64)
    for sign in (-1, 1):
        d_in = torch.from_numpy(x.copy()).to("cuda")
        d_out = torch.empty_like(d_in)
        fft2d(d_in, d_out, nx, ny, batch, sign, FFT_FP64 if fp64 else FFT_FP32)
        torch.cuda.synchronize()
        got = d_out.cpu().numpy()
        xd = x.astype(np.complex128)
        ref = np.fft.fft2(xd) if sign < 0 else np.fft.ifft2(xd) * (nx * ny)
        err = max(rel_l2(got[b], ref[b]) for b in range(batch))
        print(f"fft2d {nx}x{ny} x {batch} {'fp64' if fp64 else 'fp32'} sign {sign:+d}: {err:.2e}", flush=True)
        assert err < (1e-13 if fp64 else 1e-6)


@pytest.mark.gpu
def test_fft2d_24576_by_properties():
    """The largest reference size (one 24576^2 FP32 plane, 4.8 GB; also the largest 3 x 2^k side that a plane of at
    most 2^31 - 1 points admits), where numpy's FP64 transform is too slow for a test: (1) 48 unit impulses at seeded
    positions with seeded complex amplitudes transform to a sum of plane waves, checked in FP64 at 4096 seeded output
    positions per sign (catches index, sign and transposition mistakes of every pass); (2) a dense seeded plane
    transformed forward then inverse returns N^2 times itself (rounding of the full pipeline on dense data), and
    Parseval's sum |X|^2 = N^2 sum |x|^2 holds."""
    import torch
    from idgx import FFT_FP32, fft2d

    N = 24576
    rng = np.random.default_rng(N)
    K = 48
    rows, cols = rng.integers(0, N, K), rng.integers(0, N, K)
    assert len(set(zip(rows.tolist(), cols.tolist()))) == K
    amp = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    P, Q = rng.integers(0, N, 4096), rng.integers(0, N, 4096)
    x = torch.zeros((N, N), dtype=torch.complex64, device="cuda")
    x[torch.from_numpy(rows).cuda(), torch.from_numpy(cols).cuda()] = torch.from_numpy(amp.astype(np.complex64)).cuda()
    amp32 = amp.astype(np.complex64).astype(np.complex128)
    for sign in (-1, 1):
        y = torch.empty_like(x)
        fft2d(x, y, N, N, 1, sign, FFT_FP32)
        torch.cuda.synchronize()
        got = y[torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()].cpu().numpy().astype(np.complex128)
        # X[p][q] = sum_k a_k exp(sign 2 pi i (p r_k + q c_k) / N), phases reduced mod N exactly in integers
        ph = (np.outer(P, rows) + np.outer(Q, cols)) % N
        ref = (np.exp(sign * 2j * np.pi * ph / N) * amp32).sum(axis=1)
        err = rel_l2(got, ref)
        print(f"fft2d {N}^2 impulses sign {sign:+d}: {err:.2e}", flush=True)
        assert err < 1e-6
        del y
    gen = torch.Generator(device="cuda").manual_seed(7)
    x = torch.randn((N, N), dtype=torch.complex64, device="cuda", generator=gen)
    y = torch.empty_like(x)
    fft2d(x, y, N, N, 1, -1, FFT_FP32)
    e_in = torch.sum(x.abs().double() ** 2).item()
    e_out = torch.sum(y.abs().double() ** 2).item()
    parseval = abs(e_out / (float(N) ** 2 * e_in) - 1)
    fft2d(y, y, N, N, 1, 1, FFT_FP32)
    torch.cuda.synchronize()
    r = (torch.linalg.vector_norm((y / (float(N) ** 2) - x).to(torch.complex128)).item() /
         torch.linalg.vector_norm(x.to(torch.complex128)).item())
    print(f"fft2d {N}^2 dense: Parseval {parseval:.2e}, round trip {r:.2e}", flush=True)
    # each direction to the numpy bar of a single transform (1e-6), with slack for the second
    assert parseval < 1e-5 and r < 3e-6
