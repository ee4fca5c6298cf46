"""The gridding-benchmarks driver (C++, drop-in for the reference executable): YAML handling and error behaviour
on the CPU, a real run and the CSV layout on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ska-spc-tender-benchmarks_amd")
DRIVER = os.path.join(PKG, "bin", "gridding-benchmarks")

BASE = """grid_size: 1024
operation_types: [gridding, degridding]
niterations: 2
nstations: 3
ntimeslots: 1
ntimesteps_per_subgrid: 32
nchannels: 8
subgrid_size: 32
hardware_types: [gpu]
"""


def _run(tmp_path, text, out="result.csv"):
    cfg = tmp_path / "c.yaml"
    cfg.write_text(text)
    return subprocess.run([DRIVER, str(cfg), out], cwd=tmp_path, capture_output=True, text=True, timeout=120)


def test_usage_message():
    r = subprocess.run([DRIVER], capture_output=True, text=True)
    assert r.returncode != 0 and "gridding-benchmarks <configuration file> <output file>" in r.stderr


def test_missing_file(tmp_path):
    r = subprocess.run([DRIVER, str(tmp_path / "nope.yaml"), "o.csv"], capture_output=True, text=True)
    assert r.returncode != 0 and "not found" in r.stderr


@pytest.mark.parametrize("edit,message", [
    (("hardware_types: [gpu]", "hardware_types: []"), "hardware_types section empty."),
    (("hardware_types: [gpu]", "hardware_types: [cpu]"), "CPU support is disabled"),
    (("hardware_types: [gpu]", "hardware_types: [tpu]"), "Invalid hardware_type token tpu"),
    (("[gridding, degridding]", "[]"), "operation_types section empty."),
    (("[gridding, degridding]", "[gridding, fft]"), "Invalid operation type token fft"),
    (("grid_size: 1024\n", ""), "Invalid grid_size section"),
    (("nchannels: 8\n", ""), "Invalid nchannels section"),
])
def test_configuration_errors(tmp_path, edit, message):
    r = _run(tmp_path, BASE.replace(*edit))
    assert r.returncode != 0
    assert message in r.stderr, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_driver_runs_and_writes_csv(tmp_path, precision):
    r = _run(tmp_path, BASE + f"precision: {precision}\n", out=str(tmp_path / "sub" / "result.csv"))
    assert r.returncode == 0, r.stderr
    # Like the reference, the CSV lands in the working directory under the file name only.
    rows = [line.split(";") for line in (tmp_path / "result.csv").read_text().splitlines()]
    assert len(rows) == 3
    assert all(len(row) == 15 and row[-1] == "" for row in rows)  # 14 cells + trailing ';'
    assert rows[0][0].strip() == "Configuration ID" and rows[0][9].strip() == "Mean compute time (us)"
    assert [row[1] for row in rows[1:]] == ["gridding", "degridding"]
    for row in rows[1:]:
        assert float(row[9]) > 0 and row[12] == precision and float(row[13]) > 0
