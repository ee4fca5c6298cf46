"""idgx.pmc_summary (no GPU): counter files of rocprofv3 --pmc become calibrated HBM bytes per kernel launch, with the
instruction-mix columns only when an instruction-mix run is given."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ska-spc-tender-benchmarks_amd")

FIELDS = ["Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(FIELDS, r)))


def test_traffic_only_summary_scales_by_the_copy_calibration(tmp_path):
    # A 1024 MiB copy read as 512 Mi KiB-units of FETCH_SIZE (factor 2) and 1 Gi of WRITE_SIZE (factor 1).
    mib = 1024
    _write(tmp_path / "cal_fetch", [("__amd_rocclr_copyBuffer", "FETCH_SIZE", mib * 1024 / 2)])
    _write(tmp_path / "cal_write", [("__amd_rocclr_copyBuffer", "WRITE_SIZE", mib * 1024)])
    k = "void idgx::(anonymous namespace)::wstack_row_kernel<16384, -1>(idgx::(anonymous namespace)::WsArgs)"
    _write(tmp_path / "fetch", [(k, "FETCH_SIZE", 1000.0), (k, "FETCH_SIZE", 3000.0)])
    _write(tmp_path / "write", [(k, "WRITE_SIZE", 500.0), (k, "WRITE_SIZE", 1500.0)])
    out = tmp_path / "summary"
    subprocess.run([sys.executable, "-m", "idgx.pmc_summary", "--cal-fetch", str(tmp_path / "cal_fetch"),
                    "--cal-write", str(tmp_path / "cal_write"), "--fetch", str(tmp_path / "fetch"),
                    "--write", str(tmp_path / "write"), "--out", str(out), "--command", "cmd"],
                   check=True, cwd=PKG, capture_output=True)
    res = json.load(open(str(out) + ".json"))
    assert abs(res["calibration"]["fetch_factor"] - 2.0) < 1e-12 and abs(res["calibration"]["write_factor"] - 1.0) < 1e-12
    rec = res["kernels"]["wstack_row_kernel<16384, -1>"]
    assert rec["launches"] == 2
    assert rec["read_bytes"] == 2000.0 * 1024 * 2.0 and rec["write_bytes"] == 1000.0 * 1024
    md = open(str(out) + ".md").read()
    assert "| kernel | launches | HBM read GB/launch | HBM write GB/launch |\n" in md
    assert "VALU instr" not in md and "nan" not in md
    assert "| `wstack_row_kernel<16384, -1>` | 2 | 0.00 | 0.00 |" in md
