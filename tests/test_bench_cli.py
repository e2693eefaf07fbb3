"""Host-side checks of bench.py (no GPU needed): the one-GPU refusal and the
rocprofv3 summary helpers that turn counter and kernel-stats CSVs into the
`roofline.traffic` record and the committed profile tables."""

import csv
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402

INTERIOR = "void fdtd2d::(anonymous namespace)::step_kernel<8, 2, 2, 1, 0>(fdtd2d::StepArgs, int, int)"
EDGE = "void fdtd2d::(anonymous namespace)::step_kernel<8, 1, 1, 1, 1>(fdtd2d::StepArgs, int, int)"
FUSED = "void fdtd2d::(anonymous namespace)::step_kernel<8, 2, 2, 1, 2>(fdtd2d::StepArgs, int, int)"
COPY = "__amd_rocclr_copyBuffer"


def _counter_csv(path: Path, counter: str, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for name, value in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": value})


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no bench line


def test_pmc_summary_scales_by_the_copy_calibration(tmp_path, capsys):
    # interior kernel: two dispatches of 100 and 120 KiB fetched, 50 KiB
    # written; the CPML-class kernel and other counters must be ignored; the
    # 1 MiB copy reads FETCH 512 KiB (half) and WRITE 1024 KiB (all)
    f, w, cf, cw = (tmp_path / n for n in ("f.csv", "w.csv", "cf.csv", "cw.csv"))
    _counter_csv(f, "FETCH_SIZE", [(INTERIOR, 100.0), (INTERIOR, 120.0), (EDGE, 900.0)])
    _counter_csv(w, "WRITE_SIZE", [(INTERIOR, 50.0), (EDGE, 700.0)])
    _counter_csv(cf, "FETCH_SIZE", [(COPY, 512.0)])
    _counter_csv(cw, "WRITE_SIZE", [(COPY, 1024.0)])
    out = tmp_path / "traffic.json"
    assert bench.pmc_summary([str(f), str(w), str(cf), str(cw)], 1, str(out)) == 0
    rec = json.loads(out.read_text())
    assert rec["kernel"] == INTERIOR
    assert rec["fetch_scale"] == 2.0 and rec["write_scale"] == 1.0
    assert rec["fetch_bytes_per_launch"] == 110 * 1024 * 2
    assert rec["write_bytes_per_launch"] == 50 * 1024
    assert rec["bytes_per_launch"] == 110 * 1024 * 2 + 50 * 1024
    assert "HBM traffic per launch" in capsys.readouterr().out


def test_pmc_summary_takes_the_fused_kernel(tmp_path):
    f, w, cf, cw = (tmp_path / n for n in ("f.csv", "w.csv", "cf.csv", "cw.csv"))
    _counter_csv(f, "FETCH_SIZE", [(FUSED, 300.0), (COPY, 5.0)])
    _counter_csv(w, "WRITE_SIZE", [(FUSED, 200.0)])
    _counter_csv(cf, "FETCH_SIZE", [(COPY, 512.0)])
    _counter_csv(cw, "WRITE_SIZE", [(COPY, 1024.0)])
    out = tmp_path / "traffic.json"
    assert bench.pmc_summary([str(f), str(w), str(cf), str(cw)], 1, str(out)) == 0
    rec = json.loads(out.read_text())
    assert rec["kernel"] == FUSED and rec["bytes_per_launch"] == 300 * 1024 * 2 + 200 * 1024


def test_stats_summary_table(tmp_path, capsys):
    p = tmp_path / "stats.csv"
    with open(p, "w", newline="") as fh:
        wr = csv.DictWriter(fh, fieldnames=["Name", "Calls", "AverageNs", "MinNs", "MaxNs", "Percentage"])
        wr.writeheader()
        wr.writerow({"Name": INTERIOR, "Calls": 52, "AverageNs": 671700.0, "MinNs": 643300.0, "MaxNs": 691200.0,
                     "Percentage": 56.24})
    assert bench.stats_summary(str(p)) == 0
    lines = capsys.readouterr().out.splitlines()
    assert lines[0].split()[:2] == ["calls", "avg_us"]
    assert lines[1].split()[:2] == ["52", "671.7"] and lines[1].endswith(INTERIOR)
