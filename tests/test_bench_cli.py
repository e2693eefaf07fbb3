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


def _timed_csv(path: Path, rows):
    """rows: (kernel, counter, value, start_ns, end_ns)"""
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp",
                                          "End_Timestamp"])
        w.writeheader()
        for name, counter, value, t0, t1 in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": value, "Start_Timestamp": t0,
                        "End_Timestamp": t1})


def test_sq_summary_valu_issue_and_waits(tmp_path, capsys):
    """VALU issue = instructions over one wave64 instruction per SIMD every 4
    cycles in GRBM_GUI_ACTIVE / 8 cycles (the counter runs on every XCD)."""
    a, b = tmp_path / "a.csv", tmp_path / "b.csv"
    # 2 dispatches; per dispatch 8e6 GRBM cycles (1e6 per XCD) -> 1024 * 1e6 / 4 = 2.56e8 issue slots
    _timed_csv(a, [(FUSED, "SQ_INSTS_VALU", 1.28e8, 0, 1000), (FUSED, "SQ_INSTS_VALU", 1.28e8, 2000, 3000),
                   (FUSED, "SQ_WAVE_CYCLES", 1e9, 0, 1000), (FUSED, "SQ_WAVE_CYCLES", 1e9, 2000, 3000),
                   (FUSED, "SQ_WAIT_INST_ANY", 2.5e8, 0, 1000), (FUSED, "SQ_WAIT_INST_ANY", 2.5e8, 2000, 3000),
                   (COPY, "SQ_INSTS_VALU", 5.0, 0, 10)])
    _timed_csv(b, [(FUSED, "GRBM_GUI_ACTIVE", 8e6, 0, 1000), (FUSED, "GRBM_GUI_ACTIVE", 8e6, 2000, 3000)])
    assert bench.sq_summary([str(a), str(b)]) == 0
    out = capsys.readouterr().out
    assert "step_kernel<8, 2, 2, 1, 2>" in out and "copyBuffer" not in out
    assert "VALU issue            0.500" in out
    assert "SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES 0.250" in out
    assert "dispatches 2" in out
    # any kernel by name
    assert bench.sq_summary([str(a)], match="copyBuffer") == 0
    assert "copyBuffer" in capsys.readouterr().out


def test_pmc_total_sums_every_dispatch(tmp_path, capsys):
    """Calibrated bytes summed over every dispatch of the stepper's kernels,
    over the summed dispatch time; other kernels are left out."""
    pool0 = "void fdtd2d::(anonymous namespace)::pool_kernel<8, 1, 1, 1, 0>(fdtd2d::StepArgs const*, int)"
    pset = "fdtd2d::(anonymous namespace)::pool_set(fdtd2d::PoolQueue*, fdtd2d::(anonymous namespace)::PoolLists, int)"
    f, w, cf, cw = (tmp_path / n for n in ("f.csv", "w.csv", "cf.csv", "cw.csv"))
    # 2 dispatches of 1 ms each: FETCH 1e6 KiB raw each (x2), WRITE 5e5 KiB each (x1)
    _timed_csv(f, [(pool0, "FETCH_SIZE", 1e6, 0, 1_000_000), (pool0, "FETCH_SIZE", 1e6, 2_000_000, 3_000_000),
                   (pset, "FETCH_SIZE", 0.0, 0, 100), (COPY, "FETCH_SIZE", 9e9, 0, 10)])
    _timed_csv(w, [(pool0, "WRITE_SIZE", 5e5, 0, 1_000_000), (pool0, "WRITE_SIZE", 5e5, 2_000_000, 3_000_000),
                   (pset, "WRITE_SIZE", 0.0, 0, 100)])
    mib = 1024
    _counter_csv(cf, "FETCH_SIZE", [(COPY, mib * 1024 / 2)])  # FETCH reads half: scale 2
    _counter_csv(cw, "WRITE_SIZE", [(COPY, mib * 1024)])
    assert bench.pmc_total([str(f), str(w), str(cf), str(cw)], mib) == 0
    out = capsys.readouterr().out
    assert "pool_kernel<8, 1, 1, 1, 0>" in out and "pool_set" in out and "copyBuffer" not in out
    total = (2 * 1e6 * 1024 * 2 + 2 * 5e5 * 1024) / 1e9
    assert f"total: {total:.2f} GB in 2.0 ms of dispatches" in out


def test_serving_time_slicing_defaults():
    from types import SimpleNamespace as NS

    kw = dict(slice_cells=16e6, band_small=64, band_cpml_small=32)
    mixed = bench.slicing_options(NS(mode="serve-mixed", slice_every=-1, **kw))
    assert mixed == {"slice_every": 16, "slice_cells": 16e6, "band_small": 64, "band_cpml_small": 32}
    sweep = bench.slicing_options(NS(mode="serve-sweep", slice_every=-1, **kw))
    assert sweep == {"slice_every": 1, "slice_cells": 0, "band_small": 0, "band_cpml_small": 0}
    off = bench.slicing_options(NS(mode="serve-mixed", slice_every=1, **kw))
    assert off["slice_cells"] == 0 and off["band_small"] == 0 and off["band_cpml_small"] == 0
    forced = bench.slicing_options(NS(mode="serve-sweep", slice_every=4, slice_cells=1e6, band_small=0,
                                      band_cpml_small=0))
    assert forced == {"slice_every": 4, "slice_cells": 1e6, "band_small": 0, "band_cpml_small": 0}


def test_rel_l2():
    import numpy as np

    b = np.array([3.0, 4.0], np.float32)
    assert bench.rel_l2(b, b) == 0.0
    assert abs(bench.rel_l2(np.array([3.0, 4.5]), b) - 0.1) < 1e-12
    assert bench.rel_l2(np.zeros(2), np.zeros(2)) == 0.0
    assert bench.rel_l2(np.ones(2), np.zeros(2)) == float("inf")


def test_busy_summary_unions_overlapping_kernels(tmp_path):
    path = tmp_path / "run_kernel_trace.csv"
    rows = [  # (name, start ns, end ns): two pool kernels side by side, a copy, a 5 ms gap
        ("void pool_kernel<8, 1, 1, 1, 0>(...)", 0, 4_000_000),
        ("void pool_kernel<8, 1, 1, 1, 1>(...)", 1_000_000, 5_000_000),
        ("__amd_rocclr_copyBuffer", 5_000_000, 9_000_000),
        ("void pool_kernel<8, 1, 1, 1, 0>(...)", 10_000_000, 12_000_000),
        ("void pool_kernel<8, 1, 1, 1, 0>(...)", 12_500_000, 13_000_000),
    ]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for n, a, b in rows:
            w.writerow({"Kernel_Name": n, "Start_Timestamp": a, "End_Timestamp": b})
    out = bench.busy_summary(str(path))
    assert out["kernels"] == 4
    assert abs(out["span_s"] - 0.013) < 1e-12 and abs(out["busy_s"] - 0.0075) < 1e-12
    assert abs(out["idle_frac"] - 5.5 / 13) < 1e-12
    assert out["gaps_over"]["count"] == 1 and out["gaps_over"]["largest_ms"] == [5.0]
