"""Processing line (time-of-flight search, F/processing.cu) on the CPU.

The oracle's restatement (oracle/processing.cpp) is pinned by hand-worked
series for every quirk of the reference filters, by numpy for the window
arithmetic, and by constructed crossings for the three TOF methods; the
library's own implementation (fdtd2d_amd/csrc/processing.cpp) must then equal
the oracle bit for bit, accept and reject the same inputs, and write the
reference's file formats. The line capture rule itself is pinned on the
oracle against its own fields read after every step.
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

import fdtd2d_amd as F
from fdtd2d_amd import processing as P
from oracle import oracle as O

from helpers import random_problem

GOLD = Path(__file__).resolve().parent / "golden"
f32 = np.float32


def _filtered(series, **kw):
    """The oracle's prefiltered series of one coordinate (positive values, so
    |x| = x)."""
    line = np.asarray(series, f32)[:, None]
    a = P.config(perform=1, size_i=1, **kw)
    r = O.processing_tof(a, line, 0)
    assert r is not None
    return r["filtered"][0]


# ------------------------------------------------------------ oracle: filters
def test_moving_average_edge_copies_hand_worked():
    """w = 3 on 0..5: means at 1..4, then dst[1] <- dst[0] (never written by
    the filter, so the source value) and dst[5] <- dst[4]."""
    assert _filtered([0, 1, 2, 3, 4, 5], filter_type=1, filter_win=3, filter_runs=1).tolist() == [0, 0, 2, 3, 4, 4]
    # even w = 4: means land at 1..3; the copies run from the inside out
    got = _filtered([10, 1, 2, 3, 4, 5], filter_type=1, filter_win=4, filter_runs=1)
    assert got.tolist() == [10, 4, 4, 3.5, 3.5, 3.5]


def test_moving_average_second_run_reads_the_first():
    got = _filtered([0, 1, 2, 3, 4, 5], filter_type=1, filter_win=3, filter_runs=2)
    assert got.tolist() == [0, 0, f32(5) / f32(3), 3, f32(11) / f32(3), f32(11) / f32(3)]


def test_moving_average_closed_forms_and_numpy():
    # a constant stays constant; a ramp is reproduced at the window centre
    assert np.all(_filtered(np.full(50, 7.0), filter_type=1, filter_win=9, filter_runs=3) == 7)
    t = np.arange(60, dtype=np.float64)
    got = _filtered(t, filter_type=1, filter_win=7, filter_runs=1)
    assert np.array_equal(got[4:57], t[4:57])
    # random series: the window sits centred at t - w/2 (a shifted window fails)
    rng = np.random.default_rng(3)
    for w in (1, 2, 5, 8, 31):
        x = rng.uniform(0.0, 2.0, 200)
        got = _filtered(x, filter_type=1, filter_win=w, filter_runs=1).astype(np.float64)
        x32 = x.astype(f32).astype(np.float64)
        ref = np.convolve(x32, np.ones(w) / w, mode="valid")  # ref[k] = mean(x[k : k + w])
        h = w // 2
        off = w - 1 - h  # the mean of the window starting at k is stored at k + off
        for k in range(len(ref)):
            j = k + off
            if h < j < len(x) - h:  # not overwritten by the edge copies
                assert math.isclose(got[j], ref[k], rel_tol=2e-6, abs_tol=2e-6), (w, k)


def test_median_filter_against_numpy():
    rng = np.random.default_rng(4)
    for w in (1, 3, 4, 9, 10):
        x = rng.uniform(0.0, 1.0, 120).astype(f32)
        got = _filtered(x, filter_type=2, filter_win=w, filter_runs=1)
        h = w // 2
        for t in range(w - 1, len(x)):
            if t - h <= h:  # overwritten by the head copies
                continue
            if t - h >= len(x) - h:
                continue
            win = np.sort(x[t - w + 1:t + 1])
            assert got[t - h] == win[h], (w, t)  # rank w/2: the upper median for even w
            if w % 2:
                assert got[t - h] == np.median(x[t - w + 1:t + 1])
    # hand-worked, with the head copy from the unfiltered value at 0
    assert _filtered([5, 1, 4, 2, 3, 0], filter_type=2, filter_win=3, filter_runs=1).tolist() == [5, 5, 2, 3, 2, 2]


def test_other_filter_types_leave_the_series():
    x = np.arange(1, 20, dtype=np.float64)
    for ft in (0, 3, -1):
        assert np.array_equal(_filtered(x, filter_type=ft, filter_win=500, filter_runs=2), x)
    assert np.array_equal(_filtered(x, filter_type=1, filter_win=500, filter_runs=0), x)


# -------------------------------------------------------- oracle: TOF methods
def _tof(series, **kw):
    line = np.asarray(series, f32)[:, None]
    a = P.config(perform=1, size_i=1, filter_type=0, **kw)
    r = O.processing_tof(a, line, 0)
    assert r is not None
    return r


def test_second_derivative_crossing():
    """d2 of 0,0,1,3,6,8,9,9,9 is 1,1,1,-1,-1,-1,0 at 1..7: the first strict
    sign change is between 3 and 4; the threshold gates on the series."""
    s = [0, 0, 1, 3, 6, 8, 9, 9, 9]
    kw = dict(sec_der=1, sdf_type=1, sdf_win=1, sdf_runs=1, begin_after=100)
    assert _tof(s, tau_threshold=0.25, **kw)["sd"][0] == 103
    assert _tof(s, tau_threshold=0.5, **kw)["sd"][0] == 100 + len(s) - 1  # 3 <= 4.5: no crossing taken
    # without a post-filter the sign test runs on a copy of the (non-negative)
    # series, not on its second difference: never a crossing
    assert _tof(s, tau_threshold=0.0, sec_der=1, sdf_type=0, begin_after=0)["sd"][0] == len(s) - 1


def test_constant_fraction_crossing():
    """src = t, k = -0.4, tau = 4: p(t) = 0.6 t - 1.6 changes sign between 2 and 3."""
    t = np.arange(30)
    r = _tof(t, cf=1, cf_tau=4, cf_k=-0.4, arc=1, arc_tau=4, arc_k=-0.4, tau_threshold=0.0, begin_after=7)
    assert r["cf"][0] == 9 and r["arc"][0] == 9
    # a sample at the level does not count (strictly above), so t = 2 is skipped
    r = _tof(t, cf=1, cf_tau=4, cf_k=-0.4, tau_threshold=2 / 29, begin_after=0)
    assert r["cf"][0] == 29
    # ARC is the same search with its own delay and attenuation: k = -0.25,
    # tau = 6: p(t) = 0.75 t - 1.5 is 0 at t = 2 exactly, a zero product, so
    # no crossing anywhere
    r = _tof(t, arc=1, arc_tau=6, arc_k=-0.25, tau_threshold=0.0, begin_after=0)
    assert r["arc"][0] == 29


def test_detector_simulation_box_sum():
    """Uniform weights 1/D and the division by D: the detector at i is
    sum(|line[i..i+D-1]|) / D^2."""
    rng = np.random.default_rng(5)
    line = rng.standard_normal((40, 25)).astype(f32)
    D = 6
    a = P.config(perform=1, size_i=25, detector_size=D, filter_type=0)
    r = O.processing_tof(a, line, 1)
    box = np.stack([np.abs(line[:, i:i + D].astype(np.float64)).sum(1) for i in range(25 - D)])
    np.testing.assert_allclose(r["original"], box / D / D, rtol=3e-6)
    assert np.all(r["filtered"] == 0)  # the detector pass filters into a zero matrix


def test_oracle_line_capture_rule():
    """The captured line equals row line_j, columns [line_i0, +line_size) of
    Ez read after step line_begin + t + 1, and nothing else."""
    rng = np.random.default_rng(6)
    p = random_problem(rng, nx=40, ny=30, npml=5, nsteps=20)
    p.line_size, p.line_i0, p.line_j, p.line_begin, p.line_end = 17, 11, 9, 4, 15
    o = O.Oracle(p)
    rows = []
    for n in range(1, 21):
        o.run(1)
        if 4 < n <= 15:
            rows.append(o.read_field(F.EZ)[9, 11:28])
    assert np.array_equal(o.read_line(), np.stack(rows))
    assert o.read_line().shape == (11, 17)
    # finish beyond the run: rows stop at nsteps
    p.line_end = 1000
    o = O.Oracle(p)
    assert o.read_line().shape == (16, 17)


# ------------------------------------------------ library against the oracle
def _random_config(rng, n, T):
    return dict(perform=1, sec_der=int(rng.integers(0, 2)), cf=int(rng.integers(0, 2)), arc=int(rng.integers(0, 2)),
                filter_type=int(rng.integers(0, 4)), filter_win=int(rng.integers(1, 14)),
                filter_runs=int(rng.integers(0, 4)), sdf_type=int(rng.integers(0, 3)),
                sdf_win=int(rng.integers(1, 14)), sdf_runs=int(rng.integers(0, 3)),
                cf_tau=int(rng.integers(-1, 12)), arc_tau=int(rng.integers(0, 12)),
                cf_k=float(rng.uniform(-1, 0)), arc_k=float(rng.uniform(-1, 0)),
                tau_threshold=float(rng.uniform(0, 0.6)), size_i=n, detector_size=int(rng.integers(0, n + 1)),
                begin_after=int(rng.integers(0, 1000)))


def test_library_equals_oracle_random():
    rng = np.random.default_rng(7)
    accepted = rejected = 0
    for _ in range(150):
        T, n = int(rng.integers(1, 90)), int(rng.integers(1, 24))
        a = P.config(**_random_config(rng, n, T))
        # smooth pulses (realistic crossings) or noise, with a few repeated values
        tt = np.arange(T)[:, None]
        line = (np.sin(tt * rng.uniform(0.05, 0.5) + rng.uniform(0, 6, n)) * np.exp(-((tt - T / 2) / (T / 4 + 1)) ** 2))
        if rng.random() < 0.4:
            line = rng.standard_normal((T, n))
        if rng.random() < 0.3:
            line = np.round(line * 4) / 4
        line = line.astype(f32)
        for pass_ in (0, 1):
            w = None if rng.random() < 0.5 else rng.standard_normal(max(a.detector_size, 1)).astype(f32)
            ref = O.processing_tof(a, line, pass_, w)
            if ref is None:
                with pytest.raises(F.FDTD2DError) as e:
                    P.tof(a, line, pass_, w)
                assert e.value.code == 1
                rejected += 1
                continue
            got = P.tof(a, line, pass_, w)
            for k in got:
                assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), (k, pass_)
            accepted += 1
    assert accepted > 100 and rejected > 20


@pytest.mark.parametrize("ftype", [1, 2])
def test_library_equals_oracle_reference_windows(ftype):
    """The shipped settings' windows (300-sample filters, delays 800/1200) on a
    2000-step, 48-cell line of reflected pulses."""
    rng = np.random.default_rng(8 + ftype)
    T, n = 2000, 48
    t = np.arange(T)[:, None]
    arrival = 600 + 8 * np.arange(n)[None, :] + rng.integers(-5, 5, n)[None, :]
    env = np.clip((t - arrival) / 300.0, 0, 1)
    line = (env * np.sin(0.2 * t) + 0.01 * rng.standard_normal((T, n))).astype(f32)
    a = P.config(perform=1, sec_der=1, sdf_type=1, sdf_win=300, sdf_runs=1, cf=1, cf_tau=800, cf_k=-0.5, arc=1,
                 arc_tau=600, arc_k=-0.6, tau_threshold=0.25, filter_type=ftype, filter_win=300, filter_runs=2,
                 size_i=n, detector_size=10, begin_after=12588)
    for pass_ in (0, 1):
        ref = O.processing_tof(a, line, pass_)
        got = P.tof(a, line, pass_)
        for k in got:
            assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), (k, pass_)
        if pass_ == 0:
            # the TOFs found track the planted arrivals
            assert np.all(np.abs(got["cf"] - 12588 - arrival[0]) < 400)


# ------------------------------------------------------------------ files
def _tof_lines(v, sieve, start, dt):
    return ["%d %f %f" % (start + i, float(v[i]), (dt * float(v[i])) * 1e12) for i in range(0, len(v), sieve)]


def _matrix_lines(m, t_sieve, d_sieve):
    n, T = m.shape
    return ["".join("%f\t" % float(m[i, t]) for i in range(0, n, d_sieve)) for t in range(0, T, t_sieve)]


def test_written_files_follow_reference_formats(tmp_path):
    rng = np.random.default_rng(9)
    T, n = 240, 30
    line = rng.standard_normal((T, n)).astype(f32)
    a = P.config(perform=1, sec_der=1, sdf_type=1, sdf_win=9, sdf_runs=1, cf=1, cf_tau=20, arc=1, arc_tau=15,
                 tau_threshold=0.25, filter_type=1, filter_win=11, filter_runs=2, save_original=1, save_filtered=1,
                 dim_sieve=2, time_sieve=5, sim_detect=1, save_detector=1, load_weights=0, detector_size=8,
                 begin_after=50, pos_i=40, size_i=n, dt=4.766437173827514e-13)
    P.write_outputs(a, str(tmp_path), line)
    r0, r1 = O.processing_tof(a, line, 0), O.processing_tof(a, line, 1)
    for name, v, start in (("SD", r0["sd"], 40), ("CF", r0["cf"], 40), ("ARC", r0["arc"], 40),
                           ("det_SD", r1["sd"], 44), ("det_CF", r1["cf"], 44), ("det_ARC", r1["arc"], 44)):
        assert (tmp_path / f"{name}.txt").read_text().splitlines() == _tof_lines(v, 2, start, a.dt), name
    for name, m in (("original_field", r0["original"]), ("filtered_field", r0["filtered"]),
                    ("detector_original_field", r1["original"]), ("detector_filtered_field", r1["filtered"])):
        assert (tmp_path / f"{name}.txt").read_text().splitlines() == _matrix_lines(m, 5, 2), name


@pytest.mark.parametrize("text,expect", [
    ("0.5 0.25 1 2", "full"),       # all four read
    ("0.5 0.25\n", "short_nl"),     # only whitespace left: the rest keep 1/DetSize
    ("0.5 x 1 2", "bad"),           # a value that does not parse reads as 0 and ends the loop
    ("0.5 0.25", "short_eof"),      # at end of file before all are read: all weights 1
    (None, "absent"),               # no file: 1/DetSize
])
def test_detector_weights_file(tmp_path, text, expect):
    """DetWeights.txt read as the reference's loop reads it with operator>>."""
    D = 4
    w = {"full": [0.5, 0.25, 1, 2], "short_nl": [0.5, 0.25, 1 / D, 1 / D], "bad": [0.5, 0, 1 / D, 1 / D],
         "short_eof": [1, 1, 1, 1],
         "absent": [1 / D] * D}[expect]
    if text is not None:
        (tmp_path / "DetWeights.txt").write_text(text)
    rng = np.random.default_rng(10)
    line = rng.standard_normal((50, 12)).astype(f32)
    a = P.config(perform=1, sec_der=1, sdf_type=0, cf=1, cf_tau=5, filter_type=1, filter_win=5, sim_detect=1,
                 load_weights=1, detector_size=D, size_i=12, dt=1e-12)
    P.write_outputs(a, str(tmp_path), line)
    ref = O.processing_tof(a, line, 1, np.asarray(w, f32))
    assert (tmp_path / "det_CF.txt").read_text().splitlines() == _tof_lines(ref["cf"], 1, D // 2, a.dt)


def test_writer_rejects_what_the_reference_reads_out_of_bounds(tmp_path):
    line = np.zeros((10, 5), f32)
    for kw in (dict(filter_type=1, filter_win=11), dict(cf=1, cf_tau=10), dict(sim_detect=1, detector_size=5),
               dict(dim_sieve=0), dict(time_sieve=0)):
        a = P.config(perform=1, size_i=5, **kw)
        with pytest.raises(F.FDTD2DError) as e:
            P.write_outputs(a, str(tmp_path), line)
        assert e.value.code == 1


# ------------------------------------------------------------------ setup
@pytest.mark.parametrize("name,pos_i,size_i", [("settings_full.ini", 40, 420), ("settings_release.ini", 40, 420),
                                               ("settings_build.ini", 100, 900)])
def test_processing_keys_of_shipped_settings(name, pos_i, size_i):
    """[Processing] scaled as recuclParametersToMetric does: cells = int(v / ds)
    with ds = 0.5, steps = int(ns * 1e-9 / dt) with Use_Exp_time = 1, and the
    SDF* keys taking the INI's own values."""
    s = F.Setup.from_ini(str(GOLD / name))
    a, p = s.processing(), s.problem()
    dt, n_max = s.get("dt"), int(s.get("nMax"))
    begin = 8 if name == "settings_build.ini" else 6
    finish = 19 if name == "settings_build.ini" else 18
    assert a.perform == 1 and a.pos_i == pos_i and a.pos_j == 200 and a.pos_k == 0 and a.size_i == size_i
    assert a.begin_after == int(begin * (1e-9 / dt)) and a.finish_at == int(finish * (1e-9 / dt))
    assert a.times == min(a.finish_at, n_max) - a.begin_after
    assert (a.filter_win, a.filter_runs, a.sdf_win, a.sdf_runs, a.cf_tau, a.arc_tau) == (300, 2, 300, 1, 1200, 800)
    assert math.isclose(a.cf_k, -0.5) and math.isclose(a.arc_k, -0.6, rel_tol=1e-7)
    assert a.dt == dt
    assert (p.line_size, p.line_i0, p.line_j, p.line_begin, p.line_end) == (size_i, pos_i, 200, a.begin_after,
                                                                           a.finish_at)


@pytest.mark.parametrize("edit,captured", [
    ({"Perform": "0"}, False),
    ({"PosI": "400"}, False),      # 800 + 420 > Imax = 1000: outside the grid
    ({"PosJ": "-1"}, False),
    ({"PosK": "1"}, False),        # the 2D grid has one K plane
    ({"BeginAfter": "100"}, False),  # after the last step: nothing captured
    ({"SizeI": "20"}, True),
])
def test_processing_line_geometry(tmp_path, edit, captured):
    text = (GOLD / "settings_build.ini").read_text()
    lines = []
    in_proc = False
    for ln in text.splitlines():
        st = ln.strip()
        if st.startswith("["):
            in_proc = st.lower() == "[processing]"
        key = st.split("=")[0].strip() if "=" in st else None
        if in_proc and key in edit:
            ln = f"{key} = {edit[key]}"
        lines.append(ln)
    if "PosK" in edit:
        at = next(k for k, ln in enumerate(lines) if ln.strip().lower() == "[processing]")
        lines.insert(at + 1, f"PosK = {edit['PosK']}")
    f = tmp_path / "Settings.ini"
    f.write_text("\n".join(lines) + "\n")
    p = F.Setup.from_ini(str(f)).problem()
    assert (p.line_size > 0) == captured
