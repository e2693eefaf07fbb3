"""The host-driver path on the reference's own Settings.ini files, on the GPU.

Each shipped INI (medium builders: sphere, torus, fluctuation) is set up
exactly as the fdtd2d CLI does (fdtd2d_setup_from_ini), stepped on the GPU for
a bounded number of steps and compared with the FP32 oracle: traces bitwise,
and the output files the reference's scripts read (Detect.txt, detAll.txt,
detEz/Hx/Hy.txt, detP.txt, SourceP.txt) byte for byte; the CLI runs also
write deb.txt, iniLOG.txt and the empty files the reference opens.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import fdtd2d_amd as F
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden"
STEPS = 600
FILES = ["Detect.txt", "detAll.txt", "detEz.txt", "detHx.txt", "detHy.txt", "detP.txt", "SourceP.txt"]
# deb.txt (the material id map) and the files the reference opens and leaves empty
DEBUG_FILES = ["deb.txt", "materials.txt", "test.txt", "SDetect.txt", "LineDetect.txt", "LineDetect2.txt",
               "detEx.txt", "detEy.txt", "detHz.txt", "CompDetect.txt"]


@pytest.mark.parametrize("name", ["settings_full.ini", "settings_release.ini", "settings_build.ini"])
def test_shipped_settings_on_gpu_match_oracle(gpu_device, tmp_path, name):
    s = F.Setup.from_ini(GOLDEN / name)
    p = s.c_problem.contents
    print(f"{name}: {p.nx}x{p.ny}, P={p.npml}, nsrc={p.nsrc}, ndet={p.ndet}, ncomp={p.ncomp}, "
          f"running {STEPS} of {p.nsteps} steps", flush=True)
    with F.Plan(s, gpu_device) as plan:
        plan.run(STEPS)
        plan.sync()
        det, comp = plan.read_detectors()
        ez = plan.read_field(F.EZ)
    ref = Oracle(s, 32, threads=8)
    ref.run(STEPS)
    rdet, rcomp = ref.read_detectors()
    rez = ref.read_field(F.EZ)
    ref.close()
    assert np.array_equal(det, rdet.astype(np.float32))
    assert np.array_equal(comp, rcomp.astype(np.float32))
    assert np.array_equal(ez, rez.astype(np.float32))
    assert np.abs(det).max() > 0  # the pulse reached the detector row
    g, r = tmp_path / "gpu", tmp_path / "oracle"
    g.mkdir()
    r.mkdir()
    s.write_outputs(str(g), det, comp)
    s.write_outputs(str(r), rdet.astype(np.float32), rcomp.astype(np.float32))
    for f in FILES:
        assert (g / f).read_bytes() == (r / f).read_bytes(), f


def test_shipped_settings_rc_lines_match_oracle(gpu_device, tmp_path):
    s = F.Setup.from_ini(GOLDEN / "settings_release.ini")
    with F.Plan(s, gpu_device) as plan:
        plan.run(STEPS)
        plan.sync()
        spd = plan.read_rc()
    ref = Oracle(s, 32, threads=8)
    ref.run(STEPS)
    want = ref.read_rc()
    ref.close()
    assert np.array_equal(spd, want)


def test_full_length_run_reproduces_reference_rc_and_tau_lines(gpu_device, tmp_path):
    """All 33568 steps of the reference's own Settings.ini (FDTD_2D_Full), then
    the RC tracker's post-processing: RC_Line.txt and TAU_Line.txt equal the
    files the reference holds, line for line (the end-to-end fixture: INI
    parsing, medium builders, source table, stepping, RC tracking, filter and
    formatting)."""
    import json

    fx = json.loads((GOLDEN / "reference_lines.json").read_text())
    s = F.Setup.from_ini(GOLDEN / "settings_full.ini")
    p = s.c_problem.contents
    assert p.nsteps == 33568
    with F.Plan(s, gpu_device) as plan:
        plan.run(p.nsteps)
        plan.sync()
        spd = plan.read_rc()
        line = plan.read_line()
    s.write_rc_outputs(str(tmp_path), spd)
    rc = (tmp_path / "RC_Line.txt").read_text().splitlines()
    tau = (tmp_path / "TAU_Line.txt").read_text().splitlines()
    assert rc == fx["rc_line_text"]
    assert tau == fx["tau_line_text"]
    # the processing line of the same run (20980 steps x 420 cells) through
    # processing_core: the files equal those the CPU oracle alone produced
    # (tests/golden/make_processing_fixture.py)
    import hashlib

    from fdtd2d_amd import processing as P

    pf = json.loads((GOLDEN / "processing_full.json").read_text())
    assert float(np.abs(line).max()) == pf["line_abs_max"]
    P.write_outputs(s.processing(), str(tmp_path), line)
    for name, text in pf["text"].items():
        assert (tmp_path / name).read_text().splitlines() == text, name
    for name, digest in pf["sha256"].items():
        assert hashlib.sha256((tmp_path / name).read_bytes()).hexdigest() == digest, name


PROC_FILES = ["SD.txt", "CF.txt", "ARC.txt", "original_field.txt", "filtered_field.txt", "det_SD.txt", "det_CF.txt",
              "det_ARC.txt", "detector_original_field.txt", "detector_filtered_field.txt"]


def test_cli_processing_line_matches_oracle(gpu_device, tmp_path):
    """The fdtd2d CLI on the reference's Settings.ini with the processing
    window moved to the first nanosecond (and windows and delays shortened to
    fit the 600 steps run): the TOF and field files it writes equal those the
    same post-processing writes from the oracle's captured line."""
    import subprocess

    from fdtd2d_amd import processing as P

    edits = {"BeginAfter": "0", "FinishAt": "1", "FilterWin": "50", "SDFWin": "30", "CFTau": "100",
             "ARCTau": "80", "DetWLoad": "0"}
    out, lines, in_proc = [], (GOLDEN / "settings_full.ini").read_text().splitlines(), False
    for ln in lines:
        st = ln.strip()
        if st.startswith("["):
            in_proc = st.lower() == "[processing]"
        key = st.split("=")[0].strip() if "=" in st else None
        out.append(f"{key} = {edits[key]}" if in_proc and key in edits else ln)
    ini = tmp_path / "Settings.ini"
    ini.write_text("\n".join(out) + "\n")
    g, r = tmp_path / "gpu", tmp_path / "oracle"
    g.mkdir()
    r.mkdir()
    exe = Path(F.LIB_PATH).parent / "fdtd2d"
    print(f"CLI {STEPS} steps with the processing window at (0, 2098]", flush=True)
    res = subprocess.run([str(exe), str(ini), "--out", str(g), "--steps", str(STEPS), "--device", str(gpu_device)],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    s = F.Setup.from_ini(str(ini))
    a = s.processing()
    assert (a.begin_after, a.finish_at) == (0, int(1e-9 / s.get("dt")))
    ref = Oracle(s, 32, threads=8)
    ref.run(STEPS)
    line = ref.read_line()[:STEPS].astype(np.float32)
    ref.close()
    P.write_outputs(a, str(r), line)
    for f in PROC_FILES:
        assert (g / f).read_bytes() == (r / f).read_bytes(), f
    sd = np.loadtxt(g / "CF.txt")
    assert sd.shape == (210, 3) and sd[0, 0] == 40


def _edited_ini(tmp_path, base, edits):
    """A shipped INI with [section] key = value replaced or added."""
    edits = {(sec.lower(), key): v for (sec, key), v in edits.items()}
    out, sec, seen = [], None, set()
    for ln in (GOLDEN / base).read_text().splitlines():
        st = ln.strip()
        if st.startswith("["):
            sec = st[1:st.index("]")].lower()
        key = st.split("=")[0].strip() if "=" in st and not st.startswith(";") else None
        if sec and key and (sec, key) in edits:
            ln = f"{key} = {edits[(sec, key)]}"
            seen.add((sec, key))
        out.append(ln)
    for (s_, key), v in edits.items():
        if (s_, key) not in seen:
            at = next(k for k, ln in enumerate(out) if ln.strip().lower() == f"[{s_}]")
            out.insert(at + 1, f"{key} = {v}")
    f = tmp_path / "Settings.ini"
    f.write_text("\n".join(out) + "\n")
    return f


def _cli(ini, out_dir, steps, device):
    import subprocess

    exe = Path(F.LIB_PATH).parent / "fdtd2d"
    res = subprocess.run([str(exe), str(ini), "--out", str(out_dir), "--steps", str(steps), "--device", str(device)],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


def _oracle_traces(setup, steps, snap_steps=(), snap=None):
    ref = Oracle(setup, 32, threads=8)
    done = 0
    for n in snap_steps:
        ref.run(n - done)
        done = n
        snap.add(ref.read_field(F.EZ).astype(np.float32))
    ref.run(steps - done)
    det, comp = ref.read_detectors()
    ref.close()
    return det.astype(np.float32), comp.astype(np.float32)


def test_cli_snapshots_and_chamber_figure_match_oracle(gpu_device, tmp_path):
    """The CLI with the tokamak figure (2Dpreset 7) and the plate wall drawn,
    saving Ez every 150 steps on a sieved grid: Ez.txt, its companion files
    and the detector files equal, byte for byte, those written from the
    oracle's run of the same setup."""
    ini = _edited_ini(tmp_path, "settings_release.ini", {
        ("KameraConf", "Draw2D"): 1, ("KameraConf", "2Dpreset"): 7, ("KameraConf", "DrawWall"): 1,
        ("KameraConf", "WallPos"): 2400, ("2Dsave", "From"): 0, ("2Dsave", "To"): STEPS,
        ("2Dsave", "TimeSieve"): 150, ("2Dsave", "GridSieve"): 2})
    g, r = tmp_path / "gpu", tmp_path / "oracle"
    g.mkdir()
    r.mkdir()
    print(f"CLI {STEPS} steps, figure 7 + wall, snapshots every 150 steps", flush=True)
    log = _cli(ini, g, STEPS, gpu_device)
    assert "4 Ez snapshots" in log
    s = F.Setup.from_ini(str(ini))
    assert (s.problem().material == 5).sum() > 0
    steps = s.snapshot_steps(STEPS).tolist()
    assert steps == [150, 300, 450, 600]
    w = F.Snapshots(s, str(r))
    det, comp = _oracle_traces(s, STEPS, steps, w)
    assert w.close() == 4
    s.set("SlicesSaved", 4)
    s.write_outputs(str(r), det, comp)
    for f in ["Ez.txt", "Ex.txt", "Ey.txt", "Hx.txt", "Hy.txt", "Hz.txt", "FieldSaveDat.txt"] + FILES + DEBUG_FILES:
        assert (g / f).read_bytes() == (r / f).read_bytes(), f
    assert "Slises saved: 4" in (g / "log.txt").read_text()
    assert (g / "deb.txt").stat().st_size > 0 and "5 " in (g / "deb.txt").read_text()
    # iniLOG.txt: the INI read's echo, once per run (time of the read aside)
    s.write_ini_log(str(r))
    strip = lambda t: t.split("\n", 1)[1]  # noqa: E731
    gl, rl = (g / "iniLOG.txt").read_text(), (r / "iniLOG.txt").read_text()
    assert gl.count("Config loaded from 'Settings.ini'") == 1 and strip(gl) == strip(rl)


def test_cli_source_to_detector_regime_matches_oracle(gpu_device, tmp_path):
    """[regimes] SourceToDetect = 1 with the chamber of preset 1: a free-space
    pass (no figure, no plasma) writes the *_S2D files, the main pass its
    files and DifDetEx.txt, the component traces of the two passes
    subtracted; all equal those written from the oracle's two runs."""
    ini = _edited_ini(tmp_path, "settings_release.ini", {
        ("regimes", "SourceToDetect"): 1, ("KameraConf", "Draw2D"): 1, ("KameraConf", "2Dpreset"): 1})
    g, r = tmp_path / "gpu", tmp_path / "oracle"
    g.mkdir()
    r.mkdir()
    print(f"CLI {STEPS} steps, source-to-detector regime, figure 1", flush=True)
    _cli(ini, g, STEPS, gpu_device)
    s0 = F.Setup.from_ini(str(ini), objects=False)
    s = F.Setup.from_ini(str(ini))
    assert np.all(s0.problem().material == 0) and (s.problem().material == 5).sum() > 0
    det0, comp0 = _oracle_traces(s0, STEPS)
    det, comp = _oracle_traces(s, STEPS)
    s0.write_s2d_outputs(str(r), det0, comp0)
    s.write_outputs(str(r), det, comp)
    s.write_s2d_difference(str(r), comp, comp0)
    names = ["Detect_S2D.txt", "detAll_S2D.txt", "detEz_S2D.txt", "detHx_S2D.txt", "detHy_S2D.txt", "detP_S2D.txt",
             "CompDetect_S2D.txt", "DifDetEx.txt"] + FILES + DEBUG_FILES
    for f in names:
        assert (g / f).read_bytes() == (r / f).read_bytes(), f
    dif = np.loadtxt(g / "DifDetEx.txt")
    assert np.abs(dif[:, 3]).max() > 0  # the chamber changes what the detectors see
