"""The host-driver path on the reference's own Settings.ini files, on the GPU.

Each shipped INI (medium builders: sphere, torus, fluctuation) is set up
exactly as the fdtd2d CLI does (fdtd2d_setup_from_ini), stepped on the GPU for
a bounded number of steps and compared with the FP32 oracle: traces bitwise,
and the output files the reference's scripts read (Detect.txt, detAll.txt,
detEz/Hx/Hy.txt, detP.txt, SourceP.txt) byte for byte.
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
    s.write_rc_outputs(str(tmp_path), spd)
    rc = (tmp_path / "RC_Line.txt").read_text().splitlines()
    tau = (tmp_path / "TAU_Line.txt").read_text().splitlines()
    assert rc == fx["rc_line_text"]
    assert tau == fx["tau_line_text"]


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
