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
