"""Writes tests/golden/processing_full.json: the processing files of the
reference's own Settings.ini (FDTD_2D_Full) over all its 33568 steps, made by
the CPU oracle alone (its stepper captures the line, its restatement of
processing_core searches it) and formatted here as the reference prints them
(F/processing.cu:263-284). The TOF files are kept as text, the four field
files (about 10 MB each) as SHA-256 digests. About 6 minutes on 8 cores:

    python tests/golden/make_processing_fixture.py
"""

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import fdtd2d_amd as F  # noqa: E402
from oracle.oracle import Oracle, processing_tof  # noqa: E402

out = Path(__file__).resolve().parent


def tof_lines(v, sieve, start, dt):
    return ["%d %f %f" % (start + i, float(v[i]), (dt * float(v[i])) * 1e12) for i in range(0, len(v), sieve)]


def matrix_digest(m, t_sieve, d_sieve):
    h = hashlib.sha256()
    n, T = m.shape
    for t in range(0, T, t_sieve):
        h.update(("".join("%f\t" % x for x in m[0:n:d_sieve, t].astype(np.float64)) + "\n").encode())
    return h.hexdigest()


s = F.Setup.from_ini(str(out / "settings_full.ini"))
a = s.processing()
ref = Oracle(s, 32, threads=8)
ref.run(s.c_problem.contents.nsteps)
line = ref.read_line().astype(np.float32)
ref.close()
assert line.shape == (a.times, a.size_i)
r0, r1 = processing_tof(a, line, 0), processing_tof(a, line, 1)
half = a.detector_size // 2
fx = {
    "source": "oracle run of settings_full.ini, all steps; line rows x cells = %d x %d" % line.shape,
    "line_abs_max": float(np.abs(line).max()),
    "text": {
        "SD.txt": tof_lines(r0["sd"], a.dim_sieve, a.pos_i, a.dt),
        "CF.txt": tof_lines(r0["cf"], a.dim_sieve, a.pos_i, a.dt),
        "ARC.txt": tof_lines(r0["arc"], a.dim_sieve, a.pos_i, a.dt),
        "det_SD.txt": tof_lines(r1["sd"], a.dim_sieve, a.pos_i + half, a.dt),
        "det_CF.txt": tof_lines(r1["cf"], a.dim_sieve, a.pos_i + half, a.dt),
        "det_ARC.txt": tof_lines(r1["arc"], a.dim_sieve, a.pos_i + half, a.dt),
    },
    "sha256": {
        "original_field.txt": matrix_digest(r0["original"], a.time_sieve, a.dim_sieve),
        "filtered_field.txt": matrix_digest(r0["filtered"], a.time_sieve, a.dim_sieve),
        "detector_original_field.txt": matrix_digest(r1["original"], a.time_sieve, a.dim_sieve),
        "detector_filtered_field.txt": matrix_digest(r1["filtered"], a.time_sieve, a.dim_sieve),
    },
}
(out / "processing_full.json").write_text(json.dumps(fx, indent=0))
print("wrote", out / "processing_full.json")
