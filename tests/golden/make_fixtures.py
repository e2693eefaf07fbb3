"""Writes the fixtures under tests/golden/ from data files the reference holds.

Run in a container where the reference is mounted (it is not on GPU boxes):
    python tests/golden/make_fixtures.py /root/reference
Only data is copied: the 21 (steps, picoseconds) rows of
FDTD_2D_Full/TAU_Line.txt and the 21 rows of FDTD_2D_Full/RC_Line.txt (as
numbers and as the text lines), and the three shipped Settings.ini files
(configuration inputs).
"""

import json
import shutil
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
out = Path(__file__).resolve().parent


def rows(p):
    return [[float(x) for x in line.split()] for line in p.read_text().splitlines() if line.strip()]


def text(p):
    return [line for line in p.read_text().splitlines() if line.strip()]


fx = {
    "source": "FDTD_2D_Full/TAU_Line.txt, FDTD_2D_Full/RC_Line.txt",
    "tau_line": rows(ref / "FDTD_2D_Full" / "TAU_Line.txt"),
    "rc_line": rows(ref / "FDTD_2D_Full" / "RC_Line.txt"),
    "tau_line_text": text(ref / "FDTD_2D_Full" / "TAU_Line.txt"),
    "rc_line_text": text(ref / "FDTD_2D_Full" / "RC_Line.txt"),
}
(out / "reference_lines.json").write_text(json.dumps(fx, indent=1))
for src, dst in (("FDTD_2D_Full/Settings.ini", "settings_full.ini"),
                 ("x64/Release/Settings.ini", "settings_release.ini"),
                 ("build/FDTD_2D_FULL/Settings.ini", "settings_build.ini")):
    shutil.copyfile(ref / src, out / dst)
print("wrote", sorted(p.name for p in out.iterdir()))
