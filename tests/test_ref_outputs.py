"""Host-side output files of the reference's normal run that need no device:
the iniLOG.txt configuration echo (logINIread, F/FDTD_cuda.cu:178-303), the
deb.txt material-id map (F/kernel.cu:2737-2751) and the files the reference
opens and never writes (F/kernel.cu:1948-1993, 2685).

The echo is checked against a rendering made here from the reference's own
key list: every label in its order, each value of a key the INI holds printed
as the reference's variable's type prints through an ostream (int, bool as
0/1, float and double as %g with 6 significant digits, values read through
the INI parser the tests of test_ini.py pin), and the time/date line in the
_strtime_s / _strdate_s forms."""

from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

import fdtd2d_amd as F
from fdtd2d_amd import _lib as L

GOLDEN = Path(__file__).parent / "golden"

# (label, section, key, type) in the order logINIread prints them; type i int,
# b bool, f float, d double (the reference variable's type)
ECHO = [
    ("GPU_device", "hardware", "GPU_device", "i"), ("SourceToDetect", "regimes", "SourceToDetect", "i"),
    ("Use_Exp_time", "grid", "Use_Exp_time", "b"), ("Imax", "grid", "Imax", "i"), ("Jmax", "grid", "Jmax", "i"),
    ("Max_time", "grid", "Max_time", "i"), ("Exp_time", "grid", "Exp_time", "d"), ("ds", "grid", "ds", "d"),
    ("dt_coef", "grid", "dt_coef", "d"), ("Isource", "grid", "Isource", "i"), ("Jsource", "grid", "Jsource", "i"),
    ("Idetect", "grid", "Idetect", "i"), ("Jdetect", "grid", "Jdetect", "i"), ("SIsize", "AESA", "SIsize", "i"),
    ("SAutoDelta", "AESA", "SAutoDelta", "i"), ("SIdelta", "AESA", "SIdelta", "i"),
    ("SPAngle", "AESA", "SPAngle", "f"), ("DIsize", "AESA", "DIsize", "i"),
    ("DAutoDelta", "AESA", "SAutoDelta", "i"),  # the reference reads SAutoDelta for it
    ("DIdelta", "AESA", "DIdelta", "i"), ("Chirp", "PulseParam", "Chirp", "i"),
    ("FreqA", "PulseParam", "FreqA", "f"), ("FreqB", "PulseParam", "FreqB", "f"),
    ("Fronts", "PulseParam", "Fronts", "i"), ("Plato", "PulseParam", "Plato", "i"),
    ("Fronts_ns", "PulseParam", "Fronts_ns", "d"), ("Plato_ns", "PulseParam", "Plato_ns", "d"),
    ("Pulse_amp", "PulseParam", "Pulse_amp", "d"), ("From", "2Dsave", "From", "i"), ("To", "2Dsave", "To", "i"),
    ("TimeSieve", "2Dsave", "TimeSieve", "i"), ("GridSieve", "2Dsave", "GridSieve", "i"),
    ("Build", "medium", "Build", "i"), ("Icen", "medium", "Icen", "i"), ("Jcen", "medium", "Jcen", "i"),
    ("A", "medium", "A", "f"), ("B", "medium", "B", "f"), ("C", "medium", "C", "f"), ("P", "medium", "P", "i"),
    ("Q", "medium", "Q", "i"), ("N", "medium", "N", "i"), ("Radius", "medium", "Radius", "i"),
    ("Alpha", "medium", "Alpha", "f"), ("Beta", "medium", "Beta", "f"), ("ne0", "medium", "ne0", "f"),
    ("neA", "medium", "neA", "f"), ("FlucBuild", "mediumFluc", "FlucBuild", "i"),
    ("FlucIcen", "mediumFluc", "FlucIcen", "i"), ("FlucJcen", "mediumFluc", "FlucJcen", "i"),
    ("FlucA", "mediumFluc", "FlucA", "f"), ("FlucB", "mediumFluc", "FlucB", "f"),
    ("FlucC", "mediumFluc", "FlucC", "f"), ("FlucP", "mediumFluc", "FlucP", "i"),
    ("FlucQ", "mediumFluc", "FlucQ", "i"), ("FlucN", "mediumFluc", "FlucN", "i"),
    ("FlucAngleI", "mediumFluc", "FlucAngleI", "f"), ("FlucAngleJ", "mediumFluc", "FlucAngleJ", "f"),
    ("FlucAngleK", "mediumFluc", "FlucAngleK", "f"), ("FlucSign", "mediumFluc", "FlucSign", "i"),
    ("FlucRadius", "mediumFluc", "FlucRadius", "i"), ("FlucAlpha", "mediumFluc", "FlucAlpha", "f"),
    ("FlucBeta", "mediumFluc", "FlucBeta", "f"), ("FlucNe0", "mediumFluc", "FlucNe0", "f"),
    ("TorusBR", "medium", "TorusBR", "i"), ("Draw2D", "KameraConf", "Draw2D", "i"),
    ("2Dpreset", "KameraConf", "2Dpreset", "i"), ("PatrubokL", "KameraConf", "PatrubokL", "i"),
    ("DrawWall", "KameraConf", "DrawWall", "i"), ("WallPos", "KameraConf", "WallPos", "i"),
    ("Iline", "RC_calc", "Iline", "i"), ("LIsize", "RC_calc", "LIsize", "i"), ("LIdelta", "RC_calc", "LIdelta", "i"),
    ("FilterWindow", "RC_calc", "FilterWindow", "i"), ("Rlevel", "RC_calc", "Rlevel", "f"),
    ("Perform", "Processing", "Perform", "B"), ("SecDer", "Processing", "SecDer", "B"),
    ("SDFType", "Processing", "SDFType", "i"), ("SDFWin", "Processing", "SDFWin", "i"),
    ("SDFRuns", "Processing", "SDFRuns", "i"), ("CF", "Processing", "CF", "B"), ("CFTau", "Processing", "CFTau", "i"),
    ("CFK", "Processing", "CFK", "f"), ("ARC", "Processing", "ARC", "B"), ("ARCTau", "Processing", "ARCTau", "i"),
    ("ARCK", "Processing", "ARCK", "f"), ("TauThreshold", "Processing", "TauThreshold", "f"),
    ("FilterType", "Processing", "FilterType", "i"), ("FilterWin", "Processing", "FilterWin", "i"),
    ("FilterRuns", "Processing", "FilterRuns", "i"), ("SaveFiltered", "Processing", "SaveFiltered", "B"),
    ("SaveOriginal", "Processing", "SaveOriginal", "B"), ("PosI", "Processing", "PosI", "i"),
    ("PosJ", "Processing", "PosJ", "i"), ("PosK", "Processing", "PosK", "i"), ("SizeI", "Processing", "SizeI", "i"),
    ("BeginAfter", "Processing", "BeginAfter", "i"), ("FinishAt", "Processing", "FinishAt", "i"),
    ("TimeSieve", "Processing", "TimeSieve", "i"), ("DimSieve", "Processing", "DimSieve", "i"),
    ("SimDetect", "Processing", "DetSim", "B"), ("DetSave", "Processing", "DetSave", "B"),
    ("DetSize", "Processing", "DetSize", "i"), ("DetWLoad", "Processing", "DetWLoad", "B"),
]
# text logINIread prints before "\n <label>=" of these keys (headings, blank lines)
BEFORE = {"SourceToDetect": "\nRegimes: ", "Use_Exp_time": "\nGrid param: ", "SIsize": "\nAESA",
          "Chirp": "\nPulseParam:", "From": "\n2Dsave:", "Build": "\nmedium:", "A": "\n", "Radius": "\n",
          "FlucBuild": "\nmediumFluc:", "FlucA": "\n", "FlucAngleI": "\n", "FlucRadius": "\n",
          "Draw2D": "\nKameraConf:", "Iline": "\n\nRC_calc:", "Perform": "\n\nProcessing:"}


def _fmt(v: float, t: str) -> str:
    if t == "i":
        return str(int(v))
    if t in "bB":
        return "1" if v != 0 else "0"
    x = float(np.float32(v)) if t == "f" else float(v)
    return "%g" % x


def _ini_value(ini_text: str, sec: str, key: str, t: str):
    """The key's value as the INI reader gives it, or None when absent (two
    lookups with different defaults agree only when the key is there)."""
    kind = 2 if t == "B" else 0 if t in "ib" else 1
    defaults = (0, 1) if kind == 2 else (-12345, -54321)
    v0, _ = L.ini_lookup(ini_text, sec, key, kind, defaults[0])
    v1, _ = L.ini_lookup(ini_text, sec, key, kind, defaults[1])
    return v0 if v0 == v1 else None


@pytest.mark.parametrize("name", ["settings_full.ini", "settings_release.ini", "settings_build.ini"])
def test_ini_log_echo(tmp_path, name):
    ini = GOLDEN / name
    ini_text = ini.read_text(encoding="latin-1")
    s = F.Setup.from_ini(str(ini))
    s.write_ini_log(str(tmp_path))
    text = (tmp_path / "iniLOG.txt").read_text()
    head, _, rest = text.partition("\n")
    m = re.fullmatch(r"Config loaded from 'Settings\.ini' on time (\d\d:\d\d:\d\d) date (\d\d/\d\d/\d\d):", head)
    assert m, head
    # the file's own values, for the keys the INI leaves to compile-time defaults
    got = re.findall(r"\n (\w+)=([^\n]*)", "\n" + rest)
    assert [k for k, _ in got] == [e[0] for e in ECHO]
    want = head + "\nHardware:"
    present = 0
    for (label, sec, key, t), (_, own) in zip(ECHO, got):
        v = _ini_value(ini_text, sec, key, t)
        present += v is not None
        want += BEFORE.get(label, "") + "\n " + label + "=" + (own if v is None else _fmt(v, t))
    want += "\n" + "-" * 77 + "\n"
    assert present > 40
    assert text == want
    # appended, as the reference's std::ios::app
    s.write_ini_log(str(tmp_path))
    assert (tmp_path / "iniLOG.txt").read_text() == text * 2


def test_ini_log_absent_when_the_ini_cannot_be_opened(tmp_path):
    s = F.Setup.from_ini(str(tmp_path / "missing.ini"))
    assert s.ini_status == -1
    s.write_ini_log(str(tmp_path))
    assert not (tmp_path / "iniLOG.txt").exists()


def test_deb_map_and_empty_files(tmp_path):
    s = F.Setup.from_ini(str(GOLDEN / "settings_release.ini"))
    p = s.problem()
    det = np.zeros((p.nsteps + 1, len(p.det_i)), np.float32)
    comp = np.zeros((3, p.nsteps + 1, len(p.comp_i)), np.float32)
    s.write_outputs(str(tmp_path), det, comp)
    mat = p.material  # [Jmax, Imax], ID3[j*Imax + i]
    assert (mat > 0).sum() > 0
    rows = ["".join(f"{v} " for v in mat[1:p.ny - 1, i]) + "\n" for i in range(1, p.nx - 1)]
    assert (tmp_path / "deb.txt").read_text() == "".join(rows) + "\n\n\n\n"
    for name in ("test.txt", "SDetect.txt", "LineDetect.txt", "LineDetect2.txt", "detEx.txt", "detEy.txt",
                 "detHz.txt", "DifDetEx.txt", "CompDetect.txt", "materials.txt"):
        assert (tmp_path / name).exists() and (tmp_path / name).stat().st_size == 0, name


def test_no_deb_map_without_material_ids(tmp_path):
    s = F.Setup.workload(1, seed=1, steps=4)
    p = s.problem()
    s.write_outputs(str(tmp_path), np.zeros((5, len(p.det_i)), np.float32),
                    np.zeros((3, 5, len(p.comp_i)), np.float32))
    assert not (tmp_path / "deb.txt").exists()
    assert (tmp_path / "materials.txt").stat().st_size == 0
