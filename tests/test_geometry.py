"""Metal geometry of [KameraConf] and the objects flag, on the CPU.

The rasterisers (fdtd2d_amd/csrc/geometry.cpp) are pinned against the shapes
the reference's primitives draw, written here as rectangle and disk algebra
on the node grid: a Yee cell (I, J) marks the nodes I..I+1 x J..J+1
(yeeCube, F/FDTD_cuda.cu:2771-2790), a width-line is the line and a copy one
row off (draw_width_line, :4464-4489), a ring is the disk of radius R minus
the disk of radius R - width (draw_Circl, :4525-4546).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import fdtd2d_amd as F

GOLD = Path(__file__).resolve().parent / "golden"


def _ini(tmp_path, edits, base="settings_build.ini"):
    """The shipped INI with keys replaced (or added) per section."""
    lines = (GOLD / base).read_text().splitlines()
    edits = {(sec.lower(), key): v for (sec, key), v in edits.items()}
    out, sec, seen = [], None, set()
    for ln in lines:
        st = ln.strip()
        if st.startswith("["):
            sec = st[1:st.index("]")].lower()
        key = st.split("=")[0].strip() if "=" in st and not st.startswith(";") else None
        if sec and key and (sec, key) in edits:
            ln = f"{key} = {edits[(sec, key)]}"
            seen.add((sec, key))
        out.append(ln)
    for (sec, key), v in edits.items():
        if (sec, key) not in seen:
            at = next(k for k, ln in enumerate(out) if ln.strip().lower() == f"[{sec}]")
            out.insert(at + 1, f"{key} = {v}")
    f = tmp_path / "Settings.ini"
    f.write_text("\n".join(out) + "\n")
    return str(f)


def _material(tmp_path, **kw):
    edits = {("medium", "Build"): 0, ("mediumFluc", "FlucBuild"): 0}
    edits.update({tuple(k.split("__")): v for k, v in kw.items()})
    s = F.Setup.from_ini(_ini(tmp_path, edits))
    return s.problem().material


def test_plate_wall(tmp_path):
    m = _material(tmp_path, KameraConf__DrawWall=1, KameraConf__WallPos=650)
    ny, nx = m.shape
    want = np.zeros_like(m)
    want[650:ny, 1:nx] = 5  # cells i in [1, nx-2], j in [650, ny-2] mark nodes up to nx-1, ny-1
    assert np.array_equal(m, want)
    # the wall is drawn after the medium, over it
    s = F.Setup.from_ini(_ini(tmp_path, {("KameraConf", "DrawWall"): 1, ("KameraConf", "WallPos"): 1300}))
    m2 = s.problem().material
    assert np.all(m2[1300:, 1:] == 5)
    base = F.Setup.from_ini(_ini(tmp_path, {})).problem().material
    assert np.array_equal(m2[:1300], base[:1300]) and np.any(base[1300:, 1:] != 5)


def test_chamber_preset_1_rectangles(tmp_path):
    m = _material(tmp_path, KameraConf__Draw2D=1, **{"KameraConf__2Dpreset": 1})
    ny, nx = m.shape
    want = np.zeros_like(m)
    want[179:182, 20:nx - 19] = 5                 # top: rows 180 and 179, cells 20..nx-21
    want[179:ny - 18, nx - 20:nx - 18] = 5        # right: column nx-20, rows shifted -1 and +1
    want[ny - 20:ny - 17, 21:nx - 18] = 5         # bottom, drawn right to left: rows +0, +1
    want[180:ny - 17, 20:22] = 5                  # left, drawn upwards
    want[179:182, 275:326] = 0                    # antenna hole at I_SOURCE = 300 +- 25
    assert np.array_equal(m, want)


def test_chamber_preset_7_rings_and_port(tmp_path):
    pl = 5
    m = _material(tmp_path, KameraConf__Draw2D=1, **{"KameraConf__2Dpreset": 7, "KameraConf__PatrubokL": pl})
    ny, nx = m.shape
    a, b = nx // 2, ny // 2 + pl

    def disk(R):
        x, y = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1))
        cells = (x - a) ** 2 + (y - b) ** 2 <= R * R
        nodes = np.zeros((ny, nx), bool)
        for dj in (0, 1):
            for di in (0, 1):
                nodes[dj:dj + ny - 1, di:di + nx - 1] |= cells
        return nodes

    want = np.zeros((ny, nx), np.int32)
    for R, w in ((270, 3), (60, 4)):
        want[disk(R)] = 5
        want[disk(R - w)] = 0
    # Draw_Square cells strictly inside (x1, x2) x (y1, y2) are cleared
    x1, x2, y1, y2 = a - 50, a + 50, 11, ny // 2 + pl + 15 - 270
    want[y1 + 1:y2 - 1 + 2, x1 + 1:x2 - 1 + 2] = 0
    top = ny // 2 + pl + 4 - 270
    # the port's three walls are below the square's rows only when the grid
    # puts the ring's top inside it; compare the drawn rings away from them
    away = np.ones((ny, nx), bool)
    away[:max(top, y2) + 3, x1 - 2:x2 + 3] = False
    assert np.array_equal(m[away], want[away])
    # the centre row: cells within R mark nodes up to a + R + 1, cells within
    # R - width are cleared up to a + R - width + 1
    row = m[b]
    assert row[a] == 0
    assert row[a + 268] == 0 and (row[a + 269:a + 272] == 5).all() and row[a + 272] == 0
    assert row[a + 57] == 0 and (row[a + 58:a + 62] == 5).all() and row[a + 62] == 0
    # the port: side walls at a -+ 50 from row ny/2 - PL - 270 up to the ring,
    # its floor on that row, the antenna opening (source centre +- 20) cut
    # through the floor afterwards
    bot = ny // 2 - pl - 270
    assert m[bot + 5, a - 50] == 5 and m[bot + 5, a + 50] == 5 and m[bot + 5, a] == 0
    assert m[bot, a - 45] == 5 and m[bot, a + 45] == 5
    src_i = int(F.Setup.from_ini(_ini(tmp_path, {})).get("Isource"))
    assert m[bot, src_i] == 0 and m[bot, src_i - 19] == 0 and m[bot, src_i - 21] == 5


def test_medium_overwrites_figure_and_free_space_pass(tmp_path):
    edits = {("KameraConf", "Draw2D"): 1, ("KameraConf", "2Dpreset"): 1}
    path = _ini(tmp_path, edits)
    full = F.Setup.from_ini(path).problem().material
    fig = F.Setup.from_ini(_ini(tmp_path, {**edits, ("medium", "Build"): 0, ("mediumFluc", "FlucBuild"): 0}))
    fig = fig.problem().material
    med = F.Setup.from_ini(_ini(tmp_path, {})).problem().material
    # the medium is written after the figure: where it wrote, its ids win
    plasma = med >= 8
    assert np.array_equal(full[plasma], med[plasma])
    assert np.array_equal(full[~plasma], fig[~plasma])
    # objects off: neither the figure nor the medium
    free = F.Setup.from_ini(path, objects=False)
    assert np.all(free.problem().material == 0)
    assert free.get("objects") == 0
    # everything else of the problem is the same
    a, b = free.problem(), F.Setup.from_ini(path).problem()
    assert np.array_equal(a.src_table, b.src_table) and np.array_equal(a.det_i, b.det_i)
    assert (a.nx, a.ny, a.nsteps, a.dt) == (b.nx, b.ny, b.nsteps, b.dt)


def test_wall_not_built_in_source_to_detector_regime(tmp_path):
    path = _ini(tmp_path, {("KameraConf", "DrawWall"): 1, ("regimes", "SourceToDetect"): 1,
                           ("medium", "Build"): 0, ("mediumFluc", "FlucBuild"): 0})
    assert np.all(F.Setup.from_ini(path).problem().material == 0)


@pytest.mark.parametrize("frm,to,sieve,n,want", [
    (0, 40000, 240, 1000, [240, 480, 720, 960]),       # counter starts at 1: first save at From' + sieve - 1
    (13000, 20000, 4000, 33568, [16999]),             # the shipped build INI
    (0, -2500000, 2000, 33568, []),                   # the shipped full INI: To < From
    (5, 12, 1, 20, list(range(5, 13))),               # every step of [From, To]
    (3, 100, 0, 50, []),                              # the counter never equals 0
])
def test_snapshot_cadence(tmp_path, frm, to, sieve, n, want):
    s = F.Setup.from_ini(_ini(tmp_path, {("2Dsave", "From"): frm, ("2Dsave", "To"): to,
                                         ("2Dsave", "TimeSieve"): sieve}))
    assert s.snapshot_steps(n).tolist() == want


def test_snapshot_files(tmp_path):
    s = F.Setup.from_ini(_ini(tmp_path, {("2Dsave", "GridSieve"): 3}))
    p = s.problem()
    rng = np.random.default_rng(1)
    d = tmp_path / "out"
    d.mkdir()
    w = F.Snapshots(s, str(d))
    shots = [rng.standard_normal((p.ny, p.nx)).astype(np.float32) for _ in range(2)]
    for ez in shots:
        w.add(ez)
    assert w.close() == 2
    lines = (d / "Ez.txt").read_text().split("\n")
    rows = list(range(1, p.nx - 1, 3))
    assert len(lines) == 2 * len(rows) + 1 and lines[-1] == ""
    for k, ez in enumerate(shots):
        for r in (0, 7, len(rows) - 1):
            i = rows[r]
            assert lines[k * len(rows) + r] == "".join("%12f " % v for v in ez[1:p.ny - 1:3, i].astype(np.float64))
    for name in ("Ex.txt", "Hx.txt", "Hy.txt", "Hz.txt"):
        assert (d / name).read_text() == "\n" * (2 * len(rows))
    assert (d / "Ey.txt").read_text() == ""
    assert (d / "FieldSaveDat.txt").read_text() == f"{len(rows)} 2"
