"""Host-side checks of the launch partition (no GPU needed).

Every launch splits the grid into per-wave tiles on two kernels: interior
tiles, whose loads must stay clear of the CPML strips and the outer ring
(the interior kernel carries no masks), and CPML-class tiles. The partition
must cover each cell exactly once, for every launch shape the plan accepts.
"""

import random

import numpy as np
import pytest

from fdtd2d_amd import _lib


def check_partition(nx, ny, P, T, C, band, band_cpml, events=()):
    ent, g = _lib.tile_lists(nx, ny, P, T, C, band, band_cpml, events=list(events)) if events else \
        _lib.tile_lists(nx, ny, P, T, C, band, band_cpml)
    vw, halo = g["vw"], g["halo"]
    Tn = T if T > 0 else 8
    assert halo >= Tn and vw == 64 * C - 2 * halo
    cover = np.zeros((ny, nx), np.int32)
    Pe = max(P, 1)
    cells = [0] * 6
    vw_e = 64 - 2 * Tn  # CPML-class waves: one column per lane, halo T
    for c0, c1, r0, r1, cls in ent:
        assert 0 <= r0 < r1 <= ny and 0 <= c0 < c1 <= nx
        jl0, jl1 = max(0, r0 - Tn), min(ny, r1 + Tn)  # rows the wave loads
        if cls == 0:
            assert c1 == min(nx, c0 + vw)
            cs, ce = c0 - halo, c0 + vw + halo  # columns the wave loads
            # the interior kernel carries no masks: its loads must stay clear
            assert cs > Pe and ce < nx - Pe - 1 and jl0 > Pe and jl1 < ny - Pe - 1, (c0, r0, r1)
        else:
            assert cls in (1, 2, 3, 4, 5)
            # one column per lane: 64 columns loaded from T left of the piece,
            # or from column 0; no halo is needed on a side at the grid's edge
            cs = max(0, c0 - Tn)
            ce = min(nx, cs + 64)
            assert c0 - cs >= Tn or cs == 0, (c0, c1)
            assert cs + 64 - c1 >= Tn or c1 == nx, (c0, c1)
            assert c1 - c0 <= vw_e or c0 == 0 or c1 == nx
            rows_clear = jl0 > Pe and jl1 < ny - Pe - 1
            cols_clear = cs > Pe and ce < nx - Pe - 1
            if cls in (2, 5):  # x-strip and event pieces: rows clear of the y strips and the ring
                assert rows_clear, (c0, r0, r1)
            if cls in (3, 5):  # y-strip and event pieces: columns clear of the x strips and the ring
                assert cols_clear, (c0, r0, r1)
            if cls in (2, 3, 4):  # strip and corner pieces hold no event cell
                for i, j in events:
                    assert not (cs <= i < ce and jl0 <= j < jl1), (i, j, c0, r0, cls)
        if cls == 0:
            for i, j in events:
                assert not (c0 - halo <= i < c0 + vw + halo and jl0 <= j < jl1), (i, j, c0, r0, r1)
        cover[r0:r1, c0:c1] += 1
        cells[cls] += (c1 - c0) * (r1 - r0)
    assert cover.min() == 1 and cover.max() == 1
    return ent, g, cells


@pytest.mark.parametrize("T,C", [(1, 1), (4, 1), (8, 1), (3, 2), (8, 2), (5, 4), (0, 1)])
def test_partition_400(T, C):
    check_partition(400, 400, 12, T, C, 0, 0)


def test_partition_random_shapes():
    rng = random.Random(7)
    for _ in range(60):
        nx, ny = rng.randint(20, 700), rng.randint(20, 700)
        P = rng.choice([0, 1, 4, 8, 11, 16])
        P = min(P, (min(nx, ny) - 2) // 2)
        C = rng.choice([1, 2, 4])
        T = rng.randint(1, 8)
        band = rng.choice([0, T, 17, 64, 256])
        band_cpml = rng.choice([0, T, 9, 32])
        check_partition(nx, ny, P, T, C, band, band_cpml)


def test_partition_8192_default():
    ent, g, cells = check_partition(8192, 8192, 16, 0, 1, 0, 0)
    assert g["band"] == 128 and g["band_cpml"] == 32
    # only the tiles next to the strips take the CPML-class kernel, and the
    # margins away from the y strips are x-strip pieces on whole bands
    assert sum(cells[1:]) / (8192 * 8192) < 0.03
    mid = [(c0, c1, r0, r1, c) for c0, c1, r0, r1, c in ent if c > 0 and 200 < r0 < 7900]
    assert {c for *_, c in mid} == {2}
    assert all(c1 <= 64 or c0 >= 8192 - 128 for c0, c1, *_ in mid)
    assert all(r1 - r0 == 128 for _, _, r0, r1, _ in mid)
    # interior tiles start at the first multiple of 16 columns clear of the
    # strip (32), so the left margin is one piece; the right margin here is
    # one piece too, with no halo beyond the grid (up to 56 columns)
    edge = sorted({(c0, c1) for c0, c1, *_ in mid if c0 == 0 or c1 == 8192})
    assert edge[0] == (0, 32) and edge[1][1] == 8192 and 16 + 8 < 8192 - edge[1][0] <= 56
    # the top and bottom bands' CPML-class pieces are y-strip pieces but for
    # the corners, which carry both strips
    top = [(c0, c1, c) for c0, c1, r0, r1, c in ent if c > 0 and r0 == 0]
    assert {c for c0, c1, c in top if 100 < c0 < 8000} == {3}
    assert {c for c0, c1, c in top if c0 == 0 or c1 == 8192} == {4}
    assert cells[1] == 0  # no general piece


def test_partition_8192_two_columns():
    # two columns per lane: the x-strips go to narrow one-column margins
    ent, g, cells = check_partition(8192, 8192, 16, 8, 2, 96, 32)
    assert sum(cells[1:]) / (8192 * 8192) < 0.03
    mid = sorted({(c0, c1, c) for c0, c1, r0, _, c in ent if c > 0 and 200 < r0 < 7900})
    assert all(c1 <= 96 or c0 >= 8192 - 200 for c0, c1, _ in mid)
    assert {c for *_, c in mid} == {2}
    # no general pieces: the corners carry both strips, nothing else
    assert cells[1] == 0 and 0 < cells[4] < 4 * 200 * 200
    # the default bands at two columns per lane
    assert _lib.tile_lists(8192, 8192, 16, 0, 2, 0, 0)[1]["band"] == 96
    assert _lib.tile_lists(16384, 16384, 16, 0, 2, 0, 0)[1]["band"] == 192


def test_event_cells_take_event_or_general_tiles():
    # sources, detectors and probes must lie in no interior, strip or corner
    # tile's window
    rng = random.Random(3)
    for _ in range(20):
        nx, ny = rng.randint(100, 900), rng.randint(100, 900)
        T = rng.randint(1, 8)
        C = rng.choice([1, 2, 4])
        ev = [(rng.randrange(nx), rng.randrange(ny)) for _ in range(rng.randint(1, 12))]
        _, _, cells = check_partition(nx, ny, 8, T, C, 0, 0, events=ev)
        assert cells[5] + cells[1] > 0


def test_rejects_bad_shapes():
    for args in [(8192, 8192, 16, 9, 1, 0, 0), (400, 400, 12, 4, 3, 0, 0), (10, 10, 5, 1, 1, 0, 0),
                 (400, 400, 12, 4, 1, -1, 0)]:
        with pytest.raises(_lib.FDTD2DError):
            _lib.tile_lists(*args)
