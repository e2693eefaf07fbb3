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


def check_partition(nx, ny, P, T, C, band, band_cpml):
    ent, g = _lib.tile_lists(nx, ny, P, T, C, band, band_cpml)
    vw, halo = g["vw"], g["halo"]
    Tn = T if T > 0 else 8
    assert halo >= Tn and vw == 64 * C - 2 * halo
    cover = np.zeros((ny, nx), np.int32)
    Pe = max(P, 1)
    cells = [0, 0]
    vw_e = 64 - 2 * Tn  # CPML-class waves: one column per lane, halo T
    for c0, c1, r0, r1, cls in ent:
        assert 0 <= r0 < r1 <= ny and 0 <= c0 < c1 <= nx
        if cls == 0:
            assert c1 == min(nx, c0 + vw)
        else:
            assert c1 - c0 <= vw_e
        cover[r0:r1, c0:c1] += 1
        cells[cls] += (c1 - c0) * (r1 - r0)
        h, w = (halo, vw) if cls == 0 else (Tn, vw_e)
        cs, ce = c0 - h, c0 + w + h  # columns the wave loads
        jl0, jl1 = max(0, r0 - Tn), min(ny, r1 + Tn)  # rows the wave loads
        clear = cs > Pe and ce < nx - Pe - 1 and jl0 > Pe and jl1 < ny - Pe - 1
        # the interior kernel carries no masks: its loads must stay clear
        if cls == 0:
            assert clear, (c0, r0, r1)
        # with one column per lane the CPML-class pieces are whole tiles of
        # the interior geometry, which are CPML class only when they touch
        if C == 1 and cls == 1:
            assert not clear, (c0, r0, r1)
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
    # only the tiles next to the strips take the CPML-class kernel
    assert cells[1] / (8192 * 8192) < 0.03
    ntx = -(-8192 // g["vw"])
    mid = {c0 // g["vw"] for c0, _, r0, _, c in ent if c == 1 and 200 < r0 < 7900}
    assert mid == {0, ntx - 1}


def test_partition_8192_two_columns():
    # two columns per lane: the x-strips go to narrow one-column margins, so
    # the CPML-class share stays near the one-column tiling's
    ent, g, cells = check_partition(8192, 8192, 16, 8, 2, 96, 32)
    assert cells[1] / (8192 * 8192) < 0.03
    mid = sorted({(c0, c1) for c0, c1, r0, _, c in ent if c == 1 and 200 < r0 < 7900})
    assert all(c1 <= 96 or c0 >= 8192 - 200 for c0, c1 in mid)
    # the default bands at two columns per lane
    assert _lib.tile_lists(8192, 8192, 16, 0, 2, 0, 0)[1]["band"] == 96
    assert _lib.tile_lists(16384, 16384, 16, 0, 2, 0, 0)[1]["band"] == 192


def test_event_cells_take_cpml_class():
    # sources, detectors and probes must lie in no interior tile's window
    rng = random.Random(3)
    for _ in range(20):
        nx, ny = rng.randint(100, 900), rng.randint(100, 900)
        T = rng.randint(1, 8)
        ev = [(rng.randrange(nx), rng.randrange(ny)) for _ in range(rng.randint(1, 12))]
        ent, g = _lib.tile_lists(nx, ny, 8, T, 1, 0, 0, events=ev)
        vw, halo = g["vw"], g["halo"]
        cover = np.zeros((ny, nx), np.int32)
        for c0, c1, r0, r1, cls in ent:
            cover[r0:r1, c0:c1] += 1
            if cls == 0:
                for i, j in ev:
                    assert not (c0 - halo <= i < c0 + vw + halo and r0 - T <= j < r1 + T), (i, j, c0, r0, r1)
        assert cover.min() == 1 and cover.max() == 1


def test_rejects_bad_shapes():
    for args in [(8192, 8192, 16, 9, 1, 0, 0), (400, 400, 12, 4, 3, 0, 0), (10, 10, 5, 1, 1, 0, 0),
                 (400, 400, 12, 4, 1, -1, 0)]:
        with pytest.raises(_lib.FDTD2DError):
            _lib.tile_lists(*args)
