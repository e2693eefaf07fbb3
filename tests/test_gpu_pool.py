"""Pools of independent shots (the batched and served stepper), on the GPU.

The bar is the single-grid one: every shot's fields and traces equal to the
FP32 oracle of the same problem bit for bit, whatever else runs in the pool,
with graphs or without, across reshapes, admissions and removals.
"""

from __future__ import annotations

import numpy as np
import pytest

import fdtd2d_amd as F
from oracle.oracle import Oracle

from helpers import fp64_check, random_problem

pytestmark = pytest.mark.gpu


def _pool_state(pool, sid):
    out = {n: np.asarray(pool.read_field(sid, w), np.float64)
           for n, w in (("ez", F.EZ), ("hx", F.HX), ("hy", F.HY), ("jz", F.JZ))}
    det, comp = pool.read_detectors(sid)
    out["det"], out["comp"] = np.asarray(det, np.float64), np.asarray(comp, np.float64)
    return out


def _oracle_state(p):
    ref = Oracle(p, 32, threads=4)
    ref.run(p.nsteps)
    out = {n: ref.read_field(w) for n, w in (("ez", F.EZ), ("hx", F.HX), ("hy", F.HY), ("jz", F.JZ))}
    det, comp = ref.read_detectors()
    out["det"], out["comp"] = det, comp
    ref.close()
    return out


def _same(g, r, label):
    for k in r:
        a, b = g[k], np.asarray(r[k]).astype(np.float32).astype(np.float64)
        assert a.shape == b.shape, (label, k)
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{label}: {k} differs at {len(bad)} entries, first {bad[:4].tolist()}")


def _drain(pool, nchunk=4):
    while pool.info("running") > 0:
        pool.step(nchunk)
    pool.sync()


def _problems(seed, n, table=False, sizes=(80, 330), nsteps=None, steps=(9, 45)):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx, ny = (int(rng.integers(*sizes)) for _ in range(2))
        out.append(random_problem(rng, nx=nx, ny=ny, npml=int(rng.choice([0, 5, 8, 12])),
                                  nsteps=int(rng.integers(*steps)) if nsteps is None else nsteps, table=table))
    return out


@pytest.mark.parametrize("table,chunk,tblock,fused,depth,cols", [
    (False, 4, 8, 1, 2, 1), (True, 4, 8, 0, 1, 1), (False, 0, 5, 0, 2, 1), (True, 2, 3, 1, 1, 1),
    (True, 4, 8, 1, 2, 1), (False, 4, 8, 1, 2, 2), (True, 2, 8, 0, 1, 2), (False, 0, 6, 0, 2, 2),
    (True, 4, 7, 1, 2, 2)])
def test_pool_shots_match_oracle(gpu_device, table, chunk, tblock, fused, depth, cols):
    """Uneven shots (sizes, CPML, steps not multiples of T) in one pool."""
    probs = _problems(900 + 10 * tblock + chunk + 1000 * (cols - 1), 7, table=table, steps=(45, 90))
    with F.Pool(gpu_device, tblock=tblock, graph_chunk=chunk, fused=fused, prefetch=depth, band=0,
                band_cpml=0, cols_per_lane=cols) as pool:
        ids = [pool.add(p) for p in probs]
        print(f"table={table} chunk={chunk} T={tblock} fused={fused} D={depth} C={cols}: shots "
              f"{[(p.nx, p.ny, p.npml, p.nsteps) for p in probs]}", flush=True)
        _drain(pool)
        print({k: pool.info(k) for k in ("launches", "graph_replays", "eager_launches", "tail_launches")})
        for sid, p in zip(ids, probs):
            g = _pool_state(pool, sid)
            _same(g, _oracle_state(p), f"shot {sid} {p.nx}x{p.ny}")
            fp64_check(p, g["det"], f"pooled shot {sid} {p.nx}x{p.ny}", comp32=g["comp"])
        if chunk:
            assert pool.info("graph_replays") > 0


def test_pool_reshape_midrun_and_late_admission(gpu_device):
    """tblock and columns per lane changed between launches, and shots
    admitted while others run."""
    probs = _problems(4242, 6, table=False, nsteps=40)
    with F.Pool(gpu_device, tblock=4, graph_chunk=2) as pool:
        ids = [pool.add(p) for p in probs[:3]]
        pool.step(2)
        pool.set_option("tblock", 7)
        pool.set_option("cols_per_lane", 2)
        ids += [pool.add(p) for p in probs[3:5]]
        pool.step(3)
        pool.set_option("tblock", 2)
        ids.append(pool.add(probs[5]))
        _drain(pool, 3)
        for sid, p in zip(ids, probs):
            g = _pool_state(pool, sid)
            _same(g, _oracle_state(p), f"shot {sid}")
            fp64_check(p, g["det"], f"reshaped pool shot {sid}", comp32=g["comp"])


def test_pool_time_slicing(gpu_device):
    """Shots of at least slice_cells cells run only in every slice_every-th
    launch (their descriptors count turns, tails included) while the small
    shots run in every launch; the two alternating list sets come from the
    ring of list copies without rebuilds. Shots are admitted and removed
    (slots reused) while slicing, the period changes mid-run and slicing is
    switched off (the small shots' 32-row bands and CPML sub-bands follow,
    retiled); every shot stays equal to the oracle."""
    small = _problems(6161, 6, table=False, sizes=(80, 140), steps=(40, 75))
    large = _problems(6162, 3, table=False, sizes=(300, 331), steps=(45, 70))
    quick = _problems(6163, 1, table=False, sizes=(80, 140), nsteps=21)[0]  # done in 3 launches
    probs = [quick, large[0], small[1], small[2], large[1]]
    with F.Pool(gpu_device, tblock=8, graph_chunk=4, slice_cells=250 * 250, slice_every=3, band_small=32,
                band_cpml_small=32) as pool:
        ids = [pool.add(p) for p in probs]
        pool.step(5)
        print({k: pool.info(k) for k in ("launches", "rebuilds", "list_switches", "graph_replays")}, flush=True)
        assert pool.info("list_switches") > 0 and pool.info("graph_replays") == 0
        # a finished small shot leaves (its slot comes back), two join
        pool.sync()
        done = [sid for sid, p in zip(ids, probs) if pool.state(sid)[2]]
        assert ids[0] in done
        results = {}
        for sid in done:
            results[sid] = (_pool_state(pool, sid), probs[ids.index(sid)])
            pool.remove(sid)
        for p in (small[3], large[2]):
            probs.append(p)
            ids.append(pool.add(p))
        pool.step(4)
        pool.set_option("slice_every", 2)
        ids.append(pool.add(small[4]))
        probs.append(small[4])
        pool.step(5)
        pool.set_option("slice_cells", 0)
        ids.append(pool.add(small[5]))
        probs.append(small[5])
        _drain(pool, 3)
        print({k: pool.info(k) for k in ("launches", "rebuilds", "list_switches", "tail_launches")}, flush=True)
        for sid, p in zip(ids, probs):
            if sid in results and results[sid][1] is p:
                g = results[sid][0]
            else:
                g = _pool_state(pool, sid)
            _same(g, _oracle_state(p), f"sliced pool shot {sid} {p.nx}x{p.ny} steps {p.nsteps}")
            fp64_check(p, g["det"], f"sliced pool shot {sid}", comp32=g["comp"])


def test_pool_time_slicing_large_shots_alone(gpu_device):
    """Once only sliced shots are left, launches between their turns are
    skipped (the launch counter moves on to the next turn, on the host and
    the device) instead of stalling the pool."""
    small = _problems(7171, 1, table=False, sizes=(80, 100), nsteps=8)[0]  # one launch
    large = _problems(7172, 2, table=False, sizes=(300, 331), steps=(40, 60))
    with F.Pool(gpu_device, tblock=8, graph_chunk=0, slice_cells=250 * 250, slice_every=5) as pool:
        ids = [pool.add(p) for p in [small] + large]
        pool.step(2)  # launch 0: every shot; launch 1: the small one has finished, the large ones wait
        pool.sync()
        assert pool.state(ids[0])[2]
        g0 = pool.info("launches")
        pool.step(3)
        assert pool.info("launches") > g0 + 3  # the skipped launches count
        _drain(pool, 3)
        print({k: pool.info(k) for k in ("launches", "eager_launches", "tail_launches")}, flush=True)
        for sid, p in zip(ids, [small] + large):
            g = _pool_state(pool, sid)
            _same(g, _oracle_state(p), f"large-only sliced shot {sid} {p.nx}x{p.ny}")


def test_pool_fused_toggled_midrun_table_medium(gpu_device):
    """Switching between one kernel per tile class and the fused kernel (whose
    table-medium form leaves corner and general pieces to a second kernel)
    between launches re-routes the CPML-class entries; results stay equal to
    the oracle."""
    probs = _problems(5151, 5, table=True, nsteps=56)
    with F.Pool(gpu_device, tblock=8, graph_chunk=2, fused=0) as pool:
        ids = [pool.add(p) for p in probs]
        pool.step(2)
        pool.set_option("fused", 1)
        pool.step(2)
        pool.set_option("fused", 0)
        pool.step(1)
        pool.set_option("fused", 1)
        _drain(pool, 2)
        for sid, p in zip(ids, probs):
            g = _pool_state(pool, sid)
            _same(g, _oracle_state(p), f"shot {sid}")
            fp64_check(p, g["det"], f"fused-toggled pool shot {sid}", comp32=g["comp"])


def test_pool_budget_admission_and_slot_reuse(gpu_device):
    """Byte-accurate admission against an arena; removed shots free their
    bytes and slots for new ones, and results do not depend on neighbours."""
    probs = _problems(77, 8, table=True, sizes=(120, 260))
    sizes = [p.footprint() for p in probs]
    budget = sum(sizes[:3]) + 4096
    with F.Pool(gpu_device, budget_bytes=budget, graph_chunk=0) as pool:
        ids = [pool.add(p) for p in probs[:3]]
        with pytest.raises(F.FDTD2DError) as ei:
            pool.add(probs[3])
        assert ei.value.code == 2  # FDTD2D_ENOMEM: over budget
        assert pool.info("bytes_in_use") >= sum(sizes[:3])
        _drain(pool)
        got = {sid: _pool_state(pool, sid) for sid in ids}
        pool.remove(ids[1])
        # a new shot of exactly the freed footprint fits only if the accounting is exact
        r = probs[1]
        new = random_problem(np.random.default_rng(78), nx=r.nx, ny=r.ny, npml=r.npml, nsteps=r.nsteps, table=True,
                             nsrc=len(r.src_i), ndet=len(r.det_i), ncomp=len(r.comp_i))
        assert new.footprint() == sizes[1]
        sid3 = pool.add(new)
        assert sid3 == ids[1]
        _drain(pool)
        got_new = _pool_state(pool, sid3)
    for sid, p in zip(ids, probs[:3]):
        _same(got[sid], _oracle_state(p), f"shot {sid}")
    _same(got_new, _oracle_state(new), "re-admitted shot")


def test_pool_refuses_mixed_medium_kinds(gpu_device):
    a = _problems(5, 1, table=False)[0]
    b = _problems(6, 1, table=True)[0]
    with F.Pool(gpu_device) as pool:
        pool.add(a)
        with pytest.raises(F.FDTD2DError) as ei:
            pool.add(b)
        assert ei.value.code == 6


def test_pool_processing_lines(gpu_device):
    """Shots with processing lines in one pool: each shot's captures equal
    the oracle's, whatever runs beside it."""
    probs = _problems(515, 5, table=False, steps=(30, 70))
    rng = np.random.default_rng(516)
    for p in probs[:4]:
        p.src_table = p.src_table * 50
        size = int(rng.integers(1, p.nx + 1))
        p.line_size, p.line_i0, p.line_j = size, int(rng.integers(0, p.nx - size + 1)), int(rng.integers(0, p.ny))
        p.line_begin = int(rng.integers(0, p.nsteps // 2))
        p.line_end = int(rng.integers(p.line_begin + 1, p.nsteps + 5))
    with F.Pool(gpu_device, tblock=8, graph_chunk=4) as pool:
        ids = [pool.add(p) for p in probs]
        _drain(pool)
        for sid, p in zip(ids, probs):
            ref = Oracle(p, 32, threads=4)
            ref.run(p.nsteps)
            got = pool.read_line(sid)
            assert np.array_equal(got, ref.read_line().astype(np.float32)), f"shot {sid}"
            ref.close()
