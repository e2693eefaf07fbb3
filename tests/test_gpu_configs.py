"""The stepper at the benchmark workloads' own sizes and launch shapes, on the GPU.

Each test runs the HIP path on a full-size workload and compares every field,
psi strip and trace with the FP32 CPU oracle bit for bit (the oracle on 16
host threads, which gives the single-threaded result), and prints the
detector error against the FP64 oracle (asserted within 1e-5):

* 8192^2 single grid (CPML 16, 16-element steered array, parabolic plasma
  with 2% fluctuations, 64 detectors) at the default plan shape, for 28
  steps: three launches of 8 steps and a tail launch of 4;
* the same grid for all 20000 steps under size-independent properties
  (finite, nonzero traces; field energy bounded after the pulse);
* a 16384^2 shot of the mixed serving stream (CPML 16, 32 elements), as a
  plan at the default shape and inside a pool beside two 1024^2 shots;
* the first 16 shots of the 256-shot batch mix (256-2048 cells per side)
  through one pool: a captured 8-launch graph chunk plus a tail launch;
* three frequencies of one served sweep request (640^2, CPML, the sweep's
  plasma) for their full 3000 steps through one pool at the serving shape.
"""

from __future__ import annotations

import numpy as np
import pytest

import fdtd2d_amd as F
from oracle.oracle import Oracle

from helpers import compare_fields_lean, fp64_check, rel_l2

pytestmark = pytest.mark.gpu

THREADS = 16


class _PoolShot:
    """A pool shot seen through the plan-like read calls the comparisons use."""

    def __init__(self, pool, sid):
        self.pool, self.sid = pool, sid

    def read_field(self, which):
        return self.pool.read_field(self.sid, which)

    def read_detectors(self):
        return self.pool.read_detectors(self.sid)


def _same_traces(gpu, ref, label):
    det, comp = gpu.read_detectors()
    rdet, rcomp = ref.read_detectors()
    for name, g, r in (("det", det, rdet), ("comp", comp, rcomp)):
        r = r.astype(np.float32)
        if not np.array_equal(g.view(np.uint32), r.view(np.uint32)):
            bad = np.argwhere(g.view(np.uint32) != r.view(np.uint32))
            raise AssertionError(f"{label}: {name} differs at {len(bad)} entries, first {bad[:5].tolist()}, "
                                 f"rel L2 {rel_l2(g, r):.3e}")
    return det, comp


def test_8192_bench_shape_28_steps(gpu_device):
    steps = 28
    setup = F.Setup.workload(2, seed=2, steps=steps)
    with F.Plan(setup, gpu_device) as plan:
        info = {k: plan.info(k) for k in ("tblock", "cols_per_lane", "prefetch", "band", "band_cpml", "workgroups",
                                          "workgroups_cpml", "tiles_plasma_full", "tiles_nojz", "jz_first")}
        print(f"8192^2 default shape: {info}", flush=True)
        # the bench's shape: 8 steps per launch, two columns per lane, 96-row bands
        assert info["tblock"] == 8 and info["cols_per_lane"] == 2 and info["band"] == 96
        assert info["tiles_plasma_full"] > 0 and info["jz_first"] == 1
        plan.run(steps)  # launches of 8, 8, 8 and 4 steps
        plan.sync()
        ref = Oracle(setup, 32, threads=THREADS)
        ref.run(steps)
        compare_fields_lean(plan, ref, "8192^2", npml=16)
        det, comp = _same_traces(plan, ref, "8192^2")
        ref.close()
    assert np.abs(det).max() > 0
    fp64_check(setup, det, "8192^2 28 steps", threads=THREADS, comp32=comp)


def _energy(plan, eps0=8.8541878128e-12, mu0=4e-7 * np.pi) -> float:
    e = 0.0
    for w, k in ((F.EZ, eps0), (F.HX, mu0), (F.HY, mu0)):
        v = plan.read_field(w).astype(np.float64)
        e += k * float(np.dot(v.ravel(), v.ravel()))
    return 0.5 * e


def test_8192_full_length_properties(gpu_device):
    """All 20000 steps of the 8192^2 workload: the traces stay finite and
    nonzero, and the field energy after the source pulse has ended does not
    grow (the CPML absorbs, the plasma is collisional)."""
    setup = F.Setup.workload(2, seed=2)
    nsteps = int(setup.get("nMax"))
    t_end = setup.get("t0") + 4 * setup.get("tau") + 16 * 8 * setup.get("dx") * np.sin(np.pi / 12) / 2.99792458e8
    n_quiet = int(np.ceil(t_end / setup.get("dt")))
    assert nsteps == 20000 and n_quiet < 4000
    with F.Plan(setup, gpu_device) as plan:
        plan.run(n_quiet)
        plan.sync()
        e0 = _energy(plan)
        plan.run(nsteps - n_quiet)
        plan.sync()
        e1 = _energy(plan)
        det, comp = plan.read_detectors()
    print(f"8192^2 x {nsteps}: energy {e0:.4e} after the pulse (step {n_quiet}), {e1:.4e} at the end; "
          f"max |det| {np.abs(det).max():.3e}", flush=True)
    assert np.isfinite(det).all() and np.isfinite(comp).all()
    assert (np.abs(det[1:]).max(axis=0) > 0).all(), "a detector trace is all zero"
    assert np.isfinite(e0) and np.isfinite(e1) and e0 > 0
    assert e1 <= e0 * (1 + 1e-6)


def _big_shot_index(seed=5):
    for shot in range(200):
        s = F.Setup.workload(5, seed=seed, shot=shot, steps=-1)
        big = s.get("big") > 0
        s.close()
        if big:
            return shot
    raise AssertionError("no 16384^2 shot in the first 200 draws")


def _small_shot_indices(seed=5, n=2):
    out = []
    for shot in range(200):
        s = F.Setup.workload(5, seed=seed, shot=shot, steps=-1)
        if s.get("big") == 0:
            out.append(shot)
        s.close()
        if len(out) == n:
            return out
    raise AssertionError("too few 1024^2 shots")


def test_16384_shot_plan_and_pool(gpu_device):
    steps = 16
    big = _big_shot_index()
    setup = F.Setup.workload(5, seed=5, shot=big, steps=steps)
    assert setup.get("Imax") == 16384 and setup.get("npml") == 16 and setup.get("nel") == 32
    ref = Oracle(setup, 32, threads=THREADS)
    ref.run(steps)
    with F.Plan(setup, gpu_device) as plan:
        info = {k: plan.info(k) for k in ("tblock", "cols_per_lane", "band", "workgroups", "workgroups_cpml")}
        print(f"16384^2 shot {big} default shape: {info}", flush=True)
        assert info["cols_per_lane"] == 2 and info["band"] == 192
        plan.run(steps)
        plan.sync()
        compare_fields_lean(plan, ref, "16384^2 plan", npml=16)
        det, comp = _same_traces(plan, ref, "16384^2 plan")
    fp64_check(setup, det, "16384^2 16 steps", threads=THREADS, comp32=comp)
    smalls = [F.Setup.workload(5, seed=5, shot=s, steps=steps) for s in _small_shot_indices()]
    small_refs = []
    for s in smalls:
        r = Oracle(s, 32, threads=THREADS)
        r.run(steps)
        small_refs.append(r)
    # the pool's default shape, then the shape the serving stream of these
    # shots takes from the shape cache (two columns per lane, 192-row bands)
    for label, shape in (("default", {}), ("mixed-stream", {"cols_per_lane": 2, "band": 192, "prefetch": 2})):
        with F.Pool(gpu_device, graph_chunk=2, **shape) as pool:
            ids = [pool.add(setup)] + [pool.add(s) for s in smalls]
            while pool.info("running") > 0:
                pool.step(2)
            pool.sync()
            print(f"pool ({label} shape): launches {pool.info('launches'):.0f}, "
                  f"graph replays {pool.info('graph_replays'):.0f}", flush=True)
            assert pool.info("graph_replays") >= 1
            shot = _PoolShot(pool, ids[0])
            compare_fields_lean(shot, ref, f"16384^2 in a pool ({label} shape)", psi=False)
            _same_traces(shot, ref, f"16384^2 in a pool ({label} shape)")
            for sid, s, r in zip(ids[1:], smalls, small_refs):
                small = _PoolShot(pool, sid)
                compare_fields_lean(small, r, f"1024^2 shot {sid} beside it ({label} shape)", psi=False)
                d, c = _same_traces(small, r, f"1024^2 shot {sid} ({label} shape)")
                if label == "default":
                    fp64_check(s, d, f"1024^2 shot {sid} beside the 16384^2 shot", threads=THREADS, comp32=c)
    ref.close()
    for r in small_refs:
        r.close()


def test_batch_mix_first_16_shots(gpu_device):
    """The first 16 shots of the seed-3 batch mix at 69 steps: 8 launches of
    8 steps replayed from one captured graph chunk, then a 5-step tail
    launch, at the pool's default (mix-tuned) shape."""
    steps = 69
    setups = [F.Setup.workload(3, seed=3, shot=s, steps=steps) for s in range(16)]
    sizes = [(int(s.get("Imax")), int(s.get("Jmax"))) for s in setups]
    print(f"shots {sizes}", flush=True)
    assert max(max(z) for z in sizes) >= 1024
    with F.Pool(gpu_device, graph_chunk=8) as pool:
        ids = [pool.add(s) for s in setups]
        pool.step(8)
        pool.step(1)
        pool.sync()
        st = {k: pool.info(k) for k in ("launches", "graph_replays", "eager_launches", "tail_launches", "running")}
        print(f"pool: {st}", flush=True)
        assert st["graph_replays"] == 1 and st["tail_launches"] == 1 and st["running"] == 0
        for sid, s in zip(ids, setups):
            ref = Oracle(s, 32, threads=THREADS)
            ref.run(steps)
            shot = _PoolShot(pool, sid)
            compare_fields_lean(shot, ref, f"mix shot {sid}", psi=False)
            det, comp = _same_traces(shot, ref, f"mix shot {sid}")
            ref.close()
            fp64_check(s, det, f"mix shot {sid} {sizes[sid]}", threads=THREADS, comp32=comp)


def test_sweep_request_full_length(gpu_device):
    """Three frequencies of served sweep request 12 (seed 4: a 640^2 grid, one
    plasma profile, 3000 steps each) through one pool at the serving launch
    shape (the batch mix's cached shape), with graph chunks: every field and
    trace bitwise equal to the FP32 oracle at the end of the full run."""
    shots = [F.Setup.workload(4, seed=4, shot=12 * 64 + k, steps=0) for k in range(3)]
    sizes = {(int(s.get("Imax")), int(s.get("Jmax")), int(s.get("nMax"))) for s in shots}
    print(f"sweep request 12: {sizes}", flush=True)
    assert sizes == {(640, 640, 3000)}
    shape = {"fused": 0, "tblock": 8, "cols_per_lane": 1, "band": 128, "wg_per_cu": 3, "band_cpml": 64,
             "wg_per_cu_cpml": 2, "prefetch": 1}
    with F.Pool(gpu_device, graph_chunk=8, **shape) as pool:
        ids = [pool.add(s) for s in shots]
        while pool.info("running") > 0:
            pool.step(8)
        pool.sync()
        print({k: pool.info(k) for k in ("launches", "graph_replays", "eager_launches")}, flush=True)
        assert pool.info("graph_replays") > 0
        for sid, s in zip(ids, shots):
            ref = Oracle(s, 32, threads=THREADS)
            ref.run(3000)
            shot = _PoolShot(pool, sid)
            compare_fields_lean(shot, ref, f"sweep shot {sid}", psi=False)
            det, comp = _same_traces(shot, ref, f"sweep shot {sid}")
            ref.close()
            fp64_check(s, det, f"sweep shot {sid} 640x640 x 3000 steps", threads=THREADS, comp32=comp)
