"""HIP stepper vs the CPU oracle, through the C ABI, on the GPU.

The bar: every field, psi strip, detector and component trace equal to the
FP32 oracle bit for bit (the kernels use the reference's operation order with
no FMA contraction); the relative L2 error against the FP64 oracle is printed
for every run and must stay within 1e-5 on detector traces.
"""

from __future__ import annotations

import numpy as np
import pytest

import fdtd2d_amd as F
from oracle.oracle import Oracle

from helpers import all_state, fp64_check, random_problem, rel_l2

pytestmark = pytest.mark.gpu


def _compare_exact(gpu: dict, ref: dict, label: str):
    for k in ref:
        g, r = gpu[k], ref[k]
        assert g.shape == r.shape, (label, k, g.shape, r.shape)
        if not np.array_equal(g, r.astype(np.float32).astype(np.float64)):
            bad = np.argwhere(g != r.astype(np.float32))
            raise AssertionError(f"{label}: {k} differs at {len(bad)} entries, first {bad[:5].tolist()}; "
                                 f"rel L2 {rel_l2(g, r):.3e}")


_REF400 = {}


def _ref400():
    if not _REF400:
        setup = F.Setup.workload(1, seed=1)
        ref = Oracle(setup, 32, threads=4)
        ref.run(2000)
        _REF400["setup"] = setup
        _REF400["state"] = all_state(ref, 12)
        r64 = Oracle(setup, 64, threads=4)
        r64.run(2000)
        _REF400["det64"] = r64.read_detectors()[0]
    return _REF400


@pytest.mark.parametrize("tblock,cols,depth", [(1, 1, 1), (2, 1, 2), (3, 1, 4), (4, 1, 1), (6, 1, 2), (8, 1, 4),
                                               (1, 2, 1), (3, 2, 2), (4, 2, 1), (1, 4, 1), (2, 4, 1), (4, 4, 1)])
def test_plain_400_full_length(gpu_device, tblock, cols, depth):
    """400x400, 2000 steps, CPML 12, linear plasma ramp with collisions."""
    R = _ref400()
    setup, rstate = R["setup"], R["state"]
    with F.Plan(setup, gpu_device) as plan:
        plan.set_option("cols_per_lane", cols)
        plan.set_option("tblock", tblock)
        plan.set_option("prefetch", depth)
        print(f"T={tblock} C={cols} D={depth} band={plan.info('band')} tiles={plan.info('ntx')}x{plan.info('nty')} "
              f"cpml wgs={plan.info('workgroups_cpml')}/{plan.info('workgroups')} "
              f"plasma-full tiles={plan.info('tiles_plasma_full')}", flush=True)
        plan.run(2000)
        plan.sync()
        gstate = all_state(plan, 12)
    _compare_exact(gstate, rstate, f"400^2 T={tblock} C={cols}")
    err = rel_l2(gstate["det"], R["det64"])
    print(f"T={tblock} C={cols}: detector rel L2 vs FP64 oracle = {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("seed", list(range(24)))
def test_random_small_problems(gpu_device, seed):
    rng = np.random.default_rng(1000 + seed)
    p = random_problem(rng, nsteps=int(rng.integers(5, 40)))
    cols = int(rng.choice([1, 2, 4]))
    tblock = int(rng.integers(1, 9 if cols < 4 else 5))
    band = int(rng.choice([0, tblock, 7, 16]))
    depth = int(rng.choice({1: [1, 2, 4], 2: [1, 2], 4: [1]}[cols]))
    print(f"seed {seed}: {p.nx}x{p.ny} P={p.npml} kind={p.medium_kind} T={tblock} C={cols} band={band} "
          f"D={depth} nsrc={len(p.src_i)} ndet={len(p.det_i)}", flush=True)
    ref = Oracle(p, 32)
    ref.run(p.nsteps)
    with F.Plan(p, gpu_device) as plan:
        plan.set_option("cols_per_lane", cols)
        plan.set_option("tblock", tblock)
        plan.set_option("band", band)
        plan.set_option("prefetch", depth)
        plan.set_option("band_cpml", int(rng.choice([0, 1, 5, 9])))
        plan.set_option("jz_first", int(rng.integers(0, 2)))
        plan.set_option("fused", int(rng.integers(0, 2)))
        plan.run(p.nsteps)
        plan.sync()
        g = all_state(plan, p.npml)
    _compare_exact(g, all_state(ref, p.npml), f"seed {seed}")
    fp64_check(p, g["det"], f"seed {seed}", comp32=g["comp"])


def test_resume_equals_single_run(gpu_device):
    rng = np.random.default_rng(7)
    p = random_problem(rng, nx=150, ny=133, npml=8, nsteps=41, table=True)
    with F.Plan(p, gpu_device, tblock=3) as a, F.Plan(p, gpu_device, tblock=4) as b:
        a.run(41)
        for chunk in (5, 1, 13, 22):
            b.run(chunk)
        a.sync()
        b.sync()
        sa, sb = all_state(a, 8), all_state(b, 8)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k


def test_write_field_initial_state(gpu_device):
    """A nonzero initial state (written through the ABI) evolves as in the oracle."""
    rng = np.random.default_rng(11)
    p = random_problem(rng, nx=97, ny=71, npml=5, nsteps=17, table=False)
    ref = Oracle(p, 32)
    init = {w: rng.standard_normal(ref.read_field(w).shape).astype(np.float32) for w in (F.EZ, F.HX, F.HY, F.JZ)}
    for w, v in init.items():
        ref.write_field(w, v)
    ref.run(17)
    with F.Plan(p, gpu_device, tblock=4) as plan:
        for w, v in init.items():
            plan.write_field(w, v)
        plan.run(17)
        plan.sync()
        g = all_state(plan, 5)
    _compare_exact(g, all_state(ref, 5), "initial state")


def _run_pair(p, gpu_device, opts, label, prepare=None):
    ref = Oracle(p, 32, threads=4)
    with F.Plan(p, gpu_device) as plan:
        for k, v in opts.items():
            plan.set_option(k, v)
        if prepare:
            prepare(ref, plan)
        info = {k: plan.info(k) for k in ("tiles_interior", "tiles_cpml", "tiles_nojz", "tiles_plasma_full")}
        print(f"{label}: {p.nx}x{p.ny} P={p.npml} {opts} {info}", flush=True)
        ref.run(p.nsteps)
        plan.run(p.nsteps)
        plan.sync()
        g = all_state(plan, p.npml)
    _compare_exact(g, all_state(ref, p.npml), label)
    if prepare is None:  # the FP64 oracle starts from the problem's own (zero) state
        fp64_check(p, g["det"], label, comp32=g["comp"])
    return info


@pytest.mark.parametrize("seed", list(range(10)))
def test_interior_paths(gpu_device, seed):
    """Grids large enough for interior tiles, with and without plasma in
    their windows, both media, every launch shape."""
    rng = np.random.default_rng(5000 + seed)
    P = int(rng.choice([0, 5, 11, 16]))
    cols = int(rng.choice([1, 2, 4]))
    nx = int(rng.integers(260, 640)) if cols < 4 else int(rng.integers(760, 900))
    p = random_problem(rng, nx=nx, ny=int(rng.integers(200, 520)), npml=P, nsteps=int(rng.integers(9, 30)),
                       table=bool(seed % 2))
    tblock = int(rng.integers(1, 9 if cols < 4 else 5))
    opts = {"cols_per_lane": cols, "tblock": tblock, "band": int(rng.choice([0, 16, 64])),
            "prefetch": int(rng.choice({1: [1, 2, 4], 2: [1, 2], 4: [1]}[cols])),
            "band_cpml": int(rng.choice([0, 8, 32])), "fused": seed % 3 != 2}
    info = _run_pair(p, gpu_device, opts, f"interior seed {seed}")
    assert info["tiles_interior"] > 0


@pytest.mark.parametrize("tblock,depth", [(8, 2), (3, 1)])
def test_plasma_full_tiles(gpu_device, tblock, depth):
    """Interior tiles whose whole window holds plasma (two columns per lane,
    density medium) run without the Ne > 0 test; the tiles across the ramp's
    foot and the plasma-free ones keep their variants."""
    rng = np.random.default_rng(91 + tblock)
    p = random_problem(rng, nx=900, ny=260, npml=8, nsteps=26, table=False)
    info = _run_pair(p, gpu_device, {"cols_per_lane": 2, "tblock": tblock, "prefetch": depth}, "plasma full")
    assert info["tiles_plasma_full"] > 0
    assert info["tiles_nojz"] > 0
    assert info["tiles_interior"] > info["tiles_plasma_full"] + info["tiles_nojz"]


def test_jz_written_where_no_plasma(gpu_device):
    """Jz written into plasma-free windows (nonzero and -0.0 values) puts those
    tiles back on the Jz path."""
    rng = np.random.default_rng(77)
    p = random_problem(rng, nx=520, ny=300, npml=8, nsteps=24, table=False)
    jz = np.zeros((p.ny, p.nx), np.float32)
    jz[100:140, 60:90] = rng.standard_normal((40, 30)).astype(np.float32) * 1e-3
    jz[200:210, 120:125] = -0.0

    def prepare(ref, plan):
        before = plan.info("tiles_nojz")
        ref.write_field(F.JZ, jz)
        plan.write_field(F.JZ, jz)
        assert plan.info("tiles_nojz") < before

    _run_pair(p, gpu_device, {"tblock": 8}, "jz write", prepare)


def test_nonuniform_den_inside(gpu_device):
    """den values that differ inside the grid (not only in the CPML strips)
    send the tiles that load them to the general kernel."""
    rng = np.random.default_rng(78)
    p = random_problem(rng, nx=480, ny=400, npml=6, nsteps=20, table=True)
    p.den_ex = p.den_ex.copy()
    p.den_hy = p.den_hy.copy()
    p.den_ex[230] *= np.float32(1.25)
    p.den_hy[190:193] *= np.float32(0.8)
    info = _run_pair(p, gpu_device, {"tblock": 6}, "den inside")
    q = random_problem(np.random.default_rng(78), nx=480, ny=400, npml=6, nsteps=20, table=True)
    with F.Plan(q, gpu_device) as plan:
        plan.set_option("tblock", 6)
        assert plan.info("tiles_cpml") < info["tiles_cpml"]


@pytest.mark.parametrize("seed", range(4))
def test_rc_lines_match_oracle(gpu_device, seed):
    """RC cutoff lines: probe columns captured by the stepper every step and
    reduced after each launch, bitwise against the oracle's per-step rule."""
    rng = np.random.default_rng(6000 + seed)
    p = random_problem(rng, nx=int(rng.integers(200, 420)), ny=int(rng.integers(150, 400)),
                       npml=int(rng.choice([0, 8, 11])), nsteps=int(rng.integers(30, 70)), table=bool(seed % 2),
                       nsrc=3, ndet=2, ncomp=1)
    p.src_table = p.src_table * 50
    nrc = int(rng.integers(1, 12))
    p.rc_imax = np.sort(rng.choice(np.arange(2, p.nx - 20), nrc, replace=False)).astype(np.int32)
    p.rc_i = (p.rc_imax + np.arange(nrc)).astype(np.int32)
    p.rc_jcut, p.rc_level, p.rc_edge0 = int(rng.integers(0, p.ny)), 0.07, 1e-3
    ref = Oracle(p, 32, threads=4)
    ref.run(p.nsteps)
    want = ref.read_rc()
    for tblock in (8, 3):
        with F.Plan(p, gpu_device, tblock=tblock) as plan:
            plan.run(p.nsteps)
            plan.sync()
            got = plan.read_rc()
            det, _ = plan.read_detectors()
        print(f"seed {seed} T={tblock}: {p.nx}x{p.ny} nrc={nrc} jcut={p.rc_jcut} rows {np.unique(want).size}",
              flush=True)
        assert np.array_equal(got, want)
        assert np.array_equal(det, ref.read_detectors()[0].astype(np.float32))
    fp64_check(p, det, f"rc seed {seed}")


@pytest.mark.parametrize("seed", range(8))
def test_processing_line_matches_oracle(gpu_device, seed):
    """Processing line captured by the CPML-class kernel every step of its
    window: bitwise the oracle's captures, for lines anywhere (CPML rows, the
    outer ring, the full width) and windows not aligned to the launches."""
    rng = np.random.default_rng(7000 + seed)
    cols = int(rng.choice([1, 2, 4]))
    nx = int(rng.integers(260, 640)) if cols < 4 else int(rng.integers(760, 900))
    p = random_problem(rng, nx=nx, ny=int(rng.integers(150, 420)), npml=int(rng.choice([0, 8, 11])),
                       nsteps=int(rng.integers(20, 60)), table=bool(seed % 2))
    p.src_table = p.src_table * 50
    size = int(rng.integers(1, p.nx + 1))
    p.line_size, p.line_i0 = size, int(rng.integers(0, p.nx - size + 1))
    # the outer ring rows, or the row of a source with the line through it
    p.line_j = int([0, p.ny - 1, p.src_j[0], p.src_j[0]][seed % 4])
    if p.line_j == p.src_j[0]:
        p.line_i0 = int(rng.integers(max(0, p.src_i[0] - size + 1), min(p.src_i[0], p.nx - size) + 1))
    p.line_begin = int(rng.integers(0, p.nsteps - 1))
    p.line_end = int(rng.integers(p.line_begin + 1, p.nsteps + 20))
    opts = {"cols_per_lane": cols, "tblock": int(rng.integers(1, 9 if cols < 4 else 5)),
            "band": int(rng.choice([0, 16, 64])), "band_cpml": int(rng.choice([0, 8, 32]))}
    ref = Oracle(p, 32, threads=4)
    ref.run(p.nsteps)
    want = ref.read_line()
    with F.Plan(p, gpu_device) as plan:
        for k, v in opts.items():
            plan.set_option(k, v)
        print(f"seed {seed}: {p.nx}x{p.ny} P={p.npml} line row {p.line_j} cols [{p.line_i0}, "
              f"{p.line_i0 + size}) steps ({p.line_begin}, {p.line_end}] {opts}", flush=True)
        plan.run(p.nsteps)
        plan.sync()
        got = plan.read_line()
        ez = plan.read_field(F.EZ)
    assert got.shape == want.shape
    assert np.array_equal(got, want.astype(np.float32))
    assert np.array_equal(ez, ref.read_field(F.EZ).astype(np.float32))
    if p.line_j == p.src_j[0]:
        assert np.abs(want).max() > 0
    fp64_check(p, ref.read_detectors()[0].astype(np.float32), f"line seed {seed}")
