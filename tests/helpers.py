"""Shared helpers for the test suite: random problems and comparisons."""

from __future__ import annotations

import numpy as np

import fdtd2d_amd as F
from fdtd2d_amd.problem import Problem

EPS0 = 1.0 / (2.99792458e8 ** 2 * 4.0 * 3.14159265358979 * 1e-7)
MU0 = 4.0 * 3.14159265358979 * 1e-7
CJ = 2.81794e-8


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


def random_problem(rng: np.random.Generator, nx=None, ny=None, npml=None, nsteps=60,
                   table=None, nsrc=None, ndet=None, ncomp=None) -> Problem:
    """A small problem exercising every code path: CPML of random thickness,
    TABLE (with PEC, lossy dielectrics and plasma ids) or DENSITY medium,
    sources, detectors and component detectors anywhere, including CPML
    strips and the never-updated outer ring."""
    P = int(rng.choice([0, 2, 3, 5, 8, 11])) if npml is None else npml
    nx = int(rng.integers(2 * P + 4, 2 * P + 140)) if nx is None else nx
    ny = int(rng.integers(2 * P + 4, 2 * P + 140)) if ny is None else ny
    table = bool(rng.integers(0, 2)) if table is None else table
    box = F.Setup.box(nx, ny, P, nsteps, 1e-3)
    p = box.problem()
    p.nsteps = nsteps
    dt = p.dt
    f0 = 10e9
    nc = EPS0 * (2 * np.pi * f0) ** 2 / CJ
    if table:
        nm = 12
        eps = np.full(nm, EPS0)
        sig = np.zeros(nm)
        eps[2:8] = EPS0 * rng.uniform(1.0, 9.0, 6)
        sig[2:8] = rng.uniform(0.0, 0.05, 6)
        ca = ((1 - sig * dt / (2 * eps)) / (1 + sig * dt / (2 * eps))).astype(np.float32)
        cb = ((dt / eps) / (1 + sig * dt / (2 * eps))).astype(np.float32)
        ne = np.zeros(nm, np.float32)
        ne[8:] = (rng.uniform(0.2, 2.5, nm - 8) * nc).astype(np.float32)
        mat = np.zeros((ny, nx), np.int32)
        for _ in range(int(rng.integers(1, 6))):
            i0, j0 = int(rng.integers(0, nx)), int(rng.integers(0, ny))
            w, h = int(rng.integers(1, max(2, nx // 3))), int(rng.integers(1, max(2, ny // 3)))
            mat[j0:j0 + h, i0:i0 + w] = int(rng.integers(1, nm))
        p.medium_kind = F.MEDIUM_TABLE
        p.material, p.ca, p.cb, p.ne, p.pec_material = mat, ca, cb, ne, 1
        p.density = None
    else:
        x = np.arange(nx)
        prof = np.clip((x - nx * 0.4) / (nx * 0.5), 0, None) * 2 * nc
        dens = np.tile(prof, (ny, 1)) * (1 + 0.05 * rng.standard_normal((ny, nx)))
        p.density = np.clip(dens, 0, None).astype(np.float32)
        nu = float(rng.choice([0.0, 1e9, 5e9]))
        p.jz_decay = np.float32((1 - nu * dt / 2) / (1 + nu * dt / 2))
        p.jz_dt = dt / (1 + nu * dt / 2)
    nsrc = int(rng.integers(1, 5)) if nsrc is None else nsrc
    ndet = int(rng.integers(1, 7)) if ndet is None else ndet
    ncomp = int(rng.integers(0, 4)) if ncomp is None else ncomp
    p.src_i = rng.integers(0, nx, nsrc).astype(np.int32)
    p.src_j = rng.integers(0, ny, nsrc).astype(np.int32)
    t = np.arange(nsteps + 1)[:, None] * dt
    tau = 4 * dt
    amp = rng.uniform(0.5, 2.0, nsrc)[None, :]
    p.src_table = (-amp * np.exp(-((t - 3 * tau) / tau) ** 2) * np.sin(2 * np.pi * f0 * t)).astype(np.float32)
    p.det_i = rng.integers(0, nx, ndet).astype(np.int32)
    p.det_j = rng.integers(0, ny, ndet).astype(np.int32)
    p.comp_i = rng.integers(0, nx - 1, ncomp).astype(np.int32)
    p.comp_j = rng.integers(0, ny - 1, ncomp).astype(np.int32)
    return p


def all_state(sim, npml: int):
    """Every field and psi array of a plan or oracle, as float64."""
    out = {}
    for name, w in (("ez", F.EZ), ("hx", F.HX), ("hy", F.HY), ("jz", F.JZ)):
        out[name] = np.asarray(sim.read_field(w), np.float64)
    if npml > 0:
        for name, w in (("psi_hxy", F.PSI_HXY), ("psi_hyx", F.PSI_HYX), ("psi_ezx", F.PSI_EZX),
                        ("psi_ezy", F.PSI_EZY)):
            out[name] = np.asarray(sim.read_psi(w), np.float64)
    det, comp = sim.read_detectors()
    out["det"] = np.asarray(det, np.float64)
    out["comp"] = np.asarray(comp, np.float64)
    return out


def source_peak(problem) -> float:
    """Largest |source table| value of a Problem or a Setup."""
    if hasattr(problem, "src_table"):
        t = np.asarray(problem.src_table)
    else:
        c = problem.c_problem.contents
        n = (c.nsteps + 1) * c.nsrc
        t = np.ctypeslib.as_array(c.src_table, shape=(n,)) if n else np.zeros(0)
    return float(np.abs(t).max()) if t.size else 0.0


# Traces whose FP64 peak stays below this fraction of the source peak hold
# only the numerical precursor of a wave that has not reached the detectors
# (random problems with few steps): there the FP32 relative error is set by
# rounding of near-zero values, so it is printed but not held to 1e-5.
PRECURSOR = 1e-2


def fp64_check(problem, det32, label: str, threads: int = 4, comp32=None, nsteps=None) -> float:
    """Runs the FP64 oracle on the same problem (a Problem or a Setup), prints
    the relative L2 error of the FP32 detector traces (and component traces,
    if given) against it, and asserts it is within 1e-5 wherever the FP64
    traces are nonzero (and, beyond that bound, carry signal: peak at least
    PRECURSOR of the source peak). Returns the detector error."""
    from oracle.oracle import Oracle

    ref = Oracle(problem, 64, threads=threads)
    ref.run(ref.nsteps if nsteps is None else nsteps)
    det64, comp64 = ref.read_detectors()
    ref.close()
    err = rel_l2(det32, det64)
    peak = float(np.abs(det64).max()) if det64.size else 0.0
    msg = f"{label}: detector rel L2 vs FP64 oracle = {err:.3e} (max |det| {peak:.3e})"
    if comp32 is not None and comp64.size:
        cerr = rel_l2(comp32, comp64)
        msg += f", components {cerr:.3e}"
        if np.abs(comp64).max() > 0:
            assert cerr <= 1e-5, msg
    if peak > 0 and err > 1e-5 and peak < PRECURSOR * source_peak(problem):
        print(msg + f" (precursor only: peak below {PRECURSOR:g} of the source peak, not asserted)", flush=True)
        return err
    print(msg, flush=True)
    if peak > 0:
        assert err <= 1e-5, msg
    return err


def compare_fields_lean(sim, ref, label: str, npml: int = 0, psi: bool = True) -> None:
    """Bitwise comparison of every field (and psi strip) of a plan with the
    FP32 oracle, one array at a time in float32 (for grids whose float64
    copies would not fit comfortably in host memory)."""
    names = [("ez", F.EZ), ("hx", F.HX), ("hy", F.HY), ("jz", F.JZ)]
    for name, w in names:
        g = np.asarray(sim.read_field(w), np.float32)
        r = ref.read_field(w).astype(np.float32)
        if not np.array_equal(g.view(np.uint32), r.view(np.uint32)):
            bad = np.argwhere(g.view(np.uint32) != r.view(np.uint32))
            raise AssertionError(f"{label}: {name} differs at {len(bad)} cells, first {bad[:5].tolist()}")
        del g, r
    if psi and npml > 0:
        for name, w in (("psi_hxy", F.PSI_HXY), ("psi_hyx", F.PSI_HYX), ("psi_ezx", F.PSI_EZX),
                        ("psi_ezy", F.PSI_EZY)):
            g = np.asarray(sim.read_psi(w), np.float32)
            r = ref.read_psi(w).astype(np.float32)
            if not np.array_equal(g.view(np.uint32), r.view(np.uint32)):
                bad = np.argwhere(g.view(np.uint32) != r.view(np.uint32))
                raise AssertionError(f"{label}: {name} differs at {len(bad)} entries, first {bad[:5].tolist()}")
