"""Problems, setups and plans: the Python face of include/fdtd2d.h.

``Problem`` owns numpy arrays and produces the C struct; ``Setup`` wraps the
native builders (Settings.ini, synthetic workloads, test boxes); ``Plan`` is a
device plan (the replacement of the reference's DevFieldCalculation,
F/kernel.cu:1928).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib as L


def _fp(a: Optional[np.ndarray]):
    return a.ctypes.data_as(L._f) if a is not None and a.size else L._f()


def _ip(a: Optional[np.ndarray]):
    return a.ctypes.data_as(L._i) if a is not None and a.size else L._i()


def _copy_f(ptr, n: int) -> np.ndarray:
    if n <= 0 or not ptr:
        return np.zeros(max(n, 0), np.float32)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.float32, copy=True)


def _copy_i(ptr, n: int) -> np.ndarray:
    if n <= 0 or not ptr:
        return np.zeros(max(n, 0), np.int32)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.int32, copy=True)


@dataclass
class Problem:
    """Plain-numpy mirror of ``fdtd2d_problem`` (field meanings as there)."""

    nx: int
    ny: int
    npml: int
    nsteps: int
    dt: float
    dx: float
    dy: float
    db: float
    medium_kind: int = L.MEDIUM_DENSITY
    material: Optional[np.ndarray] = None  # int32 [ny, nx]
    ca: Optional[np.ndarray] = None
    cb: Optional[np.ndarray] = None
    ne: Optional[np.ndarray] = None
    pec_material: int = -1
    density: Optional[np.ndarray] = None  # float32 [ny, nx]
    ca0: float = 1.0
    cb0: float = 0.0
    cj: float = 2.81794e-8
    jz_decay: float = 1.0
    jz_dt: float = 0.0
    be_x: list = field(default_factory=lambda: [None, None])
    ce_x: list = field(default_factory=lambda: [None, None])
    bh_x: list = field(default_factory=lambda: [None, None])
    ch_x: list = field(default_factory=lambda: [None, None])
    be_y: list = field(default_factory=lambda: [None, None])
    ce_y: list = field(default_factory=lambda: [None, None])
    bh_y: list = field(default_factory=lambda: [None, None])
    ch_y: list = field(default_factory=lambda: [None, None])
    den_ex: Optional[np.ndarray] = None
    den_hx: Optional[np.ndarray] = None
    den_ey: Optional[np.ndarray] = None
    den_hy: Optional[np.ndarray] = None
    dx_f: float = 0.0
    dy_f: float = 0.0
    src_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    src_j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    src_table: np.ndarray = field(default_factory=lambda: np.zeros((1, 0), np.float32))
    det_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    det_j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    comp_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    comp_j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    rc_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    rc_imax: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    rc_jcut: int = 0
    rc_level: float = 0.1
    rc_edge0: float = 10.0
    # processing line: Ez of row line_j, columns [line_i0, line_i0 + line_size)
    # after steps line_begin < n <= line_end (line_size 0 = off)
    line_size: int = 0
    line_i0: int = 0
    line_j: int = 0
    line_begin: int = 0
    line_end: int = 0

    # -------------------------------------------------------------- C struct
    def to_c(self):
        """Returns (struct, keepalive); keep both alive while the struct is used."""
        keep = []

        def f32(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            keep.append(a)
            return a

        def i32(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return a

        s = L.ProblemC()
        s.nx, s.ny, s.npml, s.nsteps = self.nx, self.ny, self.npml, self.nsteps
        s.dt, s.dx, s.dy = self.dt, self.dx, self.dy
        s.db = self.db
        s.medium_kind = self.medium_kind
        s.material = _ip(i32(self.material))
        s.nmaterials = 0 if self.ca is None else int(np.asarray(self.ca).size)
        s.ca, s.cb, s.ne = _fp(f32(self.ca)), _fp(f32(self.cb)), _fp(f32(self.ne))
        s.pec_material = self.pec_material
        s.density = _fp(f32(self.density))
        s.ca0, s.cb0, s.cj, s.jz_decay, s.jz_dt = self.ca0, self.cb0, self.cj, self.jz_decay, self.jz_dt
        for name in ("be_x", "ce_x", "bh_x", "ch_x", "be_y", "ce_y", "bh_y", "ch_y"):
            arr = getattr(s, name)
            for side in range(2):
                arr[side] = _fp(f32(getattr(self, name)[side]))
        s.den_ex, s.den_hx = _fp(f32(self.den_ex)), _fp(f32(self.den_hx))
        s.den_ey, s.den_hy = _fp(f32(self.den_ey)), _fp(f32(self.den_hy))
        s.dx_f, s.dy_f = self.dx_f, self.dy_f
        s.nsrc = int(len(self.src_i))
        s.src_i, s.src_j = _ip(i32(self.src_i)), _ip(i32(self.src_j))
        s.src_table = _fp(f32(self.src_table))
        s.ndet = int(len(self.det_i))
        s.det_i, s.det_j = _ip(i32(self.det_i)), _ip(i32(self.det_j))
        s.ncomp = int(len(self.comp_i))
        s.comp_i, s.comp_j = _ip(i32(self.comp_i)), _ip(i32(self.comp_j))
        s.nrc = int(len(self.rc_i))
        s.rc_i, s.rc_imax = _ip(i32(self.rc_i)), _ip(i32(self.rc_imax))
        s.rc_jcut, s.rc_level, s.rc_edge0 = self.rc_jcut, self.rc_level, self.rc_edge0
        s.line_size, s.line_i0, s.line_j = self.line_size, self.line_i0, self.line_j
        s.line_begin, s.line_end = self.line_begin, self.line_end
        return s, keep

    @classmethod
    def from_c(cls, s: L.ProblemC) -> "Problem":
        nx, ny, P, ns = s.nx, s.ny, s.npml, s.nsteps
        nm = s.nmaterials
        table = s.medium_kind == L.MEDIUM_TABLE
        p = cls(
            nx=nx, ny=ny, npml=P, nsteps=ns, dt=s.dt, dx=s.dx, dy=s.dy, db=s.db,
            medium_kind=s.medium_kind,
            material=_copy_i(s.material, nx * ny).reshape(ny, nx) if table else None,
            ca=_copy_f(s.ca, nm) if table else None,
            cb=_copy_f(s.cb, nm) if table else None,
            ne=_copy_f(s.ne, nm) if table else None,
            pec_material=s.pec_material,
            density=None if table else _copy_f(s.density, nx * ny).reshape(ny, nx),
            ca0=s.ca0, cb0=s.cb0, cj=s.cj, jz_decay=s.jz_decay, jz_dt=s.jz_dt,
            den_ex=_copy_f(s.den_ex, nx - 1), den_hx=_copy_f(s.den_hx, nx - 1),
            den_ey=_copy_f(s.den_ey, ny - 1), den_hy=_copy_f(s.den_hy, ny - 1),
            dx_f=s.dx_f, dy_f=s.dy_f,
            src_i=_copy_i(s.src_i, s.nsrc), src_j=_copy_i(s.src_j, s.nsrc),
            src_table=_copy_f(s.src_table, (ns + 1) * s.nsrc).reshape(ns + 1, s.nsrc),
            det_i=_copy_i(s.det_i, s.ndet), det_j=_copy_i(s.det_j, s.ndet),
            comp_i=_copy_i(s.comp_i, s.ncomp), comp_j=_copy_i(s.comp_j, s.ncomp),
            rc_i=_copy_i(s.rc_i, s.nrc), rc_imax=_copy_i(s.rc_imax, s.nrc),
            rc_jcut=s.rc_jcut, rc_level=s.rc_level, rc_edge0=s.rc_edge0,
            line_size=s.line_size, line_i0=s.line_i0, line_j=s.line_j,
            line_begin=s.line_begin, line_end=s.line_end,
        )
        for name, n in (("be_x", P), ("ce_x", P), ("bh_x", P - 1), ("ch_x", P - 1),
                        ("be_y", P), ("ce_y", P), ("bh_y", P - 1), ("ch_y", P - 1)):
            arr = getattr(s, name)
            setattr(p, name, [_copy_f(arr[0], n) if P > 0 else None, _copy_f(arr[1], n) if P > 0 else None])
        return p

    def footprint(self) -> int:
        s, keep = self.to_c()
        out = C.c_uint64()
        L.check(L.lib().fdtd2d_plan_footprint(C.byref(s), C.byref(out)), "fdtd2d_plan_footprint")
        return int(out.value)

    def cell_updates(self, steps: Optional[int] = None) -> int:
        return self.nx * self.ny * (self.nsteps if steps is None else steps)


def line_shape(s) -> tuple:
    """(rows, line_size) of the processing-line captures of a problem."""
    if s.line_size <= 0:
        return (0, 0)
    rows = max(0, min(s.line_end, s.nsteps) - s.line_begin)
    return (rows, s.line_size if rows else 0)


class Setup:
    """Owner of a native problem built from Settings.ini, a workload or a box."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @classmethod
    def from_ini(cls, path: str, objects: bool = True) -> "Setup":
        """objects=False: the free-space pass of the source-to-detector regime
        (no chamber figure, no plasma medium)."""
        h = C.c_void_p()
        L.check(L.lib().fdtd2d_setup_from_ini_objects(str(path).encode(), int(objects), C.byref(h)),
                "fdtd2d_setup_from_ini_objects")
        return cls(h.value)

    @classmethod
    def workload(cls, workload: int, seed: int = 0, shot: int = 0, steps: int = 0) -> "Setup":
        h = C.c_void_p()
        L.check(L.lib().fdtd2d_setup_workload(workload, seed, shot, steps, C.byref(h)), "fdtd2d_setup_workload")
        return cls(h.value)

    @classmethod
    def box(cls, nx: int, ny: int, npml: int, nsteps: int, ds: float = 1e-3) -> "Setup":
        h = C.c_void_p()
        L.check(L.lib().fdtd2d_setup_box(nx, ny, npml, nsteps, ds, C.byref(h)), "fdtd2d_setup_box")
        return cls(h.value)

    @property
    def c_problem(self):
        return L.lib().fdtd2d_setup_problem(self._h)

    @property
    def ini_status(self) -> int:
        return L.lib().fdtd2d_setup_ini_status(self._h)

    def problem(self) -> Problem:
        return Problem.from_c(self.c_problem.contents)

    def get(self, name: str) -> float:
        v = C.c_double()
        L.check(L.lib().fdtd2d_setup_get(self._h, name.encode(), C.byref(v)), f"fdtd2d_setup_get({name})")
        return v.value

    def set(self, name: str, value: float) -> None:
        L.check(L.lib().fdtd2d_setup_set(self._h, name.encode(), float(value)), f"fdtd2d_setup_set({name})")

    def snapshot_steps(self, nsteps_run: int) -> np.ndarray:
        """Steps after which [2Dsave] saves Ez, for a run of nsteps_run steps."""
        n = C.c_size_t()
        L.check(L.lib().fdtd2d_snapshot_steps(self._h, nsteps_run, L._i(), 0, C.byref(n)), "fdtd2d_snapshot_steps")
        out = np.zeros(n.value, np.int32)
        L.check(L.lib().fdtd2d_snapshot_steps(self._h, nsteps_run, _ip(out), out.size, C.byref(n)),
                "fdtd2d_snapshot_steps")
        return out

    def write_s2d_outputs(self, out_dir: str, det: np.ndarray, comp: np.ndarray) -> None:
        d = np.ascontiguousarray(det, np.float32)
        c = np.ascontiguousarray(comp, np.float32)
        L.check(L.lib().fdtd2d_write_s2d_outputs(self._h, str(out_dir).encode(), _fp(d), _fp(c)),
                "fdtd2d_write_s2d_outputs")

    def write_s2d_difference(self, out_dir: str, comp: np.ndarray, comp_free: np.ndarray) -> None:
        c = np.ascontiguousarray(comp, np.float32)
        c0 = np.ascontiguousarray(comp_free, np.float32)
        L.check(L.lib().fdtd2d_write_s2d_difference(self._h, str(out_dir).encode(), _fp(c), _fp(c0)),
                "fdtd2d_write_s2d_difference")

    def footprint(self) -> int:
        """Device bytes a plan (or pool shot) of this setup holds."""
        out = C.c_uint64()
        L.check(L.lib().fdtd2d_plan_footprint(self.c_problem, C.byref(out)), "fdtd2d_plan_footprint")
        return int(out.value)

    def processing(self) -> L.ProcessingC:
        """A copy of the setup's [Processing] parameters (cells and steps)."""
        return L.ProcessingC.from_buffer_copy(L.lib().fdtd2d_setup_processing(self._h).contents)

    def write_rc_outputs(self, out_dir: str, spd: np.ndarray) -> None:
        """RC_Line.txt, TAU_Line.txt and RcDet_data.txt from RC tracker traces."""
        v = np.ascontiguousarray(spd, np.int32)
        L.check(L.lib().fdtd2d_write_rc_outputs(self._h, str(out_dir).encode(), v.ctypes.data_as(C.POINTER(C.c_int32))),
                "fdtd2d_write_rc_outputs")

    def write_ini_log(self, out_dir: str) -> None:
        """Appends the INI read's configuration echo to out_dir/iniLOG.txt."""
        L.check(L.lib().fdtd2d_write_ini_log(self._h, str(out_dir).encode()), "fdtd2d_write_ini_log")

    def write_outputs(self, out_dir: str, det: np.ndarray, comp: np.ndarray, seconds: float = 0.0) -> None:
        d = np.ascontiguousarray(det, np.float32)
        c = np.ascontiguousarray(comp, np.float32)
        L.check(L.lib().fdtd2d_write_outputs(self._h, str(out_dir).encode(), _fp(d), _fp(c), seconds),
                "fdtd2d_write_outputs")

    def close(self) -> None:
        if self._h:
            L.lib().fdtd2d_setup_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Snapshots:
    """Ez.txt and its companion files of [2Dsave] (fdtd2d_snapshots_*)."""

    def __init__(self, setup: Setup, out_dir: str):
        h = C.c_void_p()
        L.check(L.lib().fdtd2d_snapshots_open(setup._h, str(out_dir).encode(), C.byref(h)), "fdtd2d_snapshots_open")
        self._h = h

    def add(self, ez: np.ndarray) -> None:
        v = np.ascontiguousarray(ez, np.float32)
        L.check(L.lib().fdtd2d_snapshots_add(self._h, _fp(v), v.size), "fdtd2d_snapshots_add")

    def close(self) -> int:
        n = C.c_int()
        if self._h:
            L.check(L.lib().fdtd2d_snapshots_close(self._h, C.byref(n)), "fdtd2d_snapshots_close")
            self._h = C.c_void_p()
        return n.value


FIELD_SHAPES = {
    L.EZ: lambda nx, ny: (ny, nx),
    L.HX: lambda nx, ny: (ny - 1, nx),
    L.HY: lambda nx, ny: (ny, nx - 1),
    L.JZ: lambda nx, ny: (ny, nx),
}


def psi_shape(which: int, nx: int, ny: int, P: int):
    if which == L.PSI_HXY:
        return (2, P - 1, nx)
    if which == L.PSI_EZY:
        return (2, P, nx)
    if which == L.PSI_HYX:
        return (2, ny, P - 1)
    return (2, ny, P)


class Plan:
    """A problem resident on one GPU (fdtd2d_plan_*)."""

    def __init__(self, problem, device: int = 0, tblock: int = 0):
        if isinstance(problem, Setup):
            self._setup = problem
            cp = problem.c_problem
            self._keep = None
            ptr = cp
            s = cp.contents
        else:
            s, self._keep = problem.to_c()
            ptr = C.byref(s)
            self._setup = None
        self.nx, self.ny, self.npml, self.nsteps = s.nx, s.ny, s.npml, s.nsteps
        self.ndet, self.ncomp, self.nrc = s.ndet, s.ncomp, s.nrc
        self.line_shape = line_shape(s)
        h = C.c_void_p()
        L.check(L.lib().fdtd2d_plan_create(ptr, device, C.byref(h)), "fdtd2d_plan_create")
        self._h = h
        if tblock:
            self.set_tblock(tblock)

    @classmethod
    def in_memory(cls, problem, device: int, mem_ptr: int, nbytes: int, stream_ptr: int = 0) -> "Plan":
        """A plan in device memory the caller owns (fdtd2d_plan_create_in):
        no device allocation, work on the given stream (0: the null stream);
        at most one such plan per device at a time."""
        self = cls.__new__(cls)
        if isinstance(problem, Setup):
            self._setup, self._keep = problem, None
            ptr = problem.c_problem
            s = ptr.contents
        else:
            s, self._keep = problem.to_c()
            ptr = C.byref(s)
            self._setup = None
        self.nx, self.ny, self.npml, self.nsteps = s.nx, s.ny, s.npml, s.nsteps
        self.ndet, self.ncomp, self.nrc = s.ndet, s.ncomp, s.nrc
        self.line_shape = line_shape(s)
        h = C.c_void_p()
        L.check(L.lib().fdtd2d_plan_create_in(ptr, device, C.c_void_p(mem_ptr), nbytes, C.c_void_p(stream_ptr),
                                              C.byref(h)), "fdtd2d_plan_create_in")
        self._h = h
        return self

    def set_tblock(self, t: int) -> None:
        L.check(L.lib().fdtd2d_plan_set_tblock(self._h, t), "fdtd2d_plan_set_tblock")

    def set_option(self, name: str, value: float) -> None:
        L.check(L.lib().fdtd2d_plan_set_option(self._h, name.encode(), float(value)), "fdtd2d_plan_set_option")

    def info(self, name: str) -> float:
        v = C.c_double()
        L.check(L.lib().fdtd2d_plan_get_info(self._h, name.encode(), C.byref(v)), "fdtd2d_plan_get_info")
        return v.value

    def set_stream(self, stream_ptr: int) -> None:
        L.check(L.lib().fdtd2d_plan_set_stream(self._h, C.c_void_p(stream_ptr)), "fdtd2d_plan_set_stream")

    def run(self, nsteps: int) -> None:
        L.check(L.lib().fdtd2d_run(self._h, nsteps), "fdtd2d_run")

    def sync(self) -> None:
        L.check(L.lib().fdtd2d_sync(self._h), "fdtd2d_sync")

    @property
    def steps_done(self) -> int:
        n = C.c_int()
        L.check(L.lib().fdtd2d_steps_done(self._h, C.byref(n)), "fdtd2d_steps_done")
        return n.value

    def read_detectors(self):
        det = np.zeros((self.nsteps + 1, self.ndet), np.float32)
        comp = np.zeros((3, self.nsteps + 1, self.ncomp), np.float32)
        L.check(L.lib().fdtd2d_read_detectors(self._h, _fp(det), det.size, _fp(comp), comp.size),
                "fdtd2d_read_detectors")
        return det, comp

    def read_rc(self) -> np.ndarray:
        """RC tracker traces, (nsteps+1, nrc) int32, row n = step n."""
        out = np.zeros((self.nsteps + 1, self.nrc), np.int32)
        L.check(L.lib().fdtd2d_read_rc(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), out.size), "fdtd2d_read_rc")
        return out

    def read_line(self) -> np.ndarray:
        """Processing line captures, (rows, line_size) float32; row t = step line_begin + t + 1."""
        out = np.zeros(self.line_shape, np.float32)
        L.check(L.lib().fdtd2d_read_line(self._h, _fp(out), out.size), "fdtd2d_read_line")
        return out

    def read_field(self, which: int) -> np.ndarray:
        out = np.zeros(FIELD_SHAPES[which](self.nx, self.ny), np.float32)
        L.check(L.lib().fdtd2d_read_field(self._h, which, _fp(out), out.size), "fdtd2d_read_field")
        return out

    def write_field(self, which: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, np.float32)
        if v.shape != FIELD_SHAPES[which](self.nx, self.ny):
            raise ValueError("field shape mismatch")
        L.check(L.lib().fdtd2d_write_field(self._h, which, _fp(v), v.size), "fdtd2d_write_field")

    def read_psi(self, which: int) -> np.ndarray:
        out = np.zeros(psi_shape(which, self.nx, self.ny, self.npml), np.float32)
        L.check(L.lib().fdtd2d_read_psi(self._h, which, _fp(out), out.size), "fdtd2d_read_psi")
        return out

    def write_psi(self, which: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, np.float32)
        L.check(L.lib().fdtd2d_write_psi(self._h, which, _fp(v), v.size), "fdtd2d_write_psi")

    def time_kernels(self, enable: bool = True) -> None:
        L.check(L.lib().fdtd2d_plan_time_kernels(self._h, 1 if enable else 0), "fdtd2d_plan_time_kernels")

    def kernel_ms(self, cls: int = 0):
        """(average ms, launches timed) of the interior (cls 0) or CPML-class
        (cls 1) step kernel since time_kernels(True)."""
        ms, n = C.c_double(), C.c_int()
        L.check(L.lib().fdtd2d_plan_kernel_time(self._h, cls, C.byref(ms), C.byref(n)), "fdtd2d_plan_kernel_time")
        return ms.value, n.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().fdtd2d_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pool:
    """Many independent shots advanced together (fdtd2d_pool_*).

    Shots may differ in grid size, medium, sources, detectors and step count
    (one medium kind per pool). Every shot's results are bitwise those of a
    Plan of the same problem."""

    def __init__(self, device: int = 0, **options):
        h = C.c_void_p()
        L.check(L.lib().fdtd2d_pool_create(device, C.byref(h)), "fdtd2d_pool_create")
        self._h = h
        self._shots = {}
        if "budget_bytes" in options:  # must precede any shot
            self.set_option("budget_bytes", options.pop("budget_bytes"))
        for k, v in options.items():
            self.set_option(k, v)

    def set_option(self, name: str, value: float) -> None:
        L.check(L.lib().fdtd2d_pool_set_option(self._h, name.encode(), float(value)), "fdtd2d_pool_set_option")

    def info(self, name: str) -> float:
        v = C.c_double()
        L.check(L.lib().fdtd2d_pool_get_info(self._h, name.encode(), C.byref(v)), "fdtd2d_pool_get_info")
        return v.value

    def add(self, problem) -> int:
        """Copies the problem (a Setup or a Problem) into the pool; returns the shot id."""
        if isinstance(problem, Setup):
            ptr = problem.c_problem
            s = ptr.contents
        else:
            s, keep = problem.to_c()
            ptr = C.byref(s)
        sid = C.c_int()
        L.check(L.lib().fdtd2d_pool_add(self._h, ptr, C.byref(sid)), "fdtd2d_pool_add")
        self._shots[sid.value] = (s.nx, s.ny, s.nsteps, s.ndet, s.ncomp, line_shape(s))
        return sid.value

    def step(self, nlaunch: int) -> None:
        L.check(L.lib().fdtd2d_pool_step(self._h, nlaunch), "fdtd2d_pool_step")

    def sync(self) -> None:
        L.check(L.lib().fdtd2d_pool_sync(self._h), "fdtd2d_pool_sync")

    def wait(self, max_pending: int) -> None:
        """Blocks until at most max_pending issued launches are unfinished."""
        L.check(L.lib().fdtd2d_pool_wait(self._h, max_pending), "fdtd2d_pool_wait")

    def state(self, shot: int):
        """(steps done, steps total, complete on the device)."""
        d, t, c = C.c_int(), C.c_int(), C.c_int()
        L.check(L.lib().fdtd2d_pool_shot_state(self._h, shot, C.byref(d), C.byref(t), C.byref(c)),
                "fdtd2d_pool_shot_state")
        return d.value, t.value, bool(c.value)

    def read_detectors(self, shot: int):
        nx, ny, nsteps, ndet, ncomp = self._shots[shot][:5]
        det = np.zeros((nsteps + 1, ndet), np.float32)
        comp = np.zeros((3, nsteps + 1, ncomp), np.float32)
        L.check(L.lib().fdtd2d_pool_read_detectors(self._h, shot, _fp(det), det.size, _fp(comp), comp.size),
                "fdtd2d_pool_read_detectors")
        return det, comp

    def read_field(self, shot: int, which: int) -> np.ndarray:
        nx, ny = self._shots[shot][:2]
        out = np.zeros(FIELD_SHAPES[which](nx, ny), np.float32)
        L.check(L.lib().fdtd2d_pool_read_field(self._h, shot, which, _fp(out), out.size), "fdtd2d_pool_read_field")
        return out

    def read_line(self, shot: int) -> np.ndarray:
        out = np.zeros(self._shots[shot][5], np.float32)
        L.check(L.lib().fdtd2d_pool_read_line(self._h, shot, _fp(out), out.size), "fdtd2d_pool_read_line")
        return out

    def remove(self, shot: int) -> None:
        L.check(L.lib().fdtd2d_pool_remove(self._h, shot), "fdtd2d_pool_remove")
        self._shots.pop(shot, None)

    def close(self) -> None:
        if self._h:
            L.lib().fdtd2d_pool_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
