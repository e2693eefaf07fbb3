"""ctypes binding of the C ABI declared in include/fdtd2d.h.

The shared library is built in-tree (``python -m fdtd2d_amd.build``) and
loaded from ``fdtd2d_amd/lib/libfdtd2d.so``. There is no fallback: if the
library is missing, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfdtd2d.so"

OK, EINVAL, ENOMEM, EDEVICE, EIO, ESTATE, EUNSUPPORTED = range(7)
MEDIUM_TABLE, MEDIUM_DENSITY = 0, 1
EZ, HX, HY, JZ = 0, 1, 2, 3
PSI_HXY, PSI_HYX, PSI_EZX, PSI_EZY = 0, 1, 2, 3

_f = C.POINTER(C.c_float)
_i = C.POINTER(C.c_int32)


class ProblemC(C.Structure):
    """Mirror of ``struct fdtd2d_problem``."""

    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("npml", C.c_int32),
        ("nsteps", C.c_int32),
        ("dt", C.c_double),
        ("dx", C.c_double),
        ("dy", C.c_double),
        ("db", C.c_float),
        ("medium_kind", C.c_int32),
        ("material", _i),
        ("nmaterials", C.c_int32),
        ("ca", _f),
        ("cb", _f),
        ("ne", _f),
        ("pec_material", C.c_int32),
        ("density", _f),
        ("ca0", C.c_float),
        ("cb0", C.c_float),
        ("cj", C.c_float),
        ("jz_decay", C.c_float),
        ("jz_dt", C.c_double),
        ("be_x", _f * 2),
        ("ce_x", _f * 2),
        ("bh_x", _f * 2),
        ("ch_x", _f * 2),
        ("be_y", _f * 2),
        ("ce_y", _f * 2),
        ("bh_y", _f * 2),
        ("ch_y", _f * 2),
        ("den_ex", _f),
        ("den_hx", _f),
        ("den_ey", _f),
        ("den_hy", _f),
        ("dx_f", C.c_float),
        ("dy_f", C.c_float),
        ("nsrc", C.c_int32),
        ("src_i", _i),
        ("src_j", _i),
        ("src_table", _f),
        ("ndet", C.c_int32),
        ("det_i", _i),
        ("det_j", _i),
        ("ncomp", C.c_int32),
        ("comp_i", _i),
        ("comp_j", _i),
        ("nrc", C.c_int32),
        ("rc_i", _i),
        ("rc_imax", _i),
        ("rc_jcut", C.c_int32),
        ("rc_level", C.c_float),
        ("rc_edge0", C.c_float),
        ("line_size", C.c_int32),
        ("line_i0", C.c_int32),
        ("line_j", C.c_int32),
        ("line_begin", C.c_int32),
        ("line_end", C.c_int32),
    ]


class ProcessingC(C.Structure):
    """Mirror of ``struct fdtd2d_processing``."""

    _fields_ = [
        ("perform", C.c_int32),
        ("sec_der", C.c_int32),
        ("sdf_type", C.c_int32),
        ("sdf_win", C.c_int32),
        ("sdf_runs", C.c_int32),
        ("cf", C.c_int32),
        ("cf_tau", C.c_int32),
        ("cf_k", C.c_float),
        ("arc", C.c_int32),
        ("arc_tau", C.c_int32),
        ("arc_k", C.c_float),
        ("tau_threshold", C.c_float),
        ("filter_type", C.c_int32),
        ("filter_win", C.c_int32),
        ("filter_runs", C.c_int32),
        ("save_original", C.c_int32),
        ("save_filtered", C.c_int32),
        ("dim_sieve", C.c_int32),
        ("time_sieve", C.c_int32),
        ("sim_detect", C.c_int32),
        ("save_detector", C.c_int32),
        ("load_weights", C.c_int32),
        ("detector_size", C.c_int32),
        ("begin_after", C.c_int32),
        ("finish_at", C.c_int32),
        ("pos_i", C.c_int32),
        ("pos_j", C.c_int32),
        ("pos_k", C.c_int32),
        ("size_i", C.c_int32),
        ("times", C.c_int32),
        ("dt", C.c_double),
    ]


class FDTD2DError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str):
        super().__init__(f"{where}: status {code}: {detail}")
        self.code = code


_lib = None

# (name, restype, argtypes) for every function declared in include/fdtd2d.h
SIGNATURES = [
    ("fdtd2d_strerror", C.c_char_p, [C.c_int]),
    ("fdtd2d_last_error", C.c_char_p, []),
    ("fdtd2d_abi_version", C.c_int, []),
    ("fdtd2d_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("fdtd2d_setup_from_ini", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("fdtd2d_setup_from_ini_objects", C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]),
    ("fdtd2d_setup_set", C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    ("fdtd2d_write_s2d_outputs", C.c_int, [C.c_void_p, C.c_char_p, _f, _f]),
    ("fdtd2d_write_s2d_difference", C.c_int, [C.c_void_p, C.c_char_p, _f, _f]),
    ("fdtd2d_snapshot_steps", C.c_int, [C.c_void_p, C.c_int, _i, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("fdtd2d_snapshots_open", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("fdtd2d_snapshots_add", C.c_int, [C.c_void_p, _f, C.c_size_t]),
    ("fdtd2d_snapshots_close", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("fdtd2d_setup_ini_status", C.c_int, [C.c_void_p]),
    ("fdtd2d_ini_lookup", C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_double,
                                    C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    ("fdtd2d_setup_workload", C.c_int, [C.c_int, C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("fdtd2d_setup_box", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_void_p)]),
    ("fdtd2d_setup_problem", C.POINTER(ProblemC), [C.c_void_p]),
    ("fdtd2d_setup_problem_mut", C.POINTER(ProblemC), [C.c_void_p]),
    ("fdtd2d_setup_get", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    ("fdtd2d_setup_destroy", None, [C.c_void_p]),
    ("fdtd2d_write_outputs", C.c_int, [C.c_void_p, C.c_char_p, _f, _f, C.c_double]),
    ("fdtd2d_write_ini_log", C.c_int, [C.c_void_p, C.c_char_p]),
    ("fdtd2d_write_rc_outputs", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int32)]),
    ("fdtd2d_read_rc", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_size_t]),
    ("fdtd2d_read_line", C.c_int, [C.c_void_p, _f, C.c_size_t]),
    ("fdtd2d_setup_processing", C.POINTER(ProcessingC), [C.c_void_p]),
    ("fdtd2d_processing_tof", C.c_int, [C.POINTER(ProcessingC), _f, C.c_int, _f, C.c_int, _f, _f, _f, _f, _f]),
    ("fdtd2d_write_processing_outputs", C.c_int, [C.POINTER(ProcessingC), C.c_char_p, _f, C.c_int]),
    ("fdtd2d_plan_create", C.c_int, [C.POINTER(ProblemC), C.c_int, C.POINTER(C.c_void_p)]),
    ("fdtd2d_plan_create_in", C.c_int, [C.POINTER(ProblemC), C.c_int, C.c_void_p, C.c_size_t, C.c_void_p,
                                        C.POINTER(C.c_void_p)]),
    ("fdtd2d_plan_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("fdtd2d_plan_set_tblock", C.c_int, [C.c_void_p, C.c_int]),
    ("fdtd2d_plan_set_option", C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    ("fdtd2d_plan_get_info", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    ("fdtd2d_run", C.c_int, [C.c_void_p, C.c_int]),
    ("fdtd2d_sync", C.c_int, [C.c_void_p]),
    ("fdtd2d_steps_done", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("fdtd2d_read_detectors", C.c_int, [C.c_void_p, _f, C.c_size_t, _f, C.c_size_t]),
    ("fdtd2d_read_field", C.c_int, [C.c_void_p, C.c_int, _f, C.c_size_t]),
    ("fdtd2d_write_field", C.c_int, [C.c_void_p, C.c_int, _f, C.c_size_t]),
    ("fdtd2d_read_psi", C.c_int, [C.c_void_p, C.c_int, _f, C.c_size_t]),
    ("fdtd2d_write_psi", C.c_int, [C.c_void_p, C.c_int, _f, C.c_size_t]),
    ("fdtd2d_plan_footprint", C.c_int, [C.POINTER(ProblemC), C.POINTER(C.c_uint64)]),
    ("fdtd2d_tile_lists", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                    C.POINTER(C.c_int), C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_int)]),
    ("fdtd2d_plan_destroy", C.c_int, [C.c_void_p]),
    ("fdtd2d_plan_time_kernels", C.c_int, [C.c_void_p, C.c_int]),
    ("fdtd2d_plan_kernel_ms", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    ("fdtd2d_plan_kernel_time", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    ("fdtd2d_pool_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("fdtd2d_pool_set_option", C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    ("fdtd2d_pool_get_info", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    ("fdtd2d_pool_add", C.c_int, [C.c_void_p, C.POINTER(ProblemC), C.POINTER(C.c_int)]),
    ("fdtd2d_pool_step", C.c_int, [C.c_void_p, C.c_int]),
    ("fdtd2d_pool_sync", C.c_int, [C.c_void_p]),
    ("fdtd2d_pool_wait", C.c_int, [C.c_void_p, C.c_int]),
    ("fdtd2d_pool_shot_state", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]),
    ("fdtd2d_pool_read_detectors", C.c_int, [C.c_void_p, C.c_int, _f, C.c_size_t, _f, C.c_size_t]),
    ("fdtd2d_pool_read_field", C.c_int, [C.c_void_p, C.c_int, C.c_int, _f, C.c_size_t]),
    ("fdtd2d_pool_read_line", C.c_int, [C.c_void_p, C.c_int, _f, C.c_size_t]),
    ("fdtd2d_pool_remove", C.c_int, [C.c_void_p, C.c_int]),
    ("fdtd2d_pool_destroy", C.c_int, [C.c_void_p]),
]


def lib() -> C.CDLL:
    """Loads libfdtd2d.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -m fdtd2d_amd.build` "
                "(the HIP stepper has no CPU fallback)"
            )
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ini_lookup(text: str, section: str, name: str, kind: int = 1, default: float = 0.0):
    """(value, parse_status) of one key, with the reference INIReader's rules."""
    v, st = C.c_double(), C.c_int()
    check(lib().fdtd2d_ini_lookup(text.encode("latin-1"), section.encode(), name.encode(), kind, default,
                                  C.byref(v), C.byref(st)), "fdtd2d_ini_lookup")
    return v.value, st.value


def tile_lists(nx: int, ny: int, npml: int, tblock: int = 0, cols_per_lane: int = 1, band: int = 0,
               band_cpml: int = 0, events=()):
    """(entries, geometry): the tile partition a plan of this shape launches.

    events are (i, j) cells of sources, detectors and probes. entries is a
    list of (first column, end column, first row, end row, class), class 0
    interior, 1 general CPML-class, 2 x-strip, 3 y-strip, 4 corner, 5 event
    tiles; geometry
    the dict {vw, halo, band, band_cpml} of the interior tiles. Host-only;
    needs no GPU."""
    n = C.c_size_t()
    g = (C.c_int * 4)()
    L = lib()
    ne = len(events)
    ei = (C.c_int * max(ne, 1))(*[e[0] for e in events])
    ej = (C.c_int * max(ne, 1))(*[e[1] for e in events])
    args = (nx, ny, npml, tblock, cols_per_lane, band, band_cpml, ei, ej, ne)
    check(L.fdtd2d_tile_lists(*args, None, 0, C.byref(n), g), "fdtd2d_tile_lists")
    buf = (C.c_int * (5 * max(n.value, 1)))()
    check(L.fdtd2d_tile_lists(*args, buf, n.value, C.byref(n), g), "fdtd2d_tile_lists")
    ent = [tuple(buf[5 * k:5 * k + 5]) for k in range(n.value)]
    return ent, {"vw": g[0], "halo": g[1], "band": g[2], "band_cpml": g[3]}


def device_count() -> int:
    n = C.c_int()
    rc = lib().fdtd2d_device_count(C.byref(n))
    return n.value if rc == OK else 0


def check(rc: int, where: str) -> None:
    if rc != OK:
        L = lib()
        raise FDTD2DError(rc, where, f"{L.fdtd2d_strerror(rc).decode()} ({L.fdtd2d_last_error().decode()})")
