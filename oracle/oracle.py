"""ctypes binding of the CPU oracle (oracle/oracle.h).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py as the checker. The product (fdtd2d_amd) never imports it.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import fdtd2d_amd._lib as L
from fdtd2d_amd.problem import FIELD_SHAPES, Problem, line_shape, psi_shape

LIB_PATH = Path(__file__).resolve().parent / "_build" / "liboracle.so"
_lib = None
_d = C.POINTER(C.c_double)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} missing: run `python -m fdtd2d_amd.build`")
        O = C.CDLL(str(LIB_PATH))
        sig = [
            ("oracle_create", C.c_int, [C.POINTER(L.ProblemC), C.c_int, C.POINTER(C.c_void_p)]),
            ("oracle_set_threads", C.c_int, [C.c_void_p, C.c_int]),
            ("oracle_run", C.c_int, [C.c_void_p, C.c_int]),
            ("oracle_steps_done", C.c_int, [C.c_void_p]),
            ("oracle_read_field", C.c_int, [C.c_void_p, C.c_int, _d, C.c_size_t]),
            ("oracle_write_field", C.c_int, [C.c_void_p, C.c_int, _d, C.c_size_t]),
            ("oracle_read_psi", C.c_int, [C.c_void_p, C.c_int, _d, C.c_size_t]),
            ("oracle_write_psi", C.c_int, [C.c_void_p, C.c_int, _d, C.c_size_t]),
            ("oracle_read_detectors", C.c_int, [C.c_void_p, _d, C.c_size_t, _d, C.c_size_t]),
            ("oracle_read_rc", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_size_t]),
            ("oracle_read_line", C.c_int, [C.c_void_p, _d, C.c_size_t]),
            ("oracle_processing_tof", C.c_int, [C.POINTER(L.ProcessingC), L._f, C.c_int, L._f, C.c_int, L._f, L._f,
                                                L._f, L._f, L._f]),
            ("oracle_destroy", None, [C.c_void_p]),
            ("oracle_last_error", C.c_char_p, []),
        ]
        for name, res, args in sig:
            fn = getattr(O, name)
            fn.restype = res
            fn.argtypes = args
        _lib = O
    return _lib


def _check(rc: int, where: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{where}: status {rc}: {lib().oracle_last_error().decode()}")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_d)


class Oracle:
    """CPU restatement of the reference step, FP32 (precision=32) or FP64."""

    def __init__(self, problem, precision: int = 32, threads: int = 1):
        if isinstance(problem, Problem):
            s, self._keep = problem.to_c()
            ptr = C.byref(s)
            self._s = s
        else:  # a fdtd2d_amd.Setup
            self._keep = problem
            ptr = problem.c_problem
            s = ptr.contents
        self.nx, self.ny, self.npml, self.nsteps = s.nx, s.ny, s.npml, s.nsteps
        self.ndet, self.ncomp, self.nrc = s.ndet, s.ncomp, s.nrc
        self.line_shape = line_shape(s)
        self.precision = precision
        h = C.c_void_p()
        _check(lib().oracle_create(ptr, precision, C.byref(h)), "oracle_create")
        self._h = h
        if threads > 1:
            _check(lib().oracle_set_threads(self._h, threads), "oracle_set_threads")

    def run(self, nsteps: int) -> None:
        _check(lib().oracle_run(self._h, nsteps), "oracle_run")

    @property
    def steps_done(self) -> int:
        return lib().oracle_steps_done(self._h)

    def read_field(self, which: int) -> np.ndarray:
        out = np.zeros(FIELD_SHAPES[which](self.nx, self.ny), np.float64)
        _check(lib().oracle_read_field(self._h, which, _dp(out), out.size), "oracle_read_field")
        return out

    def write_field(self, which: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, np.float64)
        _check(lib().oracle_write_field(self._h, which, _dp(v), v.size), "oracle_write_field")

    def read_psi(self, which: int) -> np.ndarray:
        out = np.zeros(psi_shape(which, self.nx, self.ny, self.npml), np.float64)
        _check(lib().oracle_read_psi(self._h, which, _dp(out), out.size), "oracle_read_psi")
        return out

    def write_psi(self, which: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, np.float64)
        _check(lib().oracle_write_psi(self._h, which, _dp(v), v.size), "oracle_write_psi")

    def read_detectors(self):
        det = np.zeros((self.nsteps + 1, self.ndet), np.float64)
        comp = np.zeros((3, self.nsteps + 1, self.ncomp), np.float64)
        _check(lib().oracle_read_detectors(self._h, _dp(det), det.size, _dp(comp), comp.size),
               "oracle_read_detectors")
        return det, comp

    def read_rc(self) -> np.ndarray:
        """RC tracker traces, (nsteps+1, nrc) int32, row n = step n."""
        out = np.zeros((self.nsteps + 1, self.nrc), np.int32)
        _check(lib().oracle_read_rc(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), out.size), "oracle_read_rc")
        return out

    def read_line(self) -> np.ndarray:
        """Processing line captures, (rows, line_size), row t = step line_begin + t + 1."""
        out = np.zeros(self.line_shape, np.float64)
        _check(lib().oracle_read_line(self._h, _dp(out), out.size), "oracle_read_line")
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def processing_tof(a, line: np.ndarray, pass_: int = 0, weights=None):
    """The oracle's processing_core pass; same outputs as fdtd2d_amd.processing.tof,
    or None when the oracle rejects the inputs (FDTD2D_EINVAL)."""
    line = np.ascontiguousarray(line, np.float32)
    times = line.shape[0]
    n = max(a.size_i if pass_ == 0 else a.size_i - a.detector_size, 0)
    out = {k: np.zeros(n, np.float32) for k in ("sd", "cf", "arc")}
    out.update({k: np.zeros((n, times), np.float32) for k in ("original", "filtered")})
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    fp = lambda x: x.ctypes.data_as(L._f) if x is not None else L._f()  # noqa: E731
    rc = lib().oracle_processing_tof(C.byref(a), fp(line), times, fp(w), pass_, fp(out["sd"]), fp(out["cf"]),
                                     fp(out["arc"]), fp(out["original"]), fp(out["filtered"]))
    return out if rc == 0 else None
