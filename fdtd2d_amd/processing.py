"""Time-of-flight search over the processing line: the reference's
processing_core (F/processing.cu:6-144), run by the native library on the
host over the line the stepper captured (Plan.read_line / Pool.read_line).
"""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L

# [Processing] defaults of the reference's INI reader (F/FDTD_cuda.cu:144-172)
DEFAULTS = dict(perform=0, sec_der=0, cf=0, arc=0, save_filtered=0, save_original=0, sim_detect=0,
                save_detector=0, load_weights=0, detector_size=30, filter_type=1, dim_sieve=1, time_sieve=1,
                begin_after=0, finish_at=0, pos_i=0, pos_j=0, pos_k=0, size_i=0, filter_win=300, filter_runs=1,
                cf_tau=300, arc_tau=300, tau_threshold=0.0, arc_k=-0.5, cf_k=-0.5, times=0, dt=0.0)


def config(**values) -> L.ProcessingC:
    """A parameter struct: the reference defaults, SDF* following the
    prefilter's values unless given, then `values`."""
    v = dict(DEFAULTS)
    v.update(values)
    for k, src in (("sdf_win", "filter_win"), ("sdf_type", "filter_type"), ("sdf_runs", "filter_runs")):
        v.setdefault(k, v[src])
    a = L.ProcessingC()
    for k, x in v.items():
        setattr(a, k, x)
    return a


def _outputs(a: L.ProcessingC, times: int, pass_: int):
    n = a.size_i if pass_ == 0 else a.size_i - a.detector_size
    n = max(n, 0)
    return {k: np.zeros(n, np.float32) for k in ("sd", "cf", "arc")} | {
        k: np.zeros((n, times), np.float32) for k in ("original", "filtered")}


def tof(a: L.ProcessingC, line: np.ndarray, pass_: int = 0, weights: Optional[np.ndarray] = None) -> dict:
    """sd, cf, arc per coordinate and the original / filtered series
    ([coordinate, step]) of one pass (0: the line, 1: the simulated detector)."""
    line = np.ascontiguousarray(line, np.float32)
    times = line.shape[0]
    out = _outputs(a, times, pass_)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    fp = lambda x: x.ctypes.data_as(L._f) if x is not None else L._f()  # noqa: E731
    L.check(L.lib().fdtd2d_processing_tof(C.byref(a), fp(line), times, fp(w), pass_, fp(out["sd"]), fp(out["cf"]),
                                          fp(out["arc"]), fp(out["original"]), fp(out["filtered"])),
            "fdtd2d_processing_tof")
    return out


def write_outputs(a: L.ProcessingC, out_dir: str, line: np.ndarray) -> None:
    """SD/CF/ARC.txt, the field files and the detector-simulation files."""
    line = np.ascontiguousarray(line, np.float32)
    L.check(L.lib().fdtd2d_write_processing_outputs(C.byref(a), str(out_dir).encode(), line.ctypes.data_as(L._f),
                                                    line.shape[0]), "fdtd2d_write_processing_outputs")
