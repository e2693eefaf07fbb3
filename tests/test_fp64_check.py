"""The FP64-oracle check every GPU parity test reports (helpers.fp64_check):
it holds FP32 traces that carry signal to rel L2 <= 1e-5, and only prints
the error of traces that hold just the numerical precursor of a wave that
has not arrived (peak below 1% of the source peak). The FP32 oracle's own
traces stand in for the device's (the GPU tests show them bitwise equal)."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import fp64_check, random_problem, source_peak
from oracle.oracle import Oracle


def _fp32_traces(p):
    r = Oracle(p, 32, threads=4)
    r.run(p.nsteps)
    det, comp = r.read_detectors()
    r.close()
    return det, comp


def _problem(seed, k):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k + 1):
        nx, ny = (int(rng.integers(80, 330)) for _ in range(2))
        out.append(random_problem(rng, nx=nx, ny=ny, npml=int(rng.choice([0, 5, 8, 12])),
                                  nsteps=int(rng.integers(45, 90)), table=True))
    return out[k]


def test_signal_traces_are_held_to_1e5():
    p = _problem(1982, 0)
    det, comp = _fp32_traces(p)
    assert np.abs(det).max() > 1e-2 * source_peak(p)
    assert fp64_check(p, det, "fp32 oracle", comp32=comp) <= 1e-5
    # a relative error of 1e-4 on the traces fails it
    bad = det * np.float32(1 + 1e-4)
    with pytest.raises(AssertionError):
        fp64_check(p, bad, "perturbed")


def test_precursor_traces_are_reported_only(capsys):
    p = _problem(1982, 6)  # 81 x 204, detectors the wave has not reached in 73 steps
    det, comp = _fp32_traces(p)
    assert 0 < np.abs(det).max() < 1e-2 * source_peak(p)
    err = fp64_check(p, det, "precursor", comp32=comp)
    assert err > 1e-5
    assert "not asserted" in capsys.readouterr().out
