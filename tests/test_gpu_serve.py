"""The continuous-batching server on the GPU: every shot's result is the same
alone and in the mix (bit for bit), under memory pressure that forces
queueing and backfill."""

from __future__ import annotations

import numpy as np
import pytest

import fdtd2d_amd as F
from fdtd2d_amd.serve import Server, latency_summary, sweep_requests

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(gpu_device):
    reqs = sweep_requests(rate=30.0, duration=0.2, seed=11, steps=40)
    assert len(reqs) >= 2
    nshots = sum(len(r.shots) for r in reqs)
    srv = Server(gpu_device, budget_bytes=600_000_000, chunk_launches=2, max_in_flight=4, builders=4,
                 max_queued=16, tblock=8)
    srv.keep_results = True
    try:
        done, st = srv.run(reqs, max_seconds=300)
    finally:
        srv.close()
    print(f"{len(reqs)} requests, {nshots} shots, {st}", latency_summary(done), flush=True)
    assert len(done) == len(reqs) and st["shots"] == nshots
    for rq in reqs:
        for pos, (w, sd, sh, steps) in enumerate(rq.shots):
            s = F.Setup.workload(w, seed=sd, shot=sh, steps=steps)
            with F.Plan(s, gpu_device) as plan:
                plan.run(steps)
                plan.sync()
                alone, _ = plan.read_detectors()
            mixed = rq.results[pos]
            assert np.array_equal(alone.view(np.uint32), mixed.view(np.uint32)), (rq.rid, pos)
