"""Serving reflectometry shots as they arrive: continuous batching on one pool.

Requests arrive on a seeded Poisson stream. A request is one or more shots
(a frequency sweep over one plasma profile, or a single shot); each shot's
problem is built on host worker threads after its request arrives, at most
`max_queued` shots ahead of admission (which bounds host memory). Built
shots wait in a FIFO; a shot is admitted into the pool as soon as its
byte-accurate footprint (fields, Jz, CPML psi strips, coefficient tables,
traces: the plan's own allocation) fits the pool's arena. When the head of
the queue does not fit, later shots that do are admitted around it
(backfill). Admitted shots join the running launches at the next launch
boundary, so their work overlaps on the device with everything already
running; finished shots are read back and removed, freeing their bytes for
the queue. A request completes when the results of all its shots are read.
With the pool's time-slicing options (slice_cells, slice_every) large shots
run only in every slice_every-th launch, so a small shot admitted beside
them is not held to a large shot's launch time at every one of its steps.

Latency = completion - arrival, reported as p50/p99 over the completed
requests (per size class where the stream has classes), next to completed
shots per second and cell updates per second.
"""

from __future__ import annotations

import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .problem import Pool, Setup
from .tuning import mix_key


@dataclass
class Request:
    rid: int
    arrival: float           # seconds from the start of the run
    shots: list              # (workload, seed, shot index, steps override)
    size_class: str = ""
    done: int = 0
    finished: float = -1.0
    results: dict = field(default_factory=dict)  # shot position -> detector traces


def poisson_arrivals(rate: float, duration: float, seed: int):
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return out
        out.append(t)


def sweep_requests(rate: float, duration: float, seed: int = 4, steps: int = 0):
    """Served frequency sweeps: request r is shots r*64 + k, k < nfreq, of the
    sweep workload (one grid and plasma per request, 8-32 frequencies)."""
    reqs = []
    for r, t in enumerate(poisson_arrivals(rate, duration, seed)):
        head = Setup.workload(4, seed=seed, shot=r * 64, steps=-1)  # the draws only
        nfreq = int(head.get("nfreq"))
        head.close()
        reqs.append(Request(r, t, [(4, seed, r * 64 + k, steps) for k in range(nfreq)], "sweep"))
    return reqs


def mixed_requests(rate: float, duration: float, seed: int = 5, steps: int = 0, big_steps: int = 0):
    """Mixed-size stream: 90% 1024^2 single shots, 10% 16384^2 shots."""
    reqs = []
    for r, t in enumerate(poisson_arrivals(rate, duration, seed)):
        head = Setup.workload(5, seed=seed, shot=r, steps=-1)  # the draws only
        big = head.get("big") > 0
        head.close()
        st = (big_steps or steps) if big else steps
        reqs.append(Request(r, t, [(5, seed, r, st)], "16384" if big else "1024"))
    return reqs


def request_mix_key(requests, sample: int = 512) -> str:
    """tuning.mix_key of the shots of (up to `sample` shots of) a request
    stream, from the workload draws alone (no arrays are built)."""
    sizes = []
    for r in requests:
        for w, seed, shot, steps in r.shots:
            if len(sizes) >= sample:
                break
            head = Setup.workload(w, seed=seed, shot=shot, steps=-1)
            n = int(head.get("nMax")) if steps <= 0 else steps
            sizes.append((int(head.get("Imax")), int(head.get("Jmax")), n, 1))  # the density medium
            head.close()
    return mix_key(sizes)


class Server:
    """Continuous-batching loop over one pool (see the module docstring)."""

    def __init__(self, device: int = 0, budget_bytes: int = 160 << 30, chunk_launches: int = 8,
                 max_in_flight: int = 16, builders: int = 8, max_queued: int = 64, backfill: bool = True,
                 **shape):
        self.pool = Pool(device, budget_bytes=budget_bytes, graph_chunk=min(chunk_launches, 16), **shape)
        self.chunk = chunk_launches
        self.max_in_flight = max_in_flight
        self.builders = builders
        self.max_queued = max_queued  # shots built or being built ahead of admission (bounds host memory)
        self.backfill = backfill
        self.keep_results = False

    def close(self):
        self.pool.close()

    def run(self, requests, max_seconds: float = 600.0, log=None):
        pool = self.pool
        pending = deque(sorted(requests, key=lambda q: q.arrival))
        to_build = deque()  # (request, position) arrived, not yet handed to a builder
        building = deque()  # (future, request, position) in arrival order
        ready = deque()     # (setup, bytes, request, position)
        running = {}        # shot id -> (request, position)
        done = []
        ex = ThreadPoolExecutor(max_workers=self.builders)
        t0 = time.perf_counter()
        cells0 = pool.info("cell_updates")
        n_shots = 0
        last_log = 0.0
        host = {"admit": 0.0, "step": 0.0, "wait": 0.0, "poll": 0.0, "complete": 0.0}
        clock = time.perf_counter
        try:
            while True:
                now = clock() - t0
                # arrivals -> problem builds
                while pending and pending[0].arrival <= now:
                    rq = pending.popleft()
                    to_build.extend((rq, pos) for pos in range(len(rq.shots)))
                while to_build and len(building) + len(ready) < self.max_queued:
                    rq, pos = to_build.popleft()
                    w, sd, sh, st = rq.shots[pos]
                    building.append((ex.submit(Setup.workload, w, sd, sh, st), rq, pos))
                while building and building[0][0].done():
                    fut, rq, pos = building.popleft()
                    s = fut.result()
                    ready.append((s, s.footprint(), rq, pos))
                # admission: FIFO, with backfill around a head that does not fit
                if ready:
                    ta = clock()
                    free = pool.info("largest_free")
                    keep = deque()
                    blocked = False
                    while ready:
                        s, nb, rq, pos = ready.popleft()
                        if (blocked and not self.backfill) or nb > free:
                            keep.append((s, nb, rq, pos))
                            blocked = True
                            continue
                        sid = pool.add(s)
                        s.close()
                        running[sid] = (rq, pos)
                        free = pool.info("largest_free")
                    ready = keep
                    host["admit"] += clock() - ta
                # device work: keep at most max_in_flight launches queued ahead
                if running:
                    ta = clock()
                    pool.step(self.chunk)
                    tb = clock()
                    pool.wait(self.max_in_flight)
                    host["step"] += tb - ta
                    host["wait"] += clock() - tb
                # completions
                ta = clock()
                finished = [sid for sid in running if pool.state(sid)[2]]
                tb = clock()
                host["poll"] += tb - ta
                for sid in finished:
                    rq, pos = running.pop(sid)
                    det, _ = pool.read_detectors(sid)
                    pool.remove(sid)
                    n_shots += 1
                    if self.keep_results:
                        rq.results[pos] = det
                    rq.done += 1
                    if rq.done == len(rq.shots):
                        rq.finished = clock() - t0
                        done.append(rq)
                host["complete"] += clock() - tb
                if not pending and not to_build and not building and not ready and not running:
                    break
                if now > max_seconds:
                    break
                if not running:
                    time.sleep(0.001)
                if log and now - last_log > 10:
                    last_log = now
                    log(f"t={now:.0f}s done {len(done)} requests, running {len(running)} shots, "
                        f"queued {len(ready)}, building {len(building)}, waiting {len(to_build)}, "
                        f"{(pool.info('cell_updates') - cells0) / max(now, 1e-9) / 1e9:.0f} Gcell/s")
            pool.sync()
        finally:
            ex.shutdown(wait=True, cancel_futures=True)
        wall = clock() - t0
        return done, {"wall": wall, "shots": n_shots, "cell_updates": pool.info("cell_updates") - cells0,
                      "unfinished_requests": len(requests) - len(done),
                      "host_s": {k: round(v, 3) for k, v in host.items()},
                      "pool_host_s": {k: round(pool.info(k), 3) for k in ("add_s", "rebuild_s", "capture_s", "blocked_s")},
                      "pool_counts": {k: int(pool.info(k)) for k in ("launches", "graph_replays", "eager_launches",
                                                                      "tail_launches", "rebuilds", "list_switches",
                                                                      "captures")}}


def latency_summary(done, classes=None):
    """p50/p99 latency (s) of completed requests, overall and per size class."""
    def pct(v):
        if not v:
            return None
        a = np.sort(np.asarray(v))
        return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99)), "n": len(a)}

    out = {"all": pct([r.finished - r.arrival for r in done])}
    for c in classes or sorted({r.size_class for r in done}):
        out[c] = pct([r.finished - r.arrival for r in done if r.size_class == c])
    return out
