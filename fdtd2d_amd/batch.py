"""Batched shots: building a shot mix, running it through one pool, and the
launch-shape autotuner.

The autotuner tunes a live pool by coordinate descent over the launch shape
(one fused kernel or two per launch, steps per launch, columns per lane,
band rows, CPML sub-band rows, persistent work-groups per CU of each kernel,
prefetch depth). Every trial advances the real shots -
results are bitwise independent of the shape - and is scored by the cell
updates it completed per second of wall time, measured with the pool synced
on both sides. Shapes found are cached per shot mix (tuning.ShapeCache), and
a pool whose mix is in the cache takes that shape with no trials.
"""

from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

from .problem import Pool, Setup
from .tuning import ShapeCache, mix_key

# coordinate-descent axes, in the order they are tuned; the first value of
# each is the starting point (the shape that won on uneven shot mixes)
TUNE_AXES = [
    ("tblock", [8, 6, 7]),
    ("cols_per_lane", [1, 2]),
    ("band", [256, 128, 64]),
    ("wg_per_cu", [3, 4]),
    ("band_cpml", [32, 16, 64]),
    ("wg_per_cu_cpml", [3, 1, 2]),
    ("prefetch", [1, 2]),
]


def build_setups(workload: int, seed: int, shots, threads: int = 16, steps: int = 0):
    """Setups of the given shots of a workload, built on host threads (the
    builders run in C++ and release the interpreter lock)."""
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda s: Setup.workload(workload, seed=seed, shot=s, steps=steps), shots))


def _trial(pool: Pool, launches: int) -> float:
    pool.sync()
    c0 = pool.info("cell_updates")
    t0 = time.perf_counter()
    pool.step(launches)
    pool.sync()
    dt = time.perf_counter() - t0
    return (pool.info("cell_updates") - c0) / dt if dt > 0 else 0.0


def _descend(pool: Pool, start: dict, axes, launches_per_trial: int, trials: list, log=None):
    best = dict(start)
    for name, value in best.items():
        pool.set_option(name, value)
    best_rate = _trial(pool, launches_per_trial)
    trials.append((dict(best), best_rate))
    for name, values in axes:
        for v in values:
            if v == best[name] or pool.info("running") == 0:
                continue
            pool.set_option(name, v)
            rate = _trial(pool, launches_per_trial)
            trials.append(({**best, name: v}, rate))
            if log:
                log(f"autotune fused={best.get('fused')} {name}={v}: {rate / 1e9:.1f} Gcell/s "
                    f"(best {best_rate / 1e9:.1f})")
            if rate > best_rate:
                best_rate, best[name] = rate, v
        pool.set_option(name, best[name])
    return best, best_rate


def autotune(pool: Pool, launches_per_trial: int = 4, axes=TUNE_AXES, log=None, margin: float = 0.95) -> dict:
    """Picks the launch shape for the shots now in the pool; returns it with
    the score of every trial. A coordinate descent with one kernel per tile
    class, then one trial of the fused kernel at the shape found, descended
    from only when it scores within `margin` of the best (every trial runs
    the real shots, so a descent at a losing arrangement costs throughput).
    Stops early when the pool runs out of work."""
    trials = []
    found = []
    if pool.info("running") > 0:
        start = {"fused": 0, **{name: values[0] for name, values in axes}}
        for name, value in start.items():
            pool.set_option(name, value)
        # the first launches after the shots are admitted run slower (graph
        # capture, cold caches): run them before scoring the starting shape
        _trial(pool, launches_per_trial)
        found.append(_descend(pool, start, axes, launches_per_trial, trials, log))
    if found and pool.info("running") > 0:
        shape = {**found[0][0], "fused": 1}
        for name, value in shape.items():
            pool.set_option(name, value)
        rate = _trial(pool, launches_per_trial)
        trials.append((dict(shape), rate))
        if log:
            log(f"autotune fused=1 at that shape: {rate / 1e9:.1f} Gcell/s (best {found[0][1] / 1e9:.1f})")
        if rate >= margin * found[0][1]:
            found.append(_descend(pool, shape, axes, launches_per_trial, trials, log))
    if not found:
        return {"shape": {}, "rate": 0.0, "trials": trials}
    best, best_rate = max(found, key=lambda f: f[1])
    for name, value in best.items():
        pool.set_option(name, value)
    return {"shape": best, "rate": best_rate, "trials": trials}


def shot_sizes(setups):
    """(nx, ny, nsteps, medium_kind) of each setup, for mix_key."""
    out = []
    for s in setups:
        p = s.c_problem.contents
        out.append((p.nx, p.ny, p.nsteps, p.medium_kind))
    return out


def cached_shape(pool, key: str, cache: ShapeCache | None, log=None) -> dict | None:
    """Sets the shape cached for the mix `key` on the pool (best before its
    shots are added: a shape change retiles every shot) and returns its
    record, or None on a miss."""
    if cache is None or not key:
        return None
    shape, src = cache.lookup("pool", key, int(pool.info("cus")))
    if shape is None:
        return None
    for name, value in shape.items():
        pool.set_option(name, value)
    if log:
        log(f"shape for mix {key} from {src}: {shape}")
    return {"shape": shape, "rate": None, "trials": [], "source": src, "key": key}


def choose_shape(pool, key: str, cache: ShapeCache | None, log=None, lookup: bool = True) -> dict:
    """The launch shape for the shots now in the pool: the cached one for
    their mix when there is one (no trials; lookup=False skips the cache),
    else the autotuner's, which is then stored. Returns the autotune record
    with "source" (the cache file or "trials") and "key"."""
    hit = cached_shape(pool, key, cache, log) if lookup else None
    if hit is not None:
        return hit
    tuned = autotune(pool, log=log)
    tuned["source"] = "trials"
    tuned["key"] = key
    if cache is not None and key and tuned["shape"]:
        cache.store("pool", key, int(pool.info("cus")), tuned["shape"], tuned["rate"])
    return tuned


def warm_up(setup, device: int, shape: dict, launches: int = 16, **options) -> None:
    """Runs one shot for a few launches in a pool of its own at `shape`, so
    that the kernels' code objects are loaded and a graph has been captured
    before a timed run."""
    with Pool(device, **options) as pool:
        for name, value in shape.items():
            pool.set_option(name, value)
        pool.add(setup)
        pool.step(launches)
        pool.sync()


def run_pool(setups, device: int = 0, tune: bool = True, chunk_launches: int = 8, log=None,
             cache: ShapeCache | None = None, lookup: bool = True, warmup: bool = False, **options):
    """All setups through one pool to completion. Returns (pool, ids, stats);
    the caller reads results and closes the pool. tune: take the shape from
    `cache` (a hit, set before the shots are added) or the autotuner (a
    miss, or lookup=False; stored in `cache`). warmup: with a cached shape,
    run the first setup briefly in a pool of its own first (untimed). The
    time in stats["seconds"] runs from the shots being resident to the last
    launch's end, trials included."""
    pool = Pool(device, **options)
    key = mix_key(shot_sizes(setups))
    tuned = cached_shape(pool, key, cache, log) if tune and lookup else None
    if warmup and tuned is not None and setups:
        warm_up(setups[0], device, tuned["shape"], **options)
    ids = [pool.add(s) for s in setups]
    pool.sync()
    t0 = time.perf_counter()
    if tune and tuned is None:
        tuned = choose_shape(pool, key, cache, log=log, lookup=False)
    while pool.info("running") > 0:
        pool.step(chunk_launches)
    pool.sync()
    dt = time.perf_counter() - t0
    stats = {"seconds": dt, "cell_updates": pool.info("cell_updates"), "tuned": tuned,
             "launches": pool.info("launches"), "graph_replays": pool.info("graph_replays"),
             "eager_launches": pool.info("eager_launches"), "tail_launches": pool.info("tail_launches"),
             "warmup": bool(warmup and tuned is not None and tuned["source"] != "trials")}
    return pool, ids, stats
