"""Launch-shape cache: the shapes the autotuner found, keyed by shot mix.

A pool's best launch shape depends on the sizes of the shots it holds (small
grids carry more CPML-class work, large ones favour wider tiles) and on the
device. A mix is keyed by its medium kinds and by the share of its cell
updates in each size bucket (the larger side rounded up to a power of two),
in tenths; an entry also records the device's compute-unit count.

Lookups read the user cache (``$FDTD2D_TUNE_CACHE``, else
``~/.cache/fdtd2d_amd/shapes.json``) and then the table shipped with the
package (``tuned_shapes.json``: shapes measured on MI355X). New results go to
the user cache only. An exact key is a hit; ``nearest`` also accepts the
entry of the same medium and device whose bucket shares are closest (L1
distance at most ``max_distance`` tenths), for streams whose mix is only
known from a sample.

Single plans are keyed the same way by their own grid (plan_key: the size
bucket of each side, the CPML thickness and the medium), under kind "plan";
tune_plan finds a plan's shape by coordinate descent on the live plan.
"""

from __future__ import annotations

import json
import os
import tempfile
import time
from pathlib import Path

from ._lib import FDTD2DError

SHIPPED = Path(__file__).with_name("tuned_shapes.json")


def default_path() -> Path:
    env = os.environ.get("FDTD2D_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "fdtd2d_amd" / "shapes.json"


def size_bucket(nx: int, ny: int) -> int:
    """log2 of the larger side, rounded up."""
    return (max(int(nx), int(ny), 2) - 1).bit_length()


def mix_key(shots) -> str:
    """Key of a shot mix; shots: (nx, ny, nsteps, medium_kind) tuples."""
    by: dict[int, float] = {}
    media = set()
    total = 0.0
    for nx, ny, nsteps, medium in shots:
        c = float(nx) * float(ny) * max(int(nsteps), 1)
        b = size_bucket(nx, ny)
        by[b] = by.get(b, 0.0) + c
        total += c
        media.add(int(medium))
    if total <= 0:
        return ""
    parts = [f"{b}:{round(10 * c / total)}" for b, c in sorted(by.items()) if round(10 * c / total) > 0]
    return "m" + "+".join(str(m) for m in sorted(media)) + "|" + ",".join(parts)


def _shares(key: str):
    media, _, parts = key.partition("|")
    return media, {int(p.split(":")[0]): int(p.split(":")[1]) for p in parts.split(",") if p}


def key_distance(a: str, b: str) -> float:
    """L1 distance between the bucket shares of two keys (tenths); infinite
    across media."""
    ma, sa = _shares(a)
    mb, sb = _shares(b)
    if ma != mb:
        return float("inf")
    return float(sum(abs(sa.get(k, 0) - sb.get(k, 0)) for k in set(sa) | set(sb)))


def _load(path: Path) -> list:
    try:
        data = json.loads(path.read_text())
    except (OSError, ValueError):
        return []
    entries = data.get("entries", []) if isinstance(data, dict) else []
    return [e for e in entries if isinstance(e, dict) and "key" in e and "shape" in e]


class ShapeCache:
    def __init__(self, path: str | os.PathLike | None = None, shipped: str | os.PathLike | None = SHIPPED):
        self.path = Path(path) if path is not None else default_path()
        self.shipped = Path(shipped) if shipped is not None else None

    def _sources(self):
        yield str(self.path), _load(self.path)
        if self.shipped is not None:
            yield "shipped:" + self.shipped.name, _load(self.shipped)

    def lookup(self, kind: str, key: str, cus: int, nearest: bool = False, max_distance: float = 2.0):
        """(shape, source) of the entry for this kind ("pool", "plan"), key
        and CU count, or (None, None)."""
        best = None
        for src, entries in self._sources():
            for e in entries:
                if e.get("kind") != kind or int(e.get("cus", -1)) != int(cus):
                    continue
                if e["key"] == key:
                    return dict(e["shape"]), src
                if nearest:
                    d = key_distance(key, e["key"])
                    if d <= max_distance and (best is None or d < best[0]):
                        best = (d, dict(e["shape"]), f"{src} (nearest mix {e['key']})")
        if best is not None:
            return best[1], best[2]
        return None, None

    def store(self, kind: str, key: str, cus: int, shape: dict, rate: float | None = None) -> bool:
        """Adds or replaces the user-cache entry; False when it cannot be
        written."""
        entries = [e for e in _load(self.path)
                   if not (e.get("kind") == kind and e["key"] == key and int(e.get("cus", -1)) == int(cus))]
        entries.append({"kind": kind, "key": key, "cus": int(cus), "shape": dict(shape),
                        "rate_gcell_s": None if rate is None else round(rate / 1e9, 1),
                        "stored": time.strftime("%Y-%m-%dT%H:%M:%S")})
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            fd, tmp = tempfile.mkstemp(dir=self.path.parent, prefix=".shapes")
            with os.fdopen(fd, "w") as f:
                json.dump({"entries": entries}, f, indent=1)
            os.replace(tmp, self.path)
        except OSError:
            return False
        return True


# Plans: the launch shape of one grid, keyed by its size buckets, CPML
# thickness and medium. Axes of the plan descent (the plan's own shape is
# the start).
PLAN_AXES = [
    ("cols_per_lane", [1, 2]),
    ("band", [32, 64, 96, 128, 192, 256]),
    ("prefetch", [1, 2]),
]


def plan_key(nx: int, ny: int, npml: int, medium: int) -> str:
    """Key of a single grid: medium, log2 of each side rounded up, CPML."""
    return f"m{int(medium)}|{size_bucket(nx, 2)}x{size_bucket(ny, 2)}|P{int(npml)}"


def _plan_trial(plan, launches: int) -> float:
    """Cell updates per second of `launches` full launches of the live plan,
    synced on both sides; 0 when the problem has too few steps left."""
    t = int(plan.info("tblock"))
    steps = launches * t
    if int(plan.info("steps")) + steps > plan.nsteps:
        return 0.0
    plan.sync()
    t0 = time.perf_counter()
    plan.run(steps)
    plan.sync()
    dt = time.perf_counter() - t0
    return float(plan.nx) * plan.ny * steps / dt if dt > 0 else 0.0


def tune_plan(plan, launches: int = 4, axes=PLAN_AXES, log=None) -> dict:
    """Coordinate descent over the plan's launch shape (columns per lane,
    band rows, prefetch depth) on the live plan: every trial advances the
    real problem (results do not depend on the shape) and is scored by cell
    updates per second. One untimed trial first (cold caches). Stops when the
    problem runs out of steps. Leaves the plan at the best shape."""
    best = {name: int(plan.info(name)) for name, _ in axes}  # the library's default shape first
    _plan_trial(plan, launches)
    best_rate = _plan_trial(plan, launches)
    trials = [(dict(best), best_rate)]
    for name, values in axes:
        for v in values:
            if v == best[name]:
                continue
            try:
                plan.set_option(name, v)
            except FDTD2DError:  # a shape this grid cannot take; the plan keeps its shape
                continue
            rate = _plan_trial(plan, launches)
            if rate == 0.0:
                break
            trials.append(({**best, name: v}, rate))
            if log:
                log(f"plan tune {name}={v}: {rate / 1e9:.1f} Gcell/s (best {best_rate / 1e9:.1f})")
            if rate > best_rate:
                best_rate, best[name] = rate, v
        plan.set_option(name, best[name])
    return {"shape": best, "rate": best_rate, "trials": trials}


def plan_shape(plan, key: str, cus: int, cache: "ShapeCache | None", tune: bool = False, log=None) -> dict:
    """Sets the plan's launch shape from the cache entry for `key` (kind
    "plan"); on a miss, with tune, runs tune_plan on the plan and stores the
    result, else leaves the library's defaults. Returns {"shape", "source",
    "key", "trials"}."""
    if cache is not None:
        shape, src = cache.lookup("plan", key, cus)
        if shape is not None:
            for name, value in shape.items():
                plan.set_option(name, value)
            return {"shape": shape, "source": src, "key": key, "trials": []}
    if not tune:
        return {"shape": None, "source": "library defaults", "key": key, "trials": []}
    rec = tune_plan(plan, log=log)
    if cache is not None and rec["rate"] > 0:
        cache.store("plan", key, cus, rec["shape"], rec["rate"])
    return {"shape": rec["shape"], "source": "trials", "key": key, "trials": rec["trials"]}
