"""The launch-shape cache (fdtd2d_amd.tuning) and the batch runner's use of
it (batch.choose_shape), on the CPU: a fake pool stands in for the device
pool, and trials are scored from the shape by a fixed function, so the tests
check which path was taken (cache hit: no trials) and what was stored."""

from __future__ import annotations

import json

import pytest

from fdtd2d_amd import batch
from fdtd2d_amd.tuning import ShapeCache, key_distance, mix_key, size_bucket


def test_size_bucket_and_mix_key():
    assert [size_bucket(n, n) for n in (256, 257, 512, 513, 2048, 16384)] == [8, 9, 9, 10, 11, 14]
    assert size_bucket(300, 5000) == 13  # the larger side
    # shares of cell updates (nx * ny * steps), in tenths, buckets under 5% dropped
    shots = [(1024, 1024, 4000, 1)] * 9 + [(2048, 2048, 1000, 1)] * 9 + [(256, 256, 10, 1)]
    assert mix_key(shots) == "m1|10:5,11:5"
    assert mix_key([(16384, 16384, 30000, 1), (1024, 1024, 4000, 1)]) == "m1|14:10"
    assert mix_key([(600, 600, 10, 0), (600, 600, 10, 1)]) == "m0+1|10:10"
    assert mix_key([]) == ""


def test_key_distance():
    assert key_distance("m1|10:1,11:9", "m1|10:1,11:9") == 0
    assert key_distance("m1|10:1,11:9", "m1|10:3,11:7") == 4
    assert key_distance("m1|10:10", "m1|11:10") == 20
    assert key_distance("m1|10:10", "m0|10:10") == float("inf")


def test_cache_store_lookup_and_shipped_fallback(tmp_path):
    shipped = tmp_path / "shipped.json"
    shipped.write_text(json.dumps({"entries": [
        {"kind": "pool", "key": "m1|10:1,11:9", "cus": 256, "shape": {"tblock": 8, "fused": 1}},
        {"kind": "pool", "key": "m1|14:10", "cus": 304, "shape": {"tblock": 6}},
    ]}))
    c = ShapeCache(tmp_path / "user" / "shapes.json", shipped=shipped)
    shape, src = c.lookup("pool", "m1|10:1,11:9", 256)
    assert shape == {"tblock": 8, "fused": 1} and src == "shipped:shipped.json"
    assert c.lookup("pool", "m1|14:10", 256) == (None, None)  # another device
    assert c.lookup("plan", "m1|10:1,11:9", 256) == (None, None)  # another kind
    # the user cache wins over the shipped table
    assert c.store("pool", "m1|10:1,11:9", 256, {"tblock": 7}, 5e11)
    shape, src = c.lookup("pool", "m1|10:1,11:9", 256)
    assert shape == {"tblock": 7} and src.endswith("shapes.json")
    # replaced, not appended
    assert c.store("pool", "m1|10:1,11:9", 256, {"tblock": 6})
    entries = json.loads((tmp_path / "user" / "shapes.json").read_text())["entries"]
    assert [e["shape"] for e in entries] == [{"tblock": 6}]
    # nearest: same medium and device, within the distance
    shape, src = c.lookup("pool", "m1|10:2,11:8", 256, nearest=True)
    assert shape == {"tblock": 6} and "nearest mix m1|10:1,11:9" in src
    assert c.lookup("pool", "m1|10:2,11:8", 256) == (None, None)
    assert c.lookup("pool", "m1|9:5,11:5", 256, nearest=True) == (None, None)
    # an unreadable cache file reads as empty
    (tmp_path / "user" / "shapes.json").write_text("{not json")
    assert c.lookup("pool", "m1|10:1,11:9", 256)[0] == {"tblock": 8, "fused": 1}


class FakePool:
    """Enough of fdtd2d_amd.Pool for choose_shape/autotune: options, a step
    counter, and shots that never finish."""

    def __init__(self):
        self.options = {}
        self.steps = 0

    def set_option(self, name, value):
        self.options[name] = value

    def info(self, name):
        return {"cus": 256, "running": 3, "cell_updates": 1e9 * self.steps}[name]

    def step(self, n):
        self.steps += n

    def sync(self):
        pass


def _score(pool, launches):
    """Trial score from the shape: best at tblock 7, cols 2, band 128, fused."""
    pool.step(launches)
    o = pool.options
    return (1000 - 10 * abs(o.get("tblock", 8) - 7) + 5 * (o.get("cols_per_lane", 1) == 2)
            - abs(o.get("band", 256) - 128) / 64 + 3 * o.get("fused", 0)) * 1e9


def test_cache_hit_skips_the_trials(tmp_path, monkeypatch):
    monkeypatch.setattr(batch, "_trial", _score)
    cache = ShapeCache(tmp_path / "shapes.json", shipped=None)
    key = "m1|10:1,11:9"
    pool = FakePool()
    first = batch.choose_shape(pool, key, cache)
    assert first["source"] == "trials" and pool.steps > 0 and len(first["trials"]) > 5
    assert first["shape"]["tblock"] == 7 and first["shape"]["cols_per_lane"] == 2 and first["shape"]["band"] == 128
    assert first["shape"]["fused"] == 1
    # a second pool of the same mix takes the stored shape with no trials
    again = FakePool()
    hit = batch.choose_shape(again, key, cache)
    assert again.steps == 0 and hit["trials"] == [] and hit["source"] == str(tmp_path / "shapes.json")
    assert hit["shape"] == first["shape"] and again.options == first["shape"]
    # lookup=False runs the trials again (and refreshes the entry)
    forced = FakePool()
    batch.choose_shape(forced, key, cache, lookup=False)
    assert forced.steps > 0
    # no cache: trials, nothing stored
    other = tmp_path / "none.json"
    batch.choose_shape(FakePool(), "m1|14:10", None)
    assert not other.exists()


def test_shipped_table_is_well_formed():
    from fdtd2d_amd.tuning import SHIPPED, _load

    raw = json.loads(SHIPPED.read_text())
    entries = _load(SHIPPED)
    assert entries and len(entries) == len(raw["entries"])
    pool_options = {"tblock", "cols_per_lane", "band", "band_cpml", "wg_per_cu", "wg_per_cu_cpml", "prefetch",
                    "fused"}
    for e in entries:
        assert e["kind"] in ("pool", "plan") and e["cus"] > 0 and e["key"].startswith("m")
        if e["kind"] == "pool":
            assert set(e["shape"]) <= pool_options, e


@pytest.mark.parametrize("nearest", [False, True])
def test_batch_mix_has_a_shipped_shape(nearest):
    """The seed-3 batch mix's key (computed from the workload draws) is in the
    shipped table for MI355X's 256 CUs."""
    import fdtd2d_amd as F

    sizes = []
    for k in range(256):
        h = F.Setup.workload(3, seed=3, shot=k, steps=-1)
        sizes.append((int(h.get("Imax")), int(h.get("Jmax")), int(h.get("nMax")), 1))
        h.close()
    key = mix_key(sizes)
    shape, src = ShapeCache("/nonexistent/shapes.json").lookup("pool", key, 256, nearest=nearest)
    assert shape is not None and src.startswith("shipped:"), key


class FakePlan:
    """Enough of fdtd2d_amd.Plan for tune_plan/plan_shape: options, a step
    counter and a step budget."""

    def __init__(self, nsteps=4000, refuse=()):
        self.options = {"tblock": 8, "cols_per_lane": 2, "band": 96, "prefetch": 2}
        self.n = 0
        self.nsteps = nsteps
        self.nx = self.ny = 1000
        self.refuse = set(refuse)

    def set_option(self, name, value):
        from fdtd2d_amd._lib import FDTD2DError

        if (name, value) in self.refuse:
            raise FDTD2DError(-1, "fdtd2d_plan_set_option", "refused")
        self.options[name] = value

    def info(self, name):
        return {"steps": self.n, **self.options}[name]

    def run(self, steps):
        self.n += steps

    def sync(self):
        pass


def test_plan_tuner_descends_and_caches(tmp_path, monkeypatch):
    from fdtd2d_amd import tuning

    def score(plan, launches):
        if plan.n + 8 * launches > plan.nsteps:
            return 0.0
        plan.run(8 * launches)
        o = plan.options
        return (1000 + 20 * (o["cols_per_lane"] == 1) - abs(o["band"] - 192) / 8 + 3 * (o["prefetch"] == 1)) * 1e9

    monkeypatch.setattr(tuning, "_plan_trial", score)
    key = tuning.plan_key(8192, 8192, 16, 1)
    assert key == "m1|13x13|P16"
    cache = ShapeCache(tmp_path / "shapes.json", shipped=None)
    plan = FakePlan(refuse={("band", 256)})
    rec = tuning.plan_shape(plan, key, 256, cache, tune=True)
    assert rec["source"] == "trials" and plan.n > 0
    assert rec["shape"] == {"cols_per_lane": 1, "band": 192, "prefetch": 1}
    assert plan.options["band"] == 192 and plan.options["cols_per_lane"] == 1
    # cached: no trials
    again = FakePlan()
    hit = tuning.plan_shape(again, key, 256, cache, tune=True)
    assert again.n == 0 and hit["shape"] == rec["shape"] and hit["trials"] == []
    # no tune on a miss: the library's defaults, nothing run
    other = FakePlan()
    miss = tuning.plan_shape(other, "m1|9x9|P12", 256, cache, tune=False)
    assert miss["shape"] is None and miss["source"] == "library defaults" and other.n == 0
    # a problem that runs out of steps stops the descent
    short = FakePlan(nsteps=8 * 4 * 3)
    tuning.tune_plan(short)
    assert short.n <= short.nsteps


def test_shipped_plan_shapes_for_the_single_grid_workloads():
    from fdtd2d_amd.tuning import plan_key

    c = ShapeCache("/nonexistent/shapes.json")
    for n in (8192, 16384):
        shape, src = c.lookup("plan", plan_key(n, n, 16, 1), 256)
        assert shape is not None and src.startswith("shipped:")
        assert shape["cols_per_lane"] == 2 and shape["tblock"] == 8
