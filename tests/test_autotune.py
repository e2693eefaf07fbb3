"""Host logic of the pool autotuner (no GPU): the descent keeps improvements
axis by axis, starts from the first value of each axis after a warm-up
trial, and descends with the fused kernel only when its one trial comes
within the margin of the best two-kernel shape."""

import pytest

from fdtd2d_amd import batch


class FakePool:
    def __init__(self):
        self.opts = {}
        self.log = []

    def set_option(self, name, value):
        self.opts[name] = value

    def info(self, name):
        assert name == "running"
        return 1


def _install(monkeypatch, score):
    trials = []

    def fake_trial(pool, launches):
        r = score(dict(pool.opts))
        trials.append((dict(pool.opts), r))
        return r

    monkeypatch.setattr(batch, "_trial", fake_trial)
    return trials


AXES = [("tblock", [8, 6, 7]), ("band", [256, 128]), ("prefetch", [1, 2])]


def test_descent_keeps_the_best_per_axis(monkeypatch):
    # best: tblock 7, band 128, prefetch 1, two kernels; fused is far worse
    def score(o):
        s = 100.0
        s += {8: 0, 6: -5, 7: 4}[o["tblock"]]
        s += {256: 0, 128: 3}[o["band"]]
        s += {1: 0, 2: -2}[o["prefetch"]]
        return s - (50 if o["fused"] else 0)

    trials = _install(monkeypatch, score)
    res = batch.autotune(FakePool(), axes=AXES)
    assert res["shape"] == {"fused": 0, "tblock": 7, "band": 128, "prefetch": 1}
    assert res["rate"] == 107.0
    # warm-up, start, 2 + 1 + 1 axis trials, one fused trial: no fused descent
    assert len(trials) == 1 + 1 + 4 + 1
    assert trials[0][0]["fused"] == 0 and trials[0][0]["tblock"] == 8  # the warm-up runs the start shape
    assert sum(1 for o, _ in trials if o["fused"] == 1) == 1


@pytest.mark.parametrize("fused_gain,expect_fused", [(-3.0, 0), (5.0, 1)])
def test_fused_descends_only_within_margin(monkeypatch, fused_gain, expect_fused):
    # fused within 5% of the best: it gets its own descent and wins when better
    def score(o):
        return 100.0 + (fused_gain if o["fused"] else 0) + (1 if o["band"] == 128 else 0)

    trials = _install(monkeypatch, score)
    res = batch.autotune(FakePool(), axes=AXES)
    assert res["shape"]["fused"] == expect_fused
    assert res["shape"]["band"] == 128
    assert sum(1 for o, _ in trials if o["fused"] == 1) > 1  # a fused descent ran


def test_fused_skipped_below_margin(monkeypatch):
    trials = _install(monkeypatch, lambda o: 100.0 - (10.0 if o["fused"] else 0.0))
    res = batch.autotune(FakePool(), axes=AXES, margin=0.95)
    assert res["shape"]["fused"] == 0
    assert sum(1 for o, _ in trials if o["fused"] == 1) == 1
