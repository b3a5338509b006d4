"""Autotuner logic (coordinate descent over the launch knobs, size classes,
settings lookup) against a stand-in engine whose timings are a known
function of the knobs; no GPU."""
import json

import pytest

from slater_stoch_amd import autotune as AT
from slater_stoch_amd import zvmc as Z


class FakeEngine:
    """analytic ms = 10 + |items - 20000| / 1000; estimator ms per run =
    npts x (1 + |split - 4| + (tiles != 1) + (chunk != 1600)) / 1000."""

    def __init__(self):
        self.knob = {k: 0 for k in AT.KNOBS}
        self.measured = []
        self._an = self._est = 0.0

    def set_tuning(self, key, value):
        self.knob[key] = value

    def set_basis(self, bas):
        self.bas = bas

    def gauss_eri(self, device_only=False):
        self._an = 10 + abs(self.knob[Z.TUNE_ERI_ITEMS] - 20000) / 1000

    def mc_blocks(self, nbl, npts, precision, seed=0, stream=0, device_only=False):
        k = self.knob
        self._est = npts * (1 + abs(k[Z.TUNE_EST_SPLIT] - 4) + (k[Z.TUNE_EST_TILES] != 1)
                            + (k[Z.TUNE_MC_CHUNK] != 1600)) / 1000
        self.measured.append(dict(k))

    def timings(self):
        return {"analytic_eri": (self._an, 0), "estimator": (self._est, 0),
                "densities": (0.0, 0), "statistics": (0.0, 0)}


class Bas:
    def __init__(self, n):
        self.n = n


def test_coordinate_descent_finds_the_separable_optimum():
    eng = FakeEngine()
    best, log = AT.tune(eng, Bas(50), npts=800, reps=1)
    assert best == {Z.TUNE_EST_TILES: 1, Z.TUNE_EST_SPLIT: 4, Z.TUNE_MC_CHUNK: 1600,
                    Z.TUNE_ERI_ITEMS: 20000}
    assert eng.knob == best                         # left at the best settings
    costs = [e["cost_s"] for e in log]
    assert min(costs) == pytest.approx(
        0.010 + (800 * 1 / 1000 / 1e3) / (2 * 800) * 40000)
    # one baseline, then every other value of every knob once
    assert len(log) == 1 + sum(len(v) - 1 for v in AT.KNOBS.values())


def test_size_classes_tune_the_largest_member_once():
    eng = FakeEngine()
    bases = [Bas(12), Bas(18), Bas(35), Bas(150)]
    table, logs = AT.tune_classes(eng, bases, npts=400, reps=1)
    assert sorted(table) == [AT.size_class(Bas(18)), AT.size_class(Bas(35)),
                             AT.size_class(Bas(150))]
    assert json.loads(AT.to_json(table))[str(AT.size_class(Bas(35)))]["est_split"] == 4
    assert AT.from_json(AT.to_json(table)) == table


def test_settings_for_falls_back_to_the_nearest_tuned_class():
    t = {1: {Z.TUNE_EST_SPLIT: 8}, 4: {Z.TUNE_EST_SPLIT: 2}}
    assert AT.settings_for(None, Bas(30))[Z.TUNE_EST_SPLIT] == 0
    assert AT.settings_for(t, Bas(30))[Z.TUNE_EST_SPLIT] == 8       # class 1
    assert AT.settings_for(t, Bas(100))[Z.TUNE_EST_SPLIT] == 2      # class 4
    assert AT.settings_for(t, Bas(50))[Z.TUNE_EST_SPLIT] == 8       # class 2 -> 1
    assert AT.settings_for(t, Bas(10))[Z.TUNE_EST_SPLIT] == 8       # below every tuned class
    assert AT.settings_for(t, Bas(200))[Z.TUNE_EST_SPLIT] == 2
    assert set(AT.settings_for(t, Bas(30))) == set(AT.KNOBS)
