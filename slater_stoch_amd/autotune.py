"""Autotuner for the engine's launch parameters (zv_set_tuning).

For one molecule it times the analytic phase and a short estimator run
over a small grid of each knob, one knob at a time from the library's
automatic choice (coordinate descent, each knob kept at its best value
before the next is tried), and returns the settings that minimize
(analytic time + estimator time per sample x the samples a job of this
size spends):

* ZV_TUNE_EST_TILES  estimator tile shapes (0: the default set, 1: tiles
  with half the units per thread along the longer side, more tiles for
  small molecules)
* ZV_TUNE_EST_SPLIT  sample splits per estimator tile (0 = automatic)
* ZV_TUNE_MC_CHUNK   samples per estimator chunk (0 = automatic)
* ZV_TUNE_ERI_ITEMS  work items per analytic Schwarz phase (0 = 40,000)

A batch of molecules of uneven sizes is tuned per size class (by number of
basis functions): the largest molecule of each class is tuned and its
settings serve the class.  Settings change speed and the estimator's
summation order, never the method; every timing is taken on the device with
the library's own event timers (zv_last_timings).
"""
from __future__ import annotations

import json

import numpy as np

from . import zvmc as Z

KNOBS = {
    Z.TUNE_EST_TILES: [0, 1],
    Z.TUNE_EST_SPLIT: [0, 1, 2, 4, 8, 16, 32],
    Z.TUNE_MC_CHUNK: [0, 1600, 4000],
    Z.TUNE_ERI_ITEMS: [0, 10000, 20000, 80000],
}
NAMES = {Z.TUNE_EST_TILES: "est_tiles", Z.TUNE_EST_SPLIT: "est_split",
         Z.TUNE_MC_CHUNK: "mc_chunk", Z.TUNE_ERI_ITEMS: "eri_items"}
SIZE_CLASSES = (20, 40, 60, 90, 120, 160, 1024)


def size_class(bas):
    """Index of the first class bound >= nbasis."""
    for k, b in enumerate(SIZE_CLASSES):
        if bas.n <= b:
            return k
    return len(SIZE_CLASSES) - 1


def _measure(eng, bas, settings, npts, samples_per_job, precision, reps):
    for key, val in settings.items():
        eng.set_tuning(key, val)
    an, est = [], []
    for _ in range(reps):
        eng.set_basis(bas)
        eng.gauss_eri(device_only=True)
        an.append(eng.timings()["analytic_eri"][0] / 1e3)
        eng.mc_blocks(2, npts, precision, seed=1, stream=0, device_only=True)
        t = eng.timings()
        mc = (t["estimator"][0] + t["densities"][0] + t["statistics"][0]) / 1e3
        est.append(mc / (2 * npts))
    a, e = float(np.min(an)), float(np.min(est))
    return a + e * samples_per_job, a, e


def tune(eng, bas, precision="fp32", npts=4000, samples_per_job=40000, reps=1, knobs=None):
    """Best settings for this molecule: (settings dict {key: value}, log of
    every trial).  The engine's settings are left at the best ones."""
    knobs = KNOBS if knobs is None else knobs
    best = {k: 0 for k in knobs}
    cost, _, _ = _measure(eng, bas, best, npts, samples_per_job, precision, reps)
    log = [dict(settings={NAMES[k]: v for k, v in best.items()}, cost_s=cost)]
    for key, values in knobs.items():
        for v in values:
            if v == best[key]:
                continue
            trial = dict(best)
            trial[key] = v
            c, a, e = _measure(eng, bas, trial, npts, samples_per_job, precision, reps)
            log.append(dict(settings={NAMES[k]: x for k, x in trial.items()}, cost_s=c,
                            analytic_s=a, estimator_s_per_sample=e))
            if c < cost:
                cost, best = c, trial
    for key, val in best.items():
        eng.set_tuning(key, val)
    return best, log


def tune_classes(eng, bases, precision="fp32", **kw):
    """Settings per size class, tuned on the largest molecule of each class
    present in ``bases``: {class index: settings}, plus the trial logs."""
    by_class = {}
    for b in bases:
        k = size_class(b)
        if k not in by_class or b.n > by_class[k].n:
            by_class[k] = b
    table, logs = {}, {}
    for k, b in sorted(by_class.items()):
        table[k], logs[k] = tune(eng, b, precision, **kw)
    return table, logs


def settings_for(table, bas):
    """The settings of ``bas``'s size class in ``table``: the tuned class
    itself, else the nearest tuned class below it, else above it; the
    library's automatic choice (every knob 0) when there is no table."""
    if not table:
        return {k: 0 for k in KNOBS}
    k = size_class(bas)
    below = [c for c in table if c <= k]
    c = max(below) if below else min(table)
    out = {key: 0 for key in KNOBS}
    out.update(table[c])
    return out


def apply(eng, settings):
    for key, val in settings.items():
        eng.set_tuning(key, val)


def to_json(table):
    return json.dumps({str(k): {NAMES[key]: v for key, v in s.items()}
                       for k, s in table.items()})


def from_json(text):
    """The table written by to_json."""
    keys = {v: k for k, v in NAMES.items()}
    return {int(c): {keys[name]: int(v) for name, v in s.items()}
            for c, s in json.loads(text).items()}
