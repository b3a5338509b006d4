"""The batch driver: sample-size policy and determinism.

CPU tests drive run_molecule with a stand-in engine whose standard error
falls as 1/sqrt(samples), and check that run_batch refuses several worker
threads; the GPU test runs real molecules and checks that the batch gives
the same bits as one-molecule runs with the same seeds.  The case of three
contexts driven from three threads at once is kept, skipped: it hung on
MI355X, and concurrent contexts are disabled."""
import numpy as np
import pytest

from slater_stoch_amd import batch as B
from slater_stoch_amd import configs


def test_next_npts_rounds_scales_and_caps():
    # already below target: no further run
    assert B.next_npts(4000, 5e-7, 1e-6, 1.2, 10**6) == 0
    # (2e-6 / 1e-6)^2 x 1.2 x 4000 = 19200, a multiple of 40
    assert B.next_npts(4000, 2e-6, 1e-6, 1.2, 10**6) == 19200
    # rounded up to a multiple of 40
    assert B.next_npts(4000, 1.01e-6, 1e-6, 1.0, 10**6) % 40 == 0
    # at least one walker block more than the last run
    assert B.next_npts(4000, 1.0000001e-6, 1e-6, 1.0, 10**6) >= 4040
    # capped by the budget, rounded down to a multiple of 40
    assert B.next_npts(4000, 1e-4, 1e-6, 1.2, 123456) == 123440
    # no run when the budget cannot exceed the last run
    assert B.next_npts(4000, 1e-4, 1e-6, 1.2, 4000) == 0


class FakeEngine:
    """err = s0 / sqrt(nbl * npts) for every quartet of a run (for the
    quartets above the target in a refinement); records the calls."""

    def __init__(self, s0, nint=5):
        self.s0, self.nint, self.calls = s0, nint, []
        self.mean, self.err = np.zeros(nint), np.zeros(nint)

    def set_basis(self, bas):
        self.bas = bas

    def gauss_eri(self, device_only=False):
        return (None if device_only else np.zeros(self.nint)), None, 0

    def mc_blocks(self, nbl, npts, precision, seed=0, stream=0, device_only=False):
        self.calls.append((nbl, npts, precision, seed, stream))
        self.err = np.full(self.nint, self.s0 / np.sqrt(nbl * npts))
        self.mean = np.full(self.nint, float(stream))
        return (None, None) if device_only else (self.mean.copy(), self.err.copy())

    def mc_refine(self, nbl, npts, precision, target=1e-6, seed=0, stream=1):
        self.calls.append((nbl, npts, precision, seed, stream))
        sel = self.err > target
        self.err[sel] = self.s0 / np.sqrt(nbl * npts)
        self.mean[sel] = float(stream)
        return int(sel.sum()), float(self.err.max())

    def result_max(self, which):
        return float(self.err.max())

    def timings(self):
        return {"estimator": (1.0, 9)}


class FakeBasis:
    n, nint = 3, 5


def test_policy_reaches_target_in_two_runs():
    # pilot 40000 samples -> 2e-6; the second run is sized for 1e-6 x sqrt(1/1.2)
    eng = FakeEngine(s0=2e-6 * np.sqrt(40000))
    r = B.run_molecule(eng, FakeBasis(), target=1e-6, precision="fp32", seed=7,
                       pilot_npts=4000)
    assert r["rounds"] == 2 and r["reached"]
    assert [c[4] for c in eng.calls] == [0, 1]          # one Philox stream per run
    assert all(c[3] == 7 for c in eng.calls)
    assert eng.calls[1][1] == 19200
    assert r["samples"] == 10 * 4000 + 10 * 19200
    assert r["max_stderr"] == pytest.approx(2e-6 * np.sqrt(40000 / 192000))
    # device time and launches summed over the three library calls
    assert r["t_device"] == pytest.approx(0.003) and r["launches"] == 27


def test_policy_stops_at_the_sample_budget():
    eng = FakeEngine(s0=1e-3 * np.sqrt(40000))          # hopeless: 1e-3 after the pilot
    r = B.run_molecule(eng, FakeBasis(), target=1e-6, max_samples=1_000_000,
                       pilot_npts=4000)
    assert not r["reached"]
    assert r["samples"] <= 1_000_000
    assert r["rounds"] == 2                              # pilot + what the budget allows
    assert eng.calls[1][1] == (1_000_000 - 40000) // 10 // 40 * 40


def test_policy_skips_production_when_the_pilot_suffices():
    eng = FakeEngine(s0=1e-7 * np.sqrt(40000))
    r = B.run_molecule(eng, FakeBasis(), target=1e-6, pilot_npts=4000)
    assert r["rounds"] == 1 and r["reached"] and r["samples"] == 40000


def test_default_pilot_is_ten_blocks_of_1000():
    # the default pilot gives 10 x 1000 samples; 1e-6 after it needs no second run
    eng = FakeEngine(s0=1e-6 * np.sqrt(10000))
    r = B.run_molecule(eng, FakeBasis(), target=1e-6)
    assert eng.calls[0][1] == 1000
    assert r["rounds"] == 1 and r["reached"] and r["samples"] == 10000


def test_batch_rejects_several_workers():
    with pytest.raises(ValueError):
        B.run_batch([FakeBasis()], workers=3)


def _batch_case():
    bases = [b for b in configs.batch() if b.n <= 30][:4]
    assert len(bases) >= 3
    pol = dict(nbl=4, pilot_npts=400, max_samples=200_000, max_rounds=3, keep_arrays=True)
    return bases, pol


@pytest.mark.gpu
def test_batch_matches_single_runs():
    from slater_stoch_amd.zvmc import Engine

    bases, pol = _batch_case()
    one = B.run_batch(bases, target=1e-7, precision="fp64", seed=11, **pol)
    with Engine(0) as e:
        for t, b in enumerate(bases):
            ref = B.run_molecule(e, b, target=1e-7, precision="fp64", seed=11 + t, **pol)
            got = one[t]
            assert got["rounds"] == ref["rounds"] and got["samples"] == ref["samples"]
            np.testing.assert_array_equal(got["eri"], ref["eri"])
            np.testing.assert_array_equal(got["stderr"], ref["stderr"])
    # the tight target forces production runs on at least one molecule
    assert max(r["rounds"] for r in one) >= 2


@pytest.mark.gpu
@pytest.mark.skip(reason="three host threads driving three contexts at once hung on MI355X "
                         "(twice, once with the estimator's side streams switched off); "
                         "concurrent contexts are disabled: run_batch rejects workers > 1 "
                         "and the library serializes calls across contexts")
def test_three_contexts_driven_at_once_match_single_runs():
    import threading

    from slater_stoch_amd.zvmc import Engine

    bases, pol = _batch_case()
    out = [None] * len(bases)

    def work(ts):
        with Engine(0) as e:
            for t in ts:
                out[t] = B.run_molecule(e, bases[t], target=1e-7, precision="fp64",
                                        seed=11 + t, **pol)

    threads = [threading.Thread(target=work, args=(list(range(k, len(bases), 3)),))
               for k in range(3)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    with Engine(0) as e:
        for t, b in enumerate(bases):
            ref = B.run_molecule(e, b, target=1e-7, precision="fp64", seed=11 + t, **pol)
            np.testing.assert_array_equal(out[t]["eri"], ref["eri"])
