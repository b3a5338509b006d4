"""Batches of molecules through the ZVMC engine, each to a target standard
error.

Per molecule the job is two_e_slater's: analytic Gaussian-fit ERIs with the
Schwarz driver (src/two_e_slater.F90:219-284), then nbl blocks of
npts_two_elec samples (src/two_e_slater.F90:386-496), mean and standard error
by the block formula (src/two_e_slater.F90:488-496).  The reference runs one
molecule with a fixed npts_two_elec; the sample-size policy and the batch
are this package's:

* a pilot of nbl x pilot_npts samples over every quartet;
* while the largest standard error exceeds the target and the sample budget
  allows, a further run of nbl blocks with npts scaled from the latest run,
  npts' = npts (err_max / target)^2 x safety, rounded up to a multiple of
  nw = 40 and capped by what is left of max_samples (the estimator's error is
  heavy-tailed, so one run sized from the pilot can fall short).  With the
  default policy "refine" that run covers only the quartets still above the
  target (zv_mc_refine: the estimator unit pairs that hold them), which
  replace their values with the new run's; with "rerun" it covers every
  quartet (the earlier policy, for comparison);
* every reported mean and standard error is one complete nbl-block estimate
  of a single run, as two_e_slater writes it.  ``samples`` counts the
  samples of every run (a refinement run's samples serve its quartets only).

Molecules run largest first through one Engine (one library context).
Results are deterministic: molecule t uses Philox seed ``seed + t`` and
stream r for its run r.  Several host threads, each driving its own context,
are not supported: on MI355X four such workers made a 150-AO molecule's
analytic phase 60x slower (169 s instead of 2.8 s: a context uses five
streams and HIP opens four hardware queues per process by default, so
unrelated streams waited behind each other's long estimator launches), and
a run with three workers hung.  The library runs calls on different
contexts one at a time, and ``workers`` must be 1.

Each molecule's result also carries the device time of its library calls
(the library's own event timers, all phases) and their kernel launches:
the wall time not spent on the device is the host's share (planning,
launch and synchronization), the most that capturing the launches in a
graph could save.  ``--autotune`` first tunes the launch knobs per molecule
size class (slater_stoch_amd/autotune.py).

Command line (the 64-molecule batch of uneven sizes, FP32 estimator):

    python -m slater_stoch_amd.batch --count 64 --precision fp32
"""
import argparse
import json
import math
import sys
import time

import numpy as np

NW = 40  # walkers per estimator block (src/two_e_slater.F90:386)


def next_npts(npts, err_max, target, safety, cap_npts):
    """Samples per block of the next run, or 0 when none is needed or none
    fits: npts (err_max / target)^2 x safety, a multiple of NW, at most
    cap_npts (itself rounded down to a multiple of NW)."""
    if not err_max > target:
        return 0
    cap = int(cap_npts) // NW * NW
    want = int(math.ceil(npts * (err_max / target) ** 2 * safety / NW)) * NW
    want = max(want, npts + NW)
    return min(want, cap) if cap > npts else 0


def run_molecule(eng, bas, target=1e-6, precision="fp32", nbl=10, pilot_npts=1000,
                 max_samples=10_000_000, safety=1.2, max_rounds=4, seed=0, keep_arrays=False,
                 policy="refine"):
    """One molecule: analytic ERIs, pilot, then further runs by the policy
    above.  Returns a dict of the run (arrays only with keep_arrays)."""
    from .zvmc import RESULT_ERI, RESULT_MEAN, RESULT_STDERR
    if policy not in ("refine", "rerun"):
        raise ValueError(f"unknown policy {policy!r}")
    dev = [0.0, 0]  # device time (s) and kernel launches of the library's calls

    def count():
        for ms, n in eng.timings().values():
            dev[0] += ms / 1e3
            dev[1] += int(n)

    t0 = time.perf_counter()
    eng.set_basis(bas)
    _, _, nscr = eng.gauss_eri(device_only=True)
    count()
    t1 = time.perf_counter()
    npts, rounds, used, refined = pilot_npts, 0, 0, 0
    eng.mc_blocks(nbl, npts, precision, seed=seed, stream=0, device_only=True)
    count()
    emax = eng.result_max(RESULT_STDERR)
    used += nbl * npts
    rounds += 1
    while rounds < max_rounds:
        nxt = next_npts(npts, emax, target, safety, (max_samples - used) // nbl)
        if not nxt:
            break
        npts = nxt
        if policy == "refine":
            nref, emax = eng.mc_refine(nbl, npts, precision, target=target, seed=seed,
                                       stream=rounds)
            count()
            refined += nref
        else:
            eng.mc_blocks(nbl, npts, precision, seed=seed, stream=rounds, device_only=True)
            count()
            emax = eng.result_max(RESULT_STDERR)
        used += nbl * npts
        rounds += 1
    t2 = time.perf_counter()
    out = dict(n=int(bas.n), nint=int(bas.nint), n_screened=int(nscr), samples=int(used),
               rounds=rounds, npts=int(npts), max_stderr=emax, reached=bool(emax <= target),
               refined_quartets=int(refined), t_analytic=t1 - t0, t_mc=t2 - t1,
               t_device=dev[0], launches=dev[1])
    if keep_arrays:
        eri = eng.get_result(RESULT_ERI)
        mean, err = eng.get_result(RESULT_MEAN), eng.get_result(RESULT_STDERR)
        out.update(eri_gauss=eri, mc_mean=mean, stderr=err, eri=eri + mean)
    return out


def run_batch(bases, target=1e-6, precision="fp32", workers=1, device=0, seed=0,
              callback=None, verbose=0, tuning=None, autotune=False, on_tuned=None, **policy):
    """Run every basis of ``bases``; returns the per-molecule dicts in input
    order.  ``callback(index, result)`` is called as each molecule finishes.
    ``verbose`` is the engine's progress level (2: a line per finished Monte
    Carlo block).  ``policy`` goes to run_molecule (nbl, pilot_npts,
    max_samples, safety, max_rounds, keep_arrays).  ``workers`` must be 1
    (concurrent contexts are not supported, see the module docstring).

    ``tuning`` is a table {size class: {knob: value}} (autotune.tune_classes)
    applied before each molecule of that class; ``autotune=True`` builds it
    first on this batch, its time counted in the results' "t_tune" of the
    first molecule run, and passes it to ``on_tuned``."""
    if int(workers) != 1:
        raise ValueError("run_batch: workers must be 1 (contexts driven from several host "
                         "threads at once are not supported)")
    from . import autotune as AT
    from .zvmc import Engine

    order = sorted(range(len(bases)), key=lambda t: -bases[t].nint)
    results = [None] * len(bases)
    with Engine(device, verbose=verbose) as eng:
        t_tune = 0.0
        if autotune and tuning is None:
            t0 = time.perf_counter()
            tuning, _ = AT.tune_classes(eng, bases, precision)
            t_tune = time.perf_counter() - t0
            if on_tuned is not None:
                on_tuned(tuning)
        for t in order:
            AT.apply(eng, AT.settings_for(tuning, bases[t]))
            r = run_molecule(eng, bases[t], target=target, precision=precision, seed=seed + t,
                             **policy)
            r["t_tune"], t_tune = t_tune, 0.0
            r["index"] = t
            results[t] = r
            if callback is not None:
                callback(t, r)
    return results


def summary(results, wall):
    ok = sum(r["reached"] for r in results)
    return dict(molecules=len(results), reached=ok, wall_s=wall,
                quartets=int(sum(r["nint"] for r in results)),
                samples=int(sum(r["samples"] for r in results)),
                quartet_samples=float(sum(float(r["nint"]) * r["samples"] for r in results)),
                tune_s=float(sum(r.get("t_tune", 0.0) for r in results)),
                analytic_s=float(sum(r["t_analytic"] for r in results)),
                mc_s=float(sum(r["t_mc"] for r in results)),
                device_s=float(sum(r.get("t_device", 0.0) for r in results)),
                launches=int(sum(r.get("launches", 0) for r in results)))


def main(argv=None):
    from . import configs

    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--batch-seed", type=int, default=64)
    ap.add_argument("--max-ao", type=int, default=150)
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "bf16"])
    ap.add_argument("--target", type=float, default=1e-6)
    ap.add_argument("--workers", type=int, default=1,
                    help="must be 1: concurrent contexts are not supported")
    ap.add_argument("--max-samples", type=int, default=10_000_000)
    ap.add_argument("--pilot-npts", type=int, default=1000,
                    help="samples per pilot block (a multiple of 40)")
    ap.add_argument("--max-rounds", type=int, default=4)
    ap.add_argument("--policy", default="refine", choices=["refine", "rerun"])
    ap.add_argument("--autotune", action="store_true",
                    help="tune the launch parameters per size class first (timed in wall_s)")
    ap.add_argument("--tuning-file", default=None,
                    help="with --autotune: write the tuned table here; without: read it")
    ap.add_argument("--verbose", type=int, default=2,
                    help="engine progress level (2: one line per Monte Carlo block)")
    args = ap.parse_args(argv)
    bases = configs.batch(seed=args.batch_seed, count=args.count, max_ao=args.max_ao)

    def show(t, r):
        print(f"[batch] molecule {t}: {r['n']} AOs, {r['nint']} quartets, analytic "
              f"{r['t_analytic']:.2f} s, MC {r['t_mc']:.2f} s, {r['samples']} samples in "
              f"{r['rounds']} runs, max stderr {r['max_stderr']:.2e}", flush=True)

    from . import autotune as AT
    tuning = None
    if args.tuning_file and not args.autotune:
        with open(args.tuning_file) as f:
            tuning = AT.from_json(f.read())
    saved = {}

    def keep_table(table):
        saved["table"] = table

    t0 = time.perf_counter()
    res = run_batch(bases, target=args.target, precision=args.precision, workers=args.workers,
                    callback=show, verbose=args.verbose, max_samples=args.max_samples,
                    max_rounds=args.max_rounds, policy=args.policy, autotune=args.autotune,
                    pilot_npts=args.pilot_npts,
                    tuning=tuning, on_tuned=keep_table)
    if args.autotune and args.tuning_file and "table" in saved:
        with open(args.tuning_file, "w") as f:
            f.write(AT.to_json(saved["table"]))
    s = summary(res, time.perf_counter() - t0)
    s.update(precision=args.precision, workers=args.workers, target=args.target,
             policy=args.policy, autotune=args.autotune, pilot_npts=args.pilot_npts,
             tuning=AT.to_json(tuning) if tuning else None)
    print(json.dumps(s), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
