#!/usr/bin/env python3
"""Time to a 1e-6 Ha standard error on every unique (ik|jl): the ZVMC ERI job.

One step is the whole two-electron job of the reference's two_e_slater for a
seeded 12-atom molecule (5 heavy atoms with 1S,2S,2S,2P,2P,3D; 7 H with
1S,1S,2P; zeta ~ U[0.8, 8]; 110 STO orbitals, 6105 pairs, 18,638,565
quartets; 20 Gaussians per STO fit):

  1. analytic Gaussian-fit ERIs with the Cauchy-Schwarz driver (FP64),
  2. a pilot of 10 blocks x 4000 Monte Carlo samples,
  3. if the largest standard error is above 1e-6 Ha, a production run of 10
     blocks sized from the pilot (stderr ~ 1/sqrt(samples)).

The timed region starts with the basis resident on the device and ends with
the ERIs, means and standard errors computed in HBM (the library keeps them
there, and the sample-size policy reads the largest standard error reduced
on the device); copying the three arrays to the host (447 MB over PCIe)
comes after it and is reported as ``host_copy_s`` / ``with_host_copy_s``.  FP64 is the plain
GPU form (``value``); the FP32 estimator and the BF16-product estimator
variants are reported beside it, each with its error against the FP64
estimator on the same samples.
A bounded sample of the same job on the CPU oracle (all host cores) gives
``cpu_baseline``.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "sec to 1e-6 Ha stderr on all (ij|kl), mixed-prec vs FP64 GPU vs CPU; MFMA %"
TARGET = 1e-6
NBL = 10
PILOT = 4000
PEAK_FP64 = 78.6  # TFLOP/s, MI355X vector FP64
PEAK_FP32 = 157.3  # TFLOP/s, MI355X vector FP32


def job(eng, bas, prec, seed):
    """One step.  The timed region holds the device work of the job: the
    per-quartet arrays stay in HBM (the sample-size policy reads the largest
    standard error reduced on the device); fetching them to the host comes
    after it and is reported separately (``host_copy_s``).  When a
    production run follows the pilot, the pilot's arrays are fetched for the
    precision report with the clock stopped."""
    from slater_stoch_amd.zvmc import RESULT_ERI, RESULT_MEAN, RESULT_STDERR
    eng.set_basis(bas)  # inputs resident; clears the primitive-pair cache
    t0 = time.perf_counter()
    _, pre, nscr = eng.gauss_eri(device_only=True)
    t1 = time.perf_counter()
    eri_stats = eng.eri_stats()
    an_timings = eng.timings()
    eng.mc_blocks(NBL, PILOT, prec, seed=seed, stream=0, device_only=True)
    mc_timings = eng.timings()
    samples = NBL * PILOT
    emax = eng.result_max(RESULT_STDERR)
    npts = PILOT
    pilot, paused = None, 0.0
    if emax > TARGET:
        tp = time.perf_counter()
        pilot = (eng.get_result(RESULT_MEAN), eng.get_result(RESULT_STDERR))
        paused = time.perf_counter() - tp
        npts = int(math.ceil(PILOT * (emax / TARGET) ** 2 * 1.2 / 40.0)) * 40
        npts = min(npts, 1_000_000)
        eng.mc_blocks(NBL, npts, prec, seed=seed, stream=1, device_only=True)
        mc_timings = eng.timings()
        samples += NBL * npts
        emax = eng.result_max(RESULT_STDERR)
    t2 = time.perf_counter()
    eri = eng.get_result(RESULT_ERI)
    mean, err = eng.get_result(RESULT_MEAN), eng.get_result(RESULT_STDERR)
    t3 = time.perf_counter()
    if pilot is None:
        pilot = (mean, err)
    return dict(total=t2 - t0 - paused, analytic=t1 - t0, mc=t2 - t1 - paused,
                host_copy=t3 - t2, max_stderr=emax, samples=samples, production_npts=npts,
                n_screened=nscr, eri_stats=eri_stats, an_timings=an_timings,
                mc_timings=mc_timings, eri=eri, mean=mean, err=err, pilot=pilot)


def precision_report(bas, m64, e64, m32):
    """FP32- (or BF16-) estimator error against the FP64 estimator on the same samples:
    per-quartet mean differences, overall and binned by the number of
    distinct nuclei of the quartet (the classes of src/compare_nuclei.F90)."""
    from slater_stoch_amd import basis as B
    q = B.quartet_mapping(bas.n)
    cls = B.nuclei_class(bas, *q)
    d = np.abs(m32 - m64)
    out = {"quartets": int(d.size), "max_abs": float(d.max()),
           "rms": float(np.sqrt(np.mean(d * d))),
           "max_abs_over_fp64_stderr": float(np.max(d / np.maximum(e64, 1e-300))),
           "by_distinct_nuclei": {}}
    for k in range(1, 5):
        sel = cls == k
        if sel.any():
            dk = d[sel]
            out["by_distinct_nuclei"][str(k)] = {"quartets": int(sel.sum()),
                                                 "max_abs": float(dk.max()),
                                                 "rms": float(np.sqrt(np.mean(dk * dk)))}
    return out


def water_precision(eng, prec="fp32"):
    """The same report on the 7-AO water molecule (10 x 4000 samples)."""
    from slater_stoch_amd import configs
    w = configs.water_minimal()
    eng.set_basis(w)
    m64, e64 = eng.mc_blocks(NBL, PILOT, "fp64", seed=3, stream=0)
    m32, _ = eng.mc_blocks(NBL, PILOT, prec, seed=3, stream=0)
    return precision_report(w, m64, e64, m32)


def kernel_name(key):
    if key[0] == "shell":
        _, lu, ll, clb, red = key
        return f"k_eri_shell{'_red' if red else ''}<{lu}, {ll}, {clb}>"
    _, a, b = key
    return f"k_eri_class<{a}, {b}>" if a < 5 else "k_eri_generic"


def _traffic(name):
    """HBM bytes per launch of a kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/kernel_traffic.json)."""
    tf = ROOT / "profiles" / "kernel_traffic.json"
    if not tf.exists():
        return None
    try:
        return json.loads(tf.read_text()).get(name)
    except Exception:
        return None


def roofline_eri(res):
    """The analytic ERI kernels of the last step: the shell-kernel buckets of
    one Schwarz phase run side by side on the library's five streams and are
    timed together (HIP events around the fork and join), like the nine
    estimator kernels of a block.  Algorithmic FP64 flops (Boys values and R
    table once per primitive quartet, Hermite contractions per component, as
    counted by the library) over the phase time; a "launch" is one phase's
    set of buckets."""
    st = res["eri_stats"]
    flops = sum(v["flops"] for v in st.values())
    ms, nset = res["an_timings"]["analytic_eri"]
    achieved = flops / (ms / 1e3) / 1e12
    heavy = max(st.items(), key=lambda kv: kv[1]["flops"])
    # HBM bytes: the committed per-kernel, per-launch traffic summed over the
    # launches of the step (None unless every kernel is in the table)
    traffic = 0.0
    for key, d in st.items():
        t = _traffic(kernel_name(key))
        if t is None or traffic is None:
            traffic = None
            continue
        traffic += t * d["launches"]
    return {"bound": "fp64-valu",
            "kernel": "k_eri_shell<LU, LL, B> / k_eri_shell_red<LU, LL, B> (the buckets of a "
                      "Schwarz phase, side by side on five streams)",
            "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64,
            "traffic": traffic / max(1, nset) if traffic is not None else None,
            "launches": nset, "kernel_launches": int(sum(v["launches"] for v in st.values())),
            "total_ms": ms, "avg_launch_ms": ms / max(1, nset),
            "flops_per_launch": flops / max(1, nset),
            "heaviest_kernel": kernel_name(heavy[0]),
            "heaviest_kernel_flop_share": heavy[1]["flops"] / flops}


def shell_firsts(bas):
    """First AO of every shell of the library (consecutive AOs with equal
    centre, fit, type and order, at most 15), then n."""
    n = bas.n
    lao = bas.npower.sum(1) + 2 * (bas.i_type == 2)
    first = []
    for i in range(n):
        same = (i > 0 and i - first[-1] < 15 and bas.i_type[i] == bas.i_type[i - 1]
                and bas.n_gauss[i] == bas.n_gauss[i - 1] and lao[i] == lao[i - 1]
                and np.array_equal(bas.center[i], bas.center[i - 1])
                and np.array_equal(bas.g_gauss[i], bas.g_gauss[i - 1])
                and np.array_equal(bas.c_gauss[:, i], bas.c_gauss[:, i - 1]))
        if not same:
            first.append(i)
    first.append(n)
    return first


def shell_units(bas):
    """The library's split of every shell pair's components into units of 6,
    3 and 1 (csrc/k_mc.hip, group_ready): returns, per shell pair (I, K), the
    first shell I and the list of units as arrays of pair indices."""
    first = shell_firsts(bas)
    out = []
    for I in range(len(first) - 1):
        for K in range(I + 1):
            comps = [i * (i + 1) // 2 + k for i in range(first[I], first[I + 1])
                     for k in range(first[K], first[K + 1]) if k <= i]
            units, at = [], 0
            for size in (6, 3, 1):
                while len(comps) - at >= size:
                    units.append(np.array(comps[at:at + size]))
                    at += size
            out.append((I, units))
    return out


def kernel_unit_pairs(bas):
    """Unit pairs (row unit U, column unit V) holding at least one quartet
    with row pair P >= column pair Q: the 1/r12 evaluations per sample."""
    sps = shell_units(bas)
    nsh = max(I for I, _ in sps) + 1
    per_I = np.zeros(nsh)
    for I, units in sps:
        per_I[I] += len(units)
    below = np.cumsum(per_I) - per_I           # units with first shell < I
    total = float(np.sum(per_I * below))
    by_I = {}
    for I, units in sps:
        by_I.setdefault(I, []).extend(units)
    for units in by_I.values():                # same first shell: per unit pair
        hi = np.array([u.max() for u in units])
        lo = np.array([u.min() for u in units])
        total += float(np.sum(hi[:, None] >= lo[None, :]))
    return total


def roofline_estimator(res, bas, prec):
    """The group estimator of one block (9 class-pair kernels running side
    by side, timed together with HIP events): algorithmic flops = 4 per
    quartet-sample (two FMAs: K (A1 A2 + B1 B2)) + 8 per unit-pair-sample
    (the three differences and r^2 of the kernel), and one reciprocal square
    root per unit-pair-sample."""
    ms, n = res["mc_timings"]["estimator"]
    samples = NBL * res["production_npts"]
    nup = kernel_unit_pairs(bas)
    flops = (4.0 * bas.nint + 8.0 * nup) * samples
    achieved = flops / (ms / 1e3) / 1e12
    peak = PEAK_FP64 if prec != "fp32" else PEAK_FP32
    t = "double" if prec != "fp32" else "float"
    traffic = None
    tf = ROOT / "profiles" / "kernel_traffic.json"
    if tf.exists():
        try:
            tj = json.loads(tf.read_text())
            parts = [v for k, v in tj.items() if k.startswith(f"k_est_group<{t}")]
            traffic = sum(parts) if parts else None
        except Exception:
            traffic = None
    return {"bound": "fp64-valu" if prec != "fp32" else "fp32-valu",
            "kernel": f"k_est_group<{t}, NR, NC, TR, TC> (9 class-pair launches per block)",
            "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": traffic, "launches": n, "total_ms": ms,
            "avg_launch_ms": ms / max(1, n), "flops_per_launch": flops / max(1, n),
            "quartet_samples_per_s": float(bas.nint) * samples / (ms / 1e3),
            "kernel_evaluations_per_s": nup * samples / (ms / 1e3)}


def density_report(res, bas, prec):
    """The density kernel (one thread per shell pair and sample): per shell
    pair and sample it evaluates the STO and its fit for the first AO of both
    shells at both points -- 4 (1 + ng) exponentials -- and writes four
    component values per pair plus six point coordinates per shell pair."""
    ms, n = res["mc_timings"]["densities"]
    samples = NBL * res["production_npts"]
    first = shell_firsts(bas)
    ng = [int(bas.n_gauss[first[I]]) for I in range(len(first) - 1)]
    nsh = len(ng)
    exps = sum(2 * ((1 + ng[I]) + (1 + ng[K])) for I in range(nsh) for K in range(I + 1))
    nsp = nsh * (nsh + 1) // 2
    width = 4 if prec == "fp32" else 8
    written = width * (4 * bas.npair + 6 * nsp) * samples
    sec = ms / 1e3
    return {"kernel": f"k_dens_group<{'float' if prec == 'fp32' else 'double'}>",
            "launches": n, "total_ms": ms, "exp_per_s": exps * samples / sec,
            "bytes_written": written, "write_GBps": written / sec / 1e9,
            "write_frac_of_hbm_peak": written / sec / 8e12}


def eri_by_class(res):
    """Per (LU, LL) shell class of the analytic phase (both lane batches and
    modes together): algorithmic FP64 flops, their share of the phase, and
    primitive quartet x lane component evaluations; heaviest first.  The
    classes run side by side, so the phase time is not split among them."""
    agg = {}
    for key, d in res["eri_stats"].items():
        k = (key[1], key[2]) if key[0] == "shell" else ("component", key[1], key[2])
        a = agg.setdefault(k, {"units": 0.0, "flops": 0.0})
        for f in ("units", "flops"):
            a[f] += d[f]
    total = sum(a["flops"] for a in agg.values()) or 1.0
    return [{"class": list(k), "flops": a["flops"], "flop_share": a["flops"] / total,
             "prim_quartet_components": a["units"]}
            for k, a in sorted(agg.items(), key=lambda kv: -kv[1]["flops"])]


def roofline_of(res, bas):
    """The dominant launch of the step (largest total time): the estimator's
    class-pair kernels of a block or the ERI buckets of a Schwarz phase, each
    set timed together."""
    eri, est = roofline_eri(res), roofline_estimator(res, bas, "fp64")
    return (est, eri) if est["total_ms"] >= eri["total_ms"] else (eri, est)


def cpu_baseline(bas, samples, nscr, budget_s=20.0):
    """The CPU oracle (C restatement of the reference, OpenMP) on a bounded
    sample of the same job, scaled to the whole job:
      analytic: gauss_ijkl on a seeded random subset of quartets, scaled by the
                number of unscreened quartets;
      MC: one 40-sample block over all quartets, scaled by the job's samples."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle
    from slater_stoch_amd import basis as B
    orc = Oracle()
    orc.set_basis(bas)
    cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    q = B.quartet_mapping(bas.n)
    nint = q[0].size
    rng = np.random.default_rng(77)
    # analytic: grow the subset until the time budget is half used
    done, t_an, nsub = 0, 0.0, 0
    perm = rng.permutation(nint)
    chunk = 64
    while t_an < budget_s / 2 and done < nint:
        pick = perm[done:done + chunk]
        t = time.perf_counter()
        orc.gauss_eri([a[pick] for a in q], None)
        t_an += time.perf_counter() - t
        done += pick.size
        chunk = min(4 * chunk, 4096)
    an_est = t_an / done * (nint - nscr)
    # MC: one block of 40 samples over every quartet
    r = np.random.default_rng(5)
    r1, r2 = r.standard_normal((40, 3)), r.standard_normal((40, 3))
    pi0 = orc.pi0(r1, r2)
    t = time.perf_counter()
    orc.mc_fixed(1, 40, r1, r2, pi0, q)
    t_mc = time.perf_counter() - t
    mc_est = t_mc * samples / 40.0
    return {"value": an_est + mc_est, "unit": "s", "cores": cores, "kind": "port",
            "sample": (f"oracle gauss_ijkl on {done} random quartets ({t_an:.1f} s) scaled to "
                       f"{nint - nscr} unscreened quartets; oracle estimator for all {nint} "
                       f"quartets x 40 samples ({t_mc:.1f} s) scaled to {samples} samples"),
            "analytic_s": an_est, "mc_s": mc_est}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp32", action="store_true",
                    help="skip the FP32- and BF16-product estimator variants")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates the rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters against known traffic)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = (1 << 30) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from slater_stoch_amd import configs
    from slater_stoch_amd.zvmc import Engine

    bas = configs.random_molecule()
    eng = Engine(0)
    for w in range(args.warmup):
        job(eng, bas, "fp64", seed=1000 + w)
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    res64 = []
    for s in range(args.steps):
        res64.append(job(eng, bas, "fp64", seed=1 + s))
        r = res64[-1]
        print(f"[bench] fp64 step {s + 1}/{args.steps}: {r['total']:.2f} s (analytic "
              f"{r['analytic']:.2f} s, MC {r['mc']:.2f} s, {r['samples']} samples, max stderr "
              f"{r['max_stderr']:.2e})", file=sys.stderr, flush=True)
    t64 = float(np.mean([r["total"] for r in res64]))
    res32 = res16 = None
    if not args.no_fp32:
        res32 = job(eng, bas, "fp32", seed=1)
        print(f"[bench] fp32 step: {res32['total']:.2f} s (MC {res32['mc']:.2f} s)",
              file=sys.stderr, flush=True)
        res16 = job(eng, bas, "bf16", seed=1)
        print(f"[bench] bf16 step: {res16['total']:.2f} s (MC {res16['mc']:.2f} s)",
              file=sys.stderr, flush=True)
    last = res64[-1]
    roof, roof_other = roofline_of(last, bas)
    precision = precision16 = None
    if res32 is not None:
        # step 1 (seed 1) and the FP32 job share the pilot samples
        precision = {"c2": precision_report(bas, res64[0]["pilot"][0], res64[0]["pilot"][1],
                                            res32["pilot"][0]),
                     "c1_water": water_precision(eng)}
        precision16 = {"c2": precision_report(bas, res64[0]["pilot"][0], res64[0]["pilot"][1],
                                              res16["pilot"][0]),
                       "c1_water": water_precision(eng, "bf16")}
        print(f"[bench] bf16 vs fp64 estimator on the same samples: max |diff| "
              f"{precision16['c2']['max_abs']:.2e} Ha (110 AOs, "
              f"{precision16['c2']['max_abs_over_fp64_stderr']:.2f} of the FP64 stderr), "
              f"{precision16['c1_water']['max_abs']:.2e} Ha (water)", file=sys.stderr, flush=True)
        print(f"[bench] fp32 vs fp64 estimator on the same samples: max |diff| "
              f"{precision['c2']['max_abs']:.2e} Ha (110 AOs), "
              f"{precision['c1_water']['max_abs']:.2e} Ha (water)", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(bas, last["samples"], last["n_screened"])
        print(f"[bench] cpu baseline estimate {cpu['value']:.0f} s on {cpu['cores']} cores",
              file=sys.stderr, flush=True)
    mct = last["mc_timings"]
    est_ms = mct["estimator"][0]
    line = {
        "metric": METRIC,
        "value": t64,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random molecule, device Philox samples)",
        "config": {"workload": "random 12-atom molecule (5 C/N/O + 7 H), 110 STO AOs, "
                               "6105 pairs, 18638565 quartets, 20-Gaussian fits; "
                               "analytic FP64 ERIs + ZVMC to max stderr <= 1e-6 Ha, 10 blocks",
                   "nint": int(bas.nint), "npair": int(bas.npair), "nbasis": int(bas.n),
                   "blocks": NBL, "samples": int(last["samples"]),
                   "max_stderr": last["max_stderr"], "estimator": "fp64"},
        "analytic_s": float(np.mean([r["analytic"] for r in res64])),
        "host_copy_s": float(np.mean([r["host_copy"] for r in res64])),
        "with_host_copy_s": float(np.mean([r["total"] + r["host_copy"] for r in res64])),
        "mc_s": float(np.mean([r["mc"] for r in res64])),
        "mixed_precision_s": None if res32 is None else res32["total"],
        "mixed_precision_max_stderr": None if res32 is None else res32["max_stderr"],
        "mixed_precision_error": precision,
        "bf16_products_s": None if res16 is None else res16["total"],
        "bf16_products_max_stderr": None if res16 is None else res16["max_stderr"],
        "bf16_products_error": precision16,
        "estimator_qsamples_per_s": (bas.nint * NBL * last["production_npts"] / (est_ms / 1e3)
                                     if est_ms > 0 else None),
        "mfma_fraction": 0.0,
        "roofline": roof,
        "roofline_other_phase": roof_other,
        "density_phase": density_report(last, bas, "fp64"),
        "eri_by_class": eri_by_class(last)[:12],
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = cpu["value"] / t64
    print(json.dumps(line))


if __name__ == "__main__":
    main()
