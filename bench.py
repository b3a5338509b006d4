#!/usr/bin/env python3
"""Time to a 1e-6 Ha standard error on every unique (ik|jl): the ZVMC ERI job.

One step is the whole two-electron job of the reference's two_e_slater for
the 200-orbital workload: 40 atoms (5 C, 35 H) at seeded positions in a 10 A
box, C 1S,2S,2S,2P,3D and H 1S,2P with the reference's own example-basis
exponents (example/VB1), 20100 pairs, 202,015,050 quartets, 20 Gaussians per
STO fit:

  1. analytic Gaussian-fit ERIs with the Cauchy-Schwarz driver (FP64),
  2. a pilot of 10 blocks x 1000 Monte Carlo samples over every quartet,
  3. while the largest standard error is above 1e-6 Ha: a further run of 10
     blocks for the quartets above it only (zv_mc_refine: the estimator
     tiles that hold them), sized from their error (stderr ~ 1/sqrt(samples)).

The timed region starts with the basis resident on the device and ends with
the ERIs, means and standard errors computed in HBM (the sample-size policy
reads the largest standard error reduced on the device); copying the three
arrays to the host comes after it (``host_copy_s`` / ``with_host_copy_s``).
FP64 is the plain GPU form (``value``); the FP32 estimator job is reported
beside it (``mixed_precision_s``) with its error against the FP64 estimator
on the same samples.  Also reported: the 110-AO workload and water as extra
keys, the configuration's Monte Carlo sample stream as a throughput, the
cleanups of the pair supermatrix (randomized Rayleigh-Ritz of rank 2000 with
an FP32 sketch, block Davidson for the 10 lowest eigenpairs), and the
reference's CPU path (the reference build, oracle/_ref/libref.so) on a
bounded sample of the same job, with the C port of the oracle beside it.
Prints one JSON line.
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

from slater_stoch_amd.shells import kernel_unit_pairs, mfma_tile_use, shell_firsts  # noqa: E402

METRIC = "sec to 1e-6 Ha stderr on all (ij|kl), mixed-prec vs FP64 GPU vs CPU; MFMA %"
TARGET = 1e-6
NBL = 10
PILOT = 1000  # samples per pilot block
MAX_ROUNDS = 6
PEAK_FP64 = 78.6  # TFLOP/s, MI355X vector FP64
PEAK_FP32 = 157.3  # TFLOP/s, MI355X vector FP32


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def job(eng, bas, prec, seed, keep_pilot=False):
    """One step.  The per-quartet arrays stay in HBM; with keep_pilot the
    pilot's means and errors are fetched with the clock stopped (for the
    precision report)."""
    from slater_stoch_amd.zvmc import RESULT_ERI, RESULT_MEAN, RESULT_STDERR
    eng.set_basis(bas)  # inputs resident; clears the primitive-pair cache
    t0 = time.perf_counter()
    _, pre, nscr = eng.gauss_eri(device_only=True)
    t1 = time.perf_counter()
    eri_stats = eng.eri_stats()
    an_timings = eng.timings()
    eng.mc_blocks(NBL, PILOT, prec, seed=seed, stream=0, device_only=True)
    pilot_timings = eng.timings()
    emax = eng.result_max(RESULT_STDERR)
    t2 = time.perf_counter()
    pilot_emax = emax
    paused = 0.0
    pilot = None
    if keep_pilot:
        tp = time.perf_counter()
        pilot = (eng.get_result(RESULT_MEAN), eng.get_result(RESULT_STDERR))
        paused = time.perf_counter() - tp
    rounds = []
    npts = PILOT
    while emax > TARGET and len(rounds) < MAX_ROUNDS:
        npts = min(int(math.ceil(npts * (emax / TARGET) ** 2 * 1.3 / 40.0)) * 40, 1_000_000)
        nref, emax = eng.mc_refine(NBL, npts, prec, target=TARGET, seed=seed,
                                   stream=1 + len(rounds))
        tm = eng.timings()
        rounds.append({"quartets": nref, "npts": npts, "max_stderr": emax,
                       "estimator_ms": tm["estimator"][0], "densities_ms": tm["densities"][0]})
    t3 = time.perf_counter()
    tc = time.perf_counter()
    eri = eng.get_result(RESULT_ERI)
    mean, err = eng.get_result(RESULT_MEAN), eng.get_result(RESULT_STDERR)
    host_copy = time.perf_counter() - tc
    return dict(total=t3 - t0 - paused, analytic=t1 - t0, pilot=t2 - t1,
                refine=t3 - t2 - paused, host_copy=host_copy, max_stderr=emax,
                pilot_max_stderr=pilot_emax, rounds=rounds,
                samples=NBL * PILOT + sum(NBL * r["npts"] for r in rounds),
                n_screened=nscr, pre=pre, eri_stats=eri_stats, an_timings=an_timings,
                pilot_timings=pilot_timings, eri=eri, mean=mean, err=err, pilot_arrays=pilot)


def job_summary(r):
    return {"total_s": r["total"], "analytic_s": r["analytic"], "pilot_s": r["pilot"],
            "refine_s": r["refine"], "max_stderr": r["max_stderr"],
            "pilot_max_stderr": r["pilot_max_stderr"], "refine_rounds": r["rounds"],
            "n_screened": int(r["n_screened"])}


def precision_report(bas, m64, e64, m32):
    """FP32-estimator error against the FP64 estimator on the same samples:
    per-quartet mean differences, overall and binned by the number of
    distinct nuclei of the quartet (the classes of src/compare_nuclei.F90)."""
    from slater_stoch_amd import basis as B
    q = B.quartet_mapping(bas.n)
    cls = B.nuclei_class(bas, *q)
    del q
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


def same_sample_precision(eng, bas, prec):
    """prec against FP64 on the same 10 x PILOT samples (a pilot run)."""
    eng.set_basis(bas)
    m64, e64 = eng.mc_blocks(NBL, PILOT, "fp64", seed=3, stream=0)
    m32, _ = eng.mc_blocks(NBL, PILOT, prec, seed=3, stream=0)
    return precision_report(bas, m64, e64, m32)


def kernel_name(key):
    if key[0] == "shell":
        _, lu, ll, clb, red = key
        return f"k_eri_shell{'_red' if red else ''}<{lu}, {ll}, {clb}>"
    _, a, b = key
    return f"k_eri_class<{a}, {b}>" if a < 5 else "k_eri_generic"


def _traffic_table():
    """HBM bytes per launch of each kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary."""
    tf = ROOT / "profiles" / "kernel_traffic.json"
    if not tf.exists():
        return {}
    try:
        return json.loads(tf.read_text())
    except ValueError:
        return {}


def roofline_eri(res):
    """The analytic ERI kernels of the step: the shell-kernel buckets of one
    Schwarz phase run side by side on the library's five streams and are
    timed together (HIP events around the fork and join on the library's
    stream).  Algorithmic FP64 flops (Boys values and R table once per
    primitive quartet, Hermite contractions per component over its nonzero
    Hermite coefficients, as counted by the library) over the phase time; a
    "launch" is one phase's set of buckets."""
    st = res["eri_stats"]
    flops = sum(v["flops"] for v in st.values())
    ms, nset = res["an_timings"]["analytic_eri"]
    achieved = flops / (ms / 1e3) / 1e12
    heavy = max(st.items(), key=lambda kv: kv[1]["flops"])
    table = _traffic_table()
    traffic = 0.0
    for key, d in st.items():
        t = table.get(kernel_name(key))
        if t is None:
            traffic = None
            break
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


def mfma_eligibility(bas):
    """Why the estimator stays on the vector unit, measured on this basis:
    the per-shell-pair-pair products that matrix cores would run
    (shells.mfma_tile_use), as the share of quartets whose two shell pairs
    both fill a 16-wide tile and the useful share of the tile area with 16 x
    16 tiles (BF16/FP16 MFMA) and with 4 x 4 tiles (FP64 MFMA)."""
    s16, u16 = mfma_tile_use(bas, 16)
    _, u4 = mfma_tile_use(bas, 4)
    return {"quartets_in_16_wide_blocks": s16, "tile_area_used_16x16": u16,
            "tile_area_used_4x4": u4}


def roofline_estimator(timings, bas, prec, samples):
    """The group estimator over every quartet (the pilot): one "launch" is
    the nine class-pair kernels of a chunk, timed together.  Algorithmic
    flops = 4 per quartet-sample (two FMAs: K (A1 A2 + B1 B2)) + 8 per
    unit-pair-sample (the three differences and r^2 of the kernel); one
    reciprocal square root per unit-pair-sample."""
    ms, n = timings["estimator"]
    nup = kernel_unit_pairs(bas)
    flops = (4.0 * bas.nint + 8.0 * nup) * samples
    achieved = flops / (ms / 1e3) / 1e12
    peak = PEAK_FP64 if prec == "fp64" else PEAK_FP32
    t = "double" if prec == "fp64" else "float"
    parts = [v for k, v in _traffic_table().items() if k.startswith(f"k_est_group<{t}")]
    return {"bound": "fp64-valu" if prec == "fp64" else "fp32-valu",
            "kernel": f"k_est_group<{t}, NR, NC, TR, TC> (9 class-pair launches per chunk)",
            "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": sum(parts) if parts else None, "launches": n, "total_ms": ms,
            "avg_launch_ms": ms / max(1, n), "flops_per_launch": flops / max(1, n),
            "quartet_samples_per_s": float(bas.nint) * samples / (ms / 1e3),
            "kernel_evaluations_per_s": nup * samples / (ms / 1e3)}


def density_report(timings, bas, prec, samples):
    """The density kernel (one thread per shell pair and sample): per shell
    pair and sample it evaluates the STO and its fit for the first AO of both
    shells at both points -- 4 (1 + ng) exponentials -- and writes four
    component values per pair plus six point coordinates per shell pair."""
    ms, n = timings["densities"]
    first = shell_firsts(bas)
    ng = [int(bas.n_gauss[first[I]]) for I in range(len(first) - 1)]
    nsh = len(ng)
    exps = sum(2 * ((1 + ng[I]) + (1 + ng[K])) for I in range(nsh) for K in range(I + 1))
    nsp = nsh * (nsh + 1) // 2
    width = 8 if prec == "fp64" else 4
    written = width * (4 * bas.npair + 6 * nsp) * samples
    sec = ms / 1e3
    return {"kernel": f"k_dens_group<{'double' if prec == 'fp64' else 'float'}>",
            "launches": n, "total_ms": ms, "exp_per_s": exps * samples / sec,
            "bytes_written": written, "write_GBps": written / sec / 1e9,
            "write_frac_of_hbm_peak": written / sec / 8e12}


def eri_by_class(res):
    """Per (LU, LL) shell class of the analytic phase: algorithmic FP64
    flops and their share, heaviest first."""
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


def sample_stream(eng, bas, npts):
    """The configuration's Monte Carlo stream (1e8 samples in 1e6-sample
    chunks, every quartet) as a throughput: one chunk (10 blocks of npts
    samples; 1e6 samples by default) per precision, quartet-samples per
    second and the time the hundred chunks of 1e8 samples would take at
    that rate."""
    eng.set_basis(bas)
    out = {}
    for prec in ("fp64", "fp32"):
        t = time.perf_counter()
        eng.mc_blocks(NBL, npts, prec, seed=9, stream=9, device_only=True)
        sec = time.perf_counter() - t
        m = NBL * npts
        out[prec] = {"samples": m, "seconds": sec,
                     "quartet_samples_per_s": bas.nint * m / sec,
                     "estimated_s_for_1e8_samples": sec * 1e8 / m,
                     "estimator_ms": eng.timings()["estimator"][0]}
    return out


def cleanups(eng, bas, eri_final):
    """Cleanups of the pair supermatrix V[P][Q] = (ik|jl) (20100 x 20100,
    3.2 GB) of the final ERIs (analytic + Monte Carlo mean): randomized
    Rayleigh-Ritz of rank 2000 with 2 power iterations and an FP32 sketch,
    then block Davidson for the 10 lowest eigenpairs (with the shift-invert
    preconditioner, and plain for comparison), each timed alone."""
    eng.set_basis(bas)
    out = {}
    t = time.perf_counter()
    _, ritz, nrem = eng.rsvd_clean(eri_final, rank=2000, q=2, seed=1, sketch="fp32")
    out["rsvd"] = {"rank": 2000, "power_iterations": 2, "sketch": "fp32",
                   "seconds": time.perf_counter() - t, "negative_removed": nrem,
                   "ritz_min": float(ritz.min()), "ritz_max": float(ritz.max())}
    for pc, mi in (("shift-invert", 400), ("none", 300)):
        t = time.perf_counter()
        _, ev, res, nrem, its = eng.davidson_lowest(eri_final, nev=10, tol=1e-9, max_iter=mi,
                                                    precond=pc)
        out[f"davidson_{pc}"] = {"nev": 10, "tol": 1e-9, "max_iter": mi, "precond": pc,
                                 "seconds": time.perf_counter() - t, "iterations": its,
                                 "lowest": ev.tolist(), "residuals": res.tolist(),
                                 "converged": int((res <= 1e-9).sum()), "negative_removed": nrem}
    return out


def _random_canonical(rng, n, count):
    """Random canonical quartets (i k | j l): pairs P = (i,k), Q = (j,l) drawn
    uniformly, P >= Q; returned as (i, k, j, l)."""
    npair = n * (n + 1) // 2
    P = rng.integers(0, npair, count)
    Q = rng.integers(0, npair, count)
    P, Q = np.maximum(P, Q), np.minimum(P, Q)

    def ik(p):
        i = np.floor((np.sqrt(8.0 * p + 1.0) - 1.0) / 2.0).astype(np.int64) + 1
        i = np.where(i * (i - 1) // 2 > p, i - 1, i)
        i = np.where(i * (i + 1) // 2 <= p, i + 1, i)
        return i, p - i * (i - 1) // 2 + 1
    i, k = ik(P)
    j, l = ik(Q)
    return tuple(a.astype(np.int32) for a in (i, k, j, l))


def cpu_baseline(bas, pre, nscr, pilot_emax, budget_s=24.0):
    """The reference's CPU path on a bounded sample of the same job.

    Analytic: the reference's gauss_ijkl (src/gauss_ijkl.F90, compiled from
    its own sources into oracle/_ref/libref.so) on random canonical quartets
    that pass the Schwarz test, one thread, until half the budget is used;
    scaled by the number of unscreened quartets.
    Monte Carlo: the reference's densities, Jacobians and estimator (the
    harness's restatement of src/two_e_slater.F90:398-461 over the
    reference's compute_densities) for one 40-sample walker block over two
    random quartet subsets, fitted as a + b nq and scaled to every quartet
    and to the samples the reference's own policy needs (one npts for all
    quartets: the pilot's samples per block times (pilot max stderr / 1e-6)^2).
    The oracle's C port (OpenMP, all host cores) is timed on the same
    quartets beside it (``port``): the calibration ratio port / reference."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle, Reference
    rng = np.random.default_rng(77)
    n = bas.n
    npts_ref = max(PILOT, int(math.ceil(PILOT * (pilot_emax / TARGET) ** 2 / 40.0)) * 40)
    samples_ref = NBL * npts_ref
    qi, qk, qj, ql = _random_canonical(rng, n, 400000)
    live = np.abs(pre[qk - 1, qi - 1] * pre[ql - 1, qj - 1]) >= 1e-30
    qi, qk, qj, ql = qi[live], qk[live], qj[live], ql[live]
    try:
        ref = Reference(bas)
        kind = "reference"
    except (FileNotFoundError, OSError):
        ref = None
        kind = "port"
    orc = Oracle()
    orc.set_basis(bas)
    out = {"kind": kind, "samples_reference_policy": samples_ref}
    done, t_an = 0, 0.0
    if ref is not None:
        t0 = time.perf_counter()
        while done < qi.size:
            ref.gauss_ijkl(qi[done], qk[done], qj[done], ql[done])
            done += 1
            if done % 16 == 0 and time.perf_counter() - t0 > budget_s / 2:
                break
        t_an = time.perf_counter() - t0
    m = done if done else min(qi.size, 20000)
    t0 = time.perf_counter()
    orc.gauss_eri([qi[:m], qj[:m], qk[:m], ql[:m]], None)
    t_port_an = time.perf_counter() - t0
    r = np.random.default_rng(5)
    r1, r2 = r.standard_normal((40, 3)), r.standard_normal((40, 3))
    pi0 = orc.pi0(r1, r2)
    nq1, nq2 = 20000, 200000
    si, sk, sj, sl = _random_canonical(rng, n, nq2)
    quartets = [si, sj, sk, sl]  # is, js, ks, ls

    def fit(fn):
        ts = []
        for nq in (nq1, nq2):
            t0 = time.perf_counter()
            fn([a[:nq] for a in quartets])
            ts.append(time.perf_counter() - t0)
        b = (ts[1] - ts[0]) / (nq2 - nq1)
        return max(ts[0] - b * nq1, 0.0), b
    a_p, b_p = fit(lambda q: orc.mc_fixed(1, 40, r1, r2, pi0, q))
    cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    an_port = t_port_an / m * (bas.nint - nscr)
    mc_port = (samples_ref / 40.0) * (a_p + b_p * bas.nint)
    out["port"] = {"cores": cores, "analytic_s": an_port, "mc_s": mc_port,
                   "value": an_port + mc_port}
    if ref is not None:
        a_r, b_r = fit(lambda q: ref.mc_fixed(1, 40, r1, r2, pi0, q))
        an_ref = t_an / done * (bas.nint - nscr)
        mc_ref = (samples_ref / 40.0) * (a_r + b_r * bas.nint)
        out.update({"value": an_ref + mc_ref, "unit": "s", "cores": 1,
                    "analytic_s": an_ref, "mc_s": mc_ref,
                    "sample": (f"reference gauss_ijkl on {done} random unscreened quartets "
                               f"({t_an:.1f} s, 1 thread) scaled to {bas.nint - nscr} "
                               f"unscreened quartets; reference estimator on 40 samples over "
                               f"{nq1} and {nq2} random quartets (fit a + b nq) scaled to "
                               f"{bas.nint} quartets x {samples_ref} samples (the reference's "
                               f"one-npts policy); the oracle's C port on the same quartets "
                               f"in 'port' ({cores} threads)")})
        out["port"]["ratio_to_reference"] = {
            "analytic_per_quartet": (t_port_an / m) / (t_an / done),
            "mc_per_quartet": b_p / b_r if b_r > 0 else None}
    else:
        out.update({"value": an_port + mc_port, "unit": "s", "cores": cores,
                    "analytic_s": an_port, "mc_s": mc_port,
                    "sample": f"oracle C port (reference build absent) on {m} quartets"})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline only: skip the FP32 job, the 110-AO and water reports, "
                         "the sample stream and the cleanups")
    ap.add_argument("--stream-npts", type=int, default=100000,
                    help="samples per block of the sample-stream throughput run (10 blocks: "
                         "the default is one 1e6-sample chunk)")
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

    bas = configs.large_ch()
    eng = Engine(0)
    for w in range(args.warmup):
        r = job(eng, bas, "fp64", seed=1000 + w)
        log(f"warmup {w + 1}/{args.warmup}: {r['total']:.2f} s")
    res64 = []
    for s in range(args.steps):
        res64.append(job(eng, bas, "fp64", seed=1 + s, keep_pilot=(s == 0)))
        r = res64[-1]
        log(f"fp64 step {s + 1}/{args.steps}: {r['total']:.2f} s (analytic {r['analytic']:.2f} s, "
            f"pilot {r['pilot']:.2f} s, refine {r['refine']:.2f} s over "
            f"{[x['quartets'] for x in r['rounds']]} quartets, max stderr {r['max_stderr']:.2e})")
    t64 = float(np.mean([r["total"] for r in res64]))
    last = res64[-1]
    extra = {}
    if not args.no_extra:
        res32 = job(eng, bas, "fp32", seed=1, keep_pilot=True)
        log(f"fp32 step: {res32['total']:.2f} s (pilot {res32['pilot']:.2f} s, refine "
            f"{res32['refine']:.2f} s)")
        # step 1 (seed 1) and the FP32 job share the pilot samples
        prec_c5 = precision_report(bas, res64[0]["pilot_arrays"][0],
                                   res64[0]["pilot_arrays"][1], res32["pilot_arrays"][0])
        log(f"fp32 vs fp64 on the same samples: max |diff| {prec_c5['max_abs']:.2e} Ha "
            f"({prec_c5['max_abs_over_fp64_stderr']:.3f} of the FP64 stderr)")
        c2 = configs.random_molecule()
        c2_64 = job(eng, c2, "fp64", seed=1)
        c2_32 = job(eng, c2, "fp32", seed=1)
        log(f"110 AO: fp64 {c2_64['total']:.2f} s, fp32 {c2_32['total']:.2f} s")
        water = configs.water_minimal()
        extra = {
            "mixed_precision_s": res32["total"],
            "mixed_precision_max_stderr": res32["max_stderr"],
            "mixed_precision_refine_rounds": res32["rounds"],
            "mixed_precision_error": {
                "c5": prec_c5,
                "c2": same_sample_precision(eng, c2, "fp32"),
                "c1_water": same_sample_precision(eng, water, "fp32")},
            "bf16_products_error": {"c2": same_sample_precision(eng, c2, "bf16"),
                                    "c1_water": same_sample_precision(eng, water, "bf16")},
            "c2_110ao": {"workload": "random 12-atom molecule (5 C/N/O + 7 H), 110 STO AOs, "
                                     "18638565 quartets", "fp64": job_summary(c2_64),
                         "fp32": job_summary(c2_32), "roofline": roofline_eri(c2_64)},
            "sample_stream": sample_stream(eng, bas, args.stream_npts),
        }
        log(f"sample stream: {extra['sample_stream']}")
        ft = time.perf_counter()
        extra["cleanup"] = cleanups(eng, bas, last["eri"] + last["mean"])
        log(f"cleanups {time.perf_counter() - ft:.1f} s: {extra['cleanup']}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(bas, last["pre"], last["n_screened"], last["pilot_max_stderr"])
        log(f"cpu baseline ({cpu['kind']}, {cpu['cores']} cores): {cpu['value']:.4g} s")
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
        "data": "synthetic (seeded 40-atom C/H geometry, reference example exponents, "
                "device Philox samples)",
        "config": {"workload": "40-atom C/H molecule (5 C: 1S,2S,2S,2P,3D; 35 H: 1S,2P; "
                               "example/VB1 exponents), 200 STO AOs, 20100 pairs, "
                               "202015050 quartets, 20-Gaussian fits; analytic FP64 ERIs + "
                               "ZVMC to max stderr <= 1e-6 Ha (10-block pilot, then targeted "
                               "refinement)",
                   "nint": int(bas.nint), "npair": int(bas.npair), "nbasis": int(bas.n),
                   "blocks": NBL, "pilot_samples": NBL * PILOT, "samples": int(last["samples"]),
                   "max_stderr": last["max_stderr"], "estimator": "fp64"},
        "analytic_s": float(np.mean([r["analytic"] for r in res64])),
        "pilot_s": float(np.mean([r["pilot"] for r in res64])),
        "refine_s": float(np.mean([r["refine"] for r in res64])),
        "refine_rounds": last["rounds"],
        "n_screened": int(last["n_screened"]),
        "host_copy_s": float(np.mean([r["host_copy"] for r in res64])),
        "with_host_copy_s": float(np.mean([r["total"] + r["host_copy"] for r in res64])),
        "mfma_fraction": 0.0,
        "mfma_eligibility": mfma_eligibility(bas),
        "roofline": roofline_eri(last),
        "roofline_other_phase": roofline_estimator(last["pilot_timings"], bas, "fp64",
                                                   NBL * PILOT),
        "density_phase": density_report(last["pilot_timings"], bas, "fp64", NBL * PILOT),
        "eri_by_class": eri_by_class(last)[:12],
        "cpu_baseline": cpu,
    }
    line.update(extra)
    if cpu:
        line["speedup_vs_cpu"] = cpu["value"] / t64
    print(json.dumps(line))


if __name__ == "__main__":
    main()
