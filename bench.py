#!/usr/bin/env python3
"""Throughput of the ICLS two-phase step on one MI355X, in Mcell-steps/s.

Workload: the 256^3 rising bubble (1 x 1 x 2 box, density ratio 100:1, viscosity
ratio 50:1, gravity -z, Eotvos ~ 10, WENO5 level-set advection, re-initialisation and
mass-correction check every step), synthetic initial fields (no data files).  A step
is one pass of the reference's time loop body (src/main.f90:105-210): level set, AB2
momentum, GFM + Poisson projection, correction, pressure update.

Modes (--mode):
  both          (default) mixed+learned, then plain FP64, on fresh contexts
  fp64          plain FP64, every field and kernel in double
  mixed         FP32 fields; FP64 pressure, Thomas, Poisson residual + one
                refinement solve, volume sums
  learned       mixed plus the learned curvature closure (trained on the device
                before the timed steps)

The bubble is resolved by 32 cells per radius, so the closure's gate finds no
under-resolved interface cell there and the learned mode runs the mixed step plus the
gate; the line says so.  A second workload, where the closure does work, is timed
beside it by default (--drops, off with --drops 0): the 256^3 box with 64 static drops
of radius 3 .. 20 cells (seeded), one level set per drop stepped in its own window,
CFL 0.3, normalised mass correction -- learned and plain FP64.

One JSON line on stdout:
  value       = itot*jtot*ktot*K / (seconds of K timed steps) / 1e6 of the first mode
                (mixed+learned by default) on the bubble; "fp64" holds the plain FP64
                run and "speedup_vs_fp64" the ratio; "drops256" the many-drop leg
  roofline    = that run's dominant kernel: algorithmic bytes per launch / its mean
                launch time (HIP events on the library stream over the timed region)
                vs the 8 TB/s HBM peak; traffic from rocprofv3 FETCH_SIZE + WRITE_SIZE
                where a committed measurement for that kernel exists under profiles/;
                in the many-drop leg the dominant kernel that sweeps the whole grid
                (the level-set kernels there cover windows, whose bytes vary)
  cpu_baseline= the CPU oracle (C restatement of the reference, FP64) on the host's
                cores (its OpenMP build, results identical to the serial one) over
                steps of the same workload; "cpu_baseline_1core" the serial build on
                one step
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "icls-release_amd"))

from iclsamd import params as P  # noqa: E402
from iclsamd import roofline as R  # noqa: E402

METRIC = "Mcell-steps/s at 256^3: mixed+learned vs plain FP64 GPU; %HBM roofline"
MODES = {"fp64": P.MODE_FP64, "mixed": P.MODE_MIXED, "learned": P.MODE_MIXED_LEARNED}
DESCR = {"fp64": "plain FP64",
         "mixed": "FP32 fields with FP64 pressure/Thomas/residual/volumes",
         "learned": "FP32 fields with FP64 pressure/Thomas/residual/volumes + learned curvature closure"}


def omp_threads():
    """Threads the OpenMP runtime of the oracle's parallel build uses."""
    import ctypes
    try:
        return int(ctypes.CDLL("libgomp.so.1").omp_get_max_threads())
    except OSError:
        return os.cpu_count() or 1


def cpu_baseline(p, nsteps=1, omp=False):
    """The oracle on nsteps steps of the same workload (after an untimed init); with
    omp its OpenMP build on the host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(p, omp=omp)
    o.init_program()
    t0 = time.perf_counter()
    for s in range(1, nsteps + 1):
        o.step(s, 0)
    el = time.perf_counter() - t0
    o.close()
    ncell = p.itot * p.jtot * p.ktot
    cores = omp_threads() if omp else 1
    return {"value": ncell * nsteps / el / 1e6, "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
            "sample": "%d step(s) of the %dx%dx%d workload (FP64) after an untimed init, %.1f s%s" %
                      (nsteps, p.itot, p.jtot, p.ktot, el,
                       ", OpenMP build on %d threads" % cores if omp else ", serial build")}


def traffic_for(kernel, workload, fp32):
    """Measured HBM bytes per launch of the kernel a profiler scope times, from
    profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    key = R.traffic_key(kernel, fp32)
    if key is None or not os.path.exists(path):
        return None
    try:
        rec = json.load(open(path)).get(workload, {}).get(key)
        return None if rec is None else rec["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def drops_params(n, ndrops, mode):
    """The many-drop workload: static drops of radius 3 .. 20 cells, ratios 1:1,
    sigma = 1, dt from chkdt at CFL 0.3, normalised mass correction."""
    p = P.drops256(n, ndrops, mode=MODES[mode])
    p.cfl = 0.3
    p.mass_correction = 2
    return p


# scopes whose launches cover level-set windows, not the whole grid, when every drop
# is stepped in its own window: their bytes per launch depend on the windows, so the
# many-drop leg names the dominant whole-grid kernel instead
WINDOWED = {"k_tube", "k_adv_stage1", "k_adv_stage23", "k_reinit_stage1", "k_reinit_stage23", "k_curvature",
            "k_recurvature", "k_gfm", "k_mlp_curv"}


def run_mode(mode, n, steps, warmup, show_kernels, drops=0):
    """Time `steps` steps of the workload in one precision mode; returns a dict."""
    from iclsamd import _lib
    fp32 = mode != "fp64"
    if drops:
        p = drops_params(n, drops, mode)
        workload = "drops%d" % n if mode == "fp64" else "drops%d_%s" % (n, mode)
    else:
        p = P.bubble256(n, mode=MODES[mode])
        workload = "bubble%d" % n if mode == "fp64" else "bubble%d_%s" % (n, mode)
    kbytes = R.kernel_bytes(fp32)
    ncell = p.itot * p.jtot * p.ktot
    g = _lib.Solver(p)
    try:
        t0 = time.perf_counter()
        g.init_program()                       # the learned mode trains its closure here
        g.sync()
        t_init = time.perf_counter() - t0
        istep = 0
        # warm-up: the first step loads every kernel's code object; the others bracket
        # every launch with events to find the dominant kernel
        nprof = max(warmup - 1, 1)
        istep += 1
        g.step(istep, 0)
        g.profile(True)
        for _ in range(nprof):
            istep += 1
            g.step(istep, 0)
        kt = g.kernel_times()
        known = {k: v for k, v in kt.items() if k in kbytes and not (drops and k in WINDOWED)}
        dominant = max(known, key=lambda k: known[k][0])
        if show_kernels:
            tot = sum(v[0] for v in kt.values())
            print("-- %s" % mode, file=sys.stderr)
            for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0]):
                print("%-22s %9.3f ms/step %6.1f%%" % (k, v[0] / nprof, 100 * v[0] / tot), file=sys.stderr)
        # timed region: only the dominant kernel is bracketed by events
        g.profile_reset()
        g.profile_only(dominant)
        g.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            istep += 1
            g.step(istep, 0)
        g.sync()
        el = time.perf_counter() - t0
        kt = g.kernel_times()
        g.check_nan()
        listed = g.mlp_listed() if mode == "learned" else None
    finally:
        g.close()
    ms, launches = kt[dominant]
    avg_s = ms / 1e3 / launches
    bytes_per_launch = kbytes[dominant] * ncell
    achieved = bytes_per_launch / avg_s / 1e9
    roof = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1), "peak": R.HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / R.HBM_PEAK_GBS, 4),
            "traffic": traffic_for(dominant, workload, fp32), "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": round(avg_s * 1e3, 4)}
    return {"p": p, "workload": workload, "value": ncell * steps / el / 1e6, "ms_per_step": 1e3 * el / steps,
            "roofline": roof, "init_s": t_init, "listed_cells": listed}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="cells per direction")
    ap.add_argument("--mode", choices=["both", "fp64", "mixed", "learned"], default="both")
    ap.add_argument("--drops", type=int, default=64,
                    help="drops of the many-drop leg (learned and FP64 at --n cells); 0 skips it")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--kernels", action="store_true", help="also print the per-kernel table to stderr")
    a = ap.parse_args()

    if a.gpus != 1:
        print("bench.py: the solver runs the whole domain on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is not implemented)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    first = "learned" if a.mode == "both" else a.mode
    r = run_mode(first, a.n, a.steps, a.warmup, a.kernels)
    p = r["p"]
    line = {
        "metric": METRIC, "value": round(r["value"], 2), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(r["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if first == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "%s: %dx%dx%d rising bubble, 1x1x2 box, %s, WENO5" %
                               (r["workload"], p.itot, p.jtot, p.ktot, DESCR[first]),
                   "mode": first, "grid": [p.itot, p.jtot, p.ktot]},
        "roofline": r["roofline"],
    }
    if first == "learned":
        line["config"]["closure_cells_last_step"] = r["listed_cells"]
        line["config"]["init_s_with_training"] = round(r["init_s"], 2)
        if not r["listed_cells"]:
            line["config"]["closure_note"] = ("no under-resolved interface cell on this workload: the learned "
                                              "mode runs the mixed step plus the closure's gate")
    if a.mode == "both":
        f = run_mode("fp64", a.n, a.steps, a.warmup, a.kernels)
        line["fp64"] = {"value": round(f["value"], 2), "ms_per_step": round(f["ms_per_step"], 4),
                        "roofline": f["roofline"]}
        line["speedup_vs_fp64"] = round(r["value"] / f["value"], 3)
    if a.drops > 0:
        dl = run_mode("learned", a.n, a.steps, a.warmup, a.kernels, drops=a.drops)
        df = run_mode("fp64", a.n, a.steps, a.warmup, a.kernels, drops=a.drops)
        line["drops%d" % a.n] = {
            "workload": "%dx%dx%d box, %d static drops of radius 3..20 cells (seeded), ratios 1:1, sigma 1, "
                        "per-drop windows, CFL 0.3, normalised mass correction" % (a.n, a.n, a.n, a.drops),
            "learned": {"value": round(dl["value"], 2), "ms_per_step": round(dl["ms_per_step"], 4),
                        "closure_cells_last_step": dl["listed_cells"], "roofline": dl["roofline"]},
            "fp64": {"value": round(df["value"], 2), "ms_per_step": round(df["ms_per_step"], 4),
                     "roofline": df["roofline"]},
            "speedup_vs_fp64": round(dl["value"] / df["value"], 3)}
    if not a.no_cpu:
        q = P.bubble256(a.n, mode=P.MODE_FP64)
        line["cpu_baseline"] = cpu_baseline(q, nsteps=3, omp=True)
        line["cpu_baseline_1core"] = cpu_baseline(q)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
