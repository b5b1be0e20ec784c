#!/usr/bin/env python3
"""Throughput of the ICLS two-phase step on one MI355X, in Mcell-steps/s.

Workload: the 256^3 rising bubble (1 x 1 x 2 box, density ratio 100:1, viscosity
ratio 50:1, gravity -z, Eotvos ~ 10, WENO5 level-set advection, re-initialisation and
mass-correction check every step), plain FP64, synthetic initial fields (no data
files).  A step is one pass of the reference's time loop body (src/main.f90:105-210):
level set, AB2 momentum, GFM + Poisson projection, correction, pressure update.

One JSON line on stdout:
  value       = itot*jtot*ktot*K / (seconds of K timed steps) / 1e6
  roofline    = the dominant kernel's algorithmic bytes per launch / its mean launch
                time (HIP events on the library stream over the timed region) vs the
                8 TB/s HBM peak; traffic from rocprofv3 FETCH_SIZE + WRITE_SIZE if a
                committed measurement for that kernel exists under profiles/
  cpu_baseline= the CPU oracle (C restatement of the reference, 1 core) on one step
                of the same workload
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


def cpu_baseline(p, nsteps=1):
    """The oracle on one step of the same workload (after an untimed init)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(p)
    o.init_program()
    t0 = time.perf_counter()
    for s in range(1, nsteps + 1):
        o.step(s, 0)
    el = time.perf_counter() - t0
    o.close()
    ncell = p.itot * p.jtot * p.ktot
    return {"value": ncell * nsteps / el / 1e6, "unit": "Mcell-steps/s", "cores": 1, "kind": "port",
            "sample": "%d step(s) of the %dx%dx%d workload after an untimed init, %.1f s" %
                      (nsteps, p.itot, p.jtot, p.ktot, el)}


def traffic_for(kernel, workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        rec = json.load(open(path)).get(workload, {}).get(kernel)
        return None if rec is None else rec["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="cells per direction")
    ap.add_argument("--mode", choices=["fp64", "mixed"], default="fp64",
                    help="fp64: plain FP64; mixed: FP32 fields, FP64 pressure, Thomas, residual and volumes")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--kernels", action="store_true", help="also print the per-kernel table to stderr")
    a = ap.parse_args()

    if a.gpus != 1:
        print("bench.py: this build runs the whole domain on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    from iclsamd import _lib
    fp32 = a.mode == "mixed"
    p = P.bubble256(a.n, mode=P.MODE_MIXED if fp32 else P.MODE_FP64)
    workload = "bubble%d" % a.n if not fp32 else "bubble%d_mixed" % a.n
    kbytes = R.kernel_bytes(fp32)
    ncell = p.itot * p.jtot * p.ktot
    g = _lib.Solver(p)
    g.init_program()
    istep = 0
    # warm-up: the first step loads every kernel's code object; the others bracket
    # every launch with events to find the dominant kernel
    nprof = max(a.warmup - 1, 1)
    istep += 1
    g.step(istep, 0)
    g.profile(True)
    for _ in range(nprof):
        istep += 1
        g.step(istep, 0)
    kt = g.kernel_times()
    known = {k: v for k, v in kt.items() if k in kbytes}
    dominant = max(known, key=lambda k: known[k][0])
    if a.kernels:
        tot = sum(v[0] for v in kt.values())
        for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0]):
            print("%-22s %9.3f ms/step %6.1f%%" % (k, v[0] / nprof, 100 * v[0] / tot), file=sys.stderr)
    # timed region: only the dominant kernel is bracketed by events
    g.profile_reset()
    g.profile_only(dominant)
    g.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        istep += 1
        g.step(istep, 0)
    g.sync()
    el = time.perf_counter() - t0
    kt = g.kernel_times()
    g.check_nan()
    g.close()

    ms, launches = kt[dominant]
    avg_s = ms / 1e3 / launches
    bytes_per_launch = kbytes[dominant] * ncell
    achieved = bytes_per_launch / avg_s / 1e9
    tr = traffic_for(dominant, workload)
    roof = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1), "peak": R.HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / R.HBM_PEAK_GBS, 4),
            "traffic": tr, "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": round(avg_s * 1e3, 4)}
    value = ncell * a.steps / el / 1e6
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * el / a.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if fp32 else "f64", "data": "synthetic",
        "config": {"workload": "%s: %dx%dx%d rising bubble, 1x1x2 box, %s, WENO5" %
                               (workload, p.itot, p.jtot, p.ktot,
                                "FP32 fields with FP64 pressure/Thomas/residual/volumes" if fp32 else "plain FP64"),
                   "mode": a.mode, "grid": [p.itot, p.jtot, p.ktot]},
        "roofline": roof,
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(p)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
