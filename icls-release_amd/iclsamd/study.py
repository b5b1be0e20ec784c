"""Many-drop study: spurious currents and per-drop volume, mode by mode.

    python -m iclsamd.study --n 256 --drops 64 --steps 200 --mass-correction 2
    python -m iclsamd.study --workload emulsion512 --steps 1000

runs a seeded many-drop workload in each precision mode -- drops256: static drops,
radii uniform in [3, 20] cells, ratios 1:1, sigma = 1, fluid at rest; emulsion512:
512 x 512 x 256 plane Couette flow with 200 drops of radius 4 .. 16 cells, Ca = 0.2
(params.emulsion512) -- and prints one JSON line per mode with

* ms_per_step    -- wall time of the stepping loop per step (dt adaptive, cfl 0.3);
* ca             -- spurious_currents (misc.f90:802-843), Ca = 0.1 max|u_c| / 1, every
                    `--every` steps as (step, time, Ca);
* dv_max/dv_mean -- max and mean over the drops of |V_l / V_l(0) - 1| at the end;
* closure_cells  -- cells the learned closure was consulted on in the last step;

and, when FP64 is among the modes, one more line with each other mode's per-drop volume
against FP64's: max and mean over the drops of |V_l - V_l(FP64)| / V_l(0).

Every drop is its own level set ('mdp', init.f90:901-986), stepped inside a window
around the drop (--ls-full 1: over the whole grid).
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from iclsamd import _lib, params as P  # noqa: E402

MODES = {"fp64": P.MODE_FP64, "mixed": P.MODE_MIXED, "learned": P.MODE_MIXED_LEARNED}


def run(n, drops, steps, mode, mass_correction, every, seed, cfl, ls_full=0, workload="drops256", nz=None,
        trigger_exp=0):
    if workload == "emulsion512":
        p = P.emulsion512(n, nz or n // 2, drops, mode=MODES[mode], seed=seed)
    else:
        p = P.drops256(n, drops, mode=MODES[mode], seed=seed)
    p.ls_full = ls_full
    p.cfl = cfl
    p.ioutchk = every
    p.mass_correction = mass_correction
    p.mass_trigger_exp = trigger_exp
    g = _lib.Solver(p)
    try:
        t0 = time.perf_counter()
        g.init_program()
        g.sync()
        t_init = time.perf_counter() - t0
        ca = []
        elapsed = 0.0
        for s in range(1, steps + 1):
            t0 = time.perf_counter()
            g.step(s, 0)
            if s % every == 0:
                g.sync()
                elapsed += time.perf_counter() - t0
                ca.append((s, g.scalar(P.S_TIME), g.spurious_currents()))
            else:
                elapsed += time.perf_counter() - t0
        g.sync()
        vols = [g.volume(l) for l in range(drops)]
        dv = [abs(v / v0 - 1.0) for v, v0 in vols]
        return {"_v_rel": [v / v0 for v, v0 in vols],"workload": workload, "grid": [p.itot, p.jtot, p.ktot], "mode": mode, "drops": drops,
                "steps": steps, "mass_correction": mass_correction,
                "mass_trigger": 10.0 ** -trigger_exp if trigger_exp else 1e-5,
                "ls_full": ls_full,
                "ms_per_step": round(elapsed / steps * 1e3, 3), "init_s": round(t_init, 3),
                "ca": [(s, t, float("%.4e" % c)) for s, t, c in ca],
                "dv_max": max(dv), "dv_mean": sum(dv) / len(dv),
                "closure_cells": g.mlp_listed() if mode == "learned" else None}
    finally:
        g.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="drops256", choices=["drops256", "emulsion512"])
    ap.add_argument("--n", type=int, default=None, help="cells along x and y (256; 512 for emulsion512)")
    ap.add_argument("--nz", type=int, default=None, help="cells along z (emulsion512: n / 2)")
    ap.add_argument("--drops", type=int, default=None, help="64; 200 for emulsion512")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--modes", default="fp64,learned")
    ap.add_argument("--mass-correction", type=int, default=2, choices=[0, 1, 2])
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cfl", type=float, default=0.3)
    ap.add_argument("--mass-trigger-exp", type=int, default=0,
                    help="n > 0: correct a drop's volume when |V/V0 - 1| > 10^-n (the reference: 1e-5)")
    ap.add_argument("--ls-full", type=int, default=0, choices=[0, 1],
                    help="1: every level set over the whole grid instead of a window around its drop")
    a = ap.parse_args()
    emu = a.workload == "emulsion512"
    n = a.n or (512 if emu else 256)
    drops = a.drops or (200 if emu else 64)
    vrel = {}
    for m in a.modes.split(","):
        r = run(n, drops, a.steps, m, a.mass_correction, a.every, a.seed, a.cfl, a.ls_full, a.workload, a.nz,
                a.mass_trigger_exp)
        vrel[m] = r.pop("_v_rel")
        print(json.dumps(r), flush=True)
    if "fp64" in vrel:
        cmp = {}
        for m, v in vrel.items():
            if m != "fp64":
                d = [abs(x - y) for x, y in zip(v, vrel["fp64"])]
                cmp[m] = {"dv_vs_fp64_max": max(d), "dv_vs_fp64_mean": sum(d) / len(d)}
        print(json.dumps({"workload": a.workload, "steps": a.steps, "per_drop_volume_vs_fp64": cmp}), flush=True)


if __name__ == "__main__":
    main()
