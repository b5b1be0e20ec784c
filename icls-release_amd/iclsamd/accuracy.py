"""Mixed-precision accuracy against the FP64 GPU run, field by field.

    python -m iclsamd.accuracy --workload drop512 --steps 200
    python -m iclsamd.accuracy --workload bubble256 --steps 500 --modes learned

runs the workload once in FP64 and once in each listed mode, on the same initial state,
and prints one JSON line per mode with

* err_linf[f] -- max |x_mode - x_fp64| / max |x_fp64| over the interior cells of the
                 field f (phi = the level set, u, v, w, p = pnew);
* err_l2[f]   -- ||x_mode - x_fp64||_2 / ||x_fp64||_2 over the same cells;
* phi_band_dz -- max |phi_mode - phi_fp64| / dz over the cells with |phi_fp64| <= 3 dz
                 (the interface band);
* phi_linf_at_dz -- |phi_fp64| / dz at the cell of phi's largest difference;
* max_abs[f]  -- max |x_fp64| and max |x_mode - x_fp64| (velocities in units of the
                 capillary velocity sigma / mu1);
* drift       -- |V / V0 - 1| of the drop at the end, for the mode and for FP64;
* dv_vs_fp64  -- |V_mode - V_fp64| / V0;
* time        -- the simulated time reached (dt adapts per run when cfl > 0);
* ms_per_step -- wall time of the stepping loop per step.

Workloads: drop512 -- one drop R = 0.2 at the centre of a 512^3 unit box, ratios 10:1,
sigma = 1, HOUC5 (the 64^3 static drop's geometry at 512^3), dt from chkdt at CFL 0.3:
the 64^3 drop's fixed dt = 1e-4 is 30 times the explicit viscous limit dx^2 Re / 12 at
512^3; bubble256 -- the 256^3 rising bubble over 1 x 1 x 2 at CFL 0.3 (params.bubble256).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from iclsamd import _lib, params as P  # noqa: E402

MODES = {"fp64": P.MODE_FP64, "mixed": P.MODE_MIXED, "learned": P.MODE_MIXED_LEARNED}
FIELDS = {"phi": P.F_LVSET, "u": P.F_UNEW, "v": P.F_VNEW, "w": P.F_WNEW, "p": P.F_PNEW}


def params_for(workload, n, mode_name, mass_correction, trigger_exp=0):
    mode = MODES[mode_name]
    if workload == "bubble256":
        p = P.bubble256(n, mode=mode)
    else:
        p = P.base_params(n)
        p.mode = mode
        p.cfl = 0.3
    p.mass_correction = mass_correction
    p.mass_trigger_exp = trigger_exp
    return p


def run(workload, n, mode, steps, mass_correction, trigger_exp=0):
    p = params_for(workload, n, mode, mass_correction, trigger_exp)
    g = _lib.Solver(p)
    try:
        g.init_program()
        g.sync()
        t0 = time.perf_counter()
        for s in range(1, steps + 1):
            g.step(s, 0)
        g.sync()
        ms = (time.perf_counter() - t0) / steps * 1e3
        v, v0 = g.volume(0)
        fields = {}
        for name, fid in FIELDS.items():
            h = (P.field_dims(p, fid)[0] - p.itot) // 2
            a = g.download(fid)
            fields[name] = np.ascontiguousarray(a[h:a.shape[0] - h, h:a.shape[1] - h, h:a.shape[2] - h])
            del a
        return {"ms_per_step": ms, "volume": v, "volume0": v0, "time": g.scalar(P.S_TIME), "dz": p.lz / p.ktot}, fields
    finally:
        g.close()


def compare(ref, got, dz):
    linf, l2 = {}, {}
    band = np.abs(ref["phi"]) <= 3 * dz
    band_dz = float(np.abs(ref["phi"][band] - got["phi"][band]).max() / dz) if band.any() else 0.0
    for name in FIELDS:
        a, b = ref[name], got[name]
        d = a - b
        scale = np.abs(a).max()
        linf[name] = float(np.abs(d).max() / scale) if scale > 0 else float(np.abs(d).max())
        na = np.sqrt(np.vdot(a, a))
        l2[name] = float(np.sqrt(np.vdot(d, d)) / na) if na > 0 else float(np.sqrt(np.vdot(d, d)))
    d = np.abs(ref["phi"] - got["phi"])
    at = float(np.abs(ref["phi"].flat[int(np.argmax(d))]) / dz)
    absmax = {name: (float("%.3e" % np.abs(ref[name]).max()), float("%.3e" % np.abs(ref[name] - got[name]).max()))
              for name in FIELDS}
    return linf, l2, band_dz, at, absmax


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="drop512", choices=["drop512", "bubble256"])
    ap.add_argument("--n", type=int, default=None, help="cells per side (512 / 256)")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--modes", default="mixed,learned")
    ap.add_argument("--mass-correction", type=int, default=2, choices=[0, 1, 2])
    ap.add_argument("--mass-trigger-exp", type=int, default=0,
                    help="n > 0: correct the volume when |V/V0 - 1| > 10^-n (the reference: 1e-5)")
    a = ap.parse_args()
    n = a.n or (512 if a.workload == "drop512" else 256)
    ref, ref_fields = run(a.workload, n, "fp64", a.steps, a.mass_correction, a.mass_trigger_exp)
    base = {"workload": a.workload, "grid": n, "steps": a.steps, "mass_correction": a.mass_correction,
            "mass_trigger": 10.0 ** -a.mass_trigger_exp if a.mass_trigger_exp else 1e-5}
    print(json.dumps(dict(base, mode="fp64", ms_per_step=round(ref["ms_per_step"], 3), time=ref["time"],
                          drift=abs(ref["volume"] / ref["volume0"] - 1.0))), flush=True)
    for m in a.modes.split(","):
        if m == "fp64":
            continue
        res, fields = run(a.workload, n, m, a.steps, a.mass_correction, a.mass_trigger_exp)
        linf, l2, band_dz, at, absmax = compare(ref_fields, fields, ref["dz"])
        del fields
        print(json.dumps(dict(base, mode=m, ms_per_step=round(res["ms_per_step"], 3), time=res["time"],
                              err_linf={k: float("%.3e" % v) for k, v in linf.items()},
                              err_l2={k: float("%.3e" % v) for k, v in l2.items()},
                              phi_band_dz=float("%.3e" % band_dz), phi_linf_at_dz=round(at, 2), max_abs=absmax,
                              drift=abs(res["volume"] / res["volume0"] - 1.0),
                              drift_fp64=abs(ref["volume"] / ref["volume0"] - 1.0),
                              dv_vs_fp64=abs(res["volume"] - ref["volume"]) / ref["volume0"])), flush=True)


if __name__ == "__main__":
    main()
