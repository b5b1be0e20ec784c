"""HBM traffic per kernel launch from rocprofv3 counter runs.

Two subcommands:

  python -m iclsamd.traffic copy [--mib M] [--reps R]
      runs a plain device-to-device copy of M MiB, R times (the calibration kernel:
      it reads exactly M MiB and writes exactly M MiB per launch);

  python -m iclsamd.traffic summarize --fetch DIR --write DIR --calib-fetch DIR
         --calib-write DIR --workload NAME --cells N [--calib-mib M] [--out FILE]
      reads the counter_collection CSVs that `rocprofv3 --pmc FETCH_SIZE` and
      `rocprofv3 --pmc WRITE_SIZE` (each its own run, --output-format csv) wrote for
      the workload and for the copy, scales each counter by
          (bytes the copy moves) / (what the same counter reads for the copy),
      and writes the mean HBM bytes per launch of every kernel (fetch + write) into
      a JSON file keyed by workload and kernel name (template arguments stripped).

The counters report KiB per dispatch.  FETCH_SIZE counts L2 fills from HBM and
WRITE_SIZE L2 write-backs, so both see only HBM-side traffic.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys


def _rows(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection*.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection CSV under %s" % d)
    for f in files:
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                yield r


def _base(name):
    """Kernel name without template arguments, argument list or launch suffix."""
    n = name.strip()
    n = n.split("(")[0]
    n = re.sub(r"<.*>", "", n)
    n = n.replace(" [clone .kd]", "").replace(".kd", "")
    return n.split("::")[-1]


def per_dispatch(d, counter):
    """{full kernel name: [KiB of each dispatch]} for one counter (summed over the
    counter's instances within a dispatch)."""
    acc = {}
    for r in _rows(d):
        cname = r.get("Counter_Name") or r.get("counter_name") or ""
        if cname != counter:
            continue
        kname = r.get("Kernel_Name") or r.get("kernel_name") or "?"
        disp = r.get("Dispatch_Id") or r.get("dispatch_id") or r.get("Correlation_Id")
        v = float(r.get("Counter_Value") or r.get("counter_value") or 0.0)
        acc.setdefault(kname, {}).setdefault(disp, 0.0)
        acc[kname][disp] += v
    return {k: list(v.values()) for k, v in acc.items()}


def calib_factor(d, counter, mib, kernel="copyBuffer"):
    """Bytes the copy really moves per dispatch / bytes the counter reports.  The copy
    kernel is the device-to-device copy that `copy` launches (HIP's copyBuffer); the
    other kernels of that run (fill, compare) are ignored."""
    best = None
    for k, vals in per_dispatch(d, counter).items():
        if kernel not in k:
            continue
        mean_kib = max(vals)
        if best is None or mean_kib > best[1]:
            best = (k, mean_kib)
    if best is None or best[1] <= 0:
        raise SystemExit("copy kernel %r not found in %s" % (kernel, d))
    return mib * 1024.0 / best[1], best[0], best[1]


def copy_main(a):
    import torch
    n = a.mib * 1024 * 1024 // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    print("copied %d MiB x %d" % (a.mib, a.reps), flush=True)


def summarize_main(a):
    ff, fk, fkib = calib_factor(a.calib_fetch, "FETCH_SIZE", a.calib_mib)
    wf, wk, wkib = calib_factor(a.calib_write, "WRITE_SIZE", a.calib_mib)
    fetch = per_dispatch(a.fetch, "FETCH_SIZE")
    write = per_dispatch(a.write, "WRITE_SIZE")
    out = {}
    for k in set(fetch) | set(write):
        fv, wv = fetch.get(k, []), write.get(k, [])
        if not fv or not wv:
            continue
        fb = ff * 1024.0 * sum(fv) / len(fv)
        wb = wf * 1024.0 * sum(wv) / len(wv)
        rec = {"bytes_per_launch": round(fb + wb), "fetch_bytes": round(fb), "write_bytes": round(wb),
               "bytes_per_cell": round((fb + wb) / a.cells, 2), "dispatches": len(fv), "kernel": k}
        out[k.split("(")[0].strip().split("::")[-1]] = rec
        # under the bare name too; several instantiations of one template keep the
        # one with the most traffic
        b = _base(k)
        prev = out.get(b)
        if prev is None or rec["bytes_per_launch"] > prev["bytes_per_launch"]:
            out[b] = rec
    meta = {"calibration": {"copy_mib": a.calib_mib, "fetch_factor": round(ff, 4), "write_factor": round(wf, 4),
                            "fetch_kernel": fk, "fetch_kib_read": fkib, "write_kernel": wk, "write_kib_read": wkib},
            "cells": a.cells}
    data = {}
    if a.out and os.path.exists(a.out):
        data = json.load(open(a.out))
    data[a.workload] = out
    data.setdefault("_meta", {})[a.workload] = meta
    txt = json.dumps(data, indent=1, sort_keys=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(txt + "\n")
    else:
        print(txt)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m iclsamd.traffic")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=512)
    c.add_argument("--reps", type=int, default=5)
    s = sub.add_parser("summarize")
    s.add_argument("--fetch", required=True)
    s.add_argument("--write", required=True)
    s.add_argument("--calib-fetch", required=True)
    s.add_argument("--calib-write", required=True)
    s.add_argument("--calib-mib", type=int, default=512)
    s.add_argument("--workload", required=True)
    s.add_argument("--cells", type=int, required=True)
    s.add_argument("--out")
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_main(a)
    else:
        summarize_main(a)


if __name__ == "__main__":
    sys.exit(main())
