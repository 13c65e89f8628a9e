"""Reductions of rocprofv3 CSV output to the short summaries kept in profiles/.

    python -m supernormal_amd.profiling kernel-means DIR OUT [--last N]
        per kernel of a --kernel-trace run: calls, mean duration over all
        dispatches and over the last N dispatches of that kernel (the late
        iterations of a training run);
    python -m supernormal_amd.profiling counter-means DIR OUT [--last N]
        per kernel and counter of a --pmc run: the mean per dispatch over the
        kernel's last N dispatches;
    python -m supernormal_amd.profiling probe-means DIR OUT
        per kernel and counter: the values of the last 5 dispatches (3 for the
        copy kernel of the calibration) and their mean -- the launches of
        bench.py --traffic-probe;
    python -m supernormal_amd.profiling timeline DIR OUT [--last N]
        the last N dispatches in start order: start offset, duration and gap
        after the previous end (us), queue, kernel.

DIR is rocprofv3's output directory (-d); the reduction runs where the trace
was taken, so only the small OUT file has to be kept.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import os


def _rows(root, suffix):
    for f in glob.glob(os.path.join(root, "**", f"*{suffix}"), recursive=True):
        with open(f) as fh:
            yield from csv.DictReader(fh)


def kernel_means(root, out, last=1000):
    durs = collections.defaultdict(list)
    for r in _rows(root, "kernel_trace.csv"):
        name = r.get("Kernel_Name") or r.get("KernelName") or ""
        durs[name].append((int(r["Start_Timestamp"]),
                           int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    rows = []
    for k, v in durs.items():
        v.sort()
        ds = [d for _, d in v]
        tail = ds[-last:]
        rows.append((sum(tail) / len(tail), k, len(ds), sum(ds) / len(ds)))
    rows.sort(reverse=True)
    with open(out, "w") as fo:
        fo.write(f"kernel\tcalls\tmean_us_all\tmean_us_last{last}\n")
        for t, k, n, a in rows:
            fo.write(f"{k[:110]}\t{n}\t{a / 1000:.1f}\t{t / 1000:.1f}\n")


def _counter_values(root):
    vals = collections.defaultdict(lambda: collections.defaultdict(float))
    for r in _rows(root, "counter_collection.csv"):
        key = (r.get("Kernel_Name", ""), r.get("Counter_Name", ""))
        vals[key][int(r.get("Dispatch_Id", 0))] += float(r.get("Counter_Value", 0))
    return vals


def counter_means(root, out, last=500):
    with open(out, "a") as fo:
        for (k, c), dv in _counter_values(root).items():
            ids = sorted(dv)[-last:]
            fo.write(f"{k[:90]}\t{c}\t{len(dv)}\t{sum(dv[i] for i in ids) / len(ids):.6g}\n")


def probe_means(root, out):
    with open(out, "a") as fo:
        for (k, c), dv in _counter_values(root).items():
            ids = sorted(dv)[-(3 if "MulFunctor" in k else 5):]
            vals = ",".join(f"{dv[i]:.1f}" for i in ids)
            fo.write(f"{k[:80]}\t{c}\t{len(dv)}\t{vals}\t{sum(dv[i] for i in ids) / len(ids):.1f}\n")


def timeline(root, out, last=400):
    rows = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                   (r.get("Kernel_Name") or "")[:70], r.get("Queue_Id", "?"))
                  for r in _rows(root, "kernel_trace.csv"))[-last:]
    if not rows:
        open(out, "w").close()
        return
    t0, prev = rows[0][0], None
    with open(out, "w") as fo:
        for s, e, k, q in rows:
            gap = (s - prev) / 1000 if prev is not None else 0.0
            fo.write(f"{(s - t0) / 1000:9.1f}\t{(e - s) / 1000:8.1f}\t{gap:8.1f}\t{q}\t{k}\n")
            prev = e if prev is None else max(prev, e)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("what", choices=["kernel-means", "counter-means", "probe-means", "timeline"])
    ap.add_argument("dir")
    ap.add_argument("out")
    ap.add_argument("--last", type=int, default=None)
    a = ap.parse_args(argv)
    if a.what == "kernel-means":
        kernel_means(a.dir, a.out, a.last or 1000)
    elif a.what == "counter-means":
        counter_means(a.dir, a.out, a.last or 500)
    elif a.what == "probe-means":
        probe_means(a.dir, a.out)
    else:
        timeline(a.dir, a.out, a.last or 400)
    print(open(a.out).read()[-6000:])
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
