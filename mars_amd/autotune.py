"""
Autotuner of the library's launch shapes (mars_hip_tuning), run on the GPU:

    python -m mars_amd.autotune [--out mars_amd/tuning.json] [--reps 2]

For each workload class below it times complete forward passes (inputs
already on the device, the context's creation untimed) under candidate launch
shapes, by coordinate descent from the library defaults: one knob at a time,
each of its candidate values with the other knobs at their best so far, the
fastest kept.  It then times the defaults and the chosen shapes once more and
writes both times with the shapes to the table that mars_amd.tuning applies.
Every trial is checked to choose the same term sequence as the defaults (a
launch shape only reorders FP64 sums).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from mars_amd import datagen, marslib, tuning  # noqa: E402
from mars_amd import mars as M  # noqa: E402

GRIDS = {
    "pass_items": [16384.0, 32768.0, 65536.0],
    "retry_items": [2048.0, 4096.0, 8192.0],
    "update_waves": [8192.0, 16384.0, 32768.0],
    "sweep_waves": [32768.0, 65536.0, 131072.0],
    "sweep_min_rows": [256, 512, 1024],
    "append_rows": [4096, 8192, 16384],
    "narrow_avg": [16, 24, 48],
}


def _batch_run(fits, kw, form):
    Xs, ys = [f[0] for f in fits], [f[1] for f in fits]
    kw = dict(kw)
    kw.pop("linear_only", None)
    max_terms = kw.pop("max_terms")

    def run(params):
        b = marslib.Batch(Xs, ys, None, max_terms, precision=form)
        try:
            b.set_tuning(params)
            t0 = time.perf_counter()
            models = M.fit_batch(Xs, ys, batch=b, max_terms=max_terms, precision=form, tuning=False, **kw)
            dt = time.perf_counter() - t0
        finally:
            b.close()
        return dt, [[(r["type"], int(r["basis"]), int(r["input"])) for r in m] for m in models]

    return run


def _single_run(X, y, kw, form):
    w = np.ones(len(X), dtype=np.float32)

    def run(params):
        a = marslib.MarsAlgo(X, y, w, kw["max_terms"], precision=form)
        try:
            a.set_tuning(params)
            t0 = time.perf_counter()
            model = M.fit(X, y, w, algo=a, precision=form, **kw)
            dt = time.perf_counter() - t0
        finally:
            a.close()
        seq = [(r["type"], int(r["basis"]), int(r["input"]), None if np.isnan(r["hinge"]) else float(r["hinge"]))
               for r in model]
        return dt, seq

    return run


def tune(run, knobs, reps, log):
    """Coordinate descent over `knobs` from the defaults; returns (params, trials)."""
    best = marslib.default_tuning()

    def timed(params):
        ts, seq = [], None
        for _ in range(reps):
            dt, s = run(params)
            ts.append(dt)
            seq = s
        return min(ts), seq

    t_best, ref = timed(best)
    trials = [dict(params=dict(best), seconds=t_best)]
    log("  defaults %.4f s" % t_best)
    for k in knobs:
        for v in GRIDS[k]:
            if v == best[k]:
                continue
            cand = dict(best, **{k: v})
            t, seq = timed(cand)
            if seq != ref:
                raise RuntimeError("launch shape %s=%s changed the term sequence" % (k, v))
            trials.append(dict(params={k: v}, seconds=t))
            log("  %-14s %-8s %.4f s" % (k, v, t))
            if t < t_best:
                t_best, best = t, cand
    return best, trials


def workloads():
    """(class, description, runner, knobs) of the tuned workload classes."""
    out = []
    fits = datagen.workload4()
    rows4 = sum(len(f[0]) for f in fits)
    for form, knobs in (("fp64", ["pass_items", "retry_items", "append_rows", "narrow_avg"]),
                        ("mixed", ["update_waves", "pass_items", "append_rows"])):
        out.append((tuning.workload_class("batch", form, False, rows4),
                    "256 independent fits, N log-uniform [2000, 200000], p uniform [5, 64], max_terms=33",
                    _batch_run(fits, datagen.FIT_KW[4], form), knobs))
    X2, y2 = datagen.workload2()
    out.append((tuning.workload_class("single", "fp64", True, len(X2)),
                "one hinge fit: N=100000, p=50, max_terms=41",
                _single_run(X2, y2, datagen.FIT_KW[2], "fp64"), ["sweep_waves", "sweep_min_rows", "pass_items"]))
    X3, y3 = datagen.workload3()
    kw3 = dict(datagen.FIT_KW[3], max_terms=15)
    out.append((tuning.workload_class("single", "fp64", True, len(X3)),
                "one hinge fit: N=1000000, p=100, first 14 terms",
                _single_run(X3, y3, kw3, "fp64"), ["sweep_waves", "sweep_min_rows"]))
    X5, y5 = datagen.workload5(n=1_000_000, p=100)
    kw5 = dict(datagen.FIT_KW[5], max_terms=41)
    for form, knobs in (("mixed", ["update_waves", "pass_items", "append_rows"]),
                        ("fp64", ["pass_items", "retry_items", "append_rows"])):
        out.append((tuning.workload_class("single", form, False, len(X5)),
                    "one linear fit: N=1000000, p=100 correlated, max_terms=41",
                    _single_run(X5, y5, kw5, form), knobs))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=tuning.TABLE)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    entries = []
    for cls, desc, run, knobs in workloads():
        log("%s: %s" % (json.dumps(cls), desc))
        params, trials = tune(run, knobs, args.reps, log)
        # final comparison, interleaved
        t_def, t_tun = [], []
        for _ in range(3):
            t_def.append(run(marslib.default_tuning())[0])
            t_tun.append(run(params)[0])
        entries.append(dict(**{"class": cls}, workload=desc, params=params, default_s=min(t_def),
                            tuned_s=min(t_tun), trials=trials))
        log("  chosen %s: %.4f s (defaults %.4f s)" % (json.dumps({k: params[k] for k in knobs}), min(t_tun),
                                                       min(t_def)))
    with open(args.out, "w") as f:
        json.dump({"device": "MI355X (gfx950)", "method": "coordinate descent from the defaults, min of %d runs"
                   % args.reps, "entries": entries}, f, indent=1)
    log("wrote %s" % args.out)


if __name__ == "__main__":
    main()
