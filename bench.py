#!/usr/bin/env python3
"""
Forward-pass benchmark on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload 4|2] [--precision fp64|mixed|both]

Default workload: 256 independent linear_only fits (N log-uniform in
[2,000, 200,000], p uniform in [5, 64], max_terms=33, one seed per fit), run
by the batching form.  A step is one complete forward pass of all 256 fits
(every epoch: candidate scoring, device argmax, GCV choice, device append),
with the inputs already resident in HBM when the timed region starts (device
upload and the per-fit constructors run before the clock; their time is
reported as `upload_s`).  Both forms run on the same inputs: the plain FP64
form and the mixed form; `value` is candidates scored per second by the mixed
form (candidates = (X column, parent) pairs summed over every fit's epochs),
`fit_s` the seconds per step of each form, `same_term_sequence` whether both
chose identical (type, parent, variable) sequences for every fit, and
`mse_rel_diff` the largest per-step relative RSS difference between them.

Workload 2 (one N=100,000 p=50 fit with the full hinge sweep) runs one fit
per step through mars.fit.

The line carries the roofline of the dominant kernel class (HIP events on the
library's stream, live in the timed region) and the CPU baseline: the
reference's own CPU path (compiled from its sources into oracle/_ref) or,
where that is absent, the oracle's restatement, timed on this host in the same
run on a bounded sample of the same workload.
"""

import argparse
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12     # MI355X FP64 (vector and matrix), FLOP/s
PEAK_FP32 = 157.3e12    # MI355X FP32 matrix, FLOP/s
PEAK_HBM = 8.0e12       # MI355X HBM3E, B/s
FP32_KERNELS = ("mixed_update",)
# Kernel classes whose event brackets span several kernels (no single-kernel roofline, no
# rocprofv3 average to compare with): the exact rescoring and DGKS retry, the append (build,
# one MGS kernel per basis column, gate, store, final) and the FP64 passes (each pass kernel
# plus its reduce kernel).  The roofline is reported for the largest single-kernel class.
COMPOSITE = ("mixed_recheck", "dgks_retry", "append", "tgemm_f64", "project_f64", "finalize", "plain_graph")
CPU_SAMPLE_S = 20.0


def workload_inputs(k):
    from mars_amd import datagen

    if k == 4:
        return datagen.workload4(), dict(datagen.FIT_KW[k])
    return datagen.GENERATORS[k](), dict(datagen.FIT_KW[k])


def run_batch(fits, fit_kw, precision, steps, warmup):
    """Batching form: `steps` complete forward passes of every fit, device upload untimed."""
    import mars_amd
    from mars_amd import marslib

    Xs = [f[0] for f in fits]
    ys = [f[1] for f in fits]
    kw = dict(fit_kw)
    kw.pop("linear_only", None)
    max_terms = kw.pop("max_terms")

    def make():
        return marslib.Batch(Xs, ys, None, max_terms, precision=precision)

    for _ in range(warmup):
        b = make()
        mars_amd.fit_batch(Xs, ys, batch=b, max_terms=max_terms, precision=precision, **kw)
        b.close()
    # one context at a time: each step's inputs are uploaded (untimed) into a fresh context, its
    # forward pass is timed, and the context is freed before the next, so device memory does not
    # grow with the step count
    models, stats = [], []
    wall = upload = 0.0
    marslib.device_bytes(reset_peak=True)
    for _ in range(steps):
        t0 = time.perf_counter()
        b = make()
        t1 = time.perf_counter()
        st = {}
        models.append(mars_amd.fit_batch(Xs, ys, batch=b, max_terms=max_terms, precision=precision,
                                         stats=st, timing=True, **kw))
        t2 = time.perf_counter()  # fit_batch returns after the stream has drained
        b.close()
        stats.append(st)
        upload += t1 - t0
        wall += t2 - t1
    upload_s = upload / steps
    peak_bytes = marslib.device_bytes()[1]
    ks = {}
    for st in stats:
        for r in st["kernels"]:
            d = ks.setdefault(r["name"], dict(launches=0, ms=0.0, flops=0.0, bytes=0.0))
            for f in ("launches", "ms", "flops", "bytes"):
                d[f] += r[f]
    seq = [[(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in m] for m in models[0]]
    st = stats[0]
    return dict(wall=wall, fit_s=wall / steps, upload_s=upload_s, cands=st["scored"], kstats=ks, seq=seq,
                mse=st["mse"], epochs=st["steps"], dgks=st["dgks_fired"], rescored=st["rescored"], gated=st["gated"],
                terms=sum(len(m) for m in models[0]), peak_device_bytes=peak_bytes)


def run_form(X, y, fit_kw, precision, steps, warmup):
    import numpy as np

    import mars_amd
    from mars_amd import marslib

    w = np.ones(len(X), dtype=np.float32)
    max_terms = fit_kw["max_terms"]

    def one(timing):
        algo = marslib.MarsAlgo(X, y, w, max_terms, precision=precision)
        algo.set_timing(timing)
        return algo

    for _ in range(warmup):
        a = one(False)
        mars_amd.fit(X, y, w, algo=a, precision=precision, **fit_kw)
        a.close()
    traces, models, ks = [], [], {}
    wall = upload = 0.0
    dgks = None
    marslib.device_bytes(reset_peak=True)
    for _ in range(steps):  # one context at a time (see run_batch)
        t0 = time.perf_counter()
        a = one(True)
        t1 = time.perf_counter()
        tr = []
        models.append(mars_amd.fit(X, y, w, algo=a, trace=tr, precision=precision, **fit_kw))
        t2 = time.perf_counter()
        traces.append(tr)
        for r in a.kernel_stats():
            d = ks.setdefault(r["name"], dict(launches=0, ms=0.0, flops=0.0, bytes=0.0))
            for f in ("launches", "ms", "flops", "bytes"):
                d[f] += r[f]
        if dgks is None:
            st0 = a.stats()
            dgks = st0["dgks_fired"]
        a.close()
        upload += t1 - t0
        wall += t2 - t1
    cands = sum(s["n_cand"] for s in traces[0])
    seq = [(r["type"].decode(), int(r["basis"]), int(r["input"]), float(r["hinge"])) for r in models[0]]
    return dict(wall=wall, fit_s=wall / steps, upload_s=upload / steps, cands=cands, kstats=ks, seq=seq,
                trace=traces[0], epochs=len(traces[0]), dgks=dgks, peak_device_bytes=marslib.device_bytes()[1],
                mse=[[v for s in traces[0] for v in s["mse"]]], rescored=st0["rescored"], gated=st0["gated"],
                terms=len(models[0]))


def binding(name, flops, nbytes):
    """The roof that bounds a kernel class at its own arithmetic intensity, as the roofline model's
    attainable rate min(peak flops, intensity x peak HBM bandwidth) picks it: the matrix peak of its
    arithmetic (FP32 for the cache update, FP64 otherwise) when its algorithmic flops per byte reach
    that peak's ridge point, HBM below it; a class that counts only one of the two is bound by it."""
    peak = PEAK_FP32 if name in FP32_KERNELS else PEAK_FP64
    if flops > 0 and (nbytes <= 0 or flops / nbytes >= peak / PEAK_HBM):
        return "mfma", peak
    if nbytes > 0:
        return "hbm", PEAK_HBM
    return None, None


def roofline(ks):
    single = [(k, v) for k, v in ks.items() if k not in COMPOSITE and not k.startswith("host_")]
    if not single or max(v["ms"] for _, v in single) <= 0:  # a form without single-kernel classes
        single = [(k, v) for k, v in ks.items() if not k.startswith("host_")]
    name, d = max(single, key=lambda kv: kv[1]["ms"])
    avg_s = d["ms"] / 1e3 / d["launches"]
    bound, peak = binding(name, d["flops"], d["bytes"])
    common = dict(traffic=None, avg_launch_ms=d["ms"] / d["launches"], launches=d["launches"],
                  intensity_flop_per_byte=d["flops"] / d["bytes"] if d["bytes"] > 0 else None)
    if bound == "mfma":
        ach = d["flops"] / d["launches"] / avg_s
        return name, dict(bound="mfma", achieved=ach / 1e12, peak=peak / 1e12, unit="TFLOP/s", frac=ach / peak,
                          **common)
    ach = d["bytes"] / d["launches"] / avg_s
    return name, dict(bound="hbm", achieved=ach / 1e9, peak=PEAK_HBM / 1e9, unit="GB/s", frac=ach / PEAK_HBM,
                      **common)


def cpu_threads():
    """Threads the CPU path uses with threads=-1: min(64, processors available), as the reference's
    binding picks them (marslib.cc:61-62)."""
    try:
        procs = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        procs = os.cpu_count() or 1
    return min(64, procs)


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(nproc=os.cpu_count(), cpu_model=model)


def class_rooflines(ks):
    """Achieved rate of every timed kernel class: summed algorithmic flops (or bytes) over the
    launches its timer brackets, divided by the bracketed device time.  Composite classes bracket
    several kernels (a pass kernel and its reduction, the append's kernels), so their rate is that of
    the whole group."""
    out = {}
    for name, d in ks.items():
        if name.startswith("host_") or d["ms"] <= 0:
            continue
        t = d["ms"] / 1e3
        e = dict(ms=d["ms"], launches=d["launches"], composite=name in COMPOSITE)
        bound, peak = binding(name, d["flops"], d["bytes"])
        if bound == "mfma":
            e.update(bound="mfma", achieved_tflops=d["flops"] / t / 1e12, frac=d["flops"] / t / peak)
        elif bound == "hbm":
            e.update(bound="hbm", achieved_gbs=d["bytes"] / t / 1e9, frac=d["bytes"] / t / PEAK_HBM)
        out[name] = e
    return out


def _cpu_lib():
    """The reference CPU path (oracle/_ref, built from the reference's sources) if built, else the
    oracle restatement."""
    import importlib.util

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    ref = glob.glob(os.path.join(ROOT, "oracle", "_ref", "marslib*.so"))
    if ref:
        spec = importlib.util.spec_from_file_location("marslib", ref[0])
        lib = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(lib)
        return lib, "reference"
    import oracle_marslib as lib

    return lib, "port"


def _epoch_sizes(trace):
    """(candidates, basis size) of each epoch of a forward-pass trace."""
    out, m = [], 1
    for s in trace:
        out.append((int(s["n_cand"]), m))
        m += sum(1 for v in s["mse"] if v >= 0)
    return out


def cpu_baseline(X, y, fit_kw, budget_s, gpu_trace=None, full=False):
    """The CPU path on the same inputs.  full: the whole fit.  Otherwise the epochs that start
    within budget_s of the fit's start; with the GPU fit's trace, the whole fit's CPU time is then
    extrapolated: each sampled epoch's seconds are fitted as a P m + c (P candidates, m basis
    functions: the orthonormalization and the sweep both cost O(n m) per candidate; c the per-epoch
    overhead), and the model is summed over the GPU fit's epochs (the same term sequence).  On
    config 2, where the whole fit is timed too, the model from its first 8 (16) epochs came within
    +41% (+10%) of the measured total on 8 cores."""
    import numpy as np

    from oracle import fit as F

    lib, kind = _cpu_lib()
    cores = cpu_threads()
    w = np.ones(len(X), dtype=np.float32)
    if full:
        t0 = time.perf_counter()
        _, tr = F.fit(lib, X, y, w, threads=-1, **fit_kw)
        spent = time.perf_counter() - t0
        cands = sum(int(s.get("n_cand", 0)) for s in tr)
        return dict(value=cands / spent, unit="candidates/s", cores=cores, kind=kind, fit_s=spent,
                    sample="the whole fit (%d epochs, %d candidates, %.1f s)" % (len(tr), cands, spent),
                    **host_info())
    # one run that starts no new epoch after the budget: every epoch that fits is sampled
    t0 = time.perf_counter()
    _, tr = F.fit(lib, X, y, w, threads=-1, max_seconds=budget_s, **fit_kw)
    spent = time.perf_counter() - t0
    cands = sum(int(s.get("n_cand", 0)) for s in tr)
    out = dict(value=cands / spent if spent > 0 else None, unit="candidates/s", cores=cores, kind=kind,
               sample="first %d epochs of the same fit (%d candidates, %.1f s)" % (len(tr), cands, spent),
               **host_info())
    if gpu_trace and len(tr) >= 3:
        sz = _epoch_sizes(tr)
        A = np.array([[p * m, 1.0] for p, m in sz])
        t = np.array([s["secs"] for s in tr])
        coef, *_ = np.linalg.lstsq(A, t, rcond=None)
        full_sz = _epoch_sizes(gpu_trace)
        est = float(sum(coef[0] * p * m + coef[1] for p, m in full_sz))
        out.update(fit_s_extrapolated=est, extrapolation=(
            "extrapolated: per-epoch seconds fitted as a P m + c on the %d sampled epochs "
            "(a=%.3g s, c=%.3g s), summed over the %d epochs of the GPU fit (same term sequence)"
            % (len(tr), coef[0], coef[1], len(full_sz))),
            candidates_per_s_extrapolated=sum(p for p, _ in full_sz) / est if est > 0 else None)
    return out


def cpu_baseline_batch(fits, fit_kw, budget_s):
    """Reference CPU path on whole fits of the batch, in order, until about budget_s of CPU time."""
    import numpy as np

    from oracle import fit as F

    lib, kind = _cpu_lib()
    cores = cpu_threads()
    spent, cands, k = 0.0, 0, 0
    while k < len(fits) and spent < budget_s:
        X, y = fits[k]
        w = np.ones(len(X), dtype=np.float32)
        t0 = time.perf_counter()
        _, tr = F.fit(lib, X, y, w, threads=-1, **fit_kw)
        spent += time.perf_counter() - t0
        cands += sum(int(s.get("n_cand", 0)) for s in tr)
        k += 1
    return dict(value=cands / spent if spent > 0 else None, unit="candidates/s", cores=cores, kind=kind,
                sample="fits 0..%d of the 256, complete (%d candidates, %.1f s)" % (k - 1, cands, spent),
                **host_info())


def copy_probe(mib=1024, reps=4):
    """A plain device copy of known size (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read_per_dispatch": n * 4, "copy_probe_bytes_written_per_dispatch": n * 4,
                      "dispatches": reps}))


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if there is one."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    e = t.get("kernels", {}).get(kernel)
    if not e or e.get("workload") != workload:
        return None
    return e["hbm_bytes_per_launch"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=4, choices=[2, 3, 4, 5])
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "mixed-bf16", "both", "all"],
                    help="both: fp64 and mixed; all: fp64, mixed and mixed-bf16 (BF16-split products on the "
                         "hinge path)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=CPU_SAMPLE_S)
    ap.add_argument("--copy-probe", action="store_true", help="run only a 1 GiB device copy (PMC calibration)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the batch step's scoring kernels directly instead of replaying captured graphs "
                         "(profiling runs: kernel traces are taken on direct launches)")
    ap.add_argument("--cache-f64", action="store_true",
                    help="run the mixed form's cache update on FP64 MFMA instead of FP32 (MARS_HIP_CACHE=f64): "
                         "separates the FP32 arithmetic's share of the mixed form's speed from the cache's")
    args = ap.parse_args()
    if args.cache_f64:
        os.environ["MARS_HIP_CACHE"] = "f64"
    if args.no_graph:
        os.environ["MARS_HIP_GRAPH"] = "0"
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a multi-GPU run "
              "would shard the candidate columns, or the independent fits of a batch, one process "
              "per GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    data, fit_kw = workload_inputs(args.workload)
    forms = {"both": ["fp64", "mixed"], "all": ["fp64", "mixed-bf16", "mixed"]}.get(args.precision, [args.precision])
    if args.workload == 4:
        res = {f: run_batch(data, fit_kw, f, args.steps, args.warmup) for f in forms}
        workload = ("256 independent fits, N log-uniform [2000, 200000], p uniform [5, 64], "
                    "max_terms=33, linear_only (batching form)")
    else:
        X, y = data
        res = {f: run_form(X, y, fit_kw, f, args.steps, args.warmup) for f in forms}
        workload = "one fit: N=%d, p=%d, %s" % (X.shape[0], X.shape[1], fit_kw)
    head = res[forms[-1]]
    kname, roof = roofline(head["kstats"])
    roof["traffic"] = measured_traffic(kname, args.workload)
    out = {
        "metric": "forward-pass fit s + candidates/s, mixed vs FP64 GPU at same term sequence",
        "value": head["cands"] / head["fit_s"],
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["fit_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32 (MFMA) + f64", "mixed-bf16": "bf16x3 (MFMA) + f32 + f64"}[forms[-1]],
        "data": "synthetic (mars_amd.datagen, seeded)",
        "config": dict(workload=workload, precision=forms[-1],
                       graphs=os.environ.get("MARS_HIP_GRAPH", "1") != "0",
                       cache="f64" if os.environ.get("MARS_HIP_CACHE") == "f64" else "f32"),
        "fit_s": {f: r["fit_s"] for f, r in res.items()},
        "upload_s": {f: r["upload_s"] for f, r in res.items()},
        "candidates_per_fit": head["cands"],
        "epochs": head["epochs"],
        "terms": head.get("terms", len(head["seq"])),
        "dgks_fired": head["dgks"],
        "same_term_sequence": (len({json.dumps(r["seq"]) for r in res.values()}) == 1) if len(res) > 1 else None,
        "roofline_kernel": kname,
        "roofline_rule": ("largest kernel class whose timer brackets exactly one kernel launch; bound: the "
                          "roof its algorithmic flops per byte reach (MFMA peak at or above the ridge point, "
                          "else HBM)"),
        "roofline": roof,
        "kernels": {k: dict(v, avg_ms=v["ms"] / max(v["launches"], 1)) for k, v in head["kstats"].items()},
        "class_rooflines": {f: class_rooflines(r["kstats"]) for f, r in res.items()},
        "peak_device_bytes": {f: r["peak_device_bytes"] for f, r in res.items()},
    }
    if "fp64" in res and "mixed" in res:
        out["speedup_mixed_vs_fp64"] = res["fp64"]["fit_s"] / res["mixed"]["fit_s"]
    if "fp64" in res and out["same_term_sequence"]:
        diffs = {f: max((abs(a - b) / b for ma, mb in zip(r["mse"], res["fp64"]["mse"]) for a, b in zip(ma, mb)),
                        default=0.0) for f, r in res.items() if f != "fp64"}
        if diffs:
            out["mse_rel_diff"] = max(diffs.values())
            out["mse_rel_diff_by_form"] = diffs
    if "rescored" in head:
        out["rescored"] = head["rescored"]
        out["rescored_gated"] = head["gated"]
    if not args.no_cpu:
        if args.workload == 4:
            out["cpu_baseline"] = cpu_baseline_batch(data, fit_kw, args.cpu_budget)
        else:
            # config 2 runs whole; configs 3 and 5 take hours on a CPU: sampled and extrapolated
            out["cpu_baseline"] = cpu_baseline(X, y, fit_kw, args.cpu_budget, gpu_trace=head["trace"],
                                               full=args.workload == 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
