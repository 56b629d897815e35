"""
Launch-shape tuning table of the device library.

`mars_amd/tuning.json` holds, per workload class, the launch shapes that
`mars_amd.autotune` measured fastest on an MI355X (work items per FP64 pass,
waves per cache update and per hinge-sweep pass, the sweep's chunk floor, the
append's row blocks, the narrow-tile threshold; see mars_hip_tuning in
include/mars_hip.h).  A class is (driver kind, precision form, hinge sweep or
not, log2 of the rows scored per step); a workload takes the entry of its
kind, form and sweep flag whose row class is nearest, or the library defaults
when there is none.  None of these shapes changes a result beyond the order of
FP64 summation.  MARS_HIP_TUNING=0 keeps the defaults.
"""

import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")


def workload_class(kind, form, hinge, rows):
    """kind: 'batch' or 'single'; form: 'fp64' or 'mixed'; rows: rows per step (summed over a batch)."""
    return {"kind": kind, "form": form, "hinge": bool(hinge), "rows_log2": int(math.log2(max(int(rows), 1)))}


def load(path=TABLE):
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return json.load(f).get("entries", [])


def lookup(kind, form, hinge, rows, entries=None):
    """Launch shapes for a workload, or None (library defaults)."""
    if os.environ.get("MARS_HIP_TUNING", "1") == "0":
        return None
    want = workload_class(kind, form, hinge, rows)
    best, dist = None, None
    for e in load() if entries is None else entries:
        c = e["class"]
        if (c["kind"], c["form"], c["hinge"]) != (want["kind"], want["form"], want["hinge"]):
            continue
        d = abs(c["rows_log2"] - want["rows_log2"])
        if dist is None or d < dist:
            best, dist = e, d
    return None if best is None else dict(best["params"])


def apply(ctx, kind, form, hinge, rows):
    """Set the table's shapes on a marslib.MarsAlgo / marslib.Batch; returns them (None: defaults)."""
    params = lookup(kind, form, hinge, rows)
    if params:
        ctx.set_tuning(params)
    return params
