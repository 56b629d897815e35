#!/usr/bin/env python3
"""
Write the golden forward-pass fixtures from the reference itself.

Runs the reference driver `mars.fit` (read from /root/reference/mars.py) on
top of the reference's own `marslib` compiled from its sources by
`make -C oracle ref` (oracle/_ref).  Nothing else pins the full term sequence
of a fit -- the reference has no end-to-end fit test -- so this is the only
place the reference is executed.  A recording proxy around `marslib.MarsAlgo`
captures every eval (arg-maxima and the base dsse) and every append (its
arguments and returned mse) at full f64 precision.

Inputs are not stored: each case names a deterministic generator of
mars_amd/datagen.py plus its arguments, and the fixture keeps a SHA-256 of the
generated X and y bytes so generator drift is detected.

Usage:  python tests/golden/make_golden.py [case ...]
"""

import hashlib
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle", "_ref"))

from mars_amd import datagen  # noqa: E402
from tests.golden.cases import CASES, make_inputs  # noqa: E402

import marslib as ref_marslib  # noqa: E402  (oracle/_ref, built from /root/reference)


class Recorder:
    """Proxy module: same MarsAlgo surface, records every call."""

    def __init__(self, log):
        self.log = log
        self.__version__ = ref_marslib.__version__

    def MarsAlgo(self, X, y, w, max_terms):
        log = self.log
        inner = ref_marslib.MarsAlgo(X, y, w, max_terms)
        log["max_terms_clamped"] = int(max_terms)

        class Algo:
            def nbasis(self):
                return inner.nbasis()

            def yvar(self):
                v = inner.yvar()
                log["yvar"] = v
                return v

            def eval(self, mask, min_span, endspan, linear_only, threads, cuda=False):
                out = inner.eval(mask, min_span, endspan, linear_only, threads, cuda)
                sse0, s1, s2, cut = out
                j1 = np.unravel_index(np.argmax(s1), s1.shape)
                j2 = np.unravel_index(np.argmax(s2), s2.shape)
                log["evals"].append(
                    dict(
                        m=int(inner.nbasis()),
                        min_span=int(min_span),
                        endspan=int(endspan),
                        n_cand=int(mask.sum()),
                        sse0=float(sse0),
                        j1=[int(j1[0]), int(j1[1])],
                        sse1=float(s1[j1]),
                        j2=[int(j2[0]), int(j2[1])],
                        sse2=float(s2[j2]),
                        cut2=float(cut[j2]),
                        sum1=float(s1.sum()),
                        sum2=float(s2.sum()),
                    )
                )
                return out

            def append(self, t, xcol, bcol, h):
                mse = inner.append(t, xcol, bcol, h)
                log["appends"].append(dict(type=t, xcol=int(xcol), bcol=int(bcol),
                                           h=float(h), mse=float(mse)))
                return mse

        return Algo()


def load_reference_driver():
    spec = importlib.util.spec_from_file_location("ref_mars", os.path.join(REF, "mars.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes(order="A")).hexdigest()


def run_case(ref_mars, name, case):
    X, y, w = make_inputs(case)
    log = {"evals": [], "appends": []}
    ref_mars.marslib = Recorder(log)
    t0 = time.time()
    model = ref_mars.fit(X, y, w, **case["fit"])
    dt = time.time() - t0
    rows = [
        dict(type=r["type"].decode(), basis=int(r["basis"]), input=int(r["input"]),
             hinge=float(r["hinge"]), r2=float(r["r2"]), r2_cv=float(r["r2_cv"]),
             order=int(r["order"]))
        for r in model
    ]
    return dict(
        name=name,
        case=case,
        x_sha256=digest(X),
        y_sha256=digest(y),
        w_sha256=None if w is None else digest(w),
        seconds=dt,
        model=rows,
        **log,
    )


def main(argv):
    ref_mars = load_reference_driver()
    names = argv or list(CASES)
    for name in names:
        out = run_case(ref_mars, name, CASES[name])
        path = os.path.join(HERE, name + ".json")
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print("%-14s %3d terms %3d evals %7.2fs -> %s" % (name, len(out["model"]), len(out["evals"]),
                                                      out["seconds"], os.path.relpath(path, ROOT)))


if __name__ == "__main__":
    main(sys.argv[1:])
