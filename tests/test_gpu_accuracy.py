"""
Accuracy of both precision forms on every benchmark configuration (at the sizes
of the golden fixtures): orthogonality loss of the final basis, max |QᵀQ - I|,
and the per-step RSS error against the reference's FP64 outputs (the golden
fixtures written by the reference built from its own sources).

Bars: plain form RSS within 1e-10 relative, the mixed forms (FP32 MFMA, and on
hinge fits the BF16-split form) within 1e-6, and
|QᵀQ - I| at the f32 storage floor (Q is stored in f32: 1e-5).
"""

import json
import os

import numpy as np
import pytest

from tests.golden.cases import CASES, make_inputs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))

# benchmark configuration -> golden case at test size
CONFIG_CASES = [("1", "w1"), ("2", "w2"), ("3", "w3_small"), ("4", "w4_fit0"), ("4", "w4_fit5"), ("5", "w5_small")]


def _fit(case, precision):
    import mars_amd
    from mars_amd import marslib

    X, y, w = make_inputs(case)
    if w is None:
        w = np.ones(len(X), dtype=np.float32)
    algo = marslib.MarsAlgo(X, y, w, case["fit"]["max_terms"], precision=precision)
    trace = []
    model = mars_amd.fit(X, y, w, algo=algo, trace=trace, precision=precision, **case["fit"])
    _, Bo = algo.basis()
    Q = np.asarray(Bo, dtype=np.float64)
    orth = float(np.abs(Q.T @ Q - np.eye(Q.shape[1])).max())
    mse = [v for s in trace for v in s["mse"]]
    return model, mse, orth


@pytest.mark.parametrize("config,name", CONFIG_CASES, ids=["cfg%s-%s" % c for c in CONFIG_CASES])
def test_accuracy_per_config(config, name):
    case = CASES[name]
    with open(os.path.join(HERE, "golden", name + ".json")) as f:
        gold = json.load(f)
    want = [a["mse"] for a in gold["appends"]]
    want_seq = [(r["type"], r["basis"], r["input"]) for r in gold["model"]]
    hinge = not case["fit"].get("linear_only", True)
    for prec in ("fp64", "mixed") + (("mixed-bf16",) if hinge else ()):
        model, mse, orth = _fit(case, prec)
        seq = [(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in model]
        assert seq == want_seq, "%s: term sequence differs" % prec
        err = max(abs(a - b) / abs(b) for a, b in zip(mse, want))
        print("config %s (%s) %-5s terms %d  max|QtQ-I| %.3g  max RSS rel err %.3g"
              % (config, name, prec, len(model), orth, err), flush=True)
        assert orth <= 1e-5
        assert err <= (1e-10 if prec == "fp64" else 1e-6)
