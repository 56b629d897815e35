"""
Edge inputs of the forward pass against the oracle, through mars.fit (the C++ driver) in the
plain and mixed forms: the term sequence, the knots and every step's mse (1e-10 relative).

* non-finite targets and zero weights: rows with a non-finite y weigh 0 (marsalgo.cc:278-307)
* an all-zero column, a constant column and a duplicated column (degenerate candidates:
  marsalgo.cc:840-846, kernels.cc:387-416)
* 20 rows (a single 32-row chunk, most of it padding) and 1001 rows (ragged last chunk)
* one X column (min_span given: the default's formula divides by zero at p = 1, mars.py:226-227,
  in the reference as here)
* heavily tied x values (knots only where x changes, marsalgo.cc:607-626)
* term budgets of 1 and 2 (no step; a step that fits only one term)

and the reference's input errors (marsalgo.cc:302, 367) with its messages.
"""

import sys

import numpy as np
import pytest

from mars_amd import datagen

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _base(n, p, seed):
    X = np.asfortranarray(datagen.normal(seed, 0, n * p).reshape(p, n).T.astype(np.float32))
    y = (np.maximum(X[:, 0] - 0.2, 0) * 2 - X[:, min(1, p - 1)] + 0.3 * datagen.normal(seed, 1, n))
    return X, y.astype(np.float32)


def _case(name):
    w = None
    if name == "nan_y_zero_w":
        X, y = _base(3001, 6, 90)
        y[[5, 77, 1000]] = np.nan
        y[2000] = np.inf
        w = np.ones(len(y), dtype=np.float32)
        w[10:40] = 0
        kw = dict(max_terms=15, linear_only=False)
    elif name == "const_dup_cols":
        X, y = _base(2500, 6, 91)
        X[:, 2] = 0
        X[:, 3] = 1.5
        X[:, 4] = X[:, 0]
        kw = dict(max_terms=15, linear_only=False)
    elif name == "tiny_n":
        X, y = _base(20, 3, 92)
        kw = dict(max_terms=9, linear_only=False, min_span=1, tail_span=0)
    elif name == "ragged_linear":
        X, y = _base(1001, 5, 96)
        kw = dict(max_terms=21, linear_only=True)
    elif name == "p1":
        X, y = _base(1001, 1, 93)
        kw = dict(max_terms=11, linear_only=False, min_span=3)
    elif name == "ties":
        X, y = _base(4000, 4, 94)
        X = np.asfortranarray(np.round(X * 2) / 2)
        kw = dict(max_terms=13, linear_only=False, min_span=2)
    elif name == "budget1":
        X, y = _base(500, 3, 95)
        kw = dict(max_terms=1)
    elif name == "budget2":
        X, y = _base(500, 3, 95)
        kw = dict(max_terms=2, linear_only=False)
    else:
        raise KeyError(name)
    return X, y, w, kw


NAMES = ["nan_y_zero_w", "const_dup_cols", "tiny_n", "ragged_linear", "p1", "ties", "budget1", "budget2"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", NAMES)
def test_edge_fit_matches_oracle(oracle, name, precision):
    import mars_amd
    from oracle import fit as F

    say("edge", name, precision)
    X, y, w, kw = _case(name)
    ww = np.ones(len(y), dtype=np.float32) if w is None else w
    want, want_trace = F.fit(oracle, X, y, ww, **kw)
    trace = []
    model = mars_amd.fit(X, y, w, trace=trace, precision=precision, **kw)
    got = [(r["type"], int(r["basis"]), int(r["input"])) for r in model]
    assert got == [(r["type"], int(r["basis"]), int(r["input"])) for r in want]
    np.testing.assert_array_equal(model["hinge"], want["hinge"])
    mse = [v for s in trace for v in s["mse"]]
    want_mse = [v for s in want_trace for v in s["mse"]]
    assert len(mse) == len(want_mse)
    for a, b in zip(mse, want_mse):
        assert (a < 0) == (b < 0)
        assert abs(a - b) <= 1e-10 * max(abs(b), 1e-300), (a, b)
    if name == "budget1":
        assert len(model) == 1 and not trace


@pytest.mark.parametrize("bad", ["y_all_nan", "y_all_zero", "w_all_zero", "x_inf"])
def test_input_errors_match_reference(bad):
    import mars_amd

    say("input error", bad)
    X, y = _base(300, 3, 97)
    w = None
    if bad == "y_all_nan":
        y[:] = np.nan
    elif bad == "y_all_zero":
        y[:] = 0
    elif bad == "w_all_zero":
        w = np.zeros(len(y), dtype=np.float32)
    else:
        X[7, 1] = np.inf
    msg = "not all columns in X are finite" if bad == "x_inf" else "target Y is all zero or NANs"
    with pytest.raises(Exception, match=msg):
        mars_amd.fit(X, y, w, max_terms=5)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_edge_batch_matches_single_fits(precision):
    """One batch of the linear versions of the edge inputs (non-finite targets and zero weights,
    degenerate columns, 20 rows, ragged rows, one X column, ties): each fit's terms and mse equal
    the same fit run alone."""
    import mars_amd

    names = ["nan_y_zero_w", "const_dup_cols", "tiny_n", "ragged_linear", "p1", "ties"]
    say("edge batch", precision)
    cases = [_case(nm) for nm in names]
    ws = [np.ones(len(y), dtype=np.float32) if w is None else w for _, y, w, _ in cases]
    stats = {}
    models = mars_amd.fit_batch([c[0] for c in cases], [c[1] for c in cases], ws, max_terms=13,
                                precision=precision, stats=stats)
    for (X, y, _, _), w, got, mse in zip(cases, ws, models, stats["mse"]):
        trace = []
        # (min_span is unused by linear fits; given, as the default's formula fails at p = 1)
        want = mars_amd.fit(X, y, w, max_terms=13, linear_only=True, min_span=1, trace=trace,
                            precision=precision)
        assert [(r["type"], int(r["basis"]), int(r["input"])) for r in got] == \
            [(r["type"], int(r["basis"]), int(r["input"])) for r in want]
        want_mse = [v for s in trace for v in s["mse"] if v >= 0]
        got_mse = [v for v in mse if v >= 0]
        assert len(got_mse) == len(want_mse)
        for a, b in zip(got_mse, want_mse):
            assert abs(a - b) <= 1e-10 * max(abs(b), 1e-300), (a, b)
