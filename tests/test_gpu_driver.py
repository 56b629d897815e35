"""
The forward-pass drivers of mars.fit on the GPU.

* The C++ loop (mars_hip_fit_forward), which runs mars.fit's default
  configuration, against the numpy loop, which mars.fit keeps for Fast-MARS
  ranking, a per-candidate filter, a logger or a callback: the same model rows
  (every field but the time stamp) and the same step records, in both forms.
* The numpy loop's Fast-MARS ranking (mars.py:308-316) and xfilter
  (mars.py:323) against the oracle's restatement of the reference driver over
  the oracle's CPU scoring: the same terms and knots, mse within 1e-10.
"""

import io
import sys

import numpy as np
import pytest

from tests.golden.cases import CASES, make_inputs

pytestmark = pytest.mark.gpu

FIELDS = ("type", "basis", "input", "hinge", "r2", "r2_cv", "order")
STEP_KEYS = ("epoch", "xcol", "bcol", "cut", "knot", "kind", "sse0", "sse1", "sse2", "n_cand", "mse")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same_steps(a, b):
    assert len(a) == len(b)
    for s, t in zip(a, b):
        for k in STEP_KEYS:
            u, v = s[k], t[k]
            if isinstance(u, float) and np.isnan(u):
                assert np.isnan(v), k
            else:
                assert u == v, (k, u, v)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["w2_small", "w4_fit0", "deg3", "hinge_edges", "weighted"])
def test_cpp_driver_equals_numpy_loop(name, precision):
    import mars_amd

    say("driver", name, precision)
    case = CASES[name]
    X, y, w = make_inputs(case)
    t_cpp, t_np = [], []
    m_cpp = mars_amd.fit(X, y, w, trace=t_cpp, precision=precision, **case["fit"])
    calls = []
    m_np = mars_amd.fit(X, y, w, trace=t_np, precision=precision, callback=lambda *a: calls.append(a[1]),
                        **case["fit"])
    assert len(calls) == len(t_np)
    assert m_cpp.dtype == m_np.dtype and len(m_cpp) == len(m_np)
    for f in FIELDS:
        np.testing.assert_array_equal(m_cpp[f], m_np[f], err_msg=f)
    assert np.all(np.isnan(m_cpp["beta"]))
    assert np.all(np.diff(m_cpp["time"]) >= 0)
    _same_steps(t_cpp, t_np)


def _oracle_compare(model, trace, want_model, want_trace):
    got = [(r["type"], int(r["basis"]), int(r["input"])) for r in model]
    want = [(r["type"], int(r["basis"]), int(r["input"])) for r in want_model]
    assert got == want, "term sequence differs"
    np.testing.assert_array_equal(model["hinge"], want_model["hinge"])
    got_mse = [v for s in trace for v in s["mse"]]
    want_mse = [v for s in want_trace for v in s["mse"]]
    assert len(got_mse) == len(want_mse)
    for a, b in zip(got_mse, want_mse):
        assert abs(a - b) <= 1e-10 * max(abs(b), 1e-300), (a, b)


@pytest.mark.parametrize("kw", [dict(max_basis=4), dict(max_inputs=3),
                                dict(max_basis=6, max_inputs=5, aging_factor=2.0)],
                         ids=["max_basis", "max_inputs", "both_aged"])
@pytest.mark.parametrize("name", ["w2_small", "w5_small"])
def test_fast_mars_matches_oracle(oracle, name, kw):
    import mars_amd
    from oracle import fit as F

    say("fast-mars", name, kw)
    case = CASES[name]
    X, y, w = make_inputs(case)
    fit_kw = dict(case["fit"], **kw)
    trace = []
    model = mars_amd.fit(X, y, w, trace=trace, **fit_kw)
    want_model, want_trace = F.fit(oracle, X, y, w, **fit_kw)
    _oracle_compare(model, trace, want_model, want_trace)
    assert [s["n_cand"] for s in trace] == [s["n_cand"] for s in want_trace]


def test_xfilter_and_logger_match_oracle(oracle):
    import mars_amd
    from oracle import fit as F

    say("xfilter + logger")
    case = CASES["w2_small"]
    X, y, w = make_inputs(case)
    xf = lambda i, b: i != 7 and len(b) < 2  # noqa: E731  (input 7 carries signal in this generator)
    log = io.StringIO()
    trace = []
    model = mars_amd.fit(X, y, w, trace=trace, xfilter=xf, logger=log, **case["fit"])
    want_model, want_trace = F.fit(oracle, X, y, w, xfilter=xf, **case["fit"])
    _oracle_compare(model, trace, want_model, want_trace)
    assert 7 not in set(int(v) for v in model["input"][1:])
    lines = log.getvalue().splitlines()
    assert len(lines) == 2 + len(trace)  # header, rule, one row per step
