"""
Benchmark workloads at their full sizes, checked through properties that hold
at any size (the oracle would take minutes to hours here):

* every step's scored delta-SSE equals the SSE decrease that the append then
  realises (two independent computations: the hinge sweep or the linear
  scoring against the modified Gram-Schmidt append), within 1e-6 of sum(y^2),
  the tolerance the reference holds its hinge delta-SSE to against least
  squares (tests/unittest.cc:18, 301-362);
* the orthonormal basis stays orthonormal at the f32 storage floor;
* every hinge knot is a value of its input column with more than `tail` rows
  on each side (the knot grid of marsalgo.cc:607-626);
* the mixed form picks the plain form's terms with the per-step mse within
  1e-6 (relative), on all 256 fits of the batching workload and on the
  4M-row linear workload.
"""

import numpy as np
import pytest

from mars_amd import datagen

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _check_scores_realised(trace, n):
    """Scored delta-SSE of each step against the decrease of the append's mse (y scaled to
    unit sum of squares, so SSE = n * mse)."""
    checked = 0
    for s in trace:
        if min(s["mse"]) < 0 or (s["kind"] == "h" and len(s["mse"]) < 2):
            continue
        want = s["sse1"] if s["kind"] == "l" else s["sse2"]
        realised = (1.0 - s["sse0"]) - n * s["mse"][-1]
        assert abs(realised - want) <= 1e-6, (s["epoch"], s["kind"], realised, want)
        checked += 1
    return checked


def _check_knots(model, X, tail):
    hinges = 0
    for r in model:
        if r["type"] not in (b"+", b"-"):
            continue
        x = X[:, int(r["input"])]
        knot = np.float32(r["hinge"])
        assert float(knot) == float(r["hinge"])
        assert np.any(x == knot)
        assert np.count_nonzero(x >= knot) > tail and np.count_nonzero(x <= knot) > tail
        hinges += 1
    return hinges


def _fit(X, y, kw, precision):
    import mars_amd
    from mars_amd import marslib

    w = np.ones(len(X), dtype=np.float32)
    algo = marslib.MarsAlgo(X, y, w, kw["max_terms"], precision=precision)
    trace = []
    model = mars_amd.fit(X, y, w, algo=algo, trace=trace, precision=precision, **kw)
    return model, trace, algo


def _seq(model):
    return [(r["type"], int(r["basis"]), int(r["input"]),
             None if np.isnan(r["hinge"]) else float(r["hinge"])) for r in model]


def test_config3_full_properties():
    """1M rows x 100 columns, hinge fit to 65 terms (plain form), and the precision forms (FP32
    and BF16-split first projections): the same terms and knots, mse within 1e-6."""
    X, y = datagen.workload3()
    kw = datagen.FIT_KW[3]
    say("config 3 full: X", X.shape)
    mixed, mtrace, malgo = _fit(X, y, kw, "mixed")
    st = malgo.stats()
    _, Bo = malgo.basis()
    Qm = np.asarray(Bo, dtype=np.float64)
    malgo.close()
    bf, btrace, balgo = _fit(X, y, kw, "mixed-bf16")
    bst = balgo.stats()
    _, Bo = balgo.basis()
    Qb = np.asarray(Bo, dtype=np.float64)
    balgo.close()
    model, trace, algo = _fit(X, y, kw, "fp64")
    assert _seq(model) == _seq(mixed) == _seq(bf)
    pm = [v for s in trace for v in s["mse"]]
    bm = [v for s in btrace for v in s["mse"]]
    assert len(pm) == len(bm) and all(abs(u - v) <= 1e-6 * abs(u) for u, v in zip(pm, bm))
    say("config 3 mixed-bf16: rescored %d of %d, window violations %d, max|QtQ-I| %.3g"
        % (bst["rescored"], bst["scored"], bst["window_violations"], np.abs(Qb.T @ Qb - np.eye(Qb.shape[1])).max()))
    assert bst["window_violations"] == 0
    mm = [v for s in mtrace for v in s["mse"]]
    assert len(pm) == len(mm) and all(abs(u - v) <= 1e-6 * abs(u) for u, v in zip(pm, mm))
    say("config 3 mixed: rescored %d of %d, window violations %d, max|QtQ-I| %.3g"
        % (st["rescored"], st["scored"], st["window_violations"], np.abs(Qm.T @ Qm - np.eye(Qm.shape[1])).max()))
    assert st["window_violations"] == 0
    n = len(X)
    assert len(model) == kw["max_terms"]
    assert _check_scores_realised(trace, n) >= len(trace) - 1
    assert _check_knots(model, X, max(int(n * 0.02), 1)) > 0
    mse = [v for s in trace for v in s["mse"]]
    assert all(b <= a for a, b in zip(mse, mse[1:]))
    _, Bo = algo.basis()
    Q = np.asarray(Bo, dtype=np.float64)
    assert np.abs(Q.T @ Q - np.eye(Q.shape[1])).max() <= 1e-5


@pytest.mark.parametrize("policy", ["ungated", "none"])
def test_config4_full_mixed_equals_plain(policy, monkeypatch):
    """256 independent fits (2k-200k rows, 5-64 columns), 33 linear terms each, batched; the
    rescoring storing the residuals of its ungated candidates (the default) or of none."""
    import mars_amd

    monkeypatch.setenv("MARS_HIP_RESCORE_STORE", policy)
    fits = datagen.workload4()
    say("config 4 full: %d fits, %d rows" % (len(fits), sum(len(x) for x, _ in fits)))
    Xs, ys = [x for x, _ in fits], [y for _, y in fits]
    sp, sm = {}, {}
    from mars_amd import marslib

    live0 = marslib.device_bytes(reset_peak=True)[0]
    plain = mars_amd.fit_batch(Xs, ys, max_terms=33, stats=sp)
    peak_plain = marslib.device_bytes(reset_peak=True)[1] - live0
    mixed = mars_amd.fit_batch(Xs, ys, max_terms=33, precision="mixed", stats=sm)
    peak_mixed = marslib.device_bytes()[1] - live0
    say("config 4 peak device bytes: plain %.2f GB, mixed %.2f GB" % (peak_plain / 1e9, peak_mixed / 1e9))
    # one context's footprint: inputs, both basis copies, work lists and partial sums; the plain
    # form keeps no per-candidate residual columns (1.5 GB of X, 3.2 GB of basis columns)
    assert peak_plain <= 10e9 and peak_mixed <= 12e9
    for i, (a, b) in enumerate(zip(plain, mixed)):
        assert _seq(a) == _seq(b), i
        for u, v in zip(sp["mse"][i], sm["mse"][i]):
            assert abs(u - v) <= 1e-6 * abs(u), (i, u, v)
    assert 0 < sm["rescored"] < sm["scored"]
    assert sm["window_violations"] == 0
    say("config 4 mixed: rescored %d of %d, host redo steps %d" % (sm["rescored"], sm["scored"], sm["host_redo_steps"]))


def test_config5_full_mixed_equals_plain():
    """4M rows x 200 correlated columns, 101 linear terms: mixed against plain, and the
    plain form's scores against its appends."""
    X, y = datagen.workload5()
    kw = datagen.FIT_KW[5]
    say("config 5 full: X", X.shape)
    from mars_amd import marslib

    live0 = marslib.device_bytes(reset_peak=True)[0]
    plain, ptrace, algo = _fit(X, y, kw, "fp64")
    fired = algo.stats()["dgks_fired"]
    algo.close()
    peak_plain = marslib.device_bytes()[1] - live0
    say("config 5 plain peak device bytes: %.2f GB, DGKS retries %d" % (peak_plain / 1e9, fired))
    # X 3.2 GB, four basis copies 6.6 GB, the DGKS retry arena 2 GiB, work lists and partial sums
    assert peak_plain <= 20e9
    del algo
    mixed, mtrace, algo = _fit(X, y, kw, "mixed")
    del algo
    assert _seq(plain) == _seq(mixed)
    pm = [v for s in ptrace for v in s["mse"]]
    mm = [v for s in mtrace for v in s["mse"]]
    assert len(pm) == len(mm)
    for u, v in zip(pm, mm):
        assert abs(u - v) <= 1e-6 * abs(u), (u, v)
    assert _check_scores_realised(ptrace, len(X)) >= len(ptrace) - 1
