"""
Fits of more than 256 basis functions (linear terms; the device append's limit is 512).

The FP64 scoring passes hold 256 basis columns per T-GEMM launch and in the
projection's LDS tile; the columns beyond take further T-GEMM launches and
are added to the projection's FP64 sum straight from memory.  These tests run
that path:

* the same 300 appends on the device and in the oracle, then per-candidate
  linear scores at m = 301 against the oracle evaluated on the device's basis,
  plain form, and the mixed form's winner;
* whole fits past 256 terms in both forms: every step's scored delta-SSE
  equals the SSE decrease its append realises, and the two forms choose the
  same terms;
* the batching form past 256 terms: each fit of a batch chooses the terms of
  the same fit run alone.
"""

import sys

import numpy as np
import pytest

from mars_amd import datagen

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _data(n, p, seed):
    X = np.asfortranarray(datagen.normal(seed, 0, n * p).reshape(p, n).T.astype(np.float32))
    y = (X[:, 0] * X[:, 1] - X[:, 2] + datagen.normal(seed, 1, n)).astype(np.float32)
    return X, y


def _build_basis(algos, X, terms):
    """The same hinge terms appended to every algo: parents among the first 24 functions (degree
    stays small), knots at deterministic quantiles of the column; every append must agree."""
    rng = np.random.default_rng(7)
    n, p = X.shape
    added, worst = 0, 0.0
    while added < terms:
        xc = int(rng.integers(p))
        bc = int(rng.integers(min(24, algos[0].nbasis())))
        h = float(np.quantile(X[:, xc], rng.uniform(0.1, 0.9)))
        t = "+" if added % 2 == 0 else "-"
        got = [a.append(t, xc, bc, h) for a in algos]
        assert all((g < 0) == (got[0] < 0) for g in got), (added, got)
        # the mse is 1 - (sum of the appended columns' squared projections), over n: its
        # relative rounding grows with the number of terms and with 1 / mse
        for g in got[1:]:
            rel = abs(g - got[0]) / abs(got[0])
            worst = max(worst, rel)
            assert rel <= 1e-8, (added, got)
        if got[0] >= 0:
            added += 1
    say("appends: worst relative mse difference %.3g" % worst)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_linear_scores_past_256_vs_oracle(oracle, precision):
    import mars_amd

    say("linear scores at m = 301", precision)
    X, y = _data(3000, 12, 41)
    w = np.ones(len(X), dtype=np.float32)
    dev = mars_amd.marslib.MarsAlgo(X, y, w, 320, precision=precision)
    cpu = oracle.MarsAlgo(X, y, w, 320)
    _build_basis([dev, cpu], X, 300)
    m = cpu.nbasis()
    assert m == dev.nbasis() == 301
    # Over 300 appends the two f32 bases drift apart by rounding flips (each flip feeds the later
    # columns); the scores below are compared on the device's own basis, loaded into the oracle.
    B, Bo = dev.basis()
    cpu.load_basis(B, Bo)
    mask = np.ones((X.shape[1], m), dtype=bool)
    d0, l0, h0, c0 = cpu.eval(mask, 1, 1, True, 4)
    d1, l1, h1, c1, best = dev.eval_best(mask, 1, 1, True)
    assert abs(d1 - d0) <= 1e-12 * abs(d0)
    j = np.unravel_index(np.argmax(l0), l0.shape)
    assert tuple(best["j1"]) == tuple(int(v) for v in j)
    assert abs(best["sse1"] - l0[j]) <= 1e-9 * l0[j] + 1e-15
    if precision == "fp64":
        # The residual is rounded to f32 per row: where the FP64 value before rounding differs in
        # its last bits (summation order), the rounding can flip, and a score whose residual is
        # nearly orthogonal to y amplifies that flip.  Most scores agree to FP64 rounding.
        rel = np.abs(l1 - l0) / np.maximum(np.abs(l0), 1e-300)
        rel[(l0 <= 1e-15) & (np.abs(l1 - l0) <= 1e-15)] = 0.0
        say("scores: median relative difference %.3g, 99th percentile %.3g, max %.3g"
            % (np.median(rel), np.percentile(rel, 99), rel.max()))
        assert np.median(rel) <= 1e-12
        assert np.percentile(rel, 99) <= 1e-9
        np.testing.assert_allclose(l1, l0, rtol=1e-7, atol=1e-15)


def _check_scores_realised(trace, n):
    checked = 0
    for s in trace:
        if min(s["mse"]) < 0:
            continue
        realised = (1.0 - s["sse0"]) - n * s["mse"][-1]
        assert abs(realised - s["sse1"]) <= 1e-6, (s["epoch"], realised, s["sse1"])
        checked += 1
    return checked


def _terms(model):
    return [(r["type"], int(r["basis"]), int(r["input"])) for r in model]


def test_fit_past_256_terms_both_forms(oracle):
    import mars_amd
    from oracle import fit as F

    X, y = _data(3000, 14, 43)
    kw = dict(max_terms=300, linear_only=True, r2_window=10**6, r2_thresh=0.0)
    say("oracle fit to 300 terms")
    want, _ = F.fit(oracle, X, y, np.ones(len(X), dtype=np.float32), threads=-1, **kw)
    assert len(want) == 300
    for precision in ("fp64", "mixed"):
        say("fit to 300 terms", precision)
        trace = []
        model = mars_amd.fit(X, y, trace=trace, precision=precision, **kw)
        assert _check_scores_realised(trace, len(X)) >= len(trace) - 1
        assert _terms(model) == _terms(want)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_past_256_terms_matches_single_fits(precision):
    import mars_amd

    say("batch of 3 fits to 300 terms", precision)
    # (these inputs reach 300 terms in the oracle; other seeds can stop earlier, where the
    # chosen column is degenerate, in the reference's driver as here)
    data = [_data(n, p, seed) for n, p, seed in [(2000, 13, 50), (3000, 13, 52), (3000, 14, 60)]]
    kw = dict(max_terms=300, r2_window=10**6, r2_thresh=0.0, precision=precision)
    models = mars_amd.fit_batch([X for X, _ in data], [y for _, y in data], **kw)
    for (X, y), got in zip(data, models):
        want = mars_amd.fit(X, y, linear_only=True, **kw)
        assert len(got) == 300
        assert _terms(got) == _terms(want)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-bf16"])
def test_hinge_eval_past_253_is_rejected(precision):
    import mars_amd

    say("hinge eval at m = 254", precision)
    X, y = _data(2000, 8, 47)
    w = np.ones(len(X), dtype=np.float32)
    dev = mars_amd.marslib.MarsAlgo(X, y, w, 260, precision=precision)
    _build_basis([dev], X, 253)
    mask = np.ones((X.shape[1], dev.nbasis()), dtype=bool)
    with pytest.raises(RuntimeError, match="253 basis functions"):
        dev.eval_best(mask, 1, 1, False)
    d, lin, _, _, best = dev.eval_best(mask, 1, 1, True)  # linear scoring still runs
    assert best["sse1"] > 0
