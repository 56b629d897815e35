"""
GPU parity: the HIP path against the oracle (CPU restatement of the reference)
and against the golden fixtures written by the reference itself.

Bars (plain FP64 form):
  * fit term sequence (type, input, basis, knot/cut) identical to the reference;
  * per-step append mse within 1e-10 relative of the reference;
  * kernel-level outputs within the f32 storage floor of the oracle.
"""

import ctypes as C
import glob
import json
import os
import sys

import numpy as np
import pytest

from tests import abi
from tests.golden.cases import CASES, make_inputs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    return abi.load()


def orthonormal_rows(n, m, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, m)))
    return np.ascontiguousarray(q.astype(np.float32))


def mirror_ctx(lib, n, max_terms, tol, B_f, Bo_c, y=None):
    ctx = C.c_void_p()
    abi.check(lib, lib.mars_hip_context_create(n, max_terms, tol, C.byref(ctx)))
    m = Bo_c.shape[1]
    abi.check(lib, lib.mars_hip_context_sync_basis(ctx, m, abi.ptr(B_f), n, abi.ptr(Bo_c), m))
    if y is not None:
        abi.check(lib, lib.mars_hip_context_set_target(ctx, abi.ptr(y)))
    return ctx


@pytest.mark.parametrize("n,m,p", [(200, 4, 3), (1000, 17, 9), (8192, 40, 32), (5000, 1, 1)])
def test_orthonormalize_vs_oracle(lib, oracle, n, m, p):
    say("orthonormalize n=%d m=%d p=%d" % (n, m, p))
    rng = np.random.default_rng(n + m)
    Bo = orthonormal_rows(n, m, n)
    B = np.asfortranarray(rng.standard_normal((n, m)).astype(np.float32))
    x = rng.standard_normal(n).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    mask = np.sort(rng.choice(m, size=p, replace=p > m)).astype(np.int32)
    tol = n * 0.02 * 2.220446049250313e-16
    ref, ref_fired = oracle.orthonormalize(B, x, mask, Bo, tol)
    ctx = mirror_ctx(lib, n, max(m, p) + 1, tol, B, Bo, y)
    try:
        Bx = np.zeros((p, n), dtype=np.float32)  # (n, p) column-major
        ybx = np.zeros(p)
        fired = C.c_long(0)
        abi.check(lib, lib.mars_hip_orthonormalize(ctx, m, p, abi.ptr(x), abi.ptr(mask), abi.ptr(Bx), n,
                                                   abi.ptr(ybx), C.byref(fired)))
    finally:
        lib.mars_hip_context_destroy(ctx)
    got = Bx.T
    # same f32 rounding points: identical up to rare 1-ulp flips of the f32 residual
    diff = np.abs(got - ref)
    assert diff.max() <= 4 * np.finfo(np.float32).eps * max(1.0, np.abs(ref).max())
    assert np.mean(got == ref) > 0.999
    want_ybx = np.array([oracle.dot_widen(np.ascontiguousarray(ref[:, j]), y) for j in range(p)])
    np.testing.assert_allclose(ybx, want_ybx, rtol=1e-6, atol=1e-9)
    assert fired.value == ref_fired


def test_dgks_fires_once_on_severe_cancellation(lib, oracle):
    """95/5 energy split against Bo[:,0] (reference test_kernels.cc:256-314)."""
    n, m = 200, 4
    say("dgks 95/5")
    Bo = orthonormal_rows(n, m, 7)
    Bod = Bo.astype(np.float64)
    rng = np.random.default_rng(11)
    v = rng.standard_normal(n)
    v -= Bod @ (Bod.T @ v)
    v /= np.linalg.norm(v)
    B = np.zeros((n, m), dtype=np.float32, order="F")
    B[:, 0] = (np.sqrt(0.95) * Bod[:, 0] + np.sqrt(0.05) * v).astype(np.float32)
    x = np.ones(n, dtype=np.float32)
    mask = np.zeros(1, dtype=np.int32)
    ctx = mirror_ctx(lib, n, m + 1, 1e-14, B, Bo, x)
    try:
        Bx = np.zeros((1, n), dtype=np.float32)
        fired = C.c_long(0)
        abi.check(lib, lib.mars_hip_orthonormalize(ctx, m, 1, abi.ptr(x), abi.ptr(mask), abi.ptr(Bx), n,
                                                   None, C.byref(fired)))
    finally:
        lib.mars_hip_context_destroy(ctx)
    assert fired.value == 1
    bx = Bx[0].astype(np.float64)
    assert np.abs(Bod.T @ bx).max() < 1e-5
    assert abs(np.linalg.norm(bx) - 1) < 1e-5


def test_batch_matches_single_column(lib, oracle):
    n, m, px = 3000, 6, 5
    say("batch vs single")
    rng = np.random.default_rng(5)
    Bo = orthonormal_rows(n, m, 3)
    B = np.asfortranarray(rng.standard_normal((n, m)).astype(np.float32))
    X = np.asfortranarray(rng.standard_normal((n, px)).astype(np.float32))
    s = (1.0 / np.sqrt((X.astype(np.float64) ** 2).mean(axis=0))).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    tol = n * 0.02 * 2.220446049250313e-16
    ctx = mirror_ctx(lib, n, m + 1, tol, B, Bo, y)
    try:
        cols = np.arange(px, dtype=np.int32)
        abi.check(lib, lib.mars_hip_context_sync_xcols(ctx, px, abi.ptr(X), n, abi.ptr(s), abi.ptr(cols), px))
        sx = np.repeat(np.arange(px), m).astype(np.int32)
        sb = np.tile(np.arange(m), px).astype(np.int32)
        ybx = np.zeros(px * m)
        abi.check(lib, lib.mars_hip_orthonormalize_batch(ctx, m, px * m, abi.ptr(sx), abi.ptr(sb),
                                                         abi.ptr(ybx), None))
        for xc in range(px):
            xs = (X[:, xc] * s[xc]).astype(np.float32)
            one = np.zeros(m)
            abi.check(lib, lib.mars_hip_orthonormalize(ctx, m, m, abi.ptr(xs), abi.ptr(np.arange(m, dtype=np.int32)),
                                                       None, n, abi.ptr(one), None))
            np.testing.assert_allclose(ybx[xc * m:(xc + 1) * m], one, rtol=1e-12, atol=1e-15)
    finally:
        lib.mars_hip_context_destroy(ctx)


def _fit(case, precision="fp64"):
    import mars_amd

    X, y, w = make_inputs(case)
    trace = []
    model = mars_amd.fit(X, y, w, trace=trace, precision=precision, **case["fit"])
    return model, trace


def _compare_to_golden(gold, model, trace, mse_rtol):
    rows = gold["model"]
    got = [(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in model]
    want = [(r["type"], r["basis"], r["input"]) for r in rows]
    assert got == want, "term sequence differs"
    for r, g in zip(rows, model):
        if r["type"] in "+-":
            assert float(g["hinge"]) == r["hinge"], "knot differs"
    got_mse = [v for s in trace for v in s["mse"]]
    want_mse = [a["mse"] for a in gold["appends"]]
    assert len(got_mse) == len(want_mse)
    for a, b in zip(got_mse, want_mse):
        assert abs(a - b) <= mse_rtol * max(abs(b), 1e-300), (a, b)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-5] for p in GOLDEN])
def test_fit_matches_reference(path):
    gold = json.load(open(path))
    name = gold["name"]
    say("fit parity", name)
    model, trace = _fit(CASES[name])
    _compare_to_golden(gold, model, trace, 1e-10)


@pytest.mark.parametrize("sweep", ["sorted", "gather"])
def test_eval_arrays_vs_oracle(oracle, sweep, monkeypatch):
    """Per-candidate linear / hinge / cut arrays of one mid-fit step against the oracle, with the
    hinge sweep streaming the per-column sorted replicas (the default) and gathering rows through
    the sort order (MARS_HIP_SORTED=0, the path for devices the replicas do not fit)."""
    import mars_amd

    say("eval arrays", sweep)
    if sweep == "gather":
        monkeypatch.setenv("MARS_HIP_SORTED", "0")
    else:
        monkeypatch.delenv("MARS_HIP_SORTED", raising=False)
    case = CASES["w2_small"]
    X, y, w = make_inputs(case)
    w = np.ones(len(X), dtype=np.float32) if w is None else w
    dev = mars_amd.marslib.MarsAlgo(X, y, w, 21)
    cpu = oracle.MarsAlgo(X, y, w, 21)
    for t, xc, bc, h in [("l", 3, 0, 0.0), ("+", 7, 0, 0.25), ("-", 7, 0, 0.25), ("l", 0, 2, 0.0)]:
        a = dev.append(t, xc, bc, h)
        b = cpu.append(t, xc, bc, h)
        assert abs(a - b) <= 1e-10 * abs(b)
    m = cpu.nbasis()
    mask = np.ones((X.shape[1], m), dtype=bool)
    mask[3, 1] = False
    d0, l0, h0, c0, k0, cond = cpu.eval_knots(mask, 5, 200, False, 4)
    d1, l1, h1, c1, best = dev.eval_best(mask, 5, 200, False)
    assert abs(d0 - d1) <= 1e-12 * abs(d0)
    np.testing.assert_allclose(l1, l0, rtol=1e-9, atol=1e-15)
    # the hinge score amplifies summation-order noise by cond = |h+|^2 / |h+ residual|^2
    # at the winning knot (the reference has the same sensitivity); bound it accordingly
    tol = 1e-9 + 1e-14 * cond
    assert np.all(np.abs(h1 - h0) <= tol * np.abs(h0) + 1e-15), np.max(np.abs(h1 - h0) / (tol * np.abs(h0) + 1e-15))
    np.testing.assert_array_equal(np.isnan(c1), np.isnan(c0))
    np.testing.assert_array_equal(c1[~np.isnan(c0)], c0[~np.isnan(c0)])
    j1 = np.unravel_index(np.argmax(l0), l0.shape)
    j2 = np.unravel_index(np.argmax(h0), h0.shape)
    assert tuple(best["j1"]) == tuple(int(v) for v in j1)
    assert tuple(best["j2"]) == tuple(int(v) for v in j2)
    assert best["knot2"] == k0[j2]


@pytest.mark.parametrize("sweep", ["sorted", "gather"])
@pytest.mark.parametrize("n,p,mt,widest", [(6000, 6, 61, 32), (8000, 6, 118, 64)])
def test_wide_parent_blocks_vs_oracle(oracle, n, p, mt, widest, sweep, monkeypatch):
    """Hinge fits whose late steps score more than 32 parents per X column (two 32-candidate
    sweep blocks per column; beyond 64, three) against the oracle: same terms and knots, mse
    within 1e-10.  The stopping rules are off so that the fit runs to max_terms.  The second
    fit stops at 118 terms: term 118 would be x2*B80, a column the basis already spans (its
    residual is 1.2e-7 of its norm, the f32 rounding of the product), whose delta-SSE fits
    rounding noise; and the step after ranks x1*B20 against x2*B27, one product (x1 x2 x3 x4 h)
    formed in two orders, so that f32 roundings alone separate the two scores."""
    import mars_amd
    from mars_amd import datagen
    from oracle import fit as F

    say("wide parent blocks n=%d p=%d max_terms=%d %s" % (n, p, mt, sweep))
    if sweep == "gather":
        monkeypatch.setenv("MARS_HIP_SORTED", "0")
    else:
        monkeypatch.delenv("MARS_HIP_SORTED", raising=False)
    X, y = datagen.workload2(n=n, p=p)
    kw = dict(max_terms=mt, linear_only=False, max_degree=5, r2_thresh=0.0, r2_window=1000)
    want, wtrace = F.fit(oracle, X, y, None, **kw)
    trace = []
    got = mars_amd.fit(X, y, trace=trace, **kw)
    variables = [set()]
    for r in got[1:]:
        variables.append(variables[int(r["basis"])] | {int(r["input"])})
    # parents scored at the last step: every term but the last two (one step appends at most two)
    eligible = max(sum(1 for v in variables[:-2] if len(v) < 5 and i not in v) for i in range(p))
    assert eligible > widest, eligible
    seq = lambda mdl: [(r["type"], int(r["basis"]), int(r["input"]),
                        None if np.isnan(r["hinge"]) else float(r["hinge"])) for r in mdl]
    assert seq(got) == seq(want)
    gm = [v for t in trace for v in t["mse"]]
    wm = [v for t in wtrace for v in t["mse"]]
    assert len(gm) == len(wm)
    # mse is (1 - sum of squared projections of y) / n with y scaled to unit sum of squares: its
    # rounding scales with 1/n, and late in the fit mse falls to 6% of that
    for a, b in zip(gm, wm):
        assert abs(a - b) <= 1e-10 * (abs(b) + 1.0 / n), (a, b)


def test_append_basis_vs_oracle(oracle):
    import mars_amd

    say("append basis")
    X, y, _ = make_inputs(CASES["deg3"])
    w = np.ones(len(X), dtype=np.float32)
    dev = mars_amd.marslib.MarsAlgo(X, y, w, 12)
    cpu = oracle.MarsAlgo(X, y, w, 12)
    steps = [("l", 0, 0, 0.0), ("+", 1, 1, 0.1), ("-", 1, 1, 0.1), ("l", 2, 2, 0.0), ("l", 0, 0, 0.0)]
    for t, xc, bc, h in steps:
        a, b = dev.append(t, xc, bc, h), cpu.append(t, xc, bc, h)
        assert (a < 0) == (b < 0)
        if b >= 0:
            assert abs(a - b) <= 1e-10 * abs(b)
    B1, Bo1 = dev.basis()
    B0, Bo0 = cpu.basis()
    assert np.mean(B1 == B0) > 0.999 and np.abs(B1 - B0).max() < 1e-6
    assert np.abs(Bo1 - Bo0).max() < 1e-6
    assert abs(dev.yvar() - cpu.yvar()) <= 1e-13 * cpu.yvar()


def test_batch_fits_match_reference():
    """Batching form: 8 independent workload-4 fits in one batch reproduce each reference fit."""
    import mars_amd

    say("batch parity")
    golds = [json.load(open(os.path.join(HERE, "golden", "w4_fit%d.json" % i))) for i in range(8)]
    inputs = [make_inputs(CASES[g["name"]]) for g in golds]
    stats = {}
    models = mars_amd.fit_batch([x for x, _, _ in inputs], [y for _, y, _ in inputs], max_terms=33,
                                stats=stats)
    for gold, model, mse in zip(golds, models, stats["mse"]):
        got = [(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in model]
        want = [(r["type"], r["basis"], r["input"]) for r in gold["model"]]
        assert got == want, gold["name"]
        want_mse = [a["mse"] for a in gold["appends"]]
        assert len(mse) == len(want_mse)
        for a, b in zip(mse, want_mse):
            assert abs(a - b) <= 1e-10 * abs(b), (gold["name"], a, b)


def test_batch_of_one_equals_single_fit():
    import mars_amd

    say("batch of one")
    X, y, _ = make_inputs(CASES["w5_small"])
    single = mars_amd.fit(X, y, max_terms=41, linear_only=True)
    (batch,) = mars_amd.fit_batch([X], [y], max_terms=41)
    assert [tuple(r)[:3] for r in single] == [tuple(r)[:3] for r in batch]
    np.testing.assert_array_equal(single["r2"], batch["r2"])


# residual storage policies of the exact rescoring: "none" sends every rescored candidate with
# s > tol through the recomputation paths that the default policy takes only for a few
STORE_POLICIES = ["ungated", "none"]


HINGE_FORMS = ["mixed", "mixed-bf16"]


@pytest.mark.parametrize("precision", HINGE_FORMS)
@pytest.mark.parametrize("case", ["w2_small", "w3_small", "hinge_edges", "deg3", "weighted", "w2"])
def test_mixed_hinge_form_matches_reference(case, precision):
    """Precision form with hinge sweeps (FP32 or BF16-split first projection, FP64 retry and
    sweep, exact rescoring of the near winners): the reference's terms, knots and per-step mse."""
    import mars_amd
    from mars_amd import marslib

    gold = json.load(open(os.path.join(HERE, "golden", case + ".json")))
    say("mixed hinge fit", case, precision)
    X, y, w = make_inputs(CASES[case])
    w = np.ones(len(X), dtype=np.float32) if w is None else w
    algo = marslib.MarsAlgo(X, y, w, CASES[case]["fit"]["max_terms"], precision=precision)
    trace = []
    model = mars_amd.fit(X, y, w, algo=algo, trace=trace, precision=precision, **CASES[case]["fit"])
    _compare_to_golden(gold, model, trace, 1e-10)
    st = algo.stats()
    assert st["window_violations"] == 0 and st["rescored"] > 0


@pytest.mark.parametrize("policy", STORE_POLICIES)
@pytest.mark.parametrize("case", ["w1", "w5_small", "w4_fit2"])
def test_mixed_form_matches_reference(case, policy, monkeypatch):
    """Mixed-precision form: same term sequence and per-step mse as the reference."""
    monkeypatch.setenv("MARS_HIP_RESCORE_STORE", policy)
    path = os.path.join(HERE, "golden", case + ".json")
    gold = json.load(open(path))
    say("mixed fit", case, policy)
    model, trace = _fit(CASES[case], precision="mixed")
    _compare_to_golden(gold, model, trace, 1e-10)


@pytest.mark.parametrize("policy", STORE_POLICIES + ["all"])
def test_mixed_eval_at_dgks_gate_boundary(policy, monkeypatch):
    """Columns on the DGKS gate boundary: x_j = 3 + c_j z with z centred and of unit variance, so
    at the intercept-only step s = n c_j^2 and |T|^2 = 9 n, and the gate 9 s < |T|^2
    (kernels.cc:387-388) holds exactly when c_j < 1.  With c_j within 6e-7 of 1 the cache
    approximation and the exact pass disagree on some columns' gate.  Every column has the same
    score (all residuals are multiples of z), so all are rescored, and the mixed form's linear
    scores equal the plain form's up to FP64 summation order, whichever residuals the
    rescoring stores."""
    from mars_amd import marslib

    monkeypatch.setenv("MARS_HIP_RESCORE_STORE", policy)
    say("gate boundary eval", policy)
    n, p = 8192, 384
    rng = np.random.default_rng(11)
    z = rng.standard_normal(n)
    z -= z.mean()
    z /= np.sqrt(np.mean(z * z))
    c = 1.0 + np.linspace(-6e-7, 6e-7, p)
    X = np.asfortranarray((3.0 + np.outer(z, c)).astype(np.float32))
    y = (z + 0.5 * rng.standard_normal(n)).astype(np.float32)
    w = np.ones(n, dtype=np.float32)
    mask = np.ones((p, 1), dtype=bool)
    out = {}
    for form in ("fp64", "mixed"):
        algo = marslib.MarsAlgo(X, y, w, 4, precision=form)
        _, lin, _, _ = algo.eval(mask, 1, 0, True, 1, cuda=True)
        out[form] = (np.asarray(lin)[:, 0].copy(), algo.stats())
        del algo
    (lp, sp), (lm, sm) = out["fp64"], out["mixed"]
    say("gated", sm["gated"], "fired", sm["dgks_fired"], "plain fired", sp["dgks_fired"])
    assert sm["rescored"] == p
    assert sm["gated"] > 0 and sm["dgks_fired"] == sp["dgks_fired"] > 0
    assert np.all(lp > 0)
    np.testing.assert_allclose(lm, lp, rtol=1e-12, atol=0)


def test_rescore_store_policy_rejects_unknown(monkeypatch):
    from mars_amd import marslib

    X, y, _ = make_inputs(CASES["w1"])
    algo = marslib.MarsAlgo(X, y, np.ones(len(X), dtype=np.float32), 8, precision="mixed")
    monkeypatch.setenv("MARS_HIP_RESCORE_STORE", "some")
    with pytest.raises(RuntimeError, match="MARS_HIP_RESCORE_STORE"):
        algo.eval(np.ones((X.shape[1], 1), dtype=bool), 1, 0, True, 1, cuda=True)
    monkeypatch.setenv("MARS_HIP_RESCORE_STORE", "all")
    d0 = algo.eval(np.ones((X.shape[1], 1), dtype=bool), 1, 0, True, 1, cuda=True)[0]
    assert d0 > 0


@pytest.mark.parametrize("policy", STORE_POLICIES)
def test_batch_mixed_matches_reference(policy, monkeypatch):
    import mars_amd

    monkeypatch.setenv("MARS_HIP_RESCORE_STORE", policy)
    say("batch mixed", policy)
    golds = [json.load(open(os.path.join(HERE, "golden", "w4_fit%d.json" % i))) for i in range(8)]
    inputs = [make_inputs(CASES[g["name"]]) for g in golds]
    stats = {}
    models = mars_amd.fit_batch([x for x, _, _ in inputs], [y for _, y, _ in inputs], max_terms=33,
                                precision="mixed", stats=stats)
    for gold, model, mse in zip(golds, models, stats["mse"]):
        got = [(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in model]
        want = [(r["type"], r["basis"], r["input"]) for r in gold["model"]]
        assert got == want, gold["name"]
        for a, b in zip(mse, [a["mse"] for a in gold["appends"]]):
            assert abs(a - b) <= 1e-10 * abs(b)
    assert 0 < stats["rescored"] < stats["scored"]
    assert stats["window_violations"] == 0 and stats["host_redo_steps"] == 0


def test_batch_mixed_fp64_cache_matches_reference(monkeypatch):
    """The cache update on FP64 MFMA (MARS_HIP_CACHE=f64, the bench's --cache-f64 variant) chooses
    the reference's terms with the same mse: the rescoring keeps every winner exact either way."""
    import mars_amd

    monkeypatch.setenv("MARS_HIP_CACHE", "f64")
    say("batch mixed, FP64 cache")
    golds = [json.load(open(os.path.join(HERE, "golden", "w4_fit%d.json" % i))) for i in range(8)]
    inputs = [make_inputs(CASES[g["name"]]) for g in golds]
    stats = {}
    models = mars_amd.fit_batch([x for x, _, _ in inputs], [y for _, y, _ in inputs], max_terms=33,
                                precision="mixed", stats=stats)
    for gold, model, mse in zip(golds, models, stats["mse"]):
        got = [(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in model]
        want = [(r["type"], r["basis"], r["input"]) for r in gold["model"]]
        assert got == want, gold["name"]
        for a, b in zip(mse, [a["mse"] for a in gold["appends"]]):
            assert abs(a - b) <= 1e-10 * abs(b)
    assert stats["window_violations"] == 0 and stats["host_redo_steps"] == 0


def _rival_inputs(seed, n, p, eps):
    """X columns 0 and 1 nearly equal (column 1 = column 0 times 1 + eps, plus eps-sized noise):
    candidates on column 0 and column 1 score within a relative ~eps of each other, below the
    f32 error of the mixed form's cached projections, while neither is in the basis's span."""
    rng = np.random.RandomState(seed)
    X = np.asfortranarray(rng.randn(n, p).astype(np.float32))
    X[:, 1] = (X[:, 0].astype(np.float64) * (1 + eps) + eps * rng.randn(n)).astype(np.float32)
    y = (1.5 * X[:, 0] + 0.8 * X[:, 2] - 0.5 * X[:, 3] * X[:, 4] + rng.randn(n)).astype(np.float32)
    return X, y


@pytest.mark.parametrize("eps", [1e-6, 3e-7])
def test_mixed_resolves_rivals_below_fp32_error(eps):
    """Near-tied rivals away from the DGKS gate: the mixed form's score intervals send both to
    the exact rescoring, so it picks the plain form's term every step."""
    import mars_amd

    say("rivals eps", eps)
    data = [_rival_inputs(s, 20000 + 3000 * s, 8, eps) for s in range(6)]
    Xs, ys = [x for x, _ in data], [y for _, y in data]
    sp, sm = {}, {}
    plain = mars_amd.fit_batch(Xs, ys, max_terms=12, stats=sp)
    mixed = mars_amd.fit_batch(Xs, ys, max_terms=12, precision="mixed", stats=sm)
    for i, (a, b) in enumerate(zip(plain, mixed)):
        sa = [(r["type"], int(r["basis"]), int(r["input"])) for r in a]
        sb = [(r["type"], int(r["basis"]), int(r["input"])) for r in b]
        assert sa == sb, (i, sa, sb)
        assert {0, 1} & {int(r["input"]) for r in a[1:2]}  # the first pick is one of the rivals
    assert sm["window_violations"] == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_graph_replays_match_direct_launches(precision, monkeypatch):
    """The batch step's scoring runs as a captured graph replayed every step (eager work in
    between); the same fits with direct launches give bitwise the same terms and mse."""
    import mars_amd

    say("graph vs direct", precision)
    golds = [json.load(open(os.path.join(HERE, "golden", "w4_fit%d.json" % i))) for i in range(8)]
    inputs = [make_inputs(CASES[g["name"]]) for g in golds]
    Xs, ys = [x for x, _, _ in inputs], [y for _, y, _ in inputs]
    runs = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("MARS_HIP_GRAPH", mode)
        st = {}
        models = mars_amd.fit_batch(Xs, ys, max_terms=33, precision=precision, stats=st)
        runs[mode] = ([[(r["type"], int(r["basis"]), int(r["input"])) for r in m] for m in models], st)
    (sa, a), (sb, b) = runs["1"], runs["0"]
    assert sa == sb
    assert a["mse"] == b["mse"]  # bitwise
    assert a["graph_captures"] >= 1 and a["graph_launches"] == a["steps"]
    assert a["graph_captures"] <= 3  # one per basis-size class (16 columns each)
    assert b["graph_launches"] == 0



@pytest.mark.parametrize("precision,rtol", [("mixed", 1e-4), ("mixed-bf16", 2e-3)])
def test_mixed_hinge_eval_close_to_plain(precision, rtol):
    """The approximate scores themselves (most candidates are not rescored): every linear and
    hinge score of a mid-fit eval within rtol of the plain form's (checks the FP32 and the
    BF16-split MFMA operand layouts and their accuracy, not only the rescored winners), and the
    same winners.  The mask leaves out self-interactions, as fit() does by default: a hinge on a
    variable its parent already holds can cancel the sweep's denominator, and its approximate
    score is then far from the exact one (the precision form rescores such candidates, whose
    upper bound is wide; its scores are then exact, which this comparison does not need)."""
    import mars_amd

    say("mixed hinge eval", precision)
    case = CASES["w2_small"]
    X, y, w = make_inputs(case)
    w = np.ones(len(X), dtype=np.float32) if w is None else w
    out, best = {}, {}
    # (no hinge pair on one knot: with its linear term it spans x exactly, and a candidate in the
    # span scores rounding noise, which no two summation orders reproduce)
    terms = [("l", 3, 0, 0.0), ("+", 7, 0, 0.25), ("l", 0, 2, 0.0), ("-", 12, 1, -0.3), ("+", 5, 0, 0.1)]
    holds = [set()]
    for t, xc, bc, h in terms:
        holds.append(holds[bc] | {xc})
    for prec in ("fp64", precision):
        a = mars_amd.marslib.MarsAlgo(X, y, w, 21, precision=prec)
        for t, xc, bc, h in terms:
            a.append(t, xc, bc, h)
        mask = np.array([[xc not in holds[b] for b in range(a.nbasis())] for xc in range(X.shape[1])])
        _, lin, hin, _, bst = a.eval_best(mask, 5, 200, False)
        out[prec] = (lin, hin)
        best[prec] = (int(bst["j1"][0]), int(bst["j1"][1]), int(bst["j2"][0]), int(bst["j2"][1]),
                      float(bst["sse1"]), float(bst["sse2"]))
        B, Bo = a.basis()
        a.close()
    assert best[precision] == best["fp64"]
    # candidates in the basis's span (a duplicate of a basis function) score rounding noise in
    # either form: they are left out, by the residual of each candidate column against Bo
    Q = np.asarray(Bo, dtype=np.float64)
    live = np.zeros(out["fp64"][0].shape, dtype=bool)
    for xc in range(X.shape[1]):
        for b in range(Q.shape[1]):
            v = np.asarray(B[:, b], dtype=np.float64) * X[:, xc]
            r = v - Q @ (Q.T @ v)
            live[xc, b] = r @ r > 1e-8 * (v @ v) and xc not in holds[b]
    for exact, approx in zip(out["fp64"], out[precision]):
        scale = float(exact.max())
        np.testing.assert_allclose(approx[live], exact[live], rtol=rtol, atol=1e-7 * scale)
