"""
CPU tests of the oracle (the C++ restatement of the reference forward pass).

Pins, for each oracle function, something it must satisfy that a plausible
mistake would break:
  * full fits against the golden fixtures written by the reference itself
    (term sequence exact, per-step mse within 1e-12 relative);
  * orthonormalize: orthogonality / unit norm / agreement with an FP64
    classical Gram-Schmidt, exact zero for a degenerate column, DGKS firing
    on the 95/5 construction and not on well-conditioned input
    (reference tests/test_kernels.cc:120-353);
  * orthonormalize_col against an explicit modified Gram-Schmidt
    (test_kernels.cc:576-662);
  * dot_widen against the exact sum (f32 products are exact in f64);
  * argsort: stable descending, -0.0 == +0.0 (unittest.cc:101-116);
  * covariates against the plain recurrence (unittest.cc:169-209);
  * linear and hinge delta-SSE against least squares, append mse against OLS
    / WLS (unittest.cc:213-448), the hinge brute force over every grid cut;
  * thread-count invariance (unittest.cc:515-577);
  * agreement with the reference module compiled here (oracle/_ref), if built.
"""

import glob
import json
import math
import os

import numpy as np
import pytest

from tests.golden.cases import CASES, make_inputs

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.json")))
EPS = 2.220446049250313e-16


def orth_basis(n, m, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, m)))
    return np.ascontiguousarray(q.astype(np.float32))


# --------------------------------------------------------------------------
# full fits against the reference's own outputs
# --------------------------------------------------------------------------
@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-5] for p in GOLDEN])
def test_fit_matches_golden(oracle, path):
    from oracle import fit as F
    from tests.golden.make_golden import digest

    gold = json.load(open(path))
    case = CASES[gold["name"]]
    X, y, w = make_inputs(case)
    assert digest(X) == gold["x_sha256"] and digest(y) == gold["y_sha256"], "generator drift"
    model, trace = F.fit(oracle, X, y, w, **case["fit"])
    got = [(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in model]
    want = [(r["type"], r["basis"], r["input"]) for r in gold["model"]]
    assert got == want
    for r, g in zip(gold["model"], model):
        if r["type"] in "+-":
            assert float(g["hinge"]) == r["hinge"]
    got_mse = [v for s in trace for v in s["mse"]]
    want_mse = [a["mse"] for a in gold["appends"]]
    np.testing.assert_allclose(got_mse, want_mse, rtol=1e-12, atol=0)
    for s, e in zip(trace, gold["evals"]):
        assert abs(s["sse0"] - e["sse0"]) <= 1e-12 * abs(e["sse0"]) + 1e-300
        assert abs(s["sse1"] - e["sse1"]) <= 1e-11 * abs(e["sse1"]) + 1e-300
        assert abs(s["sse2"] - e["sse2"]) <= 1e-11 * abs(e["sse2"]) + 1e-300


# --------------------------------------------------------------------------
# kernels
# --------------------------------------------------------------------------
def test_orthonormalize_geometry(oracle):
    n, m, p = 1500, 12, 6
    rng = np.random.default_rng(0xC0FFEE)
    Bo = orth_basis(n, m, 1)
    B = np.asfortranarray(rng.standard_normal((n, m)).astype(np.float32))
    B[:, 5] = B[:, 4]
    x = rng.standard_normal(n).astype(np.float32)
    mask = np.array([0, 2, 4, 5, 7, 11], dtype=np.int32)
    tol = 1e-14
    Bx, fired = oracle.orthonormalize(B, x, mask, Bo, tol)
    Bxd = Bx.astype(np.float64)
    Bod = Bo.astype(np.float64)
    assert np.abs(Bod.T @ Bxd).max() < 1e-5
    np.testing.assert_allclose(np.linalg.norm(Bxd, axis=0), 1.0, atol=1e-5)
    # FP64 one-pass CGS reference (test_kernels.cc:63-79)
    ref = (B[:, mask] * x[:, None]).astype(np.float64)
    ref -= Bod @ (Bod.T @ ref)
    ref /= np.sqrt((ref**2).sum(axis=0) + tol)
    np.testing.assert_allclose(Bxd, ref, atol=1e-5)
    assert fired == 0


def test_orthonormalize_degenerate_column_is_zero(oracle):
    n, m = 400, 3
    Bo = orth_basis(n, m, 2)
    B = np.asfortranarray(Bo.copy())  # B column 1 lies in span(Bo)
    x = np.ones(n, dtype=np.float32)
    Bx, _ = oracle.orthonormalize(B, x, np.array([1], dtype=np.int32), Bo, n * 0.02 * EPS)
    assert np.all(Bx == 0.0)


def _dgks_case(n, m, frac, seed):
    Bo = orth_basis(n, m, seed)
    Bod = Bo.astype(np.float64)
    v = np.random.default_rng(seed + 1).standard_normal(n)
    v -= Bod @ (Bod.T @ v)
    v /= np.linalg.norm(v)
    B = np.zeros((n, m), dtype=np.float32, order="F")
    B[:, 0] = (np.sqrt(1 - frac) * Bod[:, 0] + np.sqrt(frac) * v).astype(np.float32)
    return B, Bo


def test_dgks_fires_on_95_5_split(oracle):
    B, Bo = _dgks_case(200, 4, 0.05, 3)
    Bx, fired = oracle.orthonormalize(B, np.ones(200, dtype=np.float32), np.zeros(1, np.int32), Bo, 1e-14)
    assert fired == 1
    assert np.abs(Bo.astype(np.float64).T @ Bx.astype(np.float64)).max() < 1e-5


def test_dgks_quiet_when_well_conditioned(oracle):
    B, Bo = _dgks_case(200, 4, 0.5, 4)
    _, fired = oracle.orthonormalize(B, np.ones(200, dtype=np.float32), np.zeros(1, np.int32), Bo, 1e-14)
    assert fired == 0


def test_orthonormalize_col_matches_mgs(oracle):
    n, m = 3000, 9
    Bo = orth_basis(n, m, 5)
    v = np.random.default_rng(6).standard_normal(n)
    out, w, fired = oracle.orthonormalize_col(v, Bo, 1e-14)
    ref = v.copy()
    Bod = Bo.astype(np.float64)
    for j in range(m):
        ref -= (Bod[:, j] @ ref) * Bod[:, j]
    np.testing.assert_allclose(w, np.linalg.norm(ref), rtol=1e-12)
    np.testing.assert_allclose(out, ref / np.linalg.norm(ref), atol=1e-9)
    assert fired == 0


def test_orthonormalize_col_retry_and_degenerate(oracle):
    n, m = 500, 4
    Bo = orth_basis(n, m, 8)
    Bod = Bo.astype(np.float64)
    v = 0.985 * Bod[:, 0] + 0.01 * np.random.default_rng(9).standard_normal(n)
    _, _, fired = oracle.orthonormalize_col(v, Bo, 1e-14)
    assert fired == 1
    deg = Bod[:, 1].copy()
    out, w, _ = oracle.orthonormalize_col(deg, Bo, 1e-6)
    assert w * w <= 1e-6
    # left unnormalized: equals the raw residual
    r = deg.copy()
    for j in range(m):
        r -= (Bod[:, j] @ r) * Bod[:, j]
    assert np.abs(out - r).max() < 1e-12


@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 31, 1024, 50000])
def test_dot_widen_exact(oracle, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    exact = math.fsum((a.astype(np.float64) * b.astype(np.float64)).tolist())
    got = oracle.dot_widen(a, b)
    assert abs(got - exact) <= 4 * EPS * max(1.0, np.abs(a.astype(np.float64) * b).sum())


def test_argsort_stable_descending(oracle):
    v = np.array([0.5, -0.0, 2.0, 0.0, 0.5, -1.0, 2.0, 0.0], dtype=np.float32)
    k = oracle.argsort(v)
    # ties keep row order; -0.0 and +0.0 compare equal
    assert list(k) == [2, 6, 0, 4, 1, 3, 7, 5]
    rng = np.random.default_rng(1)
    w = rng.integers(-5, 5, size=997).astype(np.float32)
    k = oracle.argsort(w)
    ref = sorted(range(len(w)), key=lambda i: (-w[i], i))
    assert list(k) == ref


def test_covariates_recurrence(oracle):
    rng = np.random.default_rng(3)
    m = 13
    f = np.zeros(m + 1)
    g = np.zeros(m + 1)
    fr = np.zeros(m + 1)
    gr = np.zeros(m + 1)
    y = rng.standard_normal(m)
    for _ in range(200):
        q = rng.standard_normal(m).astype(np.float32)
        xm, ym, k0 = rng.standard_normal(3)
        k1 = np.float32(rng.standard_normal())
        ff, fy = oracle.covariates(f, g, q, y, xm, ym, k0, k1)
        qq = np.append(q.astype(np.float64), xm)
        fr += k0 * gr
        gr += float(k1) * qq
        np.testing.assert_allclose(f, fr, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(g, gr, rtol=1e-9, atol=1e-9)
        ref_ff = fr @ fr
        assert abs((ff / ref_ff if ref_ff > 0 else 1.0) - 1) < 1e-6
        yy = np.append(y, ym)
        assert abs(fy - fr @ yy) <= 1e-6 * (np.abs(fr) @ np.abs(yy))


def _lstsq_dsse(A, y):
    b, *_ = np.linalg.lstsq(A, y, rcond=None)
    return (A @ b) @ y / (y @ y)


def test_delta_sse_against_least_squares(oracle):
    """unittest.cc:213-399 restated with numpy least squares."""
    n = 5891
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, size=(n, 12))
    cut = 0.25
    X[100, 7] = cut
    X[:, 10] = 0.0
    x3, x7, x9 = X[:, 3], X[:, 7], X[:, 9]
    y = 0.3 * x3 - 0.2 * x9 + 0.25 * x3 * x9 - 5 * np.maximum(x7 - cut, 0) + 2 * np.maximum(cut - x7, 0)
    y = y + rng.uniform(-1, 1, size=n)
    X32 = np.asfortranarray(X.astype(np.float32))
    y32 = y.astype(np.float32)
    Xd, yd = X32.astype(np.float64), y32.astype(np.float64)
    algo = oracle.MarsAlgo(X32, y32, np.ones(n, np.float32), 8)
    cols = [np.ones(n)]
    for xc, bc in [(3, 0), (9, 0), (9, 1)]:
        mask = np.zeros((12, algo.nbasis()), dtype=bool)
        mask[xc, :] = True
        d0, lin, _, _ = algo.eval(mask, 1, 0, True, 1)
        cand = cols[bc] * Xd[:, xc]
        assert abs((d0 + lin[xc, bc]) - _lstsq_dsse(np.column_stack(cols + [cand]), yd)) < 1e-7
        if (xc, bc) == (9, 1):
            assert np.argmax(lin[9]) == 1
        mse = algo.append("l", xc, bc, 0.0)
        cols.append(cand)
        A = np.column_stack(cols)
        b, *_ = np.linalg.lstsq(A, yd, rcond=None)
        e = A @ b - yd
        assert abs(mse - (e @ e / n) / (yd @ yd)) < 1e-8
    # hinge at x7: brute force over every on-grid cut (min_span 1, end_span 0)
    mask = np.zeros((12, algo.nbasis()), dtype=bool)
    mask[7, :] = True
    d0, lin, hin, cutv, knots, _ = algo.eval_knots(mask, 1, 0, False, 1)
    assert np.argmax(hin[7]) == 0 and np.all(hin[7] > lin[7])
    for bc in range(algo.nbasis()):
        h = cutv[7, bc]
        A = np.column_stack(cols + [cols[bc] * np.maximum(Xd[:, 7] - h, 0), cols[bc] * np.maximum(h - Xd[:, 7], 0)])
        assert abs((d0 + hin[7, bc]) / n - _lstsq_dsse(A, yd) / n) < 1e-6
    # the chosen cut of the intercept parent is the planted knot's neighbourhood
    assert abs(cutv[7, 0] - cut) < 0.05
    # empty column scores exactly zero
    mask = np.zeros((12, algo.nbasis()), dtype=bool)
    mask[10, :] = True
    _, lin, hin, _ = algo.eval(mask, 1, 0, False, 1)
    assert np.all(lin[10] == 0) and np.all(hin[10] == 0)


def test_hinge_brute_force_small(oracle):
    """Every grid cut of a tiny problem solved by least squares; the oracle's knot must be the argmax."""
    n = 300
    rng = np.random.default_rng(12)
    X = np.asfortranarray(rng.standard_normal((n, 2)).astype(np.float32))
    y = (np.maximum(X[:, 0] - 0.3, 0) * 2 + X[:, 1] + 0.1 * rng.standard_normal(n)).astype(np.float32)
    algo = oracle.MarsAlgo(X, y, np.ones(n, np.float32), 6)
    algo.append("l", 1, 0, 0.0)
    span, end = 3, 10
    mask = np.zeros((2, 2), dtype=bool)
    mask[0, :] = True
    d0, lin, hin, cutv, knots, _ = algo.eval_knots(mask, span, end, False, 1)
    Xd, yd = X.astype(np.float64), y.astype(np.float64)
    order = np.array(oracle.argsort(X[:, 0]))
    cols = [np.ones(n), Xd[:, 1]]
    for bc in range(2):
        best, bi = 0.0, -1
        base = _lstsq_dsse(np.column_stack(cols + [cols[bc] * Xd[:, 0]]), yd)
        for i in range(end + 1, n - end, span):
            h = Xd[order[i], 0]
            A = np.column_stack(cols + [cols[bc] * np.maximum(Xd[:, 0] - h, 0),
                                        cols[bc] * np.maximum(h - Xd[:, 0], 0)])
            gain = _lstsq_dsse(A, yd) - base
            if gain > best + 1e-9:
                best, bi = gain, i
        assert knots[0, bc] == bi
        assert abs((hin[0, bc] - lin[0, bc]) - best) < 1e-6


def test_weighted_least_squares(oracle):
    n = 1000
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, size=(n, 4))
    y = X[:, 1] * 0.5 + X[:, 2] * 0.3 + 0.1 * rng.uniform(-1, 1, n)
    w = rng.uniform(0.1, 2.0, n)
    X32, y32, w32 = np.asfortranarray(X.astype(np.float32)), y.astype(np.float32), w.astype(np.float32)
    algo = oracle.MarsAlgo(X32, y32, w32, 4)
    algo.append("l", 1, 0, 0.0)
    mse = algo.append("l", 2, 0, 0.0)
    Xd, yd, wd = X32.astype(np.float64), y32.astype(np.float64), w32.astype(np.float64)
    A = np.column_stack([np.ones(n), Xd[:, 1], Xd[:, 2]])
    sw = np.sqrt(wd)
    b, *_ = np.linalg.lstsq(A * sw[:, None], yd * sw, rcond=None)
    e = A @ b - yd
    ref = ((wd * e * e).sum() / n) / (wd * yd * yd).sum()
    assert abs(mse - ref) < 1e-7


def test_thread_count_invariance(oracle):
    X, y, w = make_inputs(CASES["w2_small"])
    w = np.ones(len(X), np.float32)
    algo = oracle.MarsAlgo(X, y, w, 6)
    algo.append("l", 3, 0, 0.0)
    mask = np.ones((X.shape[1], 2), dtype=bool)
    a = algo.eval(mask, 5, 200, False, 1)
    b = algo.eval(mask, 5, 200, False, 4)
    for u, v in zip(a[1:], b[1:]):
        assert np.array_equal(u, v, equal_nan=True)


_REF = glob.glob(os.path.join(ROOT, "oracle", "_ref", "marslib*.so"))


@pytest.mark.skipif(not _REF, reason="reference module not built (make -C oracle ref)")
def test_eval_agrees_with_reference_build(oracle):
    import importlib.util

    spec = importlib.util.spec_from_file_location("marslib", _REF[0])
    ref = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ref)
    X, y, w = make_inputs(CASES["w3_small"])
    w = np.ones(len(X), np.float32)
    a, b = oracle.MarsAlgo(X, y, w, 10), ref.MarsAlgo(X, y, w, 10)
    for t, xc, bc, h in [("l", 2, 0, 0.0), ("+", 17, 0, 0.1), ("-", 17, 0, 0.1), ("l", 5, 1, 0.0)]:
        u, v = a.append(t, xc, bc, h), b.append(t, xc, bc, h)
        assert abs(u - v) <= 1e-12 * abs(v)
    mask = np.ones((X.shape[1], a.nbasis()), dtype=bool)
    ra = a.eval(mask, 4, 300, False, 8)
    rb = b.eval(mask, 4, 300, False, 8)
    assert abs(ra[0] - rb[0]) <= 1e-13 * abs(rb[0])
    np.testing.assert_allclose(ra[1], rb[1], rtol=1e-10, atol=1e-16)
    np.testing.assert_allclose(ra[2], rb[2], rtol=1e-10, atol=1e-16)
    np.testing.assert_array_equal(ra[3], rb[3])
