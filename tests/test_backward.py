"""
Backward-pass helpers on the CPU: compact (reference tests/test_compact.py,
restated), pprint, and the backward elimination of prune against the oracle's
brute-force path (every candidate subset refit from scratch).  No device work.
"""

import numpy as np
import pytest

from mars_amd import marslib
from mars_amd import mars as M
from mars_amd.mars import MODEL_DTYPE
from tests import abi


def _model(rows, beta):
    m = np.zeros(len(rows), dtype=MODEL_DTYPE)
    for i, (t, b, x, h) in enumerate(rows):
        m[i]["type"], m[i]["basis"], m[i]["input"], m[i]["hinge"] = t, b, x, h
    m["beta"] = beta
    return m


# ---- compact: the reference's cases (tests/test_compact.py:29-118)

def test_compact_nothing_pruned():
    m = _model([("i", 0, 0, np.nan), ("l", 0, 0, np.nan), ("+", 1, 1, 3.0)], [1.0, 2.0, 3.0])
    out = M.compact(m)
    assert len(out) == 3
    np.testing.assert_array_equal(out["beta"], [1.0, 2.0, 3.0])


def test_compact_leaf_removed():
    m = _model([("i", 0, 0, np.nan), ("l", 0, 0, np.nan), ("+", 0, 1, 2.0)], [1.0, 2.0, 0.0])
    out = M.compact(m)
    assert len(out) == 2 and out[1]["basis"] == 0
    np.testing.assert_array_equal(out["beta"], [1.0, 2.0])


def test_compact_parent_kept_for_child():
    m = _model([("i", 0, 0, np.nan), ("l", 0, 0, np.nan), ("+", 1, 1, 5.0)], [1.0, 0.0, 3.0])
    out = M.compact(m)
    assert len(out) == 3 and out[1]["beta"] == 0.0 and out[2]["basis"] == 1


def test_compact_indices_remapped():
    m = _model([("i", 0, 0, np.nan), ("l", 0, 0, np.nan), ("l", 0, 1, np.nan), ("+", 1, 2, 1.0)],
               [1.0, 2.0, 0.0, 4.0])
    out = M.compact(m)
    assert len(out) == 3 and out[2]["basis"] == 1
    np.testing.assert_array_equal(out["beta"], [1.0, 2.0, 4.0])


def test_compact_deep_chain():
    m = _model([("i", 0, 0, np.nan), ("l", 0, 0, np.nan), ("+", 1, 1, 2.0), ("-", 2, 2, 4.0)],
               [0.0, 0.0, 0.0, 5.0])
    out = M.compact(m)
    assert len(out) == 4
    np.testing.assert_array_equal(out["beta"], [0.0, 0.0, 0.0, 5.0])


def test_compact_all_pruned():
    assert len(M.compact(_model([("i", 0, 0, np.nan), ("l", 0, 0, np.nan)], [0.0, 0.0]))) == 0


def test_pprint(capsys):
    m = _model([("i", 0, 0, np.nan), ("l", 0, 2, np.nan), ("+", 1, 0, 0.5), ("-", 0, 1, 1.25)],
               [1.5, -2.0, 3.0, 0.25])
    M.pprint(m, labels=["a", "b", "c"])
    lines = capsys.readouterr().out.splitlines()
    coef = ["  %+9.4g" % v for v in (1.5, -2.0, 3.0, 0.25)]
    assert lines == [coef[0], coef[1] + " * c", coef[2] + " * MAX(a-0.5,0) * c", coef[3] + " * MAX(1.25-b,0)"]


# ---- the C ABI rejects bad arguments before any device work

def test_expand_gram_argument_errors():
    lib = abi.load()
    out = np.zeros(4, dtype=np.float32)
    basis = np.zeros(1, dtype=np.int32)
    inp = np.zeros(1, dtype=np.int32)
    knot = np.zeros(1)
    rc = lib.mars_hip_expand(None, 4, 4, 1, 1, b"x", abi.ptr(basis), abi.ptr(inp), abi.ptr(knot), abi.ptr(out), 4)
    assert rc == -1 and b"unknown basis type" in lib.mars_hip_last_error()
    rc = lib.mars_hip_expand(None, 4, 4, 1, 1, b"l", abi.ptr(basis), abi.ptr(inp), abi.ptr(knot), abi.ptr(out), 4)
    assert rc == -1 and b"parent must be an earlier term" in lib.mars_hip_last_error()
    res = np.zeros(1)
    rc = lib.mars_hip_gram(None, 4, 4, 1, None, 4, 1, abi.ptr(res), 1)
    assert rc == -1 and b"null pointer" in lib.mars_hip_last_error()


# ---- backward elimination (mars.py:537-604) against the brute-force path

def _standardized(B, y):
    from oracle_backward import gram_ref

    XX, XY, YY = gram_ref(B), gram_ref(B, y), float(gram_ref(y, y)[0, 0])
    d = np.sqrt(np.diag(XX))
    s = np.where(d > 0, d, 1.0)
    return XX / s / s[:, None], XY[:, 0] / s, YY, d > 0


@pytest.mark.parametrize("seed,n,m,ridge", [(0, 80, 5, 0.0), (1, 60, 7, 0.0), (2, 120, 8, 0.1),
                                            (3, 400, 70, 0.0)])
def test_eliminate_matches_brute_force(seed, n, m, ridge):
    from oracle_backward import eliminate_brute

    rng = np.random.RandomState(seed)
    B = np.column_stack([np.ones(n), rng.randn(n, m - 1)])
    B[:, 2] += 0.3 * B[:, 1]                        # correlated columns
    coef = np.zeros(m)
    coef[: max(3, m // 3)] = rng.uniform(0.5, 2.0, max(3, m // 3)) * rng.choice([-1, 1], max(3, m // 3))
    y = B @ coef + 0.5 * rng.randn(n)
    XX, XY, YY, active = _standardized(B, y)
    _, got = marslib.prune_gram(XX, XY, YY, active.copy(), 3.0, n, ridge)
    want = eliminate_brute(XX, XY, YY, active.copy(), 3.0, n, ridge)
    np.testing.assert_array_equal(got, want)
    assert got[0]


def test_eliminate_keeps_intercept_and_mask():
    from oracle_backward import eliminate_brute

    rng = np.random.RandomState(9)
    n = 100
    B = np.column_stack([np.ones(n), rng.randn(n, 4)])
    y = 1 + B[:, 1] - B[:, 3] + 0.1 * rng.randn(n)
    XX, XY, YY, active = _standardized(B, y)
    mask = active.copy()
    mask[3] = False
    _, got = marslib.prune_gram(XX, XY, YY, mask.copy(), 3.0, n)
    assert got[0] and not got[3]
    np.testing.assert_array_equal(got, eliminate_brute(XX, XY, YY, mask.copy(), 3.0, n))


def test_oracle_finds_true_subset():
    """One informative column plus noise: the brute-force GCV choice is {intercept, that column}."""
    from oracle_backward import eliminate_brute

    rng = np.random.RandomState(4)
    n = 200
    B = np.column_stack([np.ones(n), rng.randn(n, 5)])
    y = 2 + 3 * B[:, 2] + 0.2 * rng.randn(n)
    XX, XY, YY, active = _standardized(B, y)
    got = eliminate_brute(XX, XY, YY, active.copy(), 3.0, n)
    np.testing.assert_array_equal(got, [True, False, True, False, False, False])


# ---- the library's prune (mars_hip_prune_gram, host C++) against the oracle's prune

@pytest.mark.parametrize("seed,n,m,ridge,weights", [(11, 150, 6, 0.0, False), (12, 300, 12, 0.05, True),
                                                    (13, 90, 9, 0.0, True)])
def test_prune_gram_matches_oracle(seed, n, m, ridge, weights):
    from oracle_backward import gram_ref, prune_ref

    rng = np.random.RandomState(seed)
    B = np.column_stack([np.ones(n), rng.randn(n, m - 1)])
    B[:, 3] = 0.0                                   # zero-norm column: never active
    y = 1.5 * B[:, 1] - 0.7 * B[:, 2] + 0.3 * rng.randn(n)
    if weights:
        sw = np.sqrt(rng.uniform(0.5, 2.0, n))
        B, y = B * sw[:, None], y * sw
    XX, XY, YY = gram_ref(B), gram_ref(B, y)[:, 0], float(gram_ref(y, y)[0, 0])
    beta, keep = marslib.prune_gram(XX, XY, YY, None, 3.0, n, ridge)
    want, wkeep = prune_ref(XX, XY, YY, None, 3.0, n, ridge)
    np.testing.assert_array_equal(keep, wkeep)
    np.testing.assert_allclose(beta, want, rtol=1e-10, atol=1e-12)
    assert beta[3] == 0.0 and keep[0]


def test_prune_gram_min_norm_on_duplicate_columns():
    """Two identical columns cannot both be resolved: the final solve returns numpy lstsq's
    minimum-norm split (equal halves), as an SVD-based least squares does."""
    rng = np.random.RandomState(5)
    n = 200
    x = rng.randn(n)
    B = np.column_stack([np.ones(n), x, x])
    y = 2.0 * x + 0.01 * rng.randn(n)
    XX, XY, YY = B.T @ B, B.T @ y, float(y @ y)
    keep = np.ones(3, dtype=bool)
    d = np.sqrt(np.diag(XX))
    S, b = XX / d / d[:, None], XY / d
    want = np.linalg.lstsq(S, b, rcond=None)[0] / d
    # force both duplicates to stay: GCV with a zero penalty and n_true huge keeps everything
    beta, chosen = marslib.prune_gram(XX, XY, YY, keep, 0.0, 1e12, 0.0)
    if chosen.all():
        np.testing.assert_allclose(beta, want, rtol=1e-9, atol=1e-12)
        assert abs(beta[1] - beta[2]) <= 1e-9 * abs(beta[1])


def test_prune_gram_errors():
    with pytest.raises(RuntimeError, match="n_true"):
        marslib.prune_gram(np.eye(2), np.ones(2), 1.0, None, 3.0, 0.0, 0.0)
