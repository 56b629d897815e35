"""
Backward-pass helpers on the device: expand (mars_hip_expand) against the
oracle's restatement, gram (mars_hip_gram) against f64 products and the
reference's own known-answer tests (tests/test_prune.py:8-58 of the reference,
restated), prune on the reference's prune cases (tests/test_prune.py:61-167)
and against prune with the oracle's Gram products, and the compact / expand
consistency case (tests/test_compact.py:121-141).
"""

import numpy as np
import pytest

import mars_amd
from mars_amd import mars as M
from mars_amd.mars import MODEL_DTYPE
from tests import abi

pytestmark = pytest.mark.gpu


def _model(rows, beta=None):
    m = np.zeros(len(rows), dtype=MODEL_DTYPE)
    for i, (t, b, x, h) in enumerate(rows):
        m[i]["type"], m[i]["basis"], m[i]["input"], m[i]["hinge"] = t, b, x, h
    if beta is not None:
        m["beta"] = beta
    return m


# ---- gram: the reference's known-answer tests

def test_gram_square():
    A = np.random.RandomState(0).randn(500, 7).astype("f")
    np.testing.assert_allclose(mars_amd.gram(A), A.T.astype("d") @ A.astype("d"), atol=1e-5)


def test_gram_cross():
    rng = np.random.RandomState(1)
    A = rng.randn(500, 7).astype("f")
    B = rng.randn(500, 3).astype("f")
    np.testing.assert_allclose(mars_amd.gram(A, B), A.T.astype("d") @ B.astype("d"), atol=1e-5)


def test_gram_1d():
    rng = np.random.RandomState(2)
    A = rng.randn(300, 4).astype("f")
    y = rng.randn(300).astype("f")
    out = mars_amd.gram(A, y)
    assert out.shape == (4,)
    np.testing.assert_allclose(out, A.T.astype("d") @ y.astype("d"), atol=1e-5)


def test_gram_scalar():
    y = np.random.RandomState(3).randn(300).astype("f")
    out = mars_amd.gram(y, y)
    assert isinstance(out, float)
    np.testing.assert_allclose(out, float(y @ y), atol=1e-3)


def test_gram_precision():
    n = 100_000
    A = np.full((n, 1), 0.5, dtype="f")
    assert mars_amd.gram(A)[0, 0] == n * 0.25


# ---- gram: exact f64 products, every shape class

@pytest.mark.parametrize("n,ma,mb,dt", [(1000, 7, 3, "f"), (1237, 33, 140, "f"), (999, 200, None, "f"),
                                        (513, 17, 5, "d"), (64, 1, 1, "f"), (5, 3, None, "d"),
                                        (40000, 65, None, "f")])
def test_gram_vs_f64(n, ma, mb, dt):
    from oracle_backward import gram_ref

    rng = np.random.RandomState(n + ma)
    A = np.asfortranarray(rng.randn(n, ma).astype(dt))
    B = None if mb is None else np.asfortranarray(rng.randn(n, mb).astype(dt))
    got = mars_amd.gram(A, B)
    want = gram_ref(A, B)
    scale = np.sqrt(np.outer(np.sum(A.astype("d") ** 2, 0), np.sum((A if B is None else B).astype("d") ** 2, 0)))
    assert np.max(np.abs(got - want) / scale) <= 5e-13   # two f64 summation orders over n rows
    if B is None:
        np.testing.assert_array_equal(got, got.T)


def test_gram_mixed_dtypes_and_views():
    from oracle_backward import gram_ref

    rng = np.random.RandomState(7)
    A = rng.randn(300, 9).astype("f")
    y = rng.randn(300)
    np.testing.assert_allclose(mars_amd.gram(A, y), gram_ref(A, y)[:, 0], rtol=1e-13)
    C_order = np.ascontiguousarray(A[:, ::2])
    np.testing.assert_allclose(mars_amd.gram(C_order), gram_ref(C_order), rtol=1e-13)
    with pytest.raises(ValueError):
        mars_amd.gram(A, y[:10])


def test_gram_through_c_abi():
    lib = abi.load()
    rng = np.random.RandomState(11)
    n, ma, mb, ld = 1000, 5, 4, 1003
    A = np.zeros((ma, ld), dtype=np.float32)   # column-major (ld rows allocated per column)
    B = np.zeros((mb, ld), dtype=np.float32)
    A[:, :n] = rng.randn(ma, n)
    B[:, :n] = rng.randn(mb, n)
    out = np.zeros((mb, ma + 2))
    abi.check(lib, lib.mars_hip_gram(abi.ptr(A), ld, n, ma, abi.ptr(B), ld, mb, abi.ptr(out), ma + 2))
    want = A[:, :n].astype("d") @ B[:, :n].astype("d").T
    np.testing.assert_allclose(out[:, :ma].T, want, rtol=1e-13, atol=1e-12)


# ---- expand against the oracle

def test_expand_vs_oracle():
    from oracle_backward import expand_ref

    rng = np.random.RandomState(5)
    X = rng.randn(2000, 4).astype("f")
    X[3, 1] = np.nan
    X[7, 2] = 0.25            # exactly at a knot below
    model = _model([("i", 0, 0, np.nan), ("l", 0, 0, np.nan), ("+", 1, 2, 0.25), ("-", 0, 1, -0.5),
                    ("l", 3, 3, np.nan), ("+", 4, 0, 1.0 / 3.0), ("-", 2, 2, 0.1)])
    for dt in ("f", "d"):
        Xd = X.astype(dt)
        got = mars_amd.expand(Xd, model)
        want = expand_ref(Xd, model)
        assert got.dtype == Xd.dtype and got.flags.f_contiguous
        np.testing.assert_array_equal(got, want)


def test_expand_closed_forms():
    X = np.array([[1.5, -2.0], [0.25, 3.0], [-1.0, 0.5]], dtype="f")
    model = _model([("i", 0, 0, np.nan), ("l", 0, 1, np.nan), ("+", 1, 0, 0.5), ("-", 1, 0, 0.5)])
    B = mars_amd.expand(X, model)
    np.testing.assert_array_equal(B[:, 0], [1, 1, 1])
    np.testing.assert_array_equal(B[:, 1], X[:, 1])
    np.testing.assert_array_equal(B[:, 2], [-2.0, 0.0, 0.0])
    np.testing.assert_array_equal(B[:, 3], [0.0, 0.75, 0.75])
    with pytest.raises(AssertionError):
        mars_amd.expand(X, _model([("i", 0, 0, np.nan), ("l", 1, 0, np.nan)]))


def test_expand_compact_consistent():
    X = np.random.RandomState(7).randn(50, 3).astype("f")
    model = _model([("i", 0, 0, np.nan), ("l", 0, 0, np.nan), ("l", 0, 1, np.nan), ("+", 1, 2, 0.5)],
                   [2.0, 3.0, 0.0, -1.0])
    y_full = mars_amd.expand(X, model) @ model["beta"]
    small = mars_amd.compact(model)
    np.testing.assert_allclose(mars_amd.expand(X, small) @ small["beta"], y_full, atol=1e-6)


# ---- prune: the reference's cases

def test_prune_perfect_fit_kept():
    rng = np.random.RandomState(0)
    n = 200
    B = np.column_stack([np.ones(n), rng.randn(n), rng.randn(n)])
    true_beta = np.array([1.0, 3.0, -2.0])
    np.testing.assert_allclose(mars_amd.prune(B, B @ true_beta, n_true=n, penalty=3), true_beta, atol=1e-5)


def test_prune_noise_columns():
    rng = np.random.RandomState(1)
    n = 500
    signal = rng.randn(n)
    B = np.column_stack([np.ones(n), signal, rng.randn(n) * 0.01, rng.randn(n) * 0.01])
    beta = mars_amd.prune(B, 5.0 + 3.0 * signal, n_true=n, penalty=3)
    assert abs(beta[0]) > 0.1 and abs(beta[1]) > 0.1
    np.testing.assert_allclose(beta[2:], 0, atol=1e-10)


def test_prune_zero_variance_column():
    rng = np.random.RandomState(2)
    n = 200
    B = np.column_stack([np.ones(n), rng.randn(n), np.zeros(n)])
    beta = mars_amd.prune(B, 2.0 + 4.0 * B[:, 1], n_true=n, penalty=3)
    assert beta[2] == 0
    np.testing.assert_allclose(beta[:2], [2.0, 4.0], atol=1e-5)


def test_prune_mask():
    rng = np.random.RandomState(3)
    n = 200
    x1, x2 = rng.randn(n), rng.randn(n)
    B = np.column_stack([np.ones(n), x1, x2])
    beta = mars_amd.prune(B, 1.0 + 2.0 * x1 + 3.0 * x2, n_true=n, penalty=3, mask=np.array([True, True, False]))
    assert beta[2] == 0


def test_prune_ridge():
    rng = np.random.RandomState(4)
    n = 200
    B = np.column_stack([np.ones(n), rng.randn(n)])
    y = B @ np.array([5.0, 10.0])
    b0 = mars_amd.prune(B, y, n_true=n, penalty=3, ridge=0)
    b1 = mars_amd.prune(B, y, n_true=n, penalty=3, ridge=0.5)
    assert np.linalg.norm(b1) < np.linalg.norm(b0)


def test_prune_single_term():
    n = 100
    np.testing.assert_allclose(mars_amd.prune(np.ones((n, 1)), np.full(n, 3.0), n_true=n, penalty=3), [3.0],
                               atol=1e-5)


def test_prune_weights():
    rng = np.random.RandomState(5)
    n = 300
    B = np.column_stack([np.ones(n), rng.randn(n), rng.randn(n)])
    y = 1.0 + 2.0 * B[:, 1] - 0.5 * B[:, 2]
    w = rng.uniform(0.1, 2.0, n)
    sw = np.sqrt(w)
    np.testing.assert_allclose(mars_amd.prune(B, y, w=w, n_true=n, penalty=3),
                               mars_amd.prune(B * sw[:, None], y * sw, n_true=n, penalty=3), atol=1e-6)


def test_prune_full_elimination():
    rng = np.random.RandomState(42)
    n = 50
    x1, x2 = rng.randn(n), rng.randn(n)
    shared = rng.randn(n) * 0.15
    B = np.column_stack([np.ones(n), x1, x2, shared + rng.randn(n) * 0.01, shared + rng.randn(n) * 0.01])
    beta = mars_amd.prune(B, 1.0 + 2.0 * x1 - 1.5 * x2, n_true=n, penalty=3)
    np.testing.assert_allclose(beta[3:], 0, atol=1e-10)
    assert min(abs(beta[:3])) > 0.1


# ---- prune with the device Gram products against prune with the oracle's

@pytest.mark.parametrize("seed,dt", [(0, "f"), (1, "d"), (2, "f")])
def test_prune_vs_oracle_gram(seed, dt, monkeypatch):
    from oracle_backward import gram_ref

    rng = np.random.RandomState(seed)
    n, m = 3000, 12
    X = rng.randn(n, 4).astype(dt)
    rows = [("i", 0, 0, np.nan)]
    for i in range(1, m):
        rows.append((("l", "+", "-")[i % 3], int(rng.randint(0, i)), int(rng.randint(0, 4)), float(rng.randn())))
    model = _model(rows)
    B = mars_amd.expand(X, model)
    y = (B[:, : m // 2].astype("d") @ rng.randn(m // 2) + 0.3 * rng.randn(n)).astype(dt)
    got = mars_amd.prune(B, y, penalty=3)

    def ref_gram(A, B2=None, chunk_size=128):
        out = gram_ref(A, B2)
        A1 = np.asarray(A)
        B1 = A1 if B2 is None else np.asarray(B2)
        if A1.ndim == 1 and B1.ndim == 1:
            return float(out[0, 0])
        return out[:, 0] if B1.ndim == 1 else out

    monkeypatch.setattr(M, "gram", ref_gram)
    want = M.prune(B, y, penalty=3)
    np.testing.assert_array_equal(got != 0, want != 0)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)
