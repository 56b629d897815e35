"""
Deterministic synthetic inputs for the five forward-pass workloads.

Every random value comes from a counter-based splitmix64 stream
(`uniform(seed, stream, count)`), so the same bytes are produced on any
machine with this numpy, independent of numpy's own generators.  Gaussian
values use Box-Muller on two uniform streams; Student-t(3) is a normal divided
by sqrt(chi2_3 / 3).  The N(0, v) notation of the workload table is read as
mean 0 and *variance* v.

Workloads (seeds: workload k uses seed k; batch fit i uses seed 1000 + i):

1. N=10,000, p=2: x0 ~ N(0,1), x1 ~ Bernoulli(0.5);
   y = 2 x0 + 3 x1 + x0 x1 + N(0,1).  max_terms=17, tail_span=0, linear_only.
2. N=100,000, p=50 iid N(0,1);
   y = 3 max(0,x0-.5) max(0,.3-x3) + 2 max(0,x7) - x12 + N(0,.25).
   Full hinge sweep, max_degree=2, max_terms=41.
3. N=1,000,000, p=100: 80 columns U(-1,1), 20 columns Student-t(3);
   y = 6 random degree<=2 hinge products with random knots + N(0,.5).  max_terms=65.
4. 256 fits, N_i log-uniform in [2e3, 2e5], p_i uniform in [5, 64], X ~ N(0,1);
   y_i = 2-5 random hinge terms + N(0,1).  max_terms=33, linear_only.
5. N=4,000,000, p=200 Gaussian with AR(1) correlation 0.9 across columns;
   y = 10 random linear terms + 5 pairwise products + N(0,1).  linear_only, max_terms=101.
"""

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _key(seed, stream):
    with np.errstate(over="ignore"):
        s = np.array([seed * 0x100000001B3 + stream * 0x1F123BB5], dtype=np.uint64)
        return _mix(s + _GOLD)[0]


def raw_bits(seed, stream, count):
    """`count` 64-bit words of stream (seed, stream)."""
    with np.errstate(over="ignore"):
        ctr = np.arange(1, count + 1, dtype=np.uint64) * _GOLD + _key(seed, stream)
        return _mix(ctr)


def uniform(seed, stream, count):
    """Uniform doubles in [0, 1) with 53 random bits."""
    return (raw_bits(seed, stream, count) >> np.uint64(11)).astype(np.float64) * (2.0**-53)


def uniform_open(seed, stream, count):
    """Uniform doubles in (0, 1]."""
    return ((raw_bits(seed, stream, count) >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0**-53)


def normal(seed, stream, count):
    """Standard normal doubles (Box-Muller, cosine branch)."""
    u1 = uniform_open(seed, 2 * stream, count)
    u2 = uniform(seed, 2 * stream + 1, count)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


class _Picker:
    """Small sequential draws (term structure, knots) from a dedicated stream."""

    def __init__(self, seed):
        self.u = uniform(seed, 999_999, 4096)
        self.i = 0

    def unif(self, lo=0.0, hi=1.0):
        v = self.u[self.i]
        self.i += 1
        return lo + (hi - lo) * v

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + min(int(self.unif() * (hi - lo + 1)), hi - lo)

    def sign(self):
        return 1.0 if self.unif() < 0.5 else -1.0

    def distinct(self, k, p):
        out = []
        while len(out) < k:
            c = self.integer(0, p - 1)
            if c not in out:
                out.append(c)
        return out


def _hinge(x, knot, direction):
    return np.maximum(x - knot, 0.0) if direction > 0 else np.maximum(knot - x, 0.0)


def _finish(X, y):
    return np.asfortranarray(X, dtype=np.float32), np.ascontiguousarray(y, dtype=np.float32)


def workload1(seed=1, n=10_000):
    X = np.empty((n, 2), dtype=np.float32, order="F")
    X[:, 0] = normal(seed, 0, n)
    X[:, 1] = (uniform(seed, 10, n) < 0.5).astype(np.float64)
    x0, x1 = X[:, 0].astype(np.float64), X[:, 1].astype(np.float64)
    y = 2 * x0 + 3 * x1 + x0 * x1 + normal(seed, 20, n)
    return _finish(X, y)


def workload2(seed=2, n=100_000, p=50):
    X = np.empty((n, p), dtype=np.float32, order="F")
    for j in range(p):
        X[:, j] = normal(seed, j, n)
    x = lambda j: X[:, j % p].astype(np.float64)  # noqa: E731  (reduced p wraps the indices)
    y = (
        3 * np.maximum(0, x(0) - 0.5) * np.maximum(0, 0.3 - x(3))
        + 2 * np.maximum(0, x(7))
        - x(12)
        + 0.5 * normal(seed, 10_000, n)
    )
    return _finish(X, y)


def _student_t3(seed, stream, n):
    z = normal(seed, stream, n)
    c = sum(normal(seed, stream + 100_000 * (r + 1), n) ** 2 for r in range(3))
    return z / np.sqrt(c / 3.0)


def workload3(seed=3, n=1_000_000, p=100, n_uniform=80):
    X = np.empty((n, p), dtype=np.float32, order="F")
    for j in range(p):
        if j < n_uniform:
            X[:, j] = 2.0 * uniform(seed, j, n) - 1.0
        else:
            X[:, j] = _student_t3(seed, j, n)
    pk = _Picker(seed)
    y = np.sqrt(0.5) * normal(seed, 10_000, n)
    for _ in range(6):
        deg = pk.integer(1, 2)
        cols = pk.distinct(deg, p)
        term = np.ones(n)
        for c in cols:
            knot = pk.unif(-0.5, 0.5)
            term *= _hinge(X[:, c].astype(np.float64), knot, pk.sign())
        y += pk.unif(1.0, 3.0) * pk.sign() * term
    return _finish(X, y)


def batch_shape(i, n_lo=2_000, n_hi=200_000, p_lo=5, p_hi=64):
    """(N_i, p_i) of batch fit i: N log-uniform in [n_lo, n_hi], p uniform in [p_lo, p_hi]."""
    pk = _Picker(1000 + i)
    n = int(round(np.exp(pk.unif(np.log(n_lo), np.log(n_hi)))))
    p = pk.integer(p_lo, p_hi)
    return n, p


def workload4_fit(i, n_lo=2_000, n_hi=200_000, p_lo=5, p_hi=64):
    seed = 1000 + i
    pk = _Picker(seed)
    n = int(round(np.exp(pk.unif(np.log(n_lo), np.log(n_hi)))))
    p = pk.integer(p_lo, p_hi)
    X = np.empty((n, p), dtype=np.float32, order="F")
    for j in range(p):
        X[:, j] = normal(seed, j, n)
    y = normal(seed, 10_000, n)
    for _ in range(pk.integer(2, 5)):
        c = pk.integer(0, p - 1)
        y += pk.unif(1.0, 3.0) * pk.sign() * _hinge(X[:, c].astype(np.float64), pk.unif(-1, 1), pk.sign())
    return _finish(X, y)


def workload4(count=256, **kw):
    return [workload4_fit(i, **kw) for i in range(count)]


def workload5(seed=5, n=4_000_000, p=200, rho=0.9):
    X = np.empty((n, p), dtype=np.float32, order="F")
    prev = normal(seed, 0, n)
    X[:, 0] = prev
    a = np.sqrt(1.0 - rho * rho)
    for j in range(1, p):
        prev = rho * prev + a * normal(seed, j, n)
        X[:, j] = prev
    pk = _Picker(seed)
    y = normal(seed, 10_000, n)
    for c in pk.distinct(10, p):
        y += pk.unif(0.5, 2.0) * pk.sign() * X[:, c].astype(np.float64)
    for _ in range(5):
        a_, b_ = pk.distinct(2, p)
        y += pk.unif(0.5, 2.0) * pk.sign() * X[:, a_].astype(np.float64) * X[:, b_].astype(np.float64)
    return _finish(X, y)


# fit() keyword arguments of each workload (besides cuda / precision).
FIT_KW = {
    1: dict(max_terms=17, tail_span=0, linear_only=True),
    2: dict(max_terms=41, max_degree=2, linear_only=False),
    3: dict(max_terms=65, linear_only=False),
    4: dict(max_terms=33, linear_only=True),
    5: dict(max_terms=101, linear_only=True),
}

GENERATORS = {1: workload1, 2: workload2, 3: workload3, 5: workload5}
