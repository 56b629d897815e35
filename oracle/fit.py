"""
ORACLE -- test infrastructure only, never part of the product.

Restatement of the reference forward-pass driver `mars.fit` (mars.py:60-417)
as a function of the scoring module, so the same loop can drive the C++
restatement (`oracle_marslib`) or the reference's own compiled `marslib`
(oracle/_ref).  Semantics kept from the reference:

  * keyword defaults and the unknown-kwarg TypeError      (mars.py:190-214)
  * Friedman min_span, GCV, degree / self-interaction filter, max_terms clamp
                                                           (mars.py:225-243)
  * Fast-MARS basis / input ranking                        (mars.py:280-323)
  * first-occurrence argmax in row-major (xcol, bcol) order (mars.py:342-343)
  * linear-vs-hinge GCV choice (`mse1 <= mse2` -> linear)  (mars.py:349-357)
  * append loop with the term budget, model rows, stops    (mars.py:359-406)

`fit` returns (model, trace); trace holds per-epoch details used by the parity
tests: the chosen (type, xcol, bcol, cut, knot index) and the append mse.
"""

import time

import numpy as np

MODEL_DTYPE = [
    ("type", "S1"),
    ("basis", "i4"),
    ("input", "i4"),
    ("hinge", "f8"),
    ("beta", "f4"),
    ("r2", "f4"),
    ("r2_cv", "f4"),
    ("order", "i4"),
    ("time", "f4"),
]


def fit(lib, X, y, w=None, **kwargs):
    if w is None:
        w = np.ones(len(X), dtype="f")
    X = np.asarray(X)
    y = np.asarray(np.asarray(y).squeeze(), dtype="f")
    w = np.asarray(np.asarray(w).squeeze(), dtype="f")
    assert X.dtype == "f", "X data must be 32-bit float"
    assert X.strides[0] == X.itemsize, "X data must be column-major"
    assert y.strides[0] == y.itemsize, "y data must be column-major"
    assert len(X) == len(y) == len(w), "input dataset length mismatch"
    assert len(X) > 0, "empty dataset"

    p = X.shape[1]
    max_terms = kwargs.pop("max_terms", 1 + 2 * min(p + len(X) // 20, 15))
    max_degree = kwargs.pop("max_degree", 3)
    penalty = kwargs.pop("penalty", 3.0)
    min_span = kwargs.pop("min_span", None)
    tail_span = kwargs.pop("tail_span", 0.02)
    max_runtime = kwargs.pop("max_runtime", 48 * 3600)
    self_interact = kwargs.pop("self_interact", False)
    linear_only = kwargs.pop("linear_only", False)
    n_true = kwargs.pop("n_true", len(X))
    r2_window = kwargs.pop("r2_window", 32)
    r2_thresh = kwargs.pop("r2_thresh", 3e-5)
    aux_filter = kwargs.pop("xfilter", lambda x, b: True)
    max_basis = kwargs.pop("max_basis", None)
    max_inputs = kwargs.pop("max_inputs", None)
    aging_factor = kwargs.pop("aging_factor", 1.0)
    threads = kwargs.pop("threads", -1)
    max_epochs = kwargs.pop("max_epochs", None)  # oracle-only: truncate long fits
    max_seconds = kwargs.pop("max_seconds", None)  # oracle-only: no new epoch after this much time
    cuda = kwargs.pop("cuda", False)  # the binding's cuda= flag (mars.py:156-160, 330)
    if kwargs:
        raise TypeError("unknown argument: %s" % kwargs)

    start_t = time.time()
    if min_span is None:
        alpha = 1 / p
        min_span = int(-np.log2(-np.log(1 - alpha) / (n_true * alpha)) / 2.5)
    min_span = max(min_span, 1)

    dof = lambda m: m + penalty * (m - 1)  # noqa: E731
    gcv = lambda mse, m: mse / (1.0 - dof(m) / n_true) ** 2  # noqa: E731
    allowed = lambda i, b: (len(b) < max_degree) and (self_interact or (i not in b))  # noqa: E731

    terms = np.arange(1, max_terms + 1)
    max_terms = terms[dof(terms) < n_true].max()

    n = len(X)
    basis = [[]]
    tail = max(int(n * tail_span), 1)
    algo = lib.MarsAlgo(X, y, w, max_terms)
    var_y = algo.yvar()

    epoch = 0
    basis_sse = np.array([0.0])
    input_sse = np.full(p, np.inf)
    input_age = np.full(p, epoch)
    model = np.zeros(max_terms, dtype=MODEL_DTYPE)
    model[0] = ("i", 0, 0, np.nan, np.nan, 0, 0, 0, 0)
    trace = []
    has_knots = hasattr(algo, "eval_knots")

    def ranks(v):
        r = np.empty_like(v, dtype=np.intp)
        r[np.argsort(v)] = np.arange(len(v))
        return r

    t_start = time.perf_counter()
    while algo.nbasis() < max_terms:
        if max_epochs is not None and epoch >= max_epochs:
            break
        if max_seconds is not None and epoch > 0 and time.perf_counter() - t_start >= max_seconds:
            break
        t_epoch = time.perf_counter()
        basis_to_use = np.argsort(basis_sse)[::-1][:max_basis]
        input_to_use = np.argsort(ranks(input_sse) + aging_factor * (epoch - input_age) * (input_sse > 0))
        n_take = None if max_inputs is None else max_inputs + int(np.isinf(input_sse).sum())
        input_to_use = np.sort(input_to_use[::-1][:n_take])

        bmask = np.zeros((p, len(basis)), dtype="bool")
        bmask[np.ix_(input_to_use, basis_to_use)] = True
        for i in input_to_use:
            bmask[i] &= np.array([allowed(i, b) for b in basis])
            bmask[i] &= np.array([aux_filter(i, b) for b in basis])

        if has_knots:
            sse0, sse1, sse2, cut, knots, _ = algo.eval_knots(bmask, min_span, tail, linear_only, threads)
        else:
            sse0, sse1, sse2, cut = algo.eval(bmask, min_span, tail, linear_only, threads, cuda)
            knots = None

        active = bmask.any(axis=0)
        basis_sse[active] = sse2[:, active].max(axis=0)
        input_sse[input_to_use] = sse2[input_to_use].max(axis=1)
        input_age[input_to_use] = epoch
        epoch += 1

        j1 = np.unravel_index(np.argmax(sse1), sse1.shape)
        j2 = np.unravel_index(np.argmax(sse2), sse2.shape)
        dt = time.time() - start_t
        if sse1[j1] <= 0 and sse2[j2] <= 0:
            break

        m0 = m = algo.nbasis()
        mse1 = gcv((1.0 - sse0 - sse1[j1]) / n, m + 1)
        mse2 = gcv((1.0 - sse0 - sse2[j2]) / n, m + 2)
        if mse1 <= mse2:
            xcol, bcol, hcut, types, knot = j1[0], j1[1], np.nan, ["l"], -1
        else:
            xcol, bcol, hcut, types = j2[0], j2[1], cut[j2], ["+", "-"]
            knot = int(knots[j2]) if knots is not None else -1

        step = dict(epoch=epoch, xcol=int(xcol), bcol=int(bcol), cut=float(hcut), knot=knot,
                    kind="l" if types == ["l"] else "h", sse0=float(sse0),
                    sse1=float(sse1[j1]), sse2=float(sse2[j2]), n_cand=int(bmask.sum()), mse=[])
        for t in types:
            if algo.nbasis() >= max_terms:
                break
            mse = algo.append(t, xcol, bcol, hcut)
            step["mse"].append(float(mse))
            if mse >= 0:
                basis.append(basis[bcol] + [xcol])
                model[m] = (t, bcol, xcol, hcut, np.nan, 1.0 - mse / var_y,
                            1.0 - gcv(mse, m + 1) / var_y, len(basis[-1]), dt)
                basis_sse = np.append(basis_sse, [0.0])
                m += 1
        step["secs"] = time.perf_counter() - t_epoch  # eval + append of this epoch (timing only)
        trace.append(step)
        assert algo.nbasis() == len(basis) == m

        if m == m0:
            break
        if model[m - 1]["r2"] > 1 - r2_thresh:
            break
        if dt > max_runtime:
            break
        r2_cvs = model["r2_cv"][: algo.nbasis()]
        if len(r2_cvs) > r2_window and r2_cvs[-r2_window:].max() < r2_cvs.max() - r2_thresh:
            break

    return model[: algo.nbasis()], trace
