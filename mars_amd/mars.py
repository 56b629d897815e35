"""
Multivariate Adaptive Regression Splines -- forward pass on the MI355X.

`fit` keeps the reference's `mars.fit` interface (mars.py:60-417): the same
keyword arguments, defaults, degree / self-interaction filter, Fast-MARS
ranking, GCV linear-versus-hinge choice, term budget, stopping rules and
model record.  Candidate scoring, the hinge sweep, the argmax and the basis
append run on the GPU through `marslib.MarsAlgo` (libmars_hip.so).

Differences from the reference, all on the device route:
  * `cuda` defaults to True -- this package is the device path; `cuda=False`
    raises (there is no CPU scoring code here).
  * `precision` selects the scoring form: "fp64" (default; the reference's
    FP64 arithmetic and f32 rounding points) or "mixed".
  * the winners come from the device argmax (same first-occurrence rule as
    numpy.argmax over the row-major (xcol, bcol) array).
"""

import os
import sys
import time

import numpy as np

from . import marslib

__version__ = marslib.__version__

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


def _dump_header(logger):
    if logger:
        logger.write("time        #   n    b    x   o    r²    r²_cv\n")
        logger.write("---------- --- --- ---- ---- --  ------ -------\n")


def _dump_row(logger, epoch, nbasis, dt, row, labels):
    logger.write("%02d:%02d:%04.1f " % (dt // 3600, (dt // 60) % 60, dt % 60))
    logger.write("%3d %3d %4d %4d %2d  %6.4f %7.4f" % (epoch, nbasis - 1, row["basis"], row["input"],
                                                      row["order"], row["r2"], row["r2_cv"]))
    if labels is not None:
        fmt = "%%-%ds" % max(len(_) for _ in labels)
        if row["type"] == b"l":
            logger.write(("  l " + fmt) % labels[row["input"]])
        else:
            logger.write(("  h " + fmt + " %.4g") % (labels[row["input"]], row["hinge"]))
    logger.write("\n")


def fit(X, y, w=None, **kwargs):
    """
    Multivariate Adaptive Regression Splines, forward pass (see the reference
    mars.fit docstring for every keyword; mars.py:60-175).  Extra keywords:

    precision : "fp64" (default) or "mixed"
    trace : list or None -- if a list, one dict per forward step is appended
        (chosen term, knot index, delta-SSEs, candidate count, append mse values).
    algo : a fresh marslib.MarsAlgo already built from (X, y, w) -- lets a caller
        upload the inputs to the GPU before timing the forward pass.
    """
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

    # fmt: off
    max_terms     = kwargs.pop("max_terms", 1 + 2 * min(X.shape[1] + len(X) // 20, 15))
    max_degree    = kwargs.pop("max_degree", 3)
    penalty       = kwargs.pop("penalty", 3.0)
    min_span      = kwargs.pop("min_span", None)
    tail_span     = kwargs.pop("tail_span", 0.02)
    max_runtime   = kwargs.pop("max_runtime", 48 * 3600)
    self_interact = kwargs.pop("self_interact", False)
    linear_only   = kwargs.pop("linear_only", False)
    n_true        = kwargs.pop("n_true", len(X))
    logger        = kwargs.pop("logger", None)
    callback      = kwargs.pop("callback", None)
    r2_window     = kwargs.pop("r2_window", 32)
    r2_thresh     = kwargs.pop("r2_thresh", 3e-5)
    labels        = kwargs.pop("labels", None)
    aux_filter    = kwargs.pop("xfilter", lambda x, b: True)
    max_basis     = kwargs.pop("max_basis", None)
    max_inputs    = kwargs.pop("max_inputs", None)
    aging_factor  = kwargs.pop("aging_factor", 1.0)
    threads       = kwargs.pop("threads", -1)
    cuda          = kwargs.pop("cuda", True)
    precision     = kwargs.pop("precision", "fp64")
    trace         = kwargs.pop("trace", None)
    algo          = kwargs.pop("algo", None)
    # fmt: on
    if kwargs:
        raise TypeError("unknown argument: %s" % kwargs)
    if not cuda:
        raise ValueError("mars_amd runs the forward pass on the GPU only; pass cuda=True")
    del threads

    start_t = time.time()
    profile = bool(os.environ.get("MARS_HIP_PROFILE"))
    eval_t = 0.0

    p = X.shape[1]
    if min_span is None:
        alpha = 1 / p
        min_span = int(-np.log2(-np.log(1 - alpha) / (n_true * alpha)) / 2.5)
    min_span = max(min_span, 1)

    get_dof = lambda m: m + penalty * (m - 1)  # noqa: E731
    gcv_adj = lambda mse, m: mse / (1.0 - get_dof(m) / n_true) ** 2  # noqa: E731
    basic_filter = lambda i, b: (len(b) < max_degree) and (self_interact or (i not in b))  # noqa: E731

    terms = np.arange(1, max_terms + 1)
    max_terms = terms[get_dof(terms) < n_true].max()

    n = len(X)
    basis = [[]]
    tail = max(int(n * tail_span), 1)
    if algo is None:
        algo = marslib.MarsAlgo(X, y, w, int(max_terms), precision=precision)
    elif algo.nbasis() != 1:
        raise ValueError("algo must be a fresh MarsAlgo built from the same X, y, w")
    else:
        algo.set_precision(precision)
    var_y = algo.yvar()

    epoch = 0
    basis_sse = np.array([0.0])
    input_sse = np.full(p, np.inf)
    input_age = np.full(p, epoch)
    model = np.zeros(max_terms, dtype=MODEL_DTYPE)
    model[0] = ("i", 0, 0, np.nan, np.nan, 0, 0, 0, 0)
    _dump_header(logger)

    def _ranks(v):
        r = np.empty_like(v, dtype=np.intp)
        r[np.argsort(v)] = np.arange(len(v))
        return r

    while algo.nbasis() < max_terms:
        basis_to_use = np.argsort(basis_sse)[::-1][:max_basis]
        input_to_use = np.argsort(_ranks(input_sse) + aging_factor * (epoch - input_age) * (input_sse > 0))
        n_take = None if max_inputs is None else max_inputs + int(np.isinf(input_sse).sum())
        input_to_use = np.sort(input_to_use[::-1][:n_take])

        bmask = np.zeros((p, len(basis)), dtype="bool")
        bmask[np.ix_(input_to_use, basis_to_use)] = True
        for i in input_to_use:
            bmask[i] &= np.array([basic_filter(i, b) for b in basis])
            bmask[i] &= np.array([aux_filter(i, b) for b in basis])

        e0 = time.perf_counter()
        sse0, sse1, sse2, cut, best = algo.eval_best(bmask, min_span, tail, linear_only)
        eval_t += time.perf_counter() - e0

        active = bmask.any(axis=0)
        basis_sse[active] = sse2[:, active].max(axis=0)
        input_sse[input_to_use] = sse2[input_to_use].max(axis=1)
        input_age[input_to_use] = epoch
        epoch += 1

        j1, j2 = best["j1"], best["j2"]
        s1, s2 = best["sse1"], best["sse2"]
        dt = time.time() - start_t
        if s1 <= 0 and s2 <= 0:
            break

        m0 = m = algo.nbasis()
        mse1 = gcv_adj((1.0 - sse0 - s1) / n, m + 1)
        mse2 = gcv_adj((1.0 - sse0 - s2) / n, m + 2)
        if mse1 <= mse2:
            xcol, bcol, hcut, htypes, knot = j1[0], j1[1], np.nan, ["l"], -1
        else:
            xcol, bcol, hcut, htypes, knot = j2[0], j2[1], best["cut2"], ["+", "-"], best["knot2"]

        step = dict(epoch=epoch, xcol=int(xcol), bcol=int(bcol), cut=float(hcut), knot=int(knot),
                    kind="l" if htypes == ["l"] else "h", sse0=float(sse0), sse1=float(s1),
                    sse2=float(s2), n_cand=int(bmask.sum()), mse=[])
        for htype in htypes:
            if algo.nbasis() >= max_terms:
                break
            mse = algo.append(htype, xcol, bcol, hcut)
            step["mse"].append(float(mse))
            if mse >= 0:
                basis.append(basis[bcol] + [xcol])
                model[m] = (htype, bcol, xcol, hcut, np.nan, 1.0 - mse / var_y,
                            1.0 - gcv_adj(mse, m + 1) / var_y, len(basis[-1]), dt)
                basis_sse = np.append(basis_sse, [0.0])
                m += 1
        if trace is not None:
            trace.append(step)
        assert algo.nbasis() == len(basis) == m

        if logger:
            _dump_row(logger, epoch, len(basis), dt, model[m - 1], labels)
        if callback:
            callback(epoch, len(basis), max_terms, model[m - 1])

        if m == m0:
            break
        if model[m - 1]["r2"] > 1 - r2_thresh:
            break
        if dt > max_runtime:
            break
        r2_cvs = model["r2_cv"][: algo.nbasis()]
        if len(r2_cvs) > r2_window and r2_cvs[-r2_window:].max() < r2_cvs.max() - r2_thresh:
            break

    if profile:
        total = time.time() - start_t
        sys.stderr.write("[mars_amd] fit: %.3fs total | %.3fs in eval (%.0f%%) | %d epochs | %s\n"
                         % (total, eval_t, 100.0 * eval_t / max(total, 1e-12), epoch, algo.stats()))
    return model[: algo.nbasis()]


def expand(X, model):
    """Rebuild the basis matrix of a model (reference mars.py:423-450); test helper."""
    X = np.asarray(X)
    B = np.empty((len(X), len(model)), dtype=X.dtype, order="F")
    for i, node in enumerate(model):
        t = node["type"].decode() if isinstance(node["type"], bytes) else node["type"]
        b, x, h = node["basis"], node["input"], node["hinge"]
        if t == "i":
            B[:, i] = 1.0
        elif t == "l":
            B[:, i] = B[:, b] * X[:, x]
        elif t == "+":
            B[:, i] = B[:, b] * np.maximum(X[:, x] - h, 0)
        elif t == "-":
            B[:, i] = B[:, b] * np.maximum(h - X[:, x], 0)
        else:
            raise ValueError("unknown basis type %r" % t)
    return B


def fit_batch(Xs, ys, ws=None, **kwargs):
    """
    Batching form: fit many independent problems (each its own X, y, w of any
    N and p) with one batched scoring pass, one device argmax and one device
    append per forward step for all of them (driver restated in C++).

    Supported keywords (same meaning as `fit`): max_terms (shared by all fits;
    default: the reference default of the largest problem), max_degree,
    penalty, self_interact, r2_window, r2_thresh, precision ("fp64"/"mixed"),
    linear_only (must be True: the batch scores linear terms), stats (a dict
    that receives step count, DGKS retries and per-kernel timing when given),
    batch (a marslib.Batch already built on these same inputs, max_terms and
    precision, so that device upload happens outside the call).
    Returns one model array per problem, as `fit` would.
    """
    Xs = [np.asarray(X) for X in Xs]
    ys = [np.ascontiguousarray(np.asarray(y).squeeze(), dtype="f") for y in ys]
    if ws is not None:
        ws = [np.ascontiguousarray(np.asarray(w).squeeze(), dtype="f") for w in ws]
    for X in Xs:
        assert X.dtype == "f", "X data must be 32-bit float"
        assert X.strides[0] == X.itemsize, "X data must be column-major"
    default_terms = max(1 + 2 * min(X.shape[1] + len(X) // 20, 15) for X in Xs)
    max_terms = kwargs.pop("max_terms", default_terms)
    max_degree = kwargs.pop("max_degree", 3)
    penalty = kwargs.pop("penalty", 3.0)
    self_interact = kwargs.pop("self_interact", False)
    r2_window = kwargs.pop("r2_window", 32)
    r2_thresh = kwargs.pop("r2_thresh", 3e-5)
    precision = kwargs.pop("precision", "fp64")
    linear_only = kwargs.pop("linear_only", True)
    stats = kwargs.pop("stats", None)
    timing = kwargs.pop("timing", False)
    batch = kwargs.pop("batch", None)
    if kwargs:
        raise TypeError("unknown argument: %s" % kwargs)
    if not linear_only:
        raise ValueError("fit_batch scores linear_only fits; use fit() for hinge sweeps")
    if batch is None:
        batch = marslib.Batch(Xs, ys, ws, int(max_terms), precision=precision)
    batch.set_timing(bool(timing))
    t0 = time.perf_counter()
    rows, steps = batch.fit(int(max_degree), float(penalty), bool(self_interact), int(r2_window), float(r2_thresh))
    dt = time.perf_counter() - t0
    models = []
    for r in rows:
        model = np.zeros(len(r), dtype=MODEL_DTYPE)
        for i, (t, b, x, o, mse, r2, r2cv) in enumerate(r):
            model[i] = (t, b, x, np.nan, np.nan, r2, r2cv, o, 0.0)
        models.append(model)
    if stats is not None:
        stats.update(batch.stats())
        stats["steps"] = steps
        stats["seconds"] = dt
        stats["mse"] = [[row[4] for row in r[1:]] for r in rows]
        stats["kernels"] = batch.kernel_stats()
    return models
