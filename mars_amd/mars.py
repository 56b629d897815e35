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
    tuning : launch shapes of a context this call creates -- None (default) the
        tuning table's entry for the workload (mars_amd/tuning.json), a dict
        those shapes, False the library defaults.
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
    aux_filter    = kwargs.pop("xfilter", None)
    max_basis     = kwargs.pop("max_basis", None)
    max_inputs    = kwargs.pop("max_inputs", None)
    aging_factor  = kwargs.pop("aging_factor", 1.0)
    threads       = kwargs.pop("threads", -1)
    cuda          = kwargs.pop("cuda", True)
    precision     = kwargs.pop("precision", "fp64")
    trace         = kwargs.pop("trace", None)
    algo          = kwargs.pop("algo", None)
    tuning        = kwargs.pop("tuning", None)
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

    terms = np.arange(1, max_terms + 1)
    max_terms = terms[get_dof(terms) < n_true].max()

    n = len(X)
    basis = [[]]
    tail = max(int(n * tail_span), 1)
    if algo is None:
        algo = marslib.MarsAlgo(X, y, w, int(max_terms), precision=precision)
        _set_tuning(algo, tuning, "single", precision, not linear_only, n)
    elif algo.nbasis() != 1:
        raise ValueError("algo must be a fresh MarsAlgo built from the same X, y, w")
    else:
        algo.set_precision(precision)
    var_y = algo.yvar()

    if max_basis is None and max_inputs is None and aux_filter is None and logger is None and callback is None:
        # the default configuration: the whole loop runs in C++ (mars_hip_fit_forward)
        offset = time.time() - start_t
        recs, _, steps = algo.forward(int(max_terms), int(max_degree), float(penalty), int(min_span), int(tail),
                                      bool(self_interact), bool(linear_only), float(n_true), int(r2_window),
                                      float(r2_thresh), float(max_runtime) - offset)
        model = np.frombuffer(recs, dtype=MODEL_DTYPE).copy()
        model["time"][1:] += np.float32(offset)
        if trace is not None:
            trace.extend(steps)
        if profile:
            sys.stderr.write("[mars_amd] fit: %.3fs total (C++ driver) | %d epochs | %s\n"
                             % (time.time() - start_t, len(steps), algo.stats()))
        return model

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

    # the degree / self-interaction filter (mars.py:237-239) as arrays: each basis function's degree
    # and variable set
    degree = np.zeros(max_terms, dtype=np.intp)
    uses = np.zeros((max_terms, p), dtype=bool)

    while algo.nbasis() < max_terms:
        basis_to_use = np.argsort(basis_sse)[::-1][:max_basis]
        input_to_use = np.argsort(_ranks(input_sse) + aging_factor * (epoch - input_age) * (input_sse > 0))
        n_take = None if max_inputs is None else max_inputs + int(np.isinf(input_sse).sum())
        input_to_use = np.sort(input_to_use[::-1][:n_take])

        nb = len(basis)
        bmask = np.zeros((p, nb), dtype="bool")
        bmask[np.ix_(input_to_use, basis_to_use)] = True
        bmask &= degree[None, :nb] < max_degree
        if not self_interact:
            bmask &= ~uses[:nb].T
        if aux_filter is not None:
            for i in input_to_use:
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
                degree[m] = degree[bcol] + 1
                uses[m] = uses[bcol]
                uses[m, xcol] = True
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
    """
    Expand a feature set X into the bases of a MARS model (reference
    mars.py:423-450), on the device (mars_hip_expand).  X is float32 or
    float64; B has X's dtype and Fortran order.  Hinge terms form
    max(x - h, 0) and its product in float64 and round once to X's dtype, as
    numpy does for an f32 array against the model's f64 knot.
    """
    X = np.asarray(X)
    if X.ndim != 2:
        raise ValueError("expand: X must be 2-D")
    if X.dtype == np.float32:
        kernel = marslib.expand_f32
    elif X.dtype == np.float64:
        kernel = marslib.expand_f64
    else:
        raise TypeError("expand: X must be float32 or float64, got %s" % X.dtype)
    M = len(model)
    codes = bytearray(M)
    parent = np.zeros(M, dtype=np.int32)
    column = np.zeros(M, dtype=np.int32)
    knot = np.zeros(M, dtype=np.float64)
    for i, node in enumerate(model):
        t = node["type"]
        t = t.decode("utf8") if isinstance(t, bytes) else str(t)
        if t != "i":
            assert 0 <= node["input"] < X.shape[1]
            assert 0 <= node["basis"] < i
            assert t in ("l", "+", "-")
        codes[i] = ord(t)
        parent[i] = node["basis"]
        column[i] = node["input"]
        knot[i] = node["hinge"]
    return kernel(X, bytes(codes), parent, column, knot)


def gram(A, B=None, chunk_size=128):
    """
    A.T @ B (A.T @ A when B is None) with float64 accumulation (reference
    mars.py:456-490), on the device (mars_hip_gram).  float32 inputs are
    multiplied exactly in float64; `chunk_size` is accepted for compatibility
    and has no effect (the device sums in float64 throughout).

    1-D inputs are column vectors and the result is squeezed like numpy's
    `@`: a float if both are 1-D, 1-D if exactly one is, otherwise 2-D.
    """
    del chunk_size
    same = B is None
    A = np.asarray(A)
    B = A if same else np.asarray(B)
    if len(A) != len(B):
        raise ValueError("shapes are not aligned: %s x %s" % (A.shape, B.shape))
    A2 = A[:, np.newaxis] if A.ndim == 1 else A
    B2 = B[:, np.newaxis] if B.ndim == 1 else B
    if A2.dtype == np.float32 and B2.dtype == np.float32:
        out = marslib.gram_f32(A2, None if same else B2)
    else:
        A2 = A2.astype(np.float64, copy=False)
        B2 = B2.astype(np.float64, copy=False)
        out = marslib.gram_f64(A2, None if same else B2)
    if A.ndim == 1 and B.ndim == 1:
        return float(out[0, 0])
    if A.ndim == 1:
        return out[0]
    if B.ndim == 1:
        return out[:, 0]
    return out


def prune(B, y, w=None, n_true=None, penalty=3.0, ridge=0.0, mask=None):
    """
    Coefficients of the basis by backward stepwise selection under GCV
    (Friedman 1991, section 3.4), with the reference's interface and
    selection rule (mars.py:493-698).  Rows whose y is not finite get weight
    0; weights fold in as sqrt(w); the intercept (column 0) is never removed;
    columns are standardised to a unit Gram diagonal, where `ridge` applies.

    The Gram products B^T B, B^T y, y^T y run on the device (gram); the
    elimination and the final solve run in the library's host code
    (mars_hip_prune_gram).  Returns beta (M,) in B's scale, 0 for pruned terms.
    """
    B, y = _weighted_rows(B, y, w)
    M = B.shape[1]
    XX = gram(B)
    XY = np.atleast_1d(gram(B, y))
    YY = float(gram(y, y))
    keep = None if mask is None else np.asarray(mask, dtype=bool).reshape(M)
    beta, _ = marslib.prune_gram(XX, XY, YY, keep, float(penalty), float(len(B) if n_true is None else n_true),
                                 float(ridge))
    return beta


def _weighted_rows(B, y, w):
    """(B, y) with each row scaled by sqrt(w); rows with a non-finite y weigh 0 (their y is set
    to 0).  The products keep numpy's promotion: sqrt(w) is taken in B's dtype."""
    finite = np.isfinite(y)
    if not finite.all():
        y = np.where(finite, y, 0)
        w = np.where(finite, np.ones(len(y), dtype=B.dtype) if w is None else w, 0)
    if w is None:
        return B, y
    root = np.sqrt(np.asarray(w, dtype=B.dtype))
    return B * root[:, None], y * root


def compact(model):
    """
    The model without its zero-`beta` nodes, except those an earlier-kept node
    needs as an ancestor (they stay, with beta 0); parent indices renumbered
    (reference mars.py:701-738).
    """
    needed = set(np.flatnonzero(model["beta"] != 0).tolist())
    stack = list(needed)
    while stack:  # close the set under "parent of"
        i = stack.pop()
        if model[i]["type"] != b"i":
            p = int(model[i]["basis"])
            if p not in needed:
                needed.add(p)
                stack.append(p)
    order = sorted(needed)
    where = {old: new for new, old in enumerate(order)}
    out = model[np.array(order, dtype=np.intp)].copy() if order else model[:0].copy()
    for r in range(len(out)):
        out[r]["basis"] = where.get(int(out[r]["basis"]), -1)
    return out


def pprint(model, beta=None, labels=None):
    """
    One line per node: its coefficient (model["beta"] unless `beta` is given)
    and the product of the factors on its path to the intercept, the node's
    own factor first (reference mars.py:741-776).
    """
    coef = model["beta"] if beta is None else beta
    names = (lambda k: "X[%d]" % k) if labels is None else (lambda k: labels[k])
    for i in range(len(model)):
        parts = []
        node = i
        while model[node]["type"] != b"i":
            row = model[node]
            x, h = names(row["input"]), row["hinge"]
            parts.append({b"+": "MAX(%s-%g,0)" % (x, h), b"-": "MAX(%g-%s,0)" % (h, x)}.get(row["type"], x))
            if row["basis"] <= 0:
                break
            node = int(row["basis"])
        text = "  "
        if coef is not None and hasattr(coef, "__getitem__"):
            text += "%+9.4g" % coef[i]
        if parts:
            text += " * " + " * ".join(parts)
        print(text)


def _set_tuning(ctx, tuning, kind, form, hinge, rows):
    """Launch shapes of a new context: None takes the tuning table's entry for the workload class
    (mars_amd/tuning.json), a dict sets those shapes, False keeps the context's."""
    from . import tuning as table

    if tuning is False:
        return
    if tuning is None:
        table.apply(ctx, kind, form, hinge, rows)
    else:
        ctx.set_tuning(dict(tuning))


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
    precision, so that device upload happens outside the call), tuning (None:
    the launch shapes of the tuning table for this workload class; a dict:
    those shapes; False: leave the context's shapes as they are).
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
    tuning = kwargs.pop("tuning", None)
    if kwargs:
        raise TypeError("unknown argument: %s" % kwargs)
    if not linear_only:
        raise ValueError("fit_batch scores linear_only fits; use fit() for hinge sweeps")
    if batch is None:
        batch = marslib.Batch(Xs, ys, ws, int(max_terms), precision=precision)
    _set_tuning(batch, tuning, "batch", precision, False, sum(len(X) for X in Xs))
    batch.set_timing(bool(timing))
    t0 = time.perf_counter()
    rec, counts, mse, steps = batch.fit(int(max_degree), float(penalty), bool(self_interact), int(r2_window),
                                        float(r2_thresh))
    dt = time.perf_counter() - t0
    rows = np.frombuffer(rec, dtype=np.dtype(MODEL_DTYPE))
    ends = np.cumsum(counts)
    models = [m.copy() for m in np.split(rows, ends[:-1])]
    if stats is not None:
        stats.update(batch.stats())
        stats["steps"] = steps
        stats["seconds"] = dt
        stats["mse"] = [list(v[1:]) for v in np.split(mse, ends[:-1])]
        stats["kernels"] = batch.kernel_stats()
    return models
