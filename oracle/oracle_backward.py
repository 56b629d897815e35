"""
ORACLE -- test infrastructure only, never part of the product.

CPU restatements of the reference's backward-pass helpers, the checkers of
mars_amd.expand / gram / prune:

  expand_ref        mars.py:423-450 under numpy 2's promotion rules: 'l' is an
                    f32 (or f64) product; '+' / '-' subtract the f64 knot, clip
                    at 0 and multiply in f64, then round once to X's dtype.
  gram_ref          A^T B with both operands widened to f64 first (exact f32
                    products), summed by numpy's f64 matmul.
  prune_ref         mars.py:607-698 over given Gram products: standardise to a
                    unit diagonal, eliminate_brute, then numpy's lstsq on the
                    kept subset, back in the original scale.
  eliminate_brute   backward elimination of mars.py:537-604 by brute force:
                    every step refits each candidate subset from scratch with
                    lstsq and drops the column whose removal raises the SSE
                    least (the intercept, column 0, never), tracking the GCV.
                    The reference's Schur downdates (beta_j^2 / H_jj) compute
                    exactly this SSE increase, so the two must pick the same
                    path wherever the increases are not near-ties.

Pinning: gram_ref reduces to a library routine and is checked against the
reference's own gram known-answer tests (tests/test_prune.py:8-58 of the
reference, restated in tests/test_backward.py: A^T A, A^T B, 1-D and scalar
shapes, n * 0.25 at n = 100000 exactly).  expand_ref is checked against the
closed forms of each term type on hand-made rows and against the compact /
expand consistency case (reference tests/test_compact.py:121-141).
eliminate_brute is the textbook definition the downdates implement.
"""

import numpy as np


def expand_ref(X, model):
    X = np.asarray(X)
    B = np.empty((len(X), len(model)), dtype=X.dtype, order="F")
    for i, node in enumerate(model):
        t = node["type"].decode() if isinstance(node["type"], bytes) else node["type"]
        b, x, h = int(node["basis"]), int(node["input"]), float(node["hinge"])
        if t == "i":
            B[:, i] = 1
        elif t == "l":
            B[:, i] = B[:, b] * X[:, x]
        else:
            d = (X[:, x].astype(np.float64) - h) if t == "+" else (h - X[:, x].astype(np.float64))
            hv = np.where(np.isnan(d), d, np.where(d > 0, d, 0.0))
            B[:, i] = (B[:, b].astype(np.float64) * hv).astype(X.dtype)
    return B


def gram_ref(A, B=None):
    A = np.asarray(A, dtype=np.float64)
    B = A if B is None else np.asarray(B, dtype=np.float64)
    A2 = A[:, None] if A.ndim == 1 else A
    B2 = B[:, None] if B.ndim == 1 else B
    return A2.T @ B2


def _sse(XX, XY, YY, cols, ridge):
    if len(cols) == 0:
        return float(YY)
    G = XX[np.ix_(cols, cols)] + ridge * np.eye(len(cols))
    beta = np.linalg.lstsq(G, XY[cols], rcond=None)[0]
    return max(float(YY - XY[cols] @ beta), 0.0)


def eliminate_brute(XX, XY, YY, mask, penalty, n_true, ridge=0.0):
    """Mask of the lowest-GCV subset on the brute-force backward path (standardized XX)."""

    def gcv(sse, k):
        denom = 1.0 - (k + penalty * (k - 1)) / n_true
        return sse / (denom * denom) if denom > 0 else np.inf

    cols = list(np.flatnonzero(mask))
    cur = _sse(XX, XY, YY, cols, ridge)
    best, best_cols = gcv(cur, len(cols)), list(cols)
    while cols:
        base = _sse(XX, XY, YY, cols, ridge)
        rise = []
        for j, c in enumerate(cols):
            if j == 0 and c == 0:
                rise.append(np.inf)
                continue
            rest = cols[:j] + cols[j + 1:]
            rise.append(_sse(XX, XY, YY, rest, ridge) - base)
        j = int(np.argmin(rise))
        if not np.isfinite(rise[j]):
            break
        cur += max(rise[j], 0.0)
        del cols[j]
        g = gcv(cur, len(cols))
        if g < best:
            best, best_cols = g, list(cols)
    out = np.zeros(len(mask), dtype=bool)
    out[best_cols] = True
    return out


def prune_ref(XX, XY, YY, mask, penalty, n_true, ridge=0.0):
    """beta (M,) of mars.prune (mars.py:607-698) from raw Gram products, via eliminate_brute."""
    XX = np.asarray(XX, dtype=np.float64)
    XY = np.asarray(XY, dtype=np.float64).reshape(-1)
    d = np.sqrt(np.diag(XX))
    on = d > 0
    s = np.where(on, d, 1.0)
    S = XX / s / s[:, None]
    b = XY / s
    m = (np.ones(len(XX), dtype=bool) if mask is None else np.asarray(mask, dtype=bool)) & on
    keep = eliminate_brute(S, b, YY, m, penalty, n_true, ridge)
    if ridge > 0:
        S = S + ridge * np.eye(len(S))
    beta = np.zeros(len(XX))
    if keep.any():
        beta[keep] = np.linalg.lstsq(S[np.ix_(keep, keep)], b[keep], rcond=None)[0]
    return beta / s, keep
