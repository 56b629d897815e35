"""
Golden fit cases: deterministic generator + fit keyword arguments.

Each case's inputs come from mars_amd/datagen.py, so only the recipe is
stored; make_golden.py records the reference's outputs for each recipe.
"""

import numpy as np

from mars_amd import datagen

CASES = {
    # workload 1 exactly as specified (README example shape)
    "w1": dict(gen="workload1", args={}, fit=dict(max_terms=17, tail_span=0, linear_only=True)),
    # workload 2 exactly as specified
    "w2": dict(gen="workload2", args={}, fit=dict(max_terms=41, max_degree=2, linear_only=False)),
    # reduced workloads 2 / 3 / 5 (CPU-test sized)
    "w2_small": dict(gen="workload2", args=dict(n=20_000, p=14),
                     fit=dict(max_terms=21, max_degree=2, linear_only=False)),
    "w3_small": dict(gen="workload3", args=dict(n=40_000, p=20, n_uniform=16),
                     fit=dict(max_terms=25, linear_only=False)),
    "w5_small": dict(gen="workload5", args=dict(n=60_000, p=40),
                     fit=dict(max_terms=41, linear_only=True)),
    # hinge sweep edge cases: every sample a knot, wide tails, self interaction,
    # degree 3, ties in x (Bernoulli column), WLS weights
    "hinge_edges": dict(gen="workload1", args=dict(n=3_000),
                        fit=dict(max_terms=15, min_span=1, tail_span=0.1, linear_only=False,
                                 self_interact=True)),
    "weighted": dict(gen="workload2", args=dict(n=8_000, p=6), weights=True,
                     fit=dict(max_terms=15, linear_only=False)),
    "deg3": dict(gen="workload2", args=dict(n=5_000, p=5),
                 fit=dict(max_terms=19, max_degree=3, min_span=3, tail_span=0.01,
                          linear_only=False)),
}
for _i in range(8):
    CASES["w4_fit%d" % _i] = dict(gen="workload4_fit", args=dict(i=_i),
                                  fit=dict(max_terms=33, linear_only=True))


def make_inputs(case):
    X, y = getattr(datagen, case["gen"])(**case["args"])
    w = None
    if case.get("weights"):
        w = (0.25 + datagen.uniform(77, 0, len(y))).astype(np.float32)
    return X, y, w
