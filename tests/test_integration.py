"""
The reference-side binding: the reference's own MarsAlgo (kernels.cc, marsalgo.cc,
marslib.cc, compiled from its sources as its USE_CUDA option compiles them) with
integration/mars_hip_shim.cc in place of cuda/orthonormalize.cu, linked to
libmars_hip.so (`make -C oracle ref-hip`).

Its `cuda=True` path is then the reference's code calling this library's C ABI:
linear_only evals go through MarsAlgo::eval_batch -> orthonormalize_batch
(marsalgo.cc:664-758), hinge evals through MarsAlgo::eval -> orthonormalize with
the normalised columns downloaded and the reference's CPU hinge sweep
(marsalgo.cc:455-481, 508-657), and the append stays the reference's host MGS.
Driven by the oracle's restatement of the mars.fit loop, every such fit must give
the golden term sequence (knots included) and per-step mse within 1e-10 of the
reference's own CPU fit.
"""

import glob
import importlib.util
import json
import os
import sys

import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.golden.cases import CASES, make_inputs

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILT = glob.glob(os.path.join(ROOT, "oracle", "_ref", "hip", "marslib_hip*.so"))

needs_build = pytest.mark.skipif(
    not BUILT, reason="oracle/_ref/hip not built (make -C oracle ref-hip; needs the reference sources)")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ref_hip():
    spec = importlib.util.spec_from_file_location("marslib_hip", BUILT[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@needs_build
@pytest.mark.skipif(gpu_available(), reason="checks the error route where no device is present")
def test_shim_error_reaches_the_binding(ref_hip):
    """With no GPU the library's error comes back through the shim as the RuntimeError the
    binding's exception capture expects (marslib.cc:113-145), text included."""
    rng = np.random.default_rng(0)
    X = np.asfortranarray(rng.standard_normal((64, 2)).astype(np.float32))
    y = rng.standard_normal(64).astype(np.float32)
    algo = ref_hip.MarsAlgo(X, y, np.ones(64, np.float32), 5)
    mask = np.ones((2, 1), dtype=bool)
    algo.eval(mask, 1, 1, True, 1, False)  # the reference's CPU path is untouched
    for linear in (True, False):
        with pytest.raises(RuntimeError, match="^mars_hip: "):
            algo.eval(mask, 1, 1, linear, 1, True)


# linear fits (eval_batch route) and hinge fits (orthonormalize route, CPU sweep) of the golden
# cases, sized so the reference's single-threaded host sweep stays within seconds
FITS = ["w1", "w4_fit0", "w4_fit3", "w5_small", "w2_small", "hinge_edges", "weighted", "deg3"]


@needs_build
@pytest.mark.gpu
@pytest.mark.parametrize("name", FITS)
def test_reference_cuda_path_matches_golden(ref_hip, name):
    from oracle import fit as F

    gold = json.load(open(os.path.join(HERE, "golden", name + ".json")))
    case = CASES[name]
    say("reference MarsAlgo, cuda=True through the shim:", name)
    X, y, w = make_inputs(case)
    model, trace = F.fit(ref_hip, X, y, w, cuda=True, threads=1, **case["fit"])
    got = [(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in model]
    want = [(r["type"], r["basis"], r["input"]) for r in gold["model"]]
    assert got == want, "term sequence differs"
    for r, g in zip(gold["model"], model):
        if r["type"] in "+-":
            assert float(g["hinge"]) == r["hinge"], "knot differs"
    got_mse = [v for s in trace for v in s["mse"]]
    want_mse = [a["mse"] for a in gold["appends"]]
    assert len(got_mse) == len(want_mse)
    for a, b in zip(got_mse, want_mse):
        assert abs(a - b) <= 1e-10 * max(abs(b), 1e-300), (a, b)
