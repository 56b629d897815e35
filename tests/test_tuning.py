"""The launch-shape tuning table (mars_amd/tuning.json, written by mars_amd.autotune on an
MI355X) and its lookup; no device work."""

import pytest

from mars_amd import marslib, tuning


def test_table_entries_are_valid_shapes():
    entries = tuning.load()
    assert entries, "the tuning table is committed with the package"
    keys = set(marslib.default_tuning())
    for e in entries:
        assert set(e["params"]) == keys
        assert e["tuned_s"] > 0 and e["default_s"] > 0
        assert {"kind", "form", "hinge", "rows_log2"} <= set(e["class"])


def test_lookup_takes_nearest_row_class_of_same_kind():
    entries = [
        {"class": tuning.workload_class("batch", "mixed", False, 1 << 20), "params": {"pass_items": 1.0}},
        {"class": tuning.workload_class("batch", "mixed", False, 1 << 10), "params": {"pass_items": 2.0}},
        {"class": tuning.workload_class("single", "mixed", False, 1 << 20), "params": {"pass_items": 3.0}},
    ]
    assert tuning.lookup("batch", "mixed", False, 1 << 19, entries) == {"pass_items": 1.0}
    assert tuning.lookup("batch", "mixed", False, 1 << 12, entries) == {"pass_items": 2.0}
    assert tuning.lookup("single", "mixed", False, 5, entries) == {"pass_items": 3.0}
    assert tuning.lookup("single", "fp64", True, 5, entries) is None


def test_lookup_can_be_switched_off(monkeypatch):
    monkeypatch.setenv("MARS_HIP_TUNING", "0")
    assert tuning.lookup("batch", "mixed", False, 1 << 23) is None


def test_defaults_match_the_library():
    d = marslib.default_tuning()
    assert d["pass_items"] == 32768 and d["update_waves"] == 16384 and d["sweep_waves"] == 65536
    assert d["sweep_min_rows"] == 512 and d["append_rows"] == 0 and d["narrow_avg"] == 24


@pytest.mark.gpu
def test_tuned_shapes_reproduce_golden_fits():
    """Every table entry's shapes, applied to the reference-written fixtures of its kind, give
    the reference's term sequence and mse (a launch shape only reorders FP64 sums)."""
    import json
    import os

    import mars_amd
    from tests.golden.cases import CASES, make_inputs

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for e in tuning.load():
        c, params = e["class"], e["params"]
        print("tuned shapes", c, flush=True)
        if c["kind"] == "batch":
            golds = [json.load(open(os.path.join(here, "w4_fit%d.json" % i))) for i in range(4)]
            inp = [make_inputs(CASES[g["name"]]) for g in golds]
            st = {}
            models = mars_amd.fit_batch([x for x, _, _ in inp], [y for _, y, _ in inp], max_terms=33,
                                        precision=c["form"], tuning=params, stats=st)
            for g, mdl, mse in zip(golds, models, st["mse"]):
                assert [(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in mdl] == \
                    [(r["type"], r["basis"], r["input"]) for r in g["model"]]
                for a, b in zip(mse, [a["mse"] for a in g["appends"]]):
                    assert abs(a - b) <= 1e-10 * abs(b)
        else:
            g = json.load(open(os.path.join(here, ("w2_small" if c["hinge"] else "w5_small") + ".json")))
            X, y, w = make_inputs(CASES[g["name"]])
            trace = []
            mdl = mars_amd.fit(X, y, w, trace=trace, precision=c["form"], tuning=params, **CASES[g["name"]]["fit"])
            assert [(r["type"].decode(), int(r["basis"]), int(r["input"])) for r in mdl] == \
                [(r["type"], r["basis"], r["input"]) for r in g["model"]]
            got = [v for s in trace for v in s["mse"]]
            for a, b in zip(got, [a["mse"] for a in g["appends"]]):
                assert abs(a - b) <= 1e-10 * abs(b)


@pytest.mark.gpu
def test_set_tuning_rejects_out_of_range_shapes():
    import numpy as np

    X = np.asfortranarray(np.random.RandomState(0).randn(500, 3).astype(np.float32))
    y = X[:, 0].copy()
    a = marslib.MarsAlgo(X, y, np.ones(500, dtype=np.float32), 5, precision="fp64")
    for bad in ({"pass_items": 10.0}, {"append_rows": 1000}, {"sweep_min_rows": 1}, {"narrow_avg": -1}):
        with pytest.raises(RuntimeError, match="tuning"):
            a.set_tuning(bad)
    with pytest.raises(RuntimeError, match="unknown tuning parameter"):
        a.set_tuning({"tile": 3})
    a.set_tuning({"pass_items": 16384.0})
    assert a.get_tuning()["pass_items"] == 16384.0
    a.close()
