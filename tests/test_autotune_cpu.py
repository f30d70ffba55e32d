"""Autotuner host logic (no GPU): graph signatures and the tuning table."""
import numpy as np

import accelgraph_amd as ag
from accelgraph_amd import autotune


def test_signature_depends_on_shape_only():
    a = ag.Graph.rmat(8, seed=1)
    b = ag.Graph.rmat(8, seed=1)
    c = ag.Graph.rmat(8, seed=2)
    assert autotune.signature(a.n, a.m, a.in_degree) == autotune.signature(b.n, b.m, b.in_degree)
    assert autotune.signature(a.n, a.m, a.in_degree) != autotune.signature(c.n, c.m, c.in_degree)


def test_table_roundtrip(tmp_path):
    p = str(tmp_path / "t" / "tuning.json")
    autotune._save(p, {"k": {"params": {"order": 1}}})
    assert autotune._load(p) == {"k": {"params": {"order": 1}}}
    assert autotune._load(str(tmp_path / "missing.json")) == {}


def test_grid_candidates_are_valid_plans():
    for knob, values in autotune.GRID.items():
        for v in values:
            d = autotune._params_dict(ag.plan_params(**{knob: v}))
            if d["thread_max_deg"] <= d["group16_max_deg"] <= d["wave_max_deg"] <= d["block_max_deg"]:
                import ctypes as C
                err = C.c_char_p()
                ag.lib().ag_plan_params_check.restype = C.c_int
                ag.lib().ag_plan_params_check.argtypes = [C.POINTER(ag.PlanParams),
                                                          C.POINTER(C.c_char_p)]
                assert ag.lib().ag_plan_params_check(C.byref(ag.plan_params(**d)),
                                                     C.byref(err)) == 0, (knob, v)
