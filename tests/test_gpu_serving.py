"""Continuous batching: a specimen solved inside a packed batch (sharing
launches with other specimens that start and finish at other times) gives
exactly the result it gives alone."""
import time

import numpy as np
import pytest

from cabanapd_amd import configs, engine
from tests.parity import ctx, say

pytestmark = pytest.mark.gpu


def make(w, precision):
    d = type(w["desc"]).from_buffer_copy(w["desc"])
    d.precision = {"fp64": 0, "mixed": 1}[precision]
    s = engine.Spec(ctx(), d, w["x"].shape[0])
    configs.upload(s, w)
    configs.setup(s, w)
    return s


def workloads():
    return [
        (configs.kw_plate(num_cells=(40, 80, 9), steps=90, name="a"), "mixed", 90),
        (configs.kw_plate(num_cells=(30, 50, 6), steps=40, G0=200.0, impact_v=25.0, name="b"), "mixed", 40),
        (configs.kw_plate(num_cells=(50, 60, 5), steps=70, name="c"), "fp64", 70),
        (configs.lps_cube(n_side=16, steps=55, fracture=True), "mixed", 55),
        (configs.elastic_cube(n_side=15, steps=30), "fp64", 30),
        (configs.kw_plate(num_cells=(20, 40, 4), steps=25, name="d"), "mixed", 25),
    ]


def test_alone_equals_mix():
    say("case: six specimens alone vs packed batch with staggered admission")
    wl = workloads()
    alone = []
    for w, prec, steps in wl:
        s = make(w, prec)
        s.step(steps)
        alone.append((s.get("u"), s.get("v"), s.bonds()[2]))
        s.close()
    c = ctx()
    specs = [make(w, prec) for w, prec, _ in wl]
    tickets = {}
    # first three now, the rest after some packed steps (continuous admission)
    for k in range(3):
        tickets[k] = c.submit(specs[k], wl[k][2])
    c.progress(17)
    for k in range(3, 5):
        tickets[k] = c.submit(specs[k], wl[k][2])
    c.progress(9)
    tickets[5] = c.submit(specs[5], wl[5][2])
    c.progress(0)
    for k, t in tickets.items():
        deadline = time.time() + 60.0
        while not c.poll(t) and time.time() < deadline:
            time.sleep(1e-3)
        assert c.poll(t), k
    for k, s in enumerate(specs):
        assert s.steps_done() == wl[k][2]
        u, v, intact = s.get("u"), s.get("v"), s.bonds()[2]
        say(f"  spec {k}: u equal {np.array_equal(u, alone[k][0])} v equal {np.array_equal(v, alone[k][1])}")
        assert np.array_equal(u, alone[k][0])
        assert np.array_equal(v, alone[k][1])
        assert np.array_equal(intact, alone[k][2])
        s.close()


def test_budget_rejects_oversized_specimen():
    say("case: HBM budget admission")
    small = engine.Context(0, hbm_budget=64 << 20)
    w = configs.kw_plate(num_cells=(200, 200, 18), steps=1)
    s = None
    try:
        with pytest.raises(engine.PdkError, match="budget"):
            s = engine.Spec(small, w["desc"], w["x"].shape[0])
            configs.upload(s, w)
            configs.setup(s, w)
    finally:
        if s is not None:
            s.close()
    assert small.bytes_in_use() == 0
    small.close()
