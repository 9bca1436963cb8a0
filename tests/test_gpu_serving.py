"""Continuous batching: a specimen solved inside a packed batch (sharing
launches with other specimens that start and finish at other times) gives
exactly the result it gives alone."""
import os
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


def test_serving_stream_against_oracle_and_fp64():
    """A seeded Poisson stream of serving-stream specimens (the generator of
    the serving benchmark: log-uniform particle counts, aspect ratios 1-4,
    thickness 3-20 cells, random E, G0, 0-3 slanted pre-notches, edge impact;
    particle counts and step counts scaled to the test's budget) submitted
    through the continuous-batching engine while earlier ones run.  Every
    other specimen is plain FP64 and must equal the oracle bit for bit (u, v,
    broken bonds); the others are mixed and must match their FP64 solo run
    (u within 1e-5 relative, identical broken-bond set) and their own mixed
    solo run bit for bit (the same alone and in the mix)."""
    from oracle import binding as orc
    from tests.parity import oracle_spec, rel_err
    seed, rate, count = 20261203, 50.0, 10
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    reqs = []
    for k in range(count):
        prec = "fp64" if k % 2 == 0 else "mixed"
        reqs.append((configs.serving_request(rng, precision=prec, n_range=(2e4, 1.2e5), steps_range=(60, 240)),
                     prec))
    say(f"case: serving stream seed {seed}, {count} specimens, arrivals over {arrivals[-1]:.3f} s")
    c = ctx()
    specs, tickets = [], {}
    t0 = time.perf_counter()
    nxt = 0
    while nxt < count or any(not c.poll(t) for t in tickets.values()):
        now = time.perf_counter() - t0
        while nxt < count and arrivals[nxt] <= now:
            w, prec = reqs[nxt]
            s = make(w, prec)
            specs.append(s)
            tickets[nxt] = c.submit(s, w["steps"])
            nxt += 1
        if tickets:
            c.progress(8)
        else:
            time.sleep(1e-3)
        assert time.perf_counter() - t0 < 90.0, "stream did not drain"
    orc.set_threads(min(os.cpu_count() or 1, 16))
    for k, ((w, prec), s) in enumerate(zip(reqs, specs)):
        assert s.steps_done() == w["steps"]
        u, v, intact = s.get("u"), s.get("v"), s.bonds()[2]
        tag = f"  spec {k} {w['name']} n={w['x'].shape[0]} steps={w['steps']} {prec}"
        if prec == "fp64":
            o = oracle_spec(w)
            configs.setup(o, w)
            o.step(w["steps"])
            _, _, io = o.bonds()
            say(f"{tag}: oracle u equal {np.array_equal(u, o.get('u'))} broken {(io == 0).sum()}")
            assert np.array_equal(u, o.get("u"))
            assert np.array_equal(v, o.get("v"))
            assert np.array_equal(intact, io)
            o.close()
        else:
            ref = make(w, "fp64")
            ref.step(w["steps"])
            solo = make(w, "mixed")
            solo.step(w["steps"])
            ir = ref.bonds()[2]
            e = rel_err(u, ref.get("u"))
            say(f"{tag}: u rel err vs fp64 {e:.2e} broken {(ir == 0).sum()} mismatches {(intact != ir).sum()}")
            assert e <= 1e-5
            assert np.array_equal(intact, ir)
            assert np.array_equal(u, solo.get("u")) and np.array_equal(intact, solo.bonds()[2])
            ref.close()
            solo.close()
        s.close()
