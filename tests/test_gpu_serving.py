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


@pytest.mark.parametrize("graph_steps", [64, 5, 0])
def test_alone_equals_mix(graph_steps):
    """graph_steps: the packed step replayed as a captured graph for up to that
    many steps at a time (0: eager launches)."""
    say(f"case: six specimens alone vs packed batch with staggered admission, graph steps {graph_steps}")
    ctx().set_graph(graph_steps)
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
    ctx().set_graph(0)


def test_graph_replays_across_eager_work():
    """Graph replays of the packed step, interleaved with the program's own
    eager work and synchronisations: two specimens run in the batch engine in
    chunks of at most 11 steps (replays of the 8-step and the 1-step graph),
    each finished run checked against the same number of solo steps, then one
    specimen continued alone through two more progress calls and checked
    again."""
    say("case: graph replays across submissions and progress chunks")
    c = ctx()
    c.set_graph(11)
    wa = configs.kw_plate(num_cells=(40, 80, 9), steps=41, name="ga")
    wb = configs.lps_cube(n_side=14, steps=13, fracture=True)

    def solo(w, prec, steps):
        s = make(w, prec)
        s.step(steps)
        out = (s.get("u"), s.get("v"), s.bonds()[2])
        s.close()
        return out

    def wait(*tickets):
        deadline = time.time() + 60.0
        while not all(c.poll(t) for t in tickets) and time.time() < deadline:
            time.sleep(1e-3)
        assert all(c.poll(t) for t in tickets)

    ref_a20, ref_a41, ref_b13 = solo(wa, "mixed", 20), solo(wa, "mixed", 41), solo(wb, "mixed", 13)
    a, b = make(wa, "mixed"), make(wb, "mixed")
    ta, tb = c.submit(a, 20), c.submit(b, 13)
    c.progress(9)
    c.progress(0)
    wait(ta, tb)
    for name, s, ref in (("a@20", a, ref_a20), ("b@13", b, ref_b13)):
        ok = [np.array_equal(s.get("u"), ref[0]), np.array_equal(s.get("v"), ref[1]),
              np.array_equal(s.bonds()[2], ref[2])]
        say(f"  {name}: u/v/bonds equal {ok}")
        assert all(ok)
    ta = c.submit(a, 21)
    c.progress(2)
    c.progress(0)
    wait(ta)
    ok = [np.array_equal(a.get("u"), ref_a41[0]), np.array_equal(a.get("v"), ref_a41[1]),
          np.array_equal(a.bonds()[2], ref_a41[2])]
    say(f"  a@41: u/v/bonds equal {ok}")
    assert all(ok)
    a.close()
    b.close()
    c.set_graph(0)


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
