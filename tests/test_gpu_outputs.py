"""Region outputs on the device (pdk_series_*) against the region-output
restatement (oracle/regions.py) and the engine's own fields.

Max and Min series are exact, so they are compared bit for bit; Sum series
reduce in the device's fixed order and are compared to 1e-12 relative.  The
Kalthoff-Winkler crack-extent series, kept on the device from step 1 to 700,
reproduces the reference's regression value (7.42e-2 +- 1.1e-4,
regression_test/tstKalthoffWinkler.hpp:53) and equals the host evaluation of
the same functor at every output step.
"""
import numpy as np
import pytest

from cabanapd_amd import _abi, configs, engine
from oracle import regions
from tests.parity import gpu_spec, oracle_spec, say

pytestmark = pytest.mark.gpu


def test_gpu_series_kalthoff_winkler_crack_extent():
    say("case: KW regression 75x151x7, crack-extent series on the device")
    w = configs.kw_regression()
    g = gpu_spec(w)
    configs.setup(g, w)
    x = w["x"]
    box = _abi.box_region(x.min(axis=0), x.max(axis=0))
    crack = g.series(box, "x", component=1, reducer="max", capacity=32, gate="phi", threshold=0.3)
    energy = g.series(_abi.all_region(), "W", reducer="sum", capacity=32, weight_volume=True)
    assert crack.members() == x.shape[0]
    host = []
    totals = []
    for k in range(w["output_every"], w["steps"] + 1, w["output_every"]):
        g.step(w["output_every"])
        g.compute_energy_async()
        crack.update(k * w["desc"].dt)
        energy.update(k * w["desc"].dt)
        W, _ = g.compute_energy()
        totals.append(W)
        host.append(configs.kw_crack_y(x, g.get("phi")))
    t, v = crack.values()
    _, e = energy.values()
    say(f"  crack y series {v.tolist()}")
    assert v.size == 14 and t.size == 14
    assert np.array_equal(v, np.array(host))
    assert np.allclose(t, np.arange(1, 15) * w["output_every"] * w["desc"].dt, rtol=0, atol=0)
    assert abs(v[-1] - 7.42e-2) <= 1.1e-4
    rel = np.max(np.abs(e - np.array(totals)) / np.abs(np.array(totals)))
    say(f"  energy series vs compute_energy: max rel {rel:.3e}")
    assert rel <= 1e-12
    crack.close()
    energy.close()


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_gpu_series_regions_and_reducers(model):
    say(f"case: {model} cube 15^3 series over box/cylinder/line/ids, sum/max/min")
    w = configs.lattice_sum_cube(model=model, m=3, nc=15, field="linear", coeff=0.0, K=1e9, G=0.5e9)
    x = w["x"]
    rng = np.random.default_rng(11)
    w["u"] = rng.uniform(-1e-4, 1e-4, size=x.shape)
    w["rho"] = np.full(x.shape[0], 1000.0)
    w["desc"].dt = 1e-5
    g = gpu_spec(w)
    configs.setup(g, w)
    g.step(3)
    lo, hi = x.min(axis=0), x.max(axis=0)
    c = 0.5 * (lo + hi)
    span = hi - lo
    dx = [span[0] / 14.0] * 3
    ids = rng.choice(x.shape[0], size=200, replace=False)
    cases = [
        ("box", _abi.box_region(lo + 0.25 * span, c), None),
        ("cylinder", _abi.cylinder_region(0.1 * span[0], 0.4 * span[0], lo[2], c[2], c[0], c[1]), None),
        ("line", _abi.line_region(1, dx), None),
        ("ids", _abi.RegionDesc(), ids),
        ("all", _abi.all_region(), None),
    ]
    u = g.get("u")
    f = g.get("f")
    for name, r, sel in cases:
        mem = regions.members(x, r, ids=sel)
        for field, comp, arr in (("u", 0, u[:, 0]), ("f", 2, f[:, 2]), ("x", 1, x[:, 1])):
            for red in ("sum", "max", "min"):
                s = g.series(r, field, component=comp, reducer=red, capacity=2, ids=sel)
                assert s.members() == mem.size
                s.update(0.0)
                s.update(1.0)
                _, v = s.values()
                ref = regions.reduce(arr[mem], red)
                if red == "sum":
                    ok = abs(v[0] - ref) <= 1e-12 * max(np.sum(np.abs(arr[mem])), 1e-300)
                else:
                    ok = v[0] == ref
                assert ok and v[1] == v[0], (name, field, red, v, ref)
                gi, gv = s.gather()
                assert np.array_equal(gi, mem) and np.array_equal(gv, arr[mem])
                s.close()
        say(f"  {name}: {mem.size} members ok")


def test_gpu_line_profile_equals_oracle():
    say("case: elastic wave 41^3 line profile (gather) vs oracle after 60 steps")
    w = configs.elastic_wave()
    g = gpu_spec(w)
    configs.setup(g, w)
    o = oracle_spec(w)
    configs.setup(o, w)
    g.step(w["steps"])
    o.step(w["steps"])
    d = [w["dx"]] * 3
    prof = g.series(_abi.line_region(0, d), "u", component=0, capacity=0)
    ids, vals = prof.gather()
    mem = regions.members(w["x"], _abi.line_region(0, d))
    assert np.array_equal(ids, mem) and ids.size == 41
    assert np.array_equal(vals, o.get("u")[mem, 0])


def test_gpu_series_errors():
    say("case: series capacity, bad field, detached spec")
    w = configs.lattice_sum_cube(model="PMB", m=3, nc=7, field="linear", coeff=0.0, K=1e9)
    g = gpu_spec(w)
    configs.setup(g, w)
    s = g.series(_abi.all_region(), "u", reducer="max", capacity=1)
    s.update(0.0)
    with pytest.raises(engine.PdkError, match="full"):
        s.update(1.0)
    with pytest.raises(engine.PdkError, match="component"):
        g.series(_abi.all_region(), "W", component=1)
    with pytest.raises(engine.PdkError, match="out of range"):
        g.series(None, "u", ids=[0, w["x"].shape[0]])
    with pytest.raises(engine.PdkError, match="repeated"):
        g.series(None, "u", ids=[3, 3])
    empty = g.series(_abi.box_region([5.0, 5.0, 5.0], [6.0, 6.0, 6.0]), "u", reducer="max", capacity=1)
    assert empty.members() == 0
    empty.update(0.0)
    assert empty.values()[1].tolist() == [regions.reduce([], "max")]  # Kokkos::Max identity
    empty.close()
    s2 = g.series(_abi.all_region(), "u", reducer="sum", capacity=4)
    g.close()
    with pytest.raises(engine.PdkError, match="destroyed"):
        s2.update(0.0)
    with pytest.raises(engine.PdkError, match="destroyed"):
        s.values()  # the device values went with the spec
    s.close()
    s2.close()


def test_gpu_dogbone_full_deck_grip_force():
    """The dogbone tensile deck (examples/mechanics/dogbone_tensile_test.cpp:
    PMB ElasticPerfectlyPlastic, grips at 0.1 m/s, 50,000 steps) with the grip
    force kept as a device series (sum of f_x V over the right grip, the
    example's output_force_x, every 10,000 steps): equal to the oracle's
    curve (tests/golden/dogbone_epp_curve.json, tests/golden/make_dogbone_curve.py)
    to 1e-12 of its size, and elastic-plastic -- the tangent stiffness of the
    last interval is below a quarter of the first's while the engineering
    stress stays below the yield stress."""
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    fx = json.load(open(os.path.join(here, "golden", "dogbone_epp_curve.json")))
    say("case: dogbone deck, 50,000 steps FP64, grip force series")
    w = configs.dogbone()
    assert w["steps"] == 50000
    g = gpu_spec(w)
    configs.setup(g, w)
    ser = g.series(None, "f", component=0, reducer="sum", capacity=8, ids=w["right_grip"], weight_volume=True)
    dt = w["desc"].dt
    assert fx["steps"] == [10000, 20000, 30000, 40000, 50000]
    for k in fx["steps"]:
        g.step(10000)
        ser.update(k * dt)
    _, v = ser.values()
    ref = np.array(fx["grip_force_x"])
    err = np.max(np.abs(v - ref)) / np.max(np.abs(ref))
    say(f"  grip force {v.tolist()} (oracle {ref.tolist()}), rel {err:.2e}")
    assert v.size == 5 and err <= 1e-12
    first = abs(ref[0]) / 10000
    last = abs(ref[4] - ref[3]) / 10000
    stress = abs(ref[4]) / (0.013 * 0.0032)
    say(f"  tangent stiffness last/first {last / first:.3f}, final stress {stress / 1e6:.1f} MPa")
    assert last < 0.25 * first and stress < 30e6
    assert fx["yielded_bonds"] > 0 and fx["broken_bonds"] == 0
    ser.close()
    g.close()
