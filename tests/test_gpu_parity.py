"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

FP64 form: the engine evaluates the reference's expressions in the oracle's
order, so bonds, broken-bond sets, forces, displacements and velocities are
compared bit for bit; energy totals (a reduction in a different order) to
1e-12 relative.
"""
import numpy as np
import pytest

from cabanapd_amd import configs
from tests.parity import bonds_equal, broken_set, ctx, gpu_spec, oracle_spec, pair, rel_err, say

pytestmark = pytest.mark.gpu


def assert_same(a, b, what):
    """Bit-for-bit equality; NaN matches NaN (the reference produces NaN for
    LPS energy of a particle whose bonds are all broken, force/CabanaPD_LPS.hpp:470-487)."""
    a = np.asarray(a)
    b = np.asarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    nd = int(np.sum((a != b) & ~both_nan))
    fin = ~both_nan
    err = rel_err(a[fin], b[fin]) if fin.any() else 0.0
    say(f"  {what}: differing={nd} nan={int(both_nan.sum())} rel_err={err:.3e}")
    assert nd == 0, f"{what}: {nd} values differ, rel err {err:.3e}"


def assert_total(a, b, what, tol=1e-12):
    say(f"  {what}: gpu {a!r} oracle {b!r}")
    if np.isnan(b):
        assert np.isnan(a)
    else:
        assert abs(a - b) <= tol * max(abs(b), 1e-300)


def test_gpu_elastic_cube_small_bitwise():
    say("case: PMB elastic cube 17^3 FP64, 20 steps")
    w = configs.elastic_cube(n_side=17, steps=20)
    g, o = pair(w)
    nb_ok, _ = bonds_equal(g, o)
    assert nb_ok
    assert g.bond_stats() == o.bond_stats()
    assert_same(g.get("f"), o.get("f"), "f at init")
    Wg, _ = g.compute_energy()
    Wo, _ = o.compute_energy()
    assert abs(Wg - Wo) <= 1e-12 * abs(Wo)
    assert_same(g.get("W"), o.get("W"), "W at init")
    g.step(20)
    o.step(20)
    for fld in ("u", "v", "f"):
        assert_same(g.get(fld), o.get(fld), fld + " after 20 steps")


def test_gpu_step_splitting_is_transparent():
    say("case: PMB elastic cube 13^3, 7 steps in one call vs 3+4")
    w = configs.elastic_cube(n_side=13, steps=7)
    g1, _ = pair(w)
    g2, _ = pair(w)
    g1.step(7)
    g2.step(3)
    g2.step(4)
    for fld in ("u", "v", "f"):
        assert np.array_equal(g1.get(fld), g2.get(fld))
    assert g1.steps_done() == 7 and g2.steps_done() == 7


def test_gpu_elastic_cube_config1():
    say("case: PMB elastic cube 41^3 FP64, 500 steps (the CPU-runnable benchmark specimen)")
    w = configs.elastic_cube(n_side=41, steps=500)
    g, o = pair(w)
    total, maxn = g.bond_stats()
    say(f"  bonds {total} max {maxn}")
    assert total == 7720394 and maxn == 122
    g.step(500)
    o.step(500)
    for fld in ("u", "v", "f"):
        assert_same(g.get(fld), o.get(fld), fld + " after 500 steps")
    Wg, _ = g.compute_energy()
    Wo, _ = o.compute_energy()
    say(f"  total strain energy gpu {Wg!r} oracle {Wo!r}")
    assert abs(Wg - Wo) <= 1e-12 * abs(Wo)


@pytest.mark.parametrize("model,field,fracture", [
    ("PMB", "linear", False), ("PMB", "quadratic", False), ("PMB", "quadratic", True),
    ("LPS", "linear", False), ("LPS", "quadratic", False), ("LPS", "quadratic", True)])
def test_gpu_lattice_known_answers(model, field, fracture):
    say(f"case: lattice {model} {field} fracture={fracture}")
    nc = 11 if model == "PMB" else 15
    w = configs.lattice_sum_cube(model=model, m=3, nc=nc, field=field, coeff=0.1 if field == "linear" else 0.01,
                                 influence=1 if model == "LPS" else None, fracture=fracture)
    g, o = pair(w, init=False)
    for s in (g, o):
        s.compute_weighted_volume()
        s.compute_dilatation()
        s.compute_force(True)
    Wg, _ = g.compute_energy()
    Wo, _ = o.compute_energy()
    g.compute_stress()
    o.compute_stress()
    assert_same(g.get("f"), o.get("f"), "f")
    assert_same(g.get("W"), o.get("W"), "W")
    assert_same(g.get("stress"), o.get("stress"), "stress")
    if model == "LPS":
        assert_same(g.get("m"), o.get("m"), "m")
        assert_same(g.get("theta"), o.get("theta"), "theta")
    assert abs(Wg - Wo) <= 1e-12 * abs(Wo)


def test_gpu_kalthoff_winkler_regression():
    say("case: Kalthoff-Winkler regression 75x151x7, 700 steps FP64")
    w = configs.kw_regression()
    g, o = pair(w)
    assert broken_set(g) == broken_set(o)
    series = []
    for k in range(1, w["steps"] + 1, w["output_every"]):
        g.step(w["output_every"])
        o.step(w["output_every"])
        g.compute_energy()
        o.compute_energy()
        series.append(configs.kw_crack_y(w["x"], g.get("phi")))
    assert_same(g.get("phi"), o.get("phi"), "phi")
    assert_same(g.get("u"), o.get("u"), "u")
    assert broken_set(g) == broken_set(o)
    say(f"  max crack y {series[-1]!r}")
    assert abs(series[-1] - 7.42e-2) <= 1.1e-4


def test_gpu_pmb_fracture_plate_small():
    say("case: PMB fracture plate 40x80x9, 150 steps FP64")
    w = configs.kw_plate(num_cells=(40, 80, 9), steps=150, dt=3.0e-7 * 0.25)
    g, o = pair(w)
    g.step(150)
    o.step(150)
    bg, bo = broken_set(g), broken_set(o)
    say(f"  broken bonds gpu {len(bg)} oracle {len(bo)}")
    assert bg == bo
    for fld in ("u", "v"):
        assert_same(g.get(fld), o.get(fld), fld)
    g.compute_stress()
    o.compute_stress()
    assert_same(g.get("stress"), o.get("stress"), "stress (broken bonds contribute mu = 0)")


@pytest.mark.parametrize("fracture", [False, True])
def test_gpu_lps_cube_small(fracture):
    say(f"case: LPS cube 20^3 fracture={fracture}, 40 steps FP64")
    w = configs.lps_cube(n_side=20, steps=40, fracture=fracture)
    g, o = pair(w)
    g.step(40)
    o.step(40)
    for fld in ("u", "v", "m", "theta"):
        assert_same(g.get(fld), o.get(fld), fld)
    if fracture:
        assert broken_set(g) == broken_set(o)
    Wg, Dg = g.compute_energy()
    Wo, Do = o.compute_energy()
    assert_same(g.get("W"), o.get("W"), "W")
    assert_total(Wg, Wo, "total strain energy")
    if fracture:
        assert_same(g.get("phi"), o.get("phi"), "phi")
        assert_total(Dg, Do, "total damage")
    g.compute_stress()
    o.compute_stress()
    assert_same(g.get("stress"), o.get("stress"), "stress")


@pytest.mark.parametrize("model,fracture,jitter", [
    ("PMB", True, False), ("PMB", True, True), ("PMB", False, True),
    ("LPS", True, False), ("LPS", True, True), ("LPS", False, False),
    ("LINEAR_PMB", False, True), ("LINEAR_LPS", False, False)])
def test_gpu_output_passes_on_row_windows(model, fracture, jitter):
    """The energy / damage and stress passes on lattice-row windows (the
    stencil-slot layout of a lattice specimen) against the oracle's
    computeEnergyFull / computeStressFull (force/CabanaPD_PMB.hpp:127-202,
    289-401; force/CabanaPD_LPS.hpp:183-267, 449-572): W, phi and the stress
    tensor bit for bit, totals to 1e-12, after 40 steps of the impact plate
    with bonds breaking; jitter: per-particle volumes scaled by 1 +- 5 %
    (the windows then stage the volumes)."""
    say(f"case: output passes on row windows, {model} fracture={fracture} volume jitter={jitter}")
    from cabanapd_amd._abi import make_desc
    w = configs.kw_plate(num_cells=(24, 48, 6), size=(0.06, 0.12, 0.015), steps=40)
    bd = w["desc"]
    w["desc"] = make_desc(model, fracture=fracture, horizon=bd.horizon, K=bd.K,
                          G=0.6 * bd.K if "LPS" in model else 0.0, G0=bd.G0, dt=bd.dt)
    if not fracture:
        w["notches"] = None
    if jitter:
        w["vol"] = w["vol"] * (1.0 + 0.05 * np.random.default_rng(11).uniform(-1.0, 1.0, w["vol"].shape[0]))
    g, o = pair(w)
    name = g.kernel_name(2)
    say(f"  energy kernel {name}")
    assert name.startswith("kEnergyRows<")
    g.step(40)
    o.step(40)
    if fracture:
        nb = len(broken_set(o))
        say(f"  broken bonds {nb}")
        assert nb > 0 and broken_set(g) == broken_set(o)
    Wg, Dg = g.compute_energy()
    Wo, Do = o.compute_energy()
    assert_same(g.get("W"), o.get("W"), "W")
    assert_total(Wg, Wo, "total strain energy")
    if fracture:
        assert_same(g.get("phi"), o.get("phi"), "phi")
        assert_total(Dg, Do, "total damage")
    g.compute_stress()
    o.compute_stress()
    assert_same(g.get("stress"), o.get("stress"), "stress")


def test_gpu_device_order_is_transparent():
    """Morton order, lattice rows with stored neighbor ids and lattice rows in
    the stencil-slot form give bit-identical FP64 results (each particle sums
    its bonds in caller-id order every way)."""
    say("case: PMB fracture plate 30x60x7, Morton vs lattice compact vs lattice slots, 60 steps")
    from cabanapd_amd import engine
    from tests.parity import ctx
    w = configs.kw_plate(num_cells=(30, 60, 7), size=(0.09, 0.18, 0.021), steps=60)
    out = {}
    for name, order, layout in (("morton", 1, 0), ("compact", 2, 1), ("slots", 2, 0)):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.ordering = order
        d.bond_layout = layout
        s = engine.Spec(ctx(), d, w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        lay = s.layout()
        say(f"  {name}: {lay}")
        assert lay["lattice_rows"] == (order == 2)
        assert lay["stencil_slots"] == (name == "slots")
        s.step(60)
        out[name] = (s.get("u"), s.get("v"), s.bonds()[2])
        s.close()
    for name in ("compact", "slots"):
        for a, b, what in zip(out["morton"], out[name], ("u", "v", "intact")):
            assert_same(a, b, f"{what} ({name} vs morton)")


@pytest.mark.parametrize("unroll", [0, 1, 2, 4, 8])
def test_gpu_unroll_is_transparent(unroll):
    """Every force-pass variant (0: row windows staged in LDS; 1-8: gathers
    in slot groups) is bit-identical to the oracle."""
    say(f"case: PMB fracture plate 24x48x6 FP64, force variant {unroll} vs oracle")
    w = configs.kw_plate(num_cells=(24, 48, 6), size=(0.06, 0.12, 0.015), steps=40)
    g, o = pair(w)
    assert g.layout()["stencil_slots"]
    g.set_tuning(unroll)
    say(f"  kernel {g.kernel_name(0)}")
    g.step(40)
    o.step(40)
    assert_same(g.get("u"), o.get("u"), "u")
    assert broken_set(g) == broken_set(o)


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_gpu_step_graph_matches_eager(model):
    """pdk_step replays captured two-step graphs; between replays the host
    downloads fields (eager kernels and copies on the same stream).  Every
    chunk must equal the same chunk run with every kernel launched eagerly."""
    from cabanapd_amd import engine
    from tests.parity import ctx
    say(f"case: {model} graph-captured steps vs eager launches, 4 chunks of 7 steps")
    if model == "PMB":
        w = configs.kw_plate(num_cells=(24, 48, 8), size=(0.06, 0.12, 0.02), steps=28)
    else:
        w = configs.lps_cube(n_side=16, steps=28, fracture=True)
    specs = []
    for graph in (True, False):
        s = engine.Spec(ctx(), w["desc"], w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        s.set_graph(graph)
        specs.append(s)
    for chunk in range(4):
        for s in specs:
            s.step(7)
        for fld in ("u", "v", "f"):
            assert_same(specs[0].get(fld), specs[1].get(fld), f"{fld} after chunk {chunk}")
        assert np.array_equal(specs[0].bonds()[2], specs[1].bonds()[2])
    for s in specs:
        s.close()


@pytest.mark.parametrize("model", ["LINEAR_PMB", "LINEAR_LPS"])
def test_gpu_linearized_models_bitwise(model):
    """LinearPMB / LinearLPS (force/CabanaPD_PMB.hpp:406-538,
    force/CabanaPD_LPS.hpp:577-724): a cube under a rigid rotation plus a
    seeded perturbation, init and 30 steps, FP64 bit for bit against the
    oracle (bonds, f, u, v, W, m, theta); the mixed form within 1e-5."""
    say(f"case: {model} cube, rotation + perturbation, 30 steps")
    nc = 11 if model == "LINEAR_PMB" else 15
    w = configs.lattice_sum_cube(model=model, m=3, nc=nc, field="linear", coeff=0.0, K=1e9, G=0.5e9,
                                 influence=1 if model == "LINEAR_LPS" else None)
    x = w["x"]
    ang = 0.05
    c, sn = np.cos(ang), np.sin(ang)
    u = np.zeros_like(x)
    u[:, 0] = (c - 1.0) * x[:, 0] - sn * x[:, 1]
    u[:, 1] = sn * x[:, 0] + (c - 1.0) * x[:, 1]
    rng = np.random.default_rng(5)
    u += rng.uniform(-1e-4, 1e-4, size=u.shape)
    w["u"] = u
    w["rho"] = np.full(x.shape[0], 1000.0)
    w["desc"].dt = 1e-5
    g, o = pair(w)
    nb_ok, _ = bonds_equal(g, o)
    assert nb_ok
    assert_same(g.get("f"), o.get("f"), "f at init")
    g.step(30)
    o.step(30)
    for fld in ("u", "v", "f") + (("m", "theta") if model == "LINEAR_LPS" else ()):
        assert_same(g.get(fld), o.get(fld), fld + " after 30 steps")
    Wg, _ = g.compute_energy()
    Wo, _ = o.compute_energy()
    assert_same(g.get("W"), o.get("W"), "W")
    assert_total(Wg, Wo, "total strain energy")
    g.compute_stress()
    o.compute_stress()
    assert_same(g.get("stress"), o.get("stress"), "stress")
    # mixed form of the same run
    from cabanapd_amd import engine
    from tests.parity import ctx
    d = type(w["desc"]).from_buffer_copy(w["desc"])
    d.precision = 1
    mx = engine.Spec(ctx(), d, x.shape[0])
    configs.upload(mx, w)
    configs.setup(mx, w)
    mx.step(30)
    e = rel_err(mx.get("u"), g.get("u"))
    say(f"  mixed vs FP64 u rel err {e:.3e} ({mx.kernel_name(0)})")
    assert e <= 1e-5
    mx.compute_stress()
    es = rel_err(mx.get("stress"), g.get("stress"))
    say(f"  mixed vs FP64 stress rel err {es:.3e}")
    assert es <= 1e-4
    mx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_elastic_wave_regression(precision):
    """regression_test/tstElasticWave.hpp on its own deck (41^3, 60 steps):
    FP64 equals the oracle bit for bit, the mixed form stays within 1e-5 of
    it, and u_x along the x axis meets the reference's polynomial check
    (relative error < 3.5 %, L2 deviation < 3e-5)."""
    import json
    import os
    say(f"case: elastic wave 41^3 60 steps {precision}")
    w = configs.elastic_wave(precision=precision)
    g = gpu_spec(w)
    configs.setup(g, w)
    wo = configs.elastic_wave()
    o = oracle_spec(wo)
    configs.setup(o, wo)
    g.step(w["steps"])
    o.step(wo["steps"])
    ug, uo = g.get("u"), o.get("u")
    if precision == "fp64":
        assert_same(ug, uo, "u")
        assert_same(g.get("theta"), o.get("theta"), "theta")
    else:
        e = rel_err(ug, uo)
        say(f"  u rel err vs FP64 oracle {e:.3e}")
        assert e <= 1e-5
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "elastic_wave_polyfit.json")))
    line = configs.line_profile(w["x"], w["dx"])
    worst, l2, k = configs.polyfit_check(w["x"][line, 0], ug[line, 0], fx)
    say(f"  polyfit: worst rel {worst:.4f} l2 {l2:.3e} points {k}")
    assert k >= 20 and worst < fx["max_rel_err"] and l2 < fx["max_l2_err"]


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_gpu_multi_material_bitwise(model):
    """createMultiForceModel(AverageTag) with two materials scattered at
    random over a cube with fracture: FP64 bonds, u, v, f, W, phi and stress
    equal the oracle's bit for bit after 20 steps."""
    say(f"case: {model} two materials, cube 11^3, 20 steps")
    w = configs.lattice_sum_cube(model=model, m=3, nc=11, field="linear", coeff=0.0, K=1e9, G=0.5e9,
                                 fracture=True, G0=1.0)
    x = w["x"]
    rng = np.random.default_rng(4)
    w["u"] = rng.uniform(-2e-3, 2e-3, size=x.shape) * (2.0 / 11)
    w["rho"] = np.full(x.shape[0], 1000.0)
    w["desc"].dt = 1e-6
    w["materials"] = ([(1e9, 0.5e9, 1.0), (3e9, 1.2e9, 4.0)], rng.integers(0, 2, x.shape[0]).astype(np.int32))
    g, o = pair(w)
    g.step(20)
    o.step(20)
    assert broken_set(g) == broken_set(o)
    for fld in ("u", "v", "f"):
        assert_same(g.get(fld), o.get(fld), fld)
    Wg, Dg = g.compute_energy()
    Wo, Do = o.compute_energy()
    assert_same(g.get("W"), o.get("W"), "W")
    assert_same(g.get("phi"), o.get("phi"), "phi")
    assert_total(Wg, Wo, "total strain energy")
    g.compute_stress()
    o.compute_stress()
    assert_same(g.get("stress"), o.get("stress"), "stress")
    say(f"  broken bonds {len(broken_set(g))}")
    assert len(broken_set(g)) > 0


def test_gpu_crack_inclusion_regression():
    """regression_test/tstCrackInclusion.hpp:55 on its own deck (two
    materials, LPS fracture, 661 steps): the crack-extent series kept on the
    device equals the oracle's at every output step; the last value is
    5.75e-3 +- 5e-5."""
    from cabanapd_amd import _abi
    say("case: crack inclusion 80x40x4, 661 steps FP64")
    w = configs.crack_inclusion()
    g, o = pair(w)
    x = w["x"]
    box = _abi.box_region(x.min(axis=0), x.max(axis=0))
    crack = g.series(box, "x", component=1, reducer="max", capacity=16, gate="phi", threshold=0.3)
    host = []
    for k in range(1, w["steps"] + 1):
        g.step(1)
        o.step(1)
        if k % w["output_every"] == 0:
            g.compute_energy_async()
            crack.update(k * w["desc"].dt)
            o.compute_energy()
            host.append(configs.kw_crack_y(x, o.get("phi")))
    _, v = crack.values()
    say(f"  crack y series {v.tolist()}")
    assert np.array_equal(v, np.array(host))
    assert abs(v[-1] - 5.75e-3) <= 5e-5
    assert_same(g.get("u"), o.get("u"), "u")
    assert broken_set(g) == broken_set(o)
    crack.close()


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_gpu_multi_material_mixed_matches_fp64(model):
    """The mixed form of a two-material specimen (per-pair c, critical
    stretch and guard band in the gather pass): against the FP64 form on the
    same fracturing cube after 20 steps, the same broken bonds and u within
    1e-5 relative."""
    from cabanapd_amd import engine
    say(f"case: {model} two materials, mixed vs FP64, cube 11^3, 20 steps")
    out = {}
    for prec in (0, 1):
        w = configs.lattice_sum_cube(model=model, m=3, nc=11, field="linear", coeff=0.0, K=1e9, G=0.5e9,
                                     fracture=True, G0=1.0)
        x = w["x"]
        rng = np.random.default_rng(4)
        w["u"] = rng.uniform(-2e-3, 2e-3, size=x.shape) * (2.0 / 11)
        w["rho"] = np.full(x.shape[0], 1000.0)
        w["desc"].dt = 1e-6
        w["desc"].precision = prec
        w["materials"] = ([(1e9, 0.5e9, 1.0), (3e9, 1.2e9, 4.0)], rng.integers(0, 2, x.shape[0]).astype(np.int32))
        s = engine.Spec(ctx(), w["desc"], x.shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        s.step(20)
        out[prec] = (s.get("u"), broken_set(s), s.kernel_name(0) if hasattr(s, "kernel_name") else "")
        s.close()
    err = rel_err(out[1][0], out[0][0])
    say(f"  kernel {out[1][2]}: u rel err {err:.3e}, broken {len(out[0][1])}, mismatches "
        f"{len(out[0][1] ^ out[1][1])}")
    assert out[0][1] == out[1][1] and len(out[0][1]) > 0
    assert err <= 1e-5


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_gpu_multi_material_guard_band(model):
    """Two materials with different critical stretches (three pair values:
    each material's and the averaged one): a uniform stretch swept through
    each pair's own critical stretch in steps far below FP32 resolution puts
    that pair's bonds inside its guard band; the mixed form decides every
    bond exactly as the FP64 form."""
    from cabanapd_amd import engine
    from cabanapd_amd._abi import make_desc
    from oracle import binding as orc
    say(f"case: guard band, {model}, two materials")
    nc = 12
    x, vol, d = configs.grid_cell_centred([-1, -1, -1], [1, 1, 1], (nc, nc, nc))
    horizon = 3.01 * d[0]
    mats = [(1e9, 0.5e9, 1e-3), (1e9, 0.5e9, 3e-3)]
    mtype = np.random.default_rng(7).integers(0, 2, x.shape[0]).astype(np.int32)
    kw = dict(fracture=True, horizon=horizon, K=1e9, G0=1e-3, dt=1e-9)
    if model == "LPS":
        kw["G"] = 0.5e9
    import math
    s0s = []
    for (K, G, G0) in mats:
        dd = make_desc(model, **dict(kw, K=K, G=G, G0=G0))
        s0s.append(orc.model_constants(dd)["s0"])
    s0ab = math.sqrt((s0s[0] ** 2 * mats[0][0] + s0s[1] ** 2 * mats[1][0]) / (mats[0][0] + mats[1][0]))
    for crit in (s0s[0], s0ab, s0s[1]):
        for rel in (-1e-7, -3e-9, 0.0, 3e-9, 1e-7):
            res = {}
            for prec in ("fp64", "mixed"):
                desc = make_desc(model, precision=prec, **kw)
                s = engine.Spec(ctx(), desc, x.shape[0])
                s.upload(x, vol, u=(crit * (1 + rel)) * x)
                s.build_bonds()
                s.set_materials(mats, mtype)
                s.compute_weighted_volume()
                s.compute_dilatation()
                s.compute_force(True)
                res[prec] = s.bonds()[2]
                s.close()
            nb = int((res["fp64"] == 0).sum())
            say(f"  crit {crit:.6e} rel {rel:+.0e}: fp64 broken {nb}/{res['fp64'].size} "
                f"mismatches {(res['fp64'] != res['mixed']).sum()}")
            assert np.array_equal(res["fp64"], res["mixed"])


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_gpu_pair_models_bitwise(model):
    """createMultiForceModel(particles, indexing, models...): three material
    types scattered over a fracturing cube.  PMB: a BinaryIndexing-style
    table (one model for like pairs, an independent stiffer, tougher one for
    unlike pairs); LPS: the FullIndexing table of three materials and their
    averages.  Bonds, u, v, f, W, phi and stress equal the oracle's bit for bit
    after 20 steps."""
    from tests.test_oracle_pair_models import pair as pm
    say(f"case: {model} explicit pair table, three types, cube 11^3, 20 steps")
    w = configs.lattice_sum_cube(model=model, m=3, nc=11, field="linear", coeff=0.0, K=1e9, G=0.5e9,
                                 fracture=True, G0=1.0)
    x = w["x"]
    rng = np.random.default_rng(6)
    w["u"] = rng.uniform(-2e-3, 2e-3, size=x.shape) * (2.0 / 11)
    w["rho"] = np.full(x.shape[0], 1000.0)
    w["desc"].dt = 1e-6
    mtype = rng.integers(0, 3, x.shape[0]).astype(np.int32)
    mats = [(1e9, 0.5e9, 1.0), (3e9, 1.2e9, 4.0), (2e9, 0.8e9, 2.0)]
    if model == "PMB":
        from oracle import binding as orc
        k = orc.model_constants(w["desc"])
        like = pm(c=k["c"], bcc=k["bond_break_coeff"])
        unlike = pm(c=1.5 * k["c"], bcc=(1.0 + 2.0 * k["s0"]) ** 2)
        table = [like if a == b else unlike for a in range(3) for b in range(3)]
    else:
        table = averaged_pairs_for(w["desc"], mats)
    w["pair_models"] = (table, mtype)
    g, o = pair(w)
    g.step(20)
    o.step(20)
    _full_compare(g, o)
    say(f"  broken bonds {len(broken_set(g))}")
    assert len(broken_set(g)) > 0


def averaged_pairs_for(desc, mats):
    """The averaged FullIndexing table of mats for this spec's horizon and influence."""
    import math
    from oracle import binding as orc
    from tests.test_oracle_pair_models import pair as pm
    ks = []
    for (K, G, G0) in mats:
        d = type(desc).from_buffer_copy(desc)
        d.K, d.G, d.G0 = K, G, G0
        ks.append((K, G, orc.model_constants(d)))
    out = []
    for a in range(len(mats)):
        for b in range(len(mats)):
            (Ka, Ga, ka), (Kb, Gb, kb) = ks[a], ks[b]
            if a == b:
                c, bcc = ka["c"], ka["bond_break_coeff"]
            else:
                c = (ka["c"] + kb["c"]) / 2.0
                s0 = math.sqrt((ka["s0"] * ka["s0"] * Ka + kb["s0"] * kb["s0"] * Kb) / (Ka + Kb))
                bcc = (1.0 + s0) * (1.0 + s0)
            out.append(pm(c=c, bcc=bcc, tc=(3.0 * Ka - 5.0 * Ga, 3.0 * Kb - 5.0 * Gb), sc=(15.0 * Ga, 15.0 * Gb)))
    return out


def test_gpu_pair_models_rejects_unfoldable_lps():
    """An LPS pair table whose i-term coefficients depend on the partner's
    type cannot be folded into per-particle terms: refused with an error."""
    from cabanapd_amd import engine
    from tests.test_oracle_pair_models import pair as pm
    w = configs.lattice_sum_cube(model="LPS", m=3, nc=7, field="linear", coeff=0.0, K=1e9, G=0.5e9,
                                 fracture=True, G0=1.0)
    s = engine.Spec(ctx(), w["desc"], w["x"].shape[0])
    configs.upload(s, w)
    same = pm(c=0.0, bcc=1.1, tc=(1.0, 1.0), sc=(2.0, 2.0))
    odd = pm(c=0.0, bcc=1.1, tc=(3.0, 1.0), sc=(2.0, 2.0))
    with pytest.raises(engine.PdkError, match="own type"):
        s.set_pair_models([same, odd, odd, same], np.zeros(w["x"].shape[0], dtype=np.int32))
    s.close()


def _full_compare(g, o, fields=("u", "v", "f")):
    assert broken_set(g) == broken_set(o)
    for fld in fields:
        assert_same(g.get(fld), o.get(fld), fld)
    Wg, Dg = g.compute_energy()
    Wo, Do = o.compute_energy()
    assert_same(g.get("W"), o.get("W"), "W")
    assert_same(g.get("phi"), o.get("phi"), "phi")
    assert_total(Wg, Wo, "total strain energy")
    g.compute_stress()
    o.compute_stress()
    assert_same(g.get("stress"), o.get("stress"), "stress")


def _mixed_matches_fp64(w, g, steps):
    """The mixed form of the run in g (FP64, `steps` steps): the same broken
    bonds, the same temperatures (the conduction sums run over the intact
    bonds), u within 1e-5 and W, stress within 1e-4 of the FP64 run."""
    from cabanapd_amd import engine
    from tests.parity import ctx
    d = type(w["desc"]).from_buffer_copy(w["desc"])
    d.precision = 1
    mx = engine.Spec(ctx(), d, w["x"].shape[0])
    configs.upload(mx, w)
    configs.setup(mx, w)
    mx.step(steps)
    diff = broken_set(mx) ^ broken_set(g)
    e = rel_err(mx.get("u"), g.get("u"))
    say(f"  mixed ({mx.kernel_name(0)}): u rel err {e:.3e}, broken-set mismatches {len(diff)}")
    assert len(diff) == 0
    assert e <= 1e-5
    if g.desc.thermal:
        assert rel_err(mx.get("temperature"), g.get("temperature")) <= 1e-12
    mx.compute_energy()
    g.compute_energy()
    ew = rel_err(mx.get("W"), g.get("W"))
    say(f"  mixed W rel err {ew:.3e}")
    # The EPP energy (force_models/CabanaPD_PMB.hpp:123-146) switches branch at
    # s = s_p + s_Y exactly, where every bond that yielded in the last force
    # pass sits: a 1e-7 change of s flips the branch (a jump of s_p^2), so the
    # field is compared for the elastic models only; stress, which goes
    # through the plastic stretch continuously, is compared for all.
    if g.desc.mechanics == 0:
        assert ew <= 1e-4
    mx.compute_stress()
    g.compute_stress()
    assert rel_err(mx.get("stress"), g.get("stress")) <= 1e-4
    mx.close()


def test_gpu_dogbone_epp_bitwise():
    """ElasticPerfectlyPlastic PMB with fracture on the dogbone specimen of
    examples/mechanics/dogbone_tensile_test.cpp, grips pulled at 20 m/s so
    that bonds yield and some break within 300 steps: bonds, u, v, f, W, phi
    and stress (which also updates the plastic stretch) equal the oracle's
    bit for bit; the right grip sits at u_x = v0 t exactly."""
    say("case: dogbone EPP 251x29x5 cut, v0 20 m/s, 300 steps FP64")
    w = configs.dogbone(grip_velocity=20.0, steps=300)
    g, o = pair(w)
    g.step(w["steps"])
    o.step(w["steps"])
    _full_compare(g, o)
    ux = g.get("u")[w["right_grip"], 0]
    assert np.all(ux == 20.0 * (w["steps"] * w["desc"].dt))
    say(f"  broken bonds {len(broken_set(g))}")
    assert len(broken_set(g)) > 0
    _mixed_matches_fp64(w, g, w["steps"])


def test_gpu_thermal_deformation_bitwise():
    """examples/thermomechanics/thermal_deformation.cpp (PMB, NoFracture,
    TemperatureDependent, T = temp0 + 5000 (y - y_low) t imposed before each
    force pass) for 400 steps: temperature, u, v, f, W and stress equal the
    oracle's bit for bit."""
    say("case: thermal deformation 101x31x3, 400 steps FP64")
    w = configs.thermal_deformation(steps=400)
    g, o = pair(w)
    g.step(w["steps"])
    o.step(w["steps"])
    assert_same(g.get("temperature"), o.get("temperature"), "temperature")
    _full_compare(g, o)
    assert np.max(np.abs(g.get("u"))) > 0.0
    _mixed_matches_fp64(w, g, w["steps"])


@pytest.mark.parametrize("model,mech,frac,thermal", [("PMB", "elastic", True, True), ("PMB", "epp", True, True),
                                                     ("LINEAR_PMB", "elastic", False, True),
                                                     ("PMB", "elastic", True, "dynamic"),
                                                     ("PMB", "epp", True, "dynamic")])
def test_gpu_thermal_cube_bitwise(model, mech, frac, thermal):
    """TemperatureDependent PMB / LinearPMB, with the thermal critical
    stretch and EPP, and DynamicTemperature (heat conduction over the intact
    bonds every 4 steps), on a cube with a random initial temperature field
    and a temperature ramp imposed on half the particles: 25 steps bit for
    bit."""
    from cabanapd_amd._abi import make_desc
    say(f"case: thermal={thermal} {model} {mech} fracture={frac}, cube 11^3, 25 steps")
    w = configs.lattice_sum_cube(model="PMB", m=3, nc=11, field="linear", coeff=0.0, K=1e9)
    x = w["x"]
    n = x.shape[0]
    rng = np.random.default_rng(9)
    kw = dict(fracture=True, G0=1.0) if frac else {}
    if mech == "epp":
        kw["mechanics"] = "epp"
        kw["sigma_y"] = 3e6
    if thermal == "dynamic":
        kw.update(kappa=5e4, cp=10.0, thermal_subcycle=4)
    w["desc"] = make_desc(model, horizon=w["desc"].horizon, K=1e9, dt=1e-6, thermal=thermal, alpha=2e-4, temp0=10.0,
                          **kw)
    w["u"] = rng.uniform(-2e-3, 2e-3, size=x.shape) * (2.0 / 11)
    w["rho"] = np.full(n, 1000.0)
    w["temperature"] = rng.uniform(0.0, 40.0, size=n)
    hot = np.flatnonzero(x[:, 2] > 0.0)
    w["prescribed"] = [("temperature", hot, np.full(hot.size, 10.0), np.full(hot.size, 4e5))]
    g, o = pair(w)
    g.step(25)
    o.step(25)
    assert_same(g.get("temperature"), o.get("temperature"), "temperature")
    _full_compare(g, o)
    if frac:
        say(f"  broken bonds {len(broken_set(g))}")
    _mixed_matches_fp64(w, g, 25)


def test_gpu_heat_transfer_bitwise():
    """examples/thermomechanics/thermal_deformation_heat_transfer.cpp's
    specimen (49^3, DynamicTemperature, the y faces held at 0) with heat
    conduction every 20 steps for 100 steps: temperature, u, v, f, W and
    stress equal the oracle's bit for bit."""
    say("case: thermal deformation with heat transfer 49^3, subcycle 20, 100 steps FP64")
    w = configs.thermal_deformation_heat_transfer(steps=100, thermal_subcycle=20)
    g, o = pair(w)
    g.step(w["steps"])
    o.step(w["steps"])
    T = g.get("temperature")
    assert_same(T, o.get("temperature"), "temperature")
    assert np.sum((T != 100.0) & (T != 0.0)) > 0
    _full_compare(g, o)
    _mixed_matches_fp64(w, g, w["steps"])


def test_gpu_plastic_thermal_rejections():
    """LinearPMB with a temperature has the gather variant only; prescribed
    specs stay out of the batch engine; prescribed ids are checked."""
    from cabanapd_amd import engine
    from cabanapd_amd._abi import make_desc
    say("case: EPP / thermal argument checks")
    w = configs.lattice_sum_cube(model="PMB", m=2, nc=7, field="linear", coeff=0.0, K=1e9)
    base = dict(horizon=w["desc"].horizon, K=1e9, dt=1e-6)
    with pytest.raises(engine.PdkError, match="fracture"):
        gpu_spec(w, desc=make_desc("PMB", mechanics="epp", sigma_y=1e6, **base))
    # LinearPMB with a temperature: the gathers only
    w["desc"] = make_desc("LINEAR_PMB", thermal=True, alpha=1e-5, **base)
    g = gpu_spec(w)
    configs.setup(g, w)
    with pytest.raises(engine.PdkError, match="gather"):
        g.set_tuning(0)
    g.close()
    w["desc"] = make_desc("PMB", fracture=True, G0=1.0, mechanics="epp", sigma_y=1e6, **base)
    g = gpu_spec(w)
    configs.setup(g, w)
    with pytest.raises(engine.PdkError, match="repeated"):
        g.set_prescribed("u", [1, 1], np.zeros((2, 3)))
    with pytest.raises(engine.PdkError, match="temperature"):
        g.set_prescribed("temperature", [1], np.zeros(1))
    with pytest.raises(engine.PdkError, match="batch"):
        ctx().submit(g, 1)
    g.close()
    # an elastic spec with prescribed values stays out of the batch engine too
    w["desc"] = make_desc("PMB", **base)
    e = gpu_spec(w)
    configs.setup(e, w)
    e.set_prescribed("u", [0], np.zeros((1, 3)))
    with pytest.raises(engine.PdkError, match="prescribed"):
        ctx().submit(e, 1)
    e.close()


def _feature_plate(case):
    """The 24x48x6 impact plate (a full lattice: row windows apply) with
    random displacements and the features of `case`: mt_pmb / mt_lps (two
    materials scattered at random), th (TemperatureDependent, random
    temperatures), ep (ElasticPerfectlyPlastic), th_ep (both), th_dyn
    (DynamicTemperature: heat conduction every other step), each with
    fracture; th_nf without fracture."""
    from cabanapd_amd._abi import make_desc
    w = configs.kw_plate(num_cells=(24, 48, 6), size=(0.06, 0.12, 0.015), steps=25)
    bd = w["desc"]
    x = w["x"]
    n = x.shape[0]
    rng = np.random.default_rng(21)
    w["u"] = rng.uniform(-1e-6, 1e-6, size=x.shape)
    kw = dict(horizon=bd.horizon, K=bd.K, dt=bd.dt)
    if case.startswith("mt"):
        model = "LPS" if case == "mt_lps" else "PMB"
        w["desc"] = make_desc(model, fracture=True, G=0.6 * bd.K, G0=bd.G0, **kw)
        w["materials"] = ([(bd.K, 0.6 * bd.K, bd.G0), (2.0 * bd.K, 1.2 * bd.K, 3.0 * bd.G0)],
                          rng.integers(0, 2, n).astype(np.int32))
        return w
    if case != "th_nf":
        kw.update(fracture=True, G0=bd.G0)
    else:
        w["notches"] = None
    if "ep" in case:
        kw.update(mechanics="epp", sigma_y=3.0 * bd.K * 1e-3)
    if "th" in case:
        kw.update(thermal="dynamic" if case == "th_dyn" else True, alpha=1e-5, temp0=10.0)
        if case == "th_dyn":
            kw.update(kappa=5e4, cp=10.0, thermal_subcycle=2)
        w["temperature"] = rng.uniform(0.0, 400.0, size=n)
    w["desc"] = make_desc("PMB", **kw)
    return w


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["mt_pmb", "mt_lps", "th", "th_nf", "ep", "th_ep", "th_dyn"])
def test_gpu_feature_row_windows_match_gathers(case, precision):
    """The row-window force pass of specimens with several materials, a
    temperature or a plastic stretch evaluates the gather pass's expressions
    in the gather pass's order: u, v, f and the broken-bond set
    bit-identical after 25 steps, in both precision forms, and the FP64 form
    equal to the oracle (the plastic
    stretch of an EPP bond feeds its force, so u and f cover it)."""
    from cabanapd_amd import engine
    say(f"case: feature row windows vs gathers, {case} {precision}, plate 24x48x6, 25 steps")
    w = _feature_plate(case)
    out = {}
    rows_spec = None
    for variant in (0, 4):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.precision = {"fp64": 0, "mixed": 1}[precision]
        s = engine.Spec(ctx(), d, w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        s.set_tuning(variant)
        name = s.kernel_name(0)
        say(f"  variant {variant}: {name}")
        assert name.startswith("kForceRowsF<") == (variant == 0)
        s.step(25)
        out[variant] = [s.get("u"), s.get("v"), s.get("f"), s.bonds()[2]]
        if "th" in case:
            out[variant].append(s.get("temperature"))
        if variant == 0:
            rows_spec = s
        else:
            s.close()
    for a, b, what in zip(out[0], out[4], ("u", "v", "f", "intact", "temperature")):
        assert_same(a, b, f"{what} (row windows vs gathers)")
    if precision == "fp64":
        o = oracle_spec(w)
        configs.setup(o, w)
        o.step(25)
        flds = ("u", "v", "f") + (("temperature",) if "th" in case else ())
        for a, fld in zip(out[0][:3] + out[0][4:], flds):
            assert_same(a, o.get(fld), f"{fld} (row windows vs oracle)")
        # the output passes on row windows with the same features
        g = rows_spec
        ename = g.kernel_name(2)
        say(f"  energy kernel {ename}")
        assert ename.startswith("kEnergyRows<")
        Wg, Dg = g.compute_energy()
        Wo, Do = o.compute_energy()
        assert_same(g.get("W"), o.get("W"), "W")
        assert_total(Wg, Wo, "total strain energy")
        if case != "th_nf":
            assert_same(g.get("phi"), o.get("phi"), "phi")
        g.compute_stress()
        o.compute_stress()
        assert_same(g.get("stress"), o.get("stress"), "stress")
        # EPP: the stress pass updated the plastic stretch of every bond
        g.step(3)
        o.step(3)
        assert_same(g.get("u"), o.get("u"), "u three steps after the stress pass")
    rows_spec.close()
    if case != "th_nf":
        nb = int(np.sum(out[0][3] == 0))
        say(f"  broken bonds {nb}")
