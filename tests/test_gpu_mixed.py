"""Mixed-precision form against the FP64 form (both on the GPU; the FP64 form
is itself bit-identical to the oracle, see test_gpu_parity.py).

Stated tolerances: relative displacement error <= 1e-5 (norm-wise,
max|u_mixed - u_fp64| / max|u_fp64|); broken-bond sets identical.
"""
import numpy as np
import pytest

from cabanapd_amd import configs, engine
from cabanapd_amd._abi import make_desc
from tests.parity import ctx, pair, rel_err, say

pytestmark = pytest.mark.gpu

U_TOL = 1e-5


def run(w, precision, steps):
    d = type(w["desc"]).from_buffer_copy(w["desc"])
    d.precision = {"fp64": 0, "mixed": 1}[precision]
    s = engine.Spec(ctx(), d, w["x"].shape[0])
    configs.upload(s, w)
    configs.setup(s, w)
    if steps:
        s.step(steps)
    return s


def test_mixed_elastic_cube():
    say("case: PMB elastic cube 41^3 mixed vs fp64, 500 steps")
    w = configs.elastic_cube(n_side=41, steps=500)
    a = run(w, "fp64", 500)
    b = run(w, "mixed", 500)
    e = rel_err(b.get("u"), a.get("u"))
    Wa, _ = a.compute_energy()
    Wb, _ = b.compute_energy()
    say(f"  u rel err {e:.3e}  energy rel err {abs(Wb - Wa) / abs(Wa):.3e}")
    assert e <= U_TOL
    assert abs(Wb - Wa) <= U_TOL * abs(Wa)


def test_mixed_fracture_plate():
    say("case: PMB fracture plate 60x120x9 mixed vs fp64, 300 steps")
    w = configs.kw_plate(num_cells=(60, 120, 9), steps=300)
    a = run(w, "fp64", 300)
    b = run(w, "mixed", 300)
    _, _, ia = a.bonds()
    _, _, ib = b.bonds()
    e = rel_err(b.get("u"), a.get("u"))
    say(f"  broken fp64 {(ia == 0).sum()} mixed {(ib == 0).sum()} mismatches {(ia != ib).sum()} u rel err {e:.3e}")
    assert np.array_equal(ia, ib)
    assert e <= U_TOL


def test_mixed_lps_cube_elastic():
    say("case: LPS cube 24^3 elastic mixed vs fp64, 60 steps")
    w = configs.lps_cube(n_side=24, steps=60, fracture=False)
    a = run(w, "fp64", 60)
    b = run(w, "mixed", 60)
    e = rel_err(b.get("u"), a.get("u"))
    say(f"  u rel err {e:.3e}")
    assert e <= U_TOL


def _lps_fracture_cube():
    # the LPS fracture benchmark cube at full size: 128^3 particles, s0 = 4.5e-5
    # under a random strain field of ~1e-3 (98.6 % of the bonds break)
    return configs.lps_cube(n_side=128, steps=100, fracture=True)


def test_mixed_lps_cube_fracture_stepwise():
    """From one state, one mixed step makes exactly the FP64 break decisions.
    The FP64 run is restarted from its own state at steps 0, 1, 10 and 100
    (pdk_spec_copy_state) in both forms; after one step each, the broken-bond
    sets are identical (248M bonds, ~1.6M of them breaking in that step at
    step 1) and the velocities agree within 1e-5.  The FP64 restart equals the
    uninterrupted FP64 run bit for bit."""
    say("case: LPS cube 128^3 fracture, one mixed vs one FP64 step from the FP64 state")
    w = _lps_fracture_cube()
    a = run(w, "fp64", 0)
    done = 0
    for k in (0, 1, 10, 100):
        if k > done:
            a.step(k - done)
            done = k
        ia = a.bonds()[2]
        forks = {}
        for prec in ("fp64", "mixed"):
            d = type(w["desc"]).from_buffer_copy(w["desc"])
            d.precision = {"fp64": 0, "mixed": 1}[prec]
            s = engine.Spec(ctx(), d, w["x"].shape[0])
            configs.upload(s, w)
            configs.setup(s, w, init=False)
            s.copy_state_from(a)
            s.step(1)
            forks[prec] = s
        i64, imx = forks["fp64"].bonds()[2], forks["mixed"].bonds()[2]
        ev = rel_err(forks["mixed"].get("v"), forks["fp64"].get("v"))
        new = int(((i64 == 0) & (ia == 1)).sum())
        say(f"  from step {k}: bonds breaking in the step {new}, mismatches {(i64 != imx).sum()}, v rel err {ev:.2e}")
        assert np.array_equal(i64, imx)
        assert ev <= U_TOL
        if k == 1:
            ref = run(w, "fp64", 2)
            assert np.array_equal(ref.get("u"), forks["fp64"].get("u"))
            assert np.array_equal(ref.get("v"), forks["fp64"].get("v"))
            assert np.array_equal(ref.bonds()[2], i64)
            ref.close()
        for s in forks.values():
            s.close()
    a.close()


def test_mixed_lps_cube_fracture_within_fp32_input_floor():
    """Over many steps this specimen amplifies any state difference: a bond
    sitting within the difference of s0 flips a step earlier or later, and the
    flip moves its neighbourhood.  The floor is what FP64 itself does when its
    initial displacement is rounded to FP32 once (a relative change <= 6e-8):
    after 100 steps that run differs from the exact-input FP64 run by ~0.08 in
    u and in ~630k bond decisions.  The mixed form (same exact input) must stay
    within twice that floor in both measures; measured at step 10 / 100:
    mixed 0.0029 / 0.060 with 8 / 458,674 flipped bonds, floor 0.025 / 0.077
    with 88 / 629,578."""
    say("case: LPS cube 128^3 fracture, mixed vs FP64 against the FP32-input floor")
    w = _lps_fracture_cube()
    a = run(w, "fp64", 0)
    b = run(w, "mixed", 0)
    w32 = dict(w)
    w32["u"] = w["u"].astype(np.float32).astype(np.float64)
    r = run(w32, "fp64", 0)
    done = 0
    for k in (10, 100):
        for s in (a, b, r):
            s.step(k - done)
        done = k
        ia, ib, ir = a.bonds()[2], b.bonds()[2], r.bonds()[2]
        ua = a.get("u")
        em, ef = rel_err(b.get("u"), ua), rel_err(r.get("u"), ua)
        mm, mf = int((ia != ib).sum()), int((ia != ir).sum())
        say(f"  step {k}: mixed u err {em:.3e} flips {mm}; FP32-input floor u err {ef:.3e} flips {mf}")
        assert em <= 2.0 * ef + U_TOL
        assert mm <= 2 * mf + 10
    for s in (a, b, r):
        s.close()


@pytest.mark.parametrize("model", ["PMB", "LPS", "PMB_THERMAL", "PMB_EPP", "PMB_EPP_THERMAL"])
def test_guard_band_break_decisions(model):
    """Every bond of a uniformly stretched lattice sits at the same stretch;
    sweeping that stretch through the critical stretch in steps far below
    FP32 resolution puts all decisions inside the guard band.  The mixed
    kernel must decide each bond exactly as the FP64 expression does for the
    same state.  Thermal cases hold a uniform temperature T, which moves the
    threshold to s0 + alpha (T - temp0) (the band grows by the thermal term);
    EPP cases use the plastic model's own s0 (its break test is the elastic
    one, on the total stretch)."""
    say(f"case: guard band, {model}")
    nc = 12
    x, vol, d = configs.grid_cell_centred([-1, -1, -1], [1, 1, 1], (nc, nc, nc))
    horizon = 3.01 * d[0]
    base = model.split("_")[0]
    thermal = "THERMAL" in model
    kw = dict(fracture=True, horizon=horizon, K=1e9, G0=1e-3, dt=1e-9)
    if base == "LPS":
        kw["G"] = 0.5e9
    if "EPP" in model:
        kw.update(mechanics="epp", sigma_y=2e5)
    temp0, T = 300.0, 340.0
    if thermal:
        kw.update(thermal=True, alpha=1.7e-6, temp0=temp0)
    for rel in (-1e-7, -3e-9, -1e-12, 0.0, 1e-12, 3e-9, 1e-7):
        res = {}
        for prec in ("fp64", "mixed"):
            desc = make_desc(base, precision=prec, **kw)
            s0 = engine.model_constants(desc)["s0"]
            crit = s0 + (kw["alpha"] * (T - temp0) if thermal else 0.0)
            s = engine.Spec(ctx(), desc, x.shape[0])
            s.upload(x, vol, u=(crit * (1 + rel)) * x)
            s.build_bonds()
            if thermal:
                s.set_field("temperature", np.full(x.shape[0], T))
            s.compute_weighted_volume()
            s.compute_dilatation()
            s.compute_force(True)
            res[prec] = s.bonds()[2]
            s.close()
        nb = int((res["fp64"] == 0).sum())
        say(f"  rel {rel:+.0e}: fp64 broken {nb}/{res['fp64'].size} mismatches {(res['fp64'] != res['mixed']).sum()}")
        assert np.array_equal(res["fp64"], res["mixed"])
        if rel in (-1e-7, 1e-7):
            # far outside the band the decision is unanimous either way
            assert nb == (res["fp64"].size if rel > 0 else 0)


def test_mixed_stencil_matches_gathered_bond_vectors():
    """Mixed form with the bond vectors from the lattice stencil (stencil-slot
    form, and compact form with stencil ids) vs the same form gathering x_j
    (Morton order): same broken bonds, displacements within rounding."""
    say("case: PMB fracture plate 40x80x9 (cubic cells) mixed, stencil vs gathered xi")
    w = configs.kw_plate(num_cells=(40, 80, 9), size=(0.1, 0.2, 0.0225), steps=150)
    res = {}
    for name, order, layout in (("morton", 1, 0), ("lattice-compact", 2, 1), ("lattice-slots", 2, 0)):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.precision = 1
        d.ordering = order
        d.bond_layout = layout
        s = engine.Spec(ctx(), d, w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        lay = s.layout()
        say(f"  {name}: {lay}")
        if order == 2:
            assert lay["stencil_entries"] == 122
            assert lay["stencil_slots"] == (layout == 0)
        s.step(150)
        res[name] = (s.get("u"), s.bonds()[2])
        s.close()
    for name in ("lattice-compact", "lattice-slots"):
        e = rel_err(res[name][0], res["morton"][0])
        say(f"  {name} vs morton: u rel err {e:.3e} mismatched bonds {(res[name][1] != res['morton'][1]).sum()}")
        assert np.array_equal(res[name][1], res["morton"][1])
        assert e <= 1e-6


@pytest.mark.parametrize("layout", ["slots", "compact"])
def test_mixed_jittered_lattice_keeps_break_decisions(layout):
    """A lattice whose x coordinates sit up to 1e-8 spacings off their sites
    (the largest jitter the lattice test accepts across 40 sites) still takes
    the stencil, whose bond vectors then differ from x_j - x_i; the guard
    band grows with the measured deviation, and the mixed form's broken bonds
    equal the FP64 form's."""
    say(f"case: PMB fracture plate 40x80x9 x jittered 1e-8 h, mixed {layout} vs FP64, 150 steps")
    w = configs.kw_plate(num_cells=(40, 80, 9), size=(0.1, 0.2, 0.0225), steps=150)
    dx = 0.1 / 40
    rng = np.random.default_rng(7)
    w["x"] = w["x"].copy()
    w["x"][:, 0] += rng.uniform(-1e-8 * dx, 1e-8 * dx, size=w["x"].shape[0])
    res = {}
    for prec in (0, 1):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.precision = prec
        d.ordering = 2
        d.bond_layout = {"slots": 0, "compact": 1}[layout]
        s = engine.Spec(ctx(), d, w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        lay = s.layout()
        say(f"  precision {prec}: {lay}")
        if prec == 1:
            assert lay["stencil_entries"] == 122
            assert lay["stencil_slots"] == (layout == "slots")
        s.step(150)
        res[prec] = (s.get("u"), s.bonds()[2])
        s.close()
    mism = int((res[0][1] != res[1][1]).sum())
    e = rel_err(res[1][0], res[0][0])
    say(f"  broken {int((res[0][1] == 0).sum())}, mismatched {mism}, u rel err {e:.3e}")
    assert mism == 0
    assert e <= 1e-5


def _holed_plate(nc=(30, 60, 7), steps=80):
    """The pre-notched plate (cubic cells of 2.5 mm) with a cylindrical hole
    of radius 6 cells through the thickness: a lattice with missing sites,
    which takes the compact form (with stencil ids in the mixed form)."""
    dx = 0.0025
    w = configs.kw_plate(num_cells=nc, size=tuple(c * dx for c in nc), steps=steps)
    keep = (w["x"][:, 0] - 0.01) ** 2 + w["x"][:, 1] ** 2 > (6 * dx) ** 2
    remap = np.cumsum(keep) - 1
    for k in ("x", "vol", "rho", "u", "v"):
        w[k] = np.ascontiguousarray(w[k][keep])
    kind, idx, val = w["bc"]
    w["bc"] = (kind, remap[idx[keep[idx]]].astype(np.int64), val)
    return w


def test_holed_lattice_compact_stencil():
    """A lattice with missing sites: FP64 bit for bit against the oracle; the
    mixed form (compact bonds, bond vectors from the stencil by stencil id)
    with FP64's broken bonds and u within 1e-5, and within 1e-6 of the mixed
    form gathering x_j (Morton order)."""
    say("case: PMB plate 30x60x7 with a hole of radius 6 cells, FP64 vs oracle, mixed vs FP64, 80 steps")
    w = _holed_plate()
    g, o = pair(w)
    say(f"  particles {w['x'].shape[0]}, FP64 layout {g.layout()}")
    assert g.layout()["lattice_rows"] and not g.layout()["stencil_slots"]
    g.step(80)
    o.step(80)
    for fld in ("u", "v", "f"):
        assert np.array_equal(g.get(fld), o.get(fld)), fld
    ref_u, ref_b = g.get("u"), g.bonds()[2]
    g.close()
    res = {}
    for name, order in (("lattice", 2), ("morton", 1)):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.precision = 1
        d.ordering = order
        s = engine.Spec(ctx(), d, w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        lay = s.layout()
        say(f"  mixed {name}: {lay} {s.kernel_name(0)}")
        if order == 2:
            assert lay["stencil_entries"] == 122 and not lay["stencil_slots"]
        s.step(80)
        res[name] = (s.get("u"), s.bonds()[2])
        s.close()
    e = rel_err(res["lattice"][0], ref_u)
    e2 = rel_err(res["lattice"][0], res["morton"][0])
    mism = int((res["lattice"][1] != ref_b).sum())
    say(f"  broken {int((ref_b == 0).sum())}, mixed vs FP64: u rel err {e:.3e} mismatched {mism}; "
        f"stencil vs x_j gathers: u rel err {e2:.3e}")
    assert mism == 0
    assert np.array_equal(res["morton"][1], ref_b)
    assert e <= 1e-5
    assert e2 <= 1e-6


def test_holed_lps_cube_compact_stencil():
    """LPS (weighted volume and dilatation from the stencil record by stencil
    id in the mixed form) on a 20^3 cube with a spherical cavity of radius 4
    spacings: FP64 bit for bit against the oracle (u, v, f, m, theta); the
    mixed form within 1e-5 of FP64 and 1e-6 of the mixed form gathering x_j."""
    say("case: LPS elastic cube 20^3 with a cavity, FP64 vs oracle, mixed vs FP64, 40 steps")
    w = configs.lps_cube(n_side=20, steps=40, fracture=False)
    dx = 1.0 / 19
    keep = np.sum((w["x"] - 0.5) ** 2, axis=1) > (4 * dx) ** 2
    for k in ("x", "vol", "rho", "u", "v", "body"):
        w[k] = np.ascontiguousarray(w[k][keep])
    g, o = pair(w)
    say(f"  particles {w['x'].shape[0]}, FP64 layout {g.layout()}")
    assert g.layout()["lattice_rows"] and not g.layout()["stencil_slots"]
    g.step(40)
    o.step(40)
    for fld in ("u", "v", "f", "m", "theta"):
        assert np.array_equal(g.get(fld), o.get(fld)), fld
    ref_u = g.get("u")
    g.close()
    res = {}
    for name, order in (("lattice", 2), ("morton", 1)):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.precision = 1
        d.ordering = order
        s = engine.Spec(ctx(), d, w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        lay = s.layout()
        say(f"  mixed {name}: {lay} {s.kernel_name(0)}")
        if order == 2:
            assert lay["stencil_entries"] > 0 and not lay["stencil_slots"]
        s.step(40)
        res[name] = s.get("u")
        s.close()
    e = rel_err(res["lattice"], ref_u)
    e2 = rel_err(res["lattice"], res["morton"])
    say(f"  mixed vs FP64: u rel err {e:.3e}; stencil vs x_j gathers: u rel err {e2:.3e}")
    assert e <= 1e-5
    assert e2 <= 1e-6


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_mixed_row_windows_match_gathers(model):
    """The mixed force pass with neighbor data staged in LDS row windows and
    the one gathering it per slot evaluate the same expressions in the same
    order: bit-identical results."""
    if model == "PMB":
        say("case: PMB fracture plate 40x80x9 mixed, row windows vs gathers, 100 steps")
        w = configs.kw_plate(num_cells=(40, 80, 9), size=(0.1, 0.2, 0.0225), steps=100)
    else:
        say("case: LPS cube 20^3 fracture mixed, row windows vs gathers, 40 steps")
        w = configs.lps_cube(n_side=20, steps=40, fracture=True)
    out = {}
    for variant in (0, 2):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.precision = 1
        s = engine.Spec(ctx(), d, w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        s.set_tuning(variant)
        say(f"  variant {variant}: {s.kernel_name(0)} {s.layout()}")
        s.step(w["steps"])
        out[variant] = (s.get("u"), s.get("v"), s.bonds()[2])
        s.close()
    for a, b in zip(out[0], out[2]):
        assert np.array_equal(a, b)
