"""Edge cases of the engine against the oracle: particle counts that do not
fill the last 64-particle slice (1, 2, 63, 65, 130), isolated particles with
no bonds, a lattice whose last slice is partial (row windows), and the empty
specimen, which is refused."""
import numpy as np
import pytest

from cabanapd_amd import configs, engine
from cabanapd_amd._abi import make_desc
from tests.parity import broken_set, ctx, pair, rel_err, say

pytestmark = pytest.mark.gpu


def cloud(n, model, seed, fracture=True):
    """n particles at random in a box sized for about 40 neighbours each
    (horizon 3 dx, s0 = 1.4e-3), a seeded displacement field that stretches
    some bonds past s0, and one particle five horizons from every other (no
    bonds) when n > 2."""
    rng = np.random.default_rng(seed)
    dx = 1e-3
    side = dx * max(n, 1) ** (1.0 / 3.0) * 0.9
    x = rng.uniform(0.0, side, size=(n, 3))
    if n > 2:
        x[n - 1] = side + 5.0 * 3.0 * dx  # isolated
    vol = np.full(n, dx ** 3)
    desc = make_desc(model, fracture=fracture, horizon=3.0 * dx, K=1e9, G=0.5e9 if model == "LPS" else 0.0,
                     G0=10.0, dt=1e-8)
    w = configs._base(desc, x, vol, 1000.0, "cloud_%d" % n, dx, 10)
    w["u"] = rng.uniform(-1e-6, 1e-6, size=(n, 3))
    w["v"] = rng.uniform(-1.0, 1.0, size=(n, 3))
    return w


@pytest.mark.parametrize("model", ["PMB", "LPS"])
@pytest.mark.parametrize("n", [1, 2, 63, 65, 130])
def test_ragged_counts_match_oracle(n, model):
    """FP64: u, v, f and the broken set equal the oracle's after 10 steps; the
    isolated particle feels no force; the mixed form meets the tolerance."""
    say(f"case: {model} cloud of {n} particles, 10 steps")
    w = cloud(n, model, seed=100 + n)
    g, o = pair(w)
    g.step(10)
    o.step(10)
    for fld in ("u", "v", "f"):
        a, b = g.get(fld), o.get(fld)
        assert np.array_equal(a, b), fld
    assert broken_set(g) == broken_set(o)
    if n > 2:
        assert np.all(g.get("f")[n - 1] == 0.0)
    nb = g.bond_stats()[0]
    say(f"  bonds {nb}, broken {len(broken_set(g))}")
    d = type(w["desc"]).from_buffer_copy(w["desc"])
    d.precision = 1
    m = engine.Spec(ctx(), d, n)
    configs.upload(m, w)
    configs.setup(m, w)
    m.step(10)
    err = rel_err(m.get("u"), g.get("u"))
    say(f"  mixed u rel err {err:.3e}, broken equal {broken_set(m) == broken_set(g)}")
    assert err <= 1e-5
    assert broken_set(m) == broken_set(g)
    m.close()
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ragged_lattice_row_windows(precision):
    """A 7x9x5 lattice (315 particles: the last slice holds 59) on the row
    windows and on the gathers: identical results, and FP64 equal to the
    oracle."""
    say(f"case: lattice 7x9x5 {precision}, row windows vs gathers vs oracle, 30 steps")
    w = configs.kw_plate(num_cells=(7, 9, 5), size=(0.0035, 0.0045, 0.0025), steps=30)
    out = []
    for variant in (0, 2):
        d = type(w["desc"]).from_buffer_copy(w["desc"])
        d.precision = {"fp64": 0, "mixed": 1}[precision]
        s = engine.Spec(ctx(), d, w["x"].shape[0])
        configs.upload(s, w)
        configs.setup(s, w)
        s.set_tuning(variant)
        say(f"  variant {variant}: {s.kernel_name(0)}")
        assert s.kernel_name(0).startswith("kForceRows<") == (variant == 0)
        s.step(30)
        out.append((s.get("u"), s.get("v"), s.bonds()[2]))
        s.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    if precision == "fp64":
        _, o = pair(w)
        o.step(30)
        assert np.array_equal(out[0][0], o.get("u"))
        assert np.array_equal(out[0][1], o.get("v"))


def test_empty_specimen_is_refused():
    say("case: a specimen with no particles")
    w = configs.elastic_cube(n_side=3, steps=1)
    with pytest.raises(engine.PdkError, match="at least one particle"):
        engine.Spec(ctx(), w["desc"], 0)
