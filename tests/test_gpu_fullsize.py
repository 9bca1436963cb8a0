"""The fracture benchmark plate at full size: 200 x 400 x 18 cells (1,440,000
particles, 163,059,312 bonds before the pre-notches), two pre-notches, edge
impact at 16.5 m/s, dt 7.5e-8 (0.85 of the critical step).

* plain FP64 against the oracle, bit for bit (u, v, f, broken-bond set,
  damage), after 50 steps;
* the mixed form against plain FP64 over the full 1000 steps: relative
  displacement error <= 1e-5 and an identical broken-bond set."""
import os

import numpy as np
import pytest

from cabanapd_amd import configs, engine
from tests.parity import ctx, oracle_spec, rel_err, say

pytestmark = pytest.mark.gpu

U_TOL = 1e-5


def plate(steps):
    return configs.kw_plate(num_cells=(200, 400, 18), steps=steps, dt=7.5e-8)


def make(w, precision):
    d = type(w["desc"]).from_buffer_copy(w["desc"])
    d.precision = {"fp64": 0, "mixed": 1}[precision]
    s = engine.Spec(ctx(), d, w["x"].shape[0])
    configs.upload(s, w)
    configs.setup(s, w)
    return s


def test_fp64_full_plate_equals_oracle():
    from oracle import binding as orc
    steps = 50
    say(f"case: full plate 200x400x18 fp64 vs oracle, {steps} steps")
    w = plate(steps)
    g = make(w, "fp64")
    g.step(steps)
    orc.set_threads(min(os.cpu_count() or 1, 16))
    o = oracle_spec(w)
    configs.setup(o, w)
    o.step(steps)
    for fld in ("u", "v", "f"):
        a, b = g.get(fld), o.get(fld)
        nd = int(np.sum(a != b))
        say(f"  {fld}: differing entries {nd}")
        assert nd == 0
    go, gn, gi = g.bonds()
    oo, on, oi = o.bonds()
    say(f"  bonds {gn.size} (oracle {on.size}), broken {(gi == 0).sum()} (oracle {(oi == 0).sum()})")
    assert np.array_equal(go, oo) and np.array_equal(gn, on) and np.array_equal(gi, oi)
    Wg, Dg = g.compute_energy()
    Wo, Do = o.compute_energy()
    assert np.array_equal(g.get("phi"), o.get("phi"))
    assert np.array_equal(g.get("W"), o.get("W"))
    say(f"  energy {Wg:.12e} / {Wo:.12e}, damage {Dg:.6f} / {Do:.6f}")
    assert abs(Wg - Wo) <= 1e-12 * abs(Wo) and abs(Dg - Do) <= 1e-12 * abs(Do)
    o.close()
    g.close()


def test_mixed_full_plate_1000_steps():
    steps = 1000
    say(f"case: full plate 200x400x18 mixed vs fp64, {steps} steps")
    w = plate(steps)
    a = make(w, "fp64")
    b = make(w, "mixed")
    a.step(steps)
    b.step(steps)
    ia, ib = a.bonds()[2], b.bonds()[2]
    e = rel_err(b.get("u"), a.get("u"))
    ev = rel_err(b.get("v"), a.get("v"))
    say(f"  broken fp64 {(ia == 0).sum()} mixed {(ib == 0).sum()} mismatches {(ia != ib).sum()}; "
        f"u rel err {e:.3e}, v rel err {ev:.3e}")
    assert np.array_equal(ia, ib)
    assert e <= U_TOL
    a.close()
    b.close()
