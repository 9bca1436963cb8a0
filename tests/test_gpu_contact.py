"""DEM contact step on the GPU: the reference's two-sphere regression goldens
(regression_test/tstHertz.hpp:135, tstHertzJKR.hpp:145) and parity with the
oracle on a polydisperse packing.

The contact models' powers (x^1.5, x^3, x^0.5 and the JKR cube roots) are
evaluated correctly rounded by the same IEEE-operation routines in the
engine and the oracle, so forces and trajectories are compared bit for bit;
only Coulomb friction (tanh from the device math library vs libm) is
compared to a tolerance.
"""
import math

import numpy as np
import pytest

from cabanapd_amd import configs, engine
from cabanapd_amd._abi import make_desc
from tests.parity import ctx, oracle_spec, rel_err, say

pytestmark = pytest.mark.gpu


def two_spheres(model, pos, rho, vol, radius, extend, nu, E, e, gamma, dt):
    x = np.array([[pos, 0.0, 0.0], [-pos, 0.0, 0.0]])
    v = np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    desc = make_desc(model, radius=radius, radius_extend=extend, nu=nu, E=E, restitution=e, gamma=gamma,
                     mu_friction=0.0, dt=dt)
    s = engine.Spec(ctx(), desc, 2)
    s.upload(x, np.full(2, vol), v=v, rho=np.full(2, rho))
    s.build_bonds()
    s.init()
    return s, desc


def test_gpu_hertz_restitution():
    say("case: Hertz two spheres e=0.5")
    rho, vol, E, nu, e = 7.95e3, 5.236e-13, 195.6e9, 0.25, 0.5
    dt, tf = 0.8e-8, 1e-5
    s, _ = two_spheres("HERTZ", 5.1e-5, rho, vol, 5e-5, 1e-8, nu, E, e, 0.0, dt)
    ke_i = 2 * 0.5 * rho * vol * 1.0
    s.step(int(tf / dt))
    v = s.get("v")
    ke_f = float(np.sum(0.5 * rho * vol * np.sum(v * v, axis=1)))
    say(f"  sqrt(KE_f/KE_i) = {math.sqrt(ke_f / ke_i)!r}")
    assert abs(math.sqrt(ke_f / ke_i) - e) <= 1e-3


def test_gpu_jkr_pulloff():
    say("case: Hertz-JKR two spheres pull-off")
    rho, vol, E, nu, e, gamma = 7.95e3, 5.236e-13, 195.6e9, 0.25, 1.0, 10.0
    dt, tf = 1e-9, 5e-7
    s, desc = two_spheres("HERTZ_JKR", 5.01e-5, rho, vol, 5e-5, 1e-8, nu, E, e, gamma, dt)
    fx = []
    for _ in range(int(tf / dt)):
        s.step(1)
        fx.append(s.get("f")[0, 0])
    fc = engine.model_constants(desc)["fc"]
    say(f"  min pull-off / fc = {min(fx) * vol / fc!r}")
    assert abs(min(fx) * vol / fc + 1.0) <= 5e-3


def compressed_packing(model, poly, speed):
    # lattice spacing below 2 R: most face neighbours overlap, so every
    # branch of the normal force (overlap, JKR tail, approach/recede) is hit
    w = configs.dem_packing(n=3000, seed=7, model=model, polydisperse=poly, dt=1e-8, steps=20, rebuild_every=10,
                            speed=speed)
    w["x"] = w["x"] * 0.93
    return w


@pytest.mark.parametrize("model,poly,mu", [("HERTZ", False, 0.0), ("HERTZ_JKR", False, 0.0),
                                           ("HERTZ_JKR", True, 0.0), ("HERTZ_JKR", True, 0.3)])
def test_gpu_contact_force_parity(model, poly, mu):
    """One contact pass on overlapping spheres: forces within 1e-12 of the
    oracle (the only differences are libm vs device pow/log/tanh ulps)."""
    say(f"case: {model} force poly={poly} friction={mu} vs oracle")
    w = compressed_packing(model, poly, speed=1.0)
    w["desc"].mu_friction = mu
    g = engine.Spec(ctx(), w["desc"], w["x"].shape[0])
    configs.upload(g, w)
    configs.setup(g, w)
    o = oracle_spec(w)
    configs.setup(o, w)
    go, gn, _ = g.bonds()
    oo, on, _ = o.bonds()
    assert np.array_equal(go, oo) and np.array_equal(gn, on)
    ef = rel_err(g.get("f"), o.get("f"))
    nd = int(np.sum(g.get("f") != o.get("f")))
    say(f"  contacts {gn.size} f rel err {ef:.3e} differing {nd}")
    assert gn.size > 1000
    if mu == 0.0:
        assert nd == 0
    else:
        # friction: tanh of the device math library vs libm
        assert ef <= 1e-12


@pytest.mark.parametrize("model,poly,e", [("HERTZ", False, 1.0), ("HERTZ", True, 1.0), ("HERTZ", True, 0.5),
                                          ("HERTZ_JKR", False, 0.5), ("HERTZ_JKR", True, 0.5)])
def test_gpu_contact_trajectory_parity(model, poly, e):
    """20 steps with two list rebuilds, damped (restitution 0.5) and undamped,
    Hertz and Hertz-JKR: u, v, f and the contact lists equal the oracle's bit
    for bit at every step (the damping term ~ |overlap|^(1/4) at contact onset
    and JKR's approach/recede switch would turn any last-ulp difference of
    the powers into a finite force jump)."""
    say(f"case: {model} trajectory poly={poly} restitution={e} vs oracle")
    w = compressed_packing(model, poly, speed=0.5)
    w["desc"].restitution = e
    g = engine.Spec(ctx(), w["desc"], w["x"].shape[0])
    configs.upload(g, w)
    configs.setup(g, w)
    o = oracle_spec(w)
    configs.setup(o, w)
    for k in range(1, w["steps"] + 1):
        g.step(1)
        o.step(1)
        nd = sum(int(np.sum(g.get(f) != o.get(f))) for f in ("u", "v", "f"))
        if k <= 3 or k % 5 == 0 or nd:
            say(f"  step {k}: differing entries {nd}, contacts {g.bond_stats()[0]}/{o.bond_stats()[0]}")
        assert nd == 0
    go, gn, _ = g.bonds()
    oo, on, _ = o.bonds()
    assert np.array_equal(go, oo) and np.array_equal(gn, on)


def test_gpu_contact_rebuild_download_capacity():
    """Regression: a contact list rebuilt with more pairs than the first
    build.  pdk_download_bonds must size itself from the current list (the
    count is refreshed after every rebuild) and refuse a caller buffer
    smaller than the current pair count instead of writing past its end."""
    import ctypes as C
    say("case: contact list growth across rebuilds, download capacity")
    w = configs.dem_packing(n=2000, seed=11, model="HERTZ", polydisperse=True, dt=1e-8, steps=4, rebuild_every=1,
                            speed=0.0, gravity=False)
    g = engine.Spec(ctx(), w["desc"], w["x"].shape[0])
    configs.upload(g, w)
    configs.setup(g, w)
    centre = w["x"].mean(axis=0)
    counts = [g.bond_stats()[0]]
    for frac in (0.01, 0.02, 0.03):
        # move every sphere toward the centre: gaps shrink, pairs appear at the next rebuild
        g.set_state(u=-(w["x"] - centre) * frac, v=np.zeros_like(w["x"]))
        g.step(1)
        total = g.bond_stats()[0]
        off, nbr, intact = g.bonds()
        assert off[-1] == total == nbr.size
        counts.append(total)
    say(f"  pair counts after each rebuild: {counts}")
    assert counts[-1] > counts[0]
    small = np.zeros(max(counts[0], 1), dtype=np.int64)
    offs = np.zeros(g.n + 1, dtype=np.int64)
    L = engine.lib()
    rc = L.pdk_download_bonds(g.h, small.size, offs.ctypes.data_as(C.POINTER(C.c_int64)),
                              small.ctypes.data_as(C.POINTER(C.c_int64)), None)
    assert rc != 0 and b"capacity" in L.pdk_last_error()
    g.close()


def test_gpu_random_packing_reaches_solid_fraction():
    """The granular-box packing generator (a melted lattice compacted by this
    engine's damped Hertz step inside frozen wall sheets): 8,000 spheres with
    radii log-normal (sigma 0.1) clipped to +-30 % end at a solid fraction >=
    0.55 of the box that holds them, the fraction recomputed here from the
    returned spheres, with no two spheres overlapping (checked over all
    pairs) and every sphere inside the container."""
    say("case: random packing at 0.55, 8000 spheres")
    y, r, frac, stages, yw, rw = configs.dem_random_packing(ctx(), n=8000, fraction=0.55, seed=5, log=say)
    say(f"  fraction {frac:.4f} after {stages} stages, radii [{r.min():.3e}, {r.max():.3e}], "
        f"{yw.shape[0]} wall spheres")
    lo = (y - r[:, None]).min(axis=0)
    hi = (y + r[:, None]).max(axis=0)
    own = float((4.0 / 3.0 * np.pi * r ** 3).sum() / np.prod(hi - lo))
    assert abs(own - frac) <= 1e-12 and frac >= 0.55
    assert r.min() >= 0.7e-3 * 0.99 and r.max() <= 1.3e-3
    assert np.all(lo >= yw.min(axis=0) - 1e-9) and np.all(hi <= yw.max(axis=0) + 1e-9)
    worst = 0.0
    for k in range(0, y.shape[0], 1000):
        d = np.linalg.norm(y[k:k + 1000, None, :] - y[None, :, :], axis=2)
        ov = (r[k:k + 1000, None] + r[None, :]) - d
        idx = np.arange(k, min(k + 1000, y.shape[0]))
        ov[np.arange(idx.size), idx] = -np.inf
        worst = max(worst, float(ov.max()))
    say(f"  largest pair overlap {worst:.3e} m (negative: a gap)")
    assert worst <= 0.0
