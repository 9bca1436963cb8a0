"""Pins of oracle behaviour that no reference-held number covers.

- LPS fracture tests nofail(i) twice (force/CabanaPD_LPS.hpp:417-418) while
  PMB tests nofail(i) and nofail(j) (force/CabanaPD_PMB.hpp:260-261): with
  only particle j no-fail, an over-stretched pair breaks i -> j but keeps
  j -> i under LPS, and keeps both under PMB.
- LPS fracture energy of a particle whose bonds are all broken
  (force/CabanaPD_LPS.hpp:470-487): 1 / num_bonds = 1/0 = inf times a
  dilatation of exactly 0 is NaN in IEEE arithmetic, and the damage is 1.
- The polydisperse contact extension (per-particle radii): for spheres of
  radii R1 != R2 the pair uses R* = R1 R2 / (R1 + R2), so at rest the Hertz
  force is (4/3) E* sqrt(R*) delta^(3/2) with E* = E / (2 (1 - nu^2)),
  equal and opposite on the two spheres (Johnson, Contact Mechanics, 1985,
  eq. 4.22), and on the receding branch the JKR force reaches the pull-off
  value -(3/2) pi R* Gamma (Johnson, Kendall and Roberts, 1971) -- equal
  radii reduce to the reference's R/2.
"""
import math

import numpy as np

from cabanapd_amd._abi import make_desc
from oracle import binding as orc


def _pair(model, dx, stretch, nofail):
    horizon = 1.01 * dx
    kw = dict(G=0.6) if model == "LPS" else {}
    d = make_desc(model, fracture=True, horizon=horizon, K=1.0, G0=1e-4, dt=1e-3, **kw)
    x = np.array([[0.0, 0.0, 0.0], [dx, 0.0, 0.0]])
    u = np.array([[0.0, 0.0, 0.0], [stretch * dx, 0.0, 0.0]])
    s = orc.OracleSpec(d, 2, x, np.full(2, 0.125), u=u, nofail=np.array(nofail, dtype=np.int32))
    s.build_bonds()
    return s, orc.model_constants(d)["s0"]


def test_lps_nofail_tests_i_twice():
    for model, expect in (("LPS", [0, 1]), ("PMB", [1, 1])):
        s, s0 = _pair(model, 0.5, 0.0, [0, 1])
        s, _ = _pair(model, 0.5, 3.0 * s0, [0, 1])
        s.compute_weighted_volume()
        s.compute_dilatation()
        s.compute_force(True)
        off, nbr, intact = s.bonds()
        # particle 0's bond to 1, then particle 1's bond to 0
        assert list(nbr) == [1, 0]
        assert list(intact) == expect, (model, list(intact))


def test_lps_energy_of_isolated_particle_is_nan():
    s, s0 = _pair("LPS", 0.5, 0.0, [0, 0])
    s, s0 = _pair("LPS", 0.5, 3.0 * s0, [0, 0])
    s.compute_weighted_volume()
    s.compute_dilatation()
    s.compute_force(True)  # both directions break
    assert list(s.bonds()[2]) == [0, 0]
    s.compute_weighted_volume()
    s.compute_dilatation()
    assert np.all(s.get("theta") == 0.0)
    W, D = s.compute_energy()
    assert np.all(np.isnan(s.get("W")))
    assert math.isnan(W)
    assert np.all(s.get("phi") == 1.0) and D == 2.0


def _spheres(model, R1, R2, delta, E=195.6e9, nu=0.25, gamma=0.0, separating=False):
    extend = 1e-6
    x = np.array([[0.0, 0.0, 0.0], [R1 + R2 - delta, 0.0, 0.0]])
    vol = np.array([4.0 / 3.0 * math.pi * R1 ** 3, 4.0 / 3.0 * math.pi * R2 ** 3])
    # restitution 1: no damping term; a separating pair (v_n > 0) is on the
    # adhesive (receding) branch of JKR (force_models/CabanaPD_HertzianJKR.hpp:109-110)
    d = make_desc(model, radius=max(R1, R2), radius_extend=extend, nu=nu, E=E, restitution=1.0, gamma=gamma,
                  mu_friction=0.0, dt=1e-9, polydisperse=True)
    v = np.zeros((2, 3))
    if separating:
        v[1, 0] = 1e-3
    s = orc.OracleSpec(d, 2, x, vol, v=v, rho=np.full(2, 7950.0), radius=np.array([R1, R2]))
    s.build_bonds()
    return s, vol


def test_polydisperse_hertz_matches_dissimilar_sphere_closed_form():
    E, nu = 195.6e9, 0.25
    Es = E / (2.0 * (1.0 - nu * nu))
    for R1, R2, delta in ((5e-5, 5e-5, 2e-7), (3e-5, 8e-5, 2e-7), (2e-5, 1e-4, 5e-8)):
        s, vol = _spheres("HERTZ", R1, R2, delta, E, nu)
        s.compute_force(True)
        F = s.get("f") * vol[:, None]
        Rstar = R1 * R2 / (R1 + R2)
        expect = 4.0 / 3.0 * Es * math.sqrt(Rstar) * delta ** 1.5
        assert abs(-F[0, 0] - expect) <= 1e-9 * expect, (R1, R2)
        assert abs(F[1, 0] - expect) <= 1e-9 * expect
        assert np.all(F[:, 1:] == 0.0)


def test_polydisperse_jkr_pulloff_matches_closed_form():
    E, nu, gamma = 195.6e9, 0.25, 10.0
    for R1, R2 in ((5e-5, 5e-5), (3e-5, 8e-5)):
        Rstar = R1 * R2 / (R1 + R2)
        fc = 1.5 * math.pi * Rstar * (2.0 * gamma)
        # the force over a sweep of overlaps, down to the tear-off separation
        fmin = 0.0
        for delta in np.linspace(-2e-8, 1e-8, 3001):
            s, vol = _spheres("HERTZ_JKR", R1, R2, delta, E, nu, gamma, separating=True)
            s.compute_force(True)
            fmin = min(fmin, s.get("f")[1, 0] * vol[1])
        assert abs(fmin / fc + 1.0) <= 2e-3, (R1, R2, fmin, fc)
