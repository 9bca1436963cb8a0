"""Oracle checks of the explicit-indexing multi-material form
(createMultiForceModel(particles, indexing, models...),
CabanaPD_ForceModelsMulti.hpp:228-250, 351-358): each bond calls the model of
its ordered type pair; an LPS pair model puts theta_coeff/s_coeff [0] on
particle i's terms and [1] on particle j's (force_models/CabanaPD_LPS.hpp:117-132).

Pinned by closed forms on two particles (every side and pair distinct, so a
transposed table or swapped [0]/[1] fails) and by bit equality of the
averaged table with the AverageTag form (orc_set_materials).
"""
import math

import numpy as np
import pytest

from cabanapd_amd._abi import make_desc
from oracle import binding as orc
from tests.test_oracle_known_answers import multi_material_run


def two_particles(model, s_bond, influence=0):
    """Particles 0 (type 0) at the origin and 1 (type 1) at dx on the x axis,
    the bond stretched by s_bond; unequal volumes."""
    dx = 0.5
    horizon = 1.01 * dx
    desc = make_desc(model, horizon=horizon, K=2.0, G=0.7, dt=1e-3, influence=influence)
    x = np.array([[0.0, 0.0, 0.0], [dx, 0.0, 0.0]])
    u = np.array([[0.0, 0.0, 0.0], [s_bond * dx, 0.0, 0.0]])
    vol = np.array([0.125, 0.25])
    s = orc.OracleSpec(desc, 2, x, vol, u=u)
    s.build_bonds()
    return s, dx, vol


def pair(c=0.0, bcc=0.0, tc=(0.0, 0.0), sc=(0.0, 0.0)):
    return {"c": c, "bond_break_coeff": bcc, "theta_coeff": tc, "s_coeff": sc}


def test_pmb_pair_table_closed_form():
    # f_0 = c(0,1) s V_1, f_1 = -c(1,0) s V_0: each particle's force uses the
    # model of (its type, the other's type); c(0,0), c(1,1) unused here
    s_bond = 1e-3
    s, dx, vol = two_particles("PMB", s_bond)
    c = {(0, 0): 3.0, (0, 1): 5.0, (1, 0): 7.0, (1, 1): 11.0}
    s.set_pair_models([pair(c=c[(a, b)]) for a in range(2) for b in range(2)], np.array([0, 1]))
    s.compute_force(True)
    f = s.get("f")
    r = dx * (1.0 + s_bond)
    st = (r - dx) / dx
    assert abs(f[0, 0] - c[(0, 1)] * st * vol[1]) <= 1e-14 * abs(f[0, 0])
    assert abs(f[1, 0] + c[(1, 0)] * st * vol[0]) <= 1e-14 * abs(f[1, 0])
    assert np.all(f[:, 1:] == 0.0)


@pytest.mark.parametrize("influence", [0, 1])
def test_lps_pair_table_closed_form(influence):
    # one bond: m_i = w xi^2 V_j, theta_i = 3 w s xi^2 V_j / m_i (w the
    # influence); particle i's force coefficient with the pair model (a, b):
    # (tc[0] theta_i / m_i + tc[1] theta_j / m_j + s (sc[0] / m_i + sc[1] / m_j)) w xi V_j
    s_bond = 2e-3
    s, dx, vol = two_particles("LPS", s_bond, influence)
    tabs = {(0, 0): pair(tc=(1.0, 1.0), sc=(1.0, 1.0)),
            (0, 1): pair(tc=(2.0, 30.0), sc=(500.0, 7000.0)),
            (1, 0): pair(tc=(11.0, 130.0), sc=(1700.0, 19000.0)),
            (1, 1): pair(tc=(1.0, 1.0), sc=(1.0, 1.0))}
    s.set_pair_models([tabs[(a, b)] for a in range(2) for b in range(2)], np.array([0, 1]))
    s.compute_weighted_volume()
    s.compute_dilatation()
    s.compute_force(True)
    f = s.get("f")
    xi = dx
    r = dx * (1.0 + s_bond)
    st = (r - xi) / xi
    w = 1.0 / xi if influence == 1 else 1.0
    m = [w * xi * xi * vol[1], w * xi * xi * vol[0]]
    th = [3.0 * w * st * xi * xi * vol[1] / m[0], 3.0 * w * st * xi * xi * vol[0] / m[1]]
    assert np.allclose(s.get("m"), m, rtol=1e-14, atol=0.0)
    assert np.allclose(s.get("theta"), th, rtol=1e-12, atol=0.0)

    def coeff(i, j):
        p = tabs[(i, j)]
        return ((p["theta_coeff"][0] * th[i] / m[i] + p["theta_coeff"][1] * th[j] / m[j] +
                 st * (p["s_coeff"][0] / m[i] + p["s_coeff"][1] / m[j])) * w * xi * vol[j])

    assert abs(f[0, 0] - coeff(0, 1)) <= 1e-12 * abs(f[0, 0])
    assert abs(f[1, 0] + coeff(1, 0)) <= 1e-12 * abs(f[1, 0])


def averaged_pairs(model, mats):
    """The AverageTag pair table, from the reference's averaging constructors
    evaluated here (force_models/CabanaPD_PMB.hpp:51-55,
    force_models/CabanaPD_LPS.hpp:55-69, CabanaPD_ForceModels.hpp:114-121)."""
    ks = []
    for (K, G, G0) in mats:
        d = make_desc(model, fracture=True, horizon=3.0 * 2.0 / 9, K=K, G=G, G0=G0)
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
            out.append(pair(c=c, bcc=bcc, tc=(3.0 * Ka - 5.0 * Ga, 3.0 * Kb - 5.0 * Gb),
                            sc=(15.0 * Ga, 15.0 * Gb)))
    return out


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_averaged_pair_table_equals_average_tag(model):
    # the full-indexing table of the averaged models runs bit for bit as the
    # AverageTag form, with bonds breaking
    mats = [(1e9, 0.5e9, 1.0), (3e9, 1.2e9, 4.0)]
    mtype = np.random.default_rng(2).integers(0, 2, 9 ** 3).astype(np.int32)
    a = multi_material_run(model, mats, mtype)
    b = multi_material_run(model, None, mtype, pairs=averaged_pairs(model, mats))
    for p, q in zip(a, b):
        assert np.array_equal(p, q)
    assert (a[2] == 0).sum() > 0


def test_binary_indexing_pmb_differs_from_average():
    # BinaryIndexing with an independent cross-material model: bonds between
    # unlike particles use c_diff; changing c_diff changes only runs with
    # unlike bonds
    mats = [(1e9, 0.5e9, 1.0), (1e9, 0.5e9, 1.0)]
    base = averaged_pairs("PMB", mats)
    same = base[0]
    n = 9 ** 3
    t_all0 = np.zeros(n, dtype=np.int32)
    t_mix = np.random.default_rng(3).integers(0, 2, n).astype(np.int32)
    stiff = dict(same, c=same["c"] * 1.5)
    table = [same, stiff, stiff, same]
    a0 = multi_material_run("PMB", None, t_all0, pairs=table)
    b0 = multi_material_run("PMB", mats, t_all0)
    for p, q in zip(a0, b0):
        assert np.array_equal(p, q)
    a1 = multi_material_run("PMB", None, t_mix, pairs=table)
    b1 = multi_material_run("PMB", mats, t_mix)
    assert not np.array_equal(a1[1], b1[1])
