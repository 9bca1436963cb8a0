"""Pins of the CPU oracle against the reference's own known answers.

unit_test/tstForce.hpp: lattice sums and closed forms for PMB/LPS force,
energy, dilatation and neighbor counts; tstComm.hpp:153-166 (6 neighbors at
delta = dx); plus brute-force neighbor sets.
"""
import numpy as np
import pytest

from cabanapd_amd import configs
from oracle import binding as orc
from tests import lattice_reference as ref


def run_lattice(model, field, coeff, m, nc, K=1.0, G=0.5, influence=None, fracture=False):
    w = configs.lattice_sum_cube(model=model, m=m, nc=nc, field=field, coeff=coeff, K=K, G=G,
                                 influence=influence, fracture=fracture)
    s = orc.OracleSpec(w["desc"], w["x"].shape[0], w["x"], w["vol"], u=w["u"], v=w["v"], rho=w["rho"])
    s.build_bonds()
    s.compute_weighted_volume()
    s.compute_dilatation()
    s.compute_force(True)
    Phi, _ = s.compute_energy()
    return w, s, Phi


def interior(x, horizon, bw):
    lo, hi = -1.0 + horizon * bw, 1.0 - horizon * bw
    return np.all((x > lo) & (x < hi), axis=1)


def float_eq(a, b):
    # gtest EXPECT_FLOAT_EQ: within 4 ulps at single precision
    return abs(np.float32(a) - np.float32(b)) <= 4 * np.spacing(np.float32(max(abs(a), abs(b))))


@pytest.mark.parametrize("fracture", [False, True])
def test_pmb_linear(fracture):
    # test_force_pmb / test_force_pmb_damage, LinearTag (tstForce.hpp:1140-1155, 1182-1199)
    m, nc, K, s0 = 3, 11, 1.0, 0.1
    w, s, Phi = run_lattice("PMB", "linear", s0, m, nc, K=K, fracture=fracture)
    horizon = w["desc"].horizon
    x, f, W, vol = w["x"], s.get("f"), s.get("W"), s.get("vol")
    sel = interior(x, horizon, 1.1)
    assert sel.sum() == 125
    ref_W = ref.pmb_W_linear(K, horizon, m, s0)
    assert np.all(f[sel] <= 1e-13)
    assert all(float_eq(Wi, ref_W) for Wi in W[sel])
    assert np.all(np.abs(W[sel] - 4.5 * K * s0 * s0) <= 0.15 * W[sel])
    assert abs(Phi - np.sum(W * vol)) <= 1e-5
    if fracture:
        phi = s.get("phi")
        assert np.all(phi == 0.0)  # G0 = 1000: nothing breaks
    # stress (tstForce.hpp:744-749): diagonal equal to the lattice sum, the
    # rest zero; the continuum value 3 K s0 within the lattice's 15 %
    s.compute_stress()
    sig = s.get("stress")[sel]
    ref_sig = ref.pmb_stress_linear(K, horizon, m, s0)
    for d in range(3):
        assert all(float_eq(v, ref_sig[d]) for v in sig[:, d, d])
        assert np.all(np.abs(sig[:, d, d] - 3.0 * K * s0) <= 0.15 * 3.0 * K * s0)
    off = sig[:, ~np.eye(3, dtype=bool)]
    assert np.all(np.abs(off) <= 1e-13)


@pytest.mark.parametrize("fracture", [False, True])
def test_pmb_quadratic(fracture):
    m, nc, K, u11 = 3, 11, 1.0, 0.01
    w, s, Phi = run_lattice("PMB", "quadratic", u11, m, nc, K=K, fracture=fracture)
    horizon = w["desc"].horizon
    x, f, W, vol = w["x"], s.get("f"), s.get("W"), s.get("vol")
    sel = np.nonzero(interior(x, horizon, 1.1))[0]
    for p in sel:
        xp = x[p, 0]
        assert float_eq(f[p, 0], ref.pmb_fx_quadratic(K, horizon, m, u11, xp))
        assert abs(f[p, 0] - 18.0 / 5.0 * K * u11) <= abs(0.10 * f[p, 0])
        assert f[p, 1] <= 1e-13 and f[p, 2] <= 1e-13
        assert abs(W[p] - ref.pmb_W_quadratic(K, horizon, m, u11, xp)) <= 1e-6
        an = 18.0 * K * u11 * u11 * (xp * xp / 5.0 + horizon * horizon / 42.0)
        assert abs(W[p] - an) <= 0.05 * W[p]
    assert abs(Phi - np.sum(W * vol)) <= 1e-5


@pytest.mark.parametrize("fracture", [False, True])
def test_lps_linear(fracture):
    # test_force_lps / test_force_lps_damage (influence 1), tstForce.hpp:1168-1180, 1217-1232
    m, nc, K, G, s0 = 3, 15, 1.0, 0.5, 0.1
    w, s, Phi = run_lattice("LPS", "linear", s0, m, nc, K=K, G=G, influence=1, fracture=fracture)
    horizon = w["desc"].horizon
    x, f, W, vol, theta = w["x"], s.get("f"), s.get("W"), s.get("vol"), s.get("theta")
    sel = interior(x, horizon, 2.1)
    assert sel.sum() > 0
    ref_W = ref.lps_W_linear(K, G, horizon, m, s0, 1)
    assert np.all(f[sel] <= 1e-13)
    assert all(float_eq(Wi, ref_W) for Wi in W[sel])
    assert all(float_eq(Wi, 4.5 * K * s0 * s0) for Wi in W[sel])  # LPS is exact
    # theta = 3 s0 is checked for every particle in the reference; it holds
    # on the interior (boundary particles have partial neighborhoods).
    assert all(float_eq(t, 3 * s0) for t in theta[sel])
    assert abs(Phi - np.sum(W * vol)) <= 1e-5
    # stress: the lattice sum of tstForce.hpp:345-379, and exactly 3 K s0 on
    # a full neighborhood (sum of infl xi_a^2 V is m/3 by symmetry)
    s.compute_stress()
    sig = s.get("stress")[sel]
    ref_sig = ref.lps_stress_linear(K, G, horizon, m, s0, 1)
    for d in range(3):
        assert all(float_eq(v, ref_sig[d]) for v in sig[:, d, d])
        assert all(float_eq(v, 3.0 * K * s0) for v in sig[:, d, d])
    off = sig[:, ~np.eye(3, dtype=bool)]
    assert np.all(np.abs(off) <= 1e-13)


@pytest.mark.parametrize("fracture", [False, True])
def test_lps_quadratic(fracture):
    m, nc, K, G, u11 = 3, 15, 1.0, 0.5, 0.01
    w, s, Phi = run_lattice("LPS", "quadratic", u11, m, nc, K=K, G=G, influence=1, fracture=fracture)
    horizon = w["desc"].horizon
    x, f, W, vol = w["x"], s.get("f"), s.get("W"), s.get("vol")
    sel = np.nonzero(interior(x, horizon, 2.1))[0]
    assert sel.size > 0
    for p in sel[:: max(1, sel.size // 8)]:
        xp = x[p, 0]
        assert float_eq(f[p, 0], ref.lps_fx_quadratic(K, G, horizon, m, u11, xp, 1))
        assert abs(f[p, 0] - 2.0 * (K + 4.0 / 3.0 * G) * u11) <= abs(0.10 * f[p, 0])
        assert abs(W[p] - ref.lps_W_quadratic(K, G, horizon, m, u11, xp, 1)) <= 1e-6
        an = u11 * u11 * ((2.0 * K + 8.0 / 3.0 * G) * xp * xp + 75.0 / 2.0 * G * horizon * horizon / 49.0)
        assert abs(W[p] - an) <= 0.20 * W[p]
    assert abs(Phi - np.sum(W * vol)) <= 1e-5


@pytest.mark.parametrize("m,nc", [(3, 11), (6, 15)])
def test_neighbor_count(m, nc):
    # test_neighbor (tstForce.hpp:983-1015): max count equals the lattice count
    w = configs.lattice_sum_cube(model="PMB", m=m, nc=nc)
    s = orc.OracleSpec(w["desc"], w["x"].shape[0], w["x"], w["vol"])
    s.build_bonds()
    total, maxn = s.bond_stats()
    assert maxn == ref.ref_neighbors(w["desc"].horizon, m)
    if m == 3:
        assert maxn == 122


def test_neighbor_count_m1():
    # tstComm.hpp:153-166: delta = dx gives 6 interior neighbors
    x, vol, d = configs.grid_cell_centred([0, 0, 0], [1, 1, 1], (5, 5, 5))
    from cabanapd_amd._abi import make_desc
    desc = make_desc("PMB", horizon=d[0], K=1.0)
    s = orc.OracleSpec(desc, x.shape[0], x, vol)
    s.build_bonds()
    _, maxn = s.bond_stats()
    assert maxn == 6


def brute_force_bonds(x, rc):
    n = x.shape[0]
    d = x[None, :, :] - x[:, None, :]
    dist2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
    np.fill_diagonal(dist2, np.inf)
    rows = [np.nonzero(dist2[i] <= rc * rc)[0] for i in range(n)]
    return rows


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neighbor_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = 700
    x = rng.uniform(-1, 1, size=(n, 3)) * np.array([1.0, 0.5, 0.25])
    vol = np.full(n, 1e-3)
    from cabanapd_amd._abi import make_desc
    horizon = 0.17
    desc = make_desc("PMB", horizon=horizon, K=1.0)
    s = orc.OracleSpec(desc, n, x, vol)
    s.build_bonds()
    off, nbr, _ = s.bonds()
    rows = brute_force_bonds(x, horizon + 1e-14)
    for i in range(n):
        assert np.array_equal(nbr[off[i]:off[i + 1]], rows[i])


def test_break_two_particle_closed_form():
    # BaseFractureModel: break iff r^2 >= (1 + s0)^2 xi^2 (CabanaPD_ForceModels.hpp:123-128)
    from cabanapd_amd._abi import make_desc
    K, horizon, G0 = 127.33e9, 1.505e-3, 42408.0
    desc = make_desc("PMB", fracture=True, horizon=horizon, K=K, G0=G0)
    k = orc.model_constants(desc)
    s0 = np.sqrt(5.0 * G0 / 9.0 / K / horizon)
    assert k["s0"] == pytest.approx(s0, rel=1e-15)
    xi = np.array([1e-3, 0.0, 0.0])
    for eps, expect in [(1e-6, True), (-1e-6, False), (0.5, True), (-0.5, False)]:
        eta = np.array([xi[0] * s0 * (1 + eps), 0.0, 0.0])
        f, w, brk = orc.bond_pmb(desc, xi, eta, 1.0)
        assert brk == expect
        s = s0 * (1 + eps)
        c = 18.0 * K / (np.pi * horizon ** 4)
        assert f[0] == pytest.approx(c * s, rel=1e-9)
        assert w == pytest.approx(0.25 * c * s * s * xi[0], rel=1e-9)


def test_prenotch_geometry():
    v1 = np.array([0.05, 0.0, 0.0])
    v2 = np.array([0.0, 0.0, 0.009])
    p0 = np.array([-0.05, 0.025, -0.0045])
    cases = [
        # bond crossing the patch
        (np.array([-0.03, 0.024, 0.0]), np.array([-0.03, 0.026, 0.0]), False),
        # crossing the plane outside the patch (beyond v1)
        (np.array([0.01, 0.024, 0.0]), np.array([0.01, 0.026, 0.0]), True),
        # below the patch in z
        (np.array([-0.03, 0.024, -0.006]), np.array([-0.03, 0.026, -0.006]), True),
        # both ends on one side
        (np.array([-0.03, 0.020, 0.0]), np.array([-0.03, 0.024, 0.0]), True),
        # parallel to the plane, not coplanar
        (np.array([-0.03, 0.024, 0.0]), np.array([-0.02, 0.024, 0.0]), True),
        # coplanar and inside the patch
        (np.array([-0.03, 0.025, 0.0]), np.array([-0.02, 0.025, 0.0]), False),
        # coplanar, outside
        (np.array([0.02, 0.025, 0.0]), np.array([0.03, 0.025, 0.0]), True),
        # endpoint on the plane inside the patch (within tol)
        (np.array([-0.03, 0.025, 0.0]), np.array([-0.03, 0.027, 0.0]), False),
    ]
    for xi, xj, keep in cases:
        assert orc.bond_prenotch(v1, v2, p0, xi, xj) == keep, (xi, xj)
        assert orc.bond_prenotch(v1, v2, p0, xj, xi) == keep, (xj, xi)


def test_verlet_constant_force_closed_form():
    # VelocityVerlet with a constant body force and no bonds:
    # u_n = u0 + n dt v0 + (n dt)^2 a / 2 exactly in real arithmetic.
    from cabanapd_amd._abi import make_desc
    n = 5
    x = np.arange(n * 3, dtype=float).reshape(n, 3) * 10.0
    vol = np.ones(n)
    rho = np.full(n, 2.0)
    v0 = np.tile([1.0, -2.0, 0.5], (n, 1))
    b = np.tile([4.0, 2.0, -6.0], (n, 1))
    dt = 1e-3
    desc = make_desc("PMB", horizon=0.5, K=1.0, dt=dt)
    s = orc.OracleSpec(desc, n, x, vol, v=v0, rho=rho)
    s.build_bonds()
    s.set_body_force(b)
    s.init()
    steps = 200
    s.step(steps)
    t = steps * dt
    a = b / rho[:, None]
    u = s.get("u")
    v = s.get("v")
    np.testing.assert_allclose(u, v0 * t + 0.5 * a * t * t, rtol=1e-12)
    np.testing.assert_allclose(v, v0 + a * t, rtol=1e-12)


def test_critical_timestep_lattice_sum():
    # Inputs::computeCriticalTimeStep (CabanaPD_Input.hpp:217-336); values for
    # the benchmark specimens as computed independently from the lattice sum.
    dt1 = orc.critical_timestep(1 / 40, 1 / 40, 1 / 40, 48e9, 2440, 3.01 / 40, 1.0)
    assert dt1 == pytest.approx(4.00e-6, rel=0.01)
    K2 = 191e9 / 1.5
    dt2 = orc.critical_timestep(5e-4, 5e-4, 5e-4, K2, 8000, 3.01 * 5e-4, 1.0)
    assert dt2 == pytest.approx(8.90e-8, rel=0.01)
    # the stored value carries the safety factor (CabanaPD_Input.hpp:387-398)
    assert orc.critical_timestep(5e-4, 5e-4, 5e-4, K2, 8000, 3.01 * 5e-4, 0.85) == pytest.approx(0.85 * dt2)


def test_model_constants():
    from cabanapd_amd._abi import make_desc
    # LPS with influence 0: s0 = sqrt(8 G0 / (15 K delta)) (force_models/CabanaPD_LPS.hpp:218-221)
    K, G, horizon = 42.857e9, 29.508e9, 0.023701
    d = make_desc("LPS", fracture=True, horizon=horizon, K=K, G=G, G0=3.8, influence=0)
    k = orc.model_constants(d)
    assert k["s0"] == pytest.approx(np.sqrt(8 * 3.8 / (15 * K * horizon)), rel=1e-14)
    assert k["theta_coeff"] == 3 * K - 5 * G
    assert k["s_coeff"] == 15 * G
    assert k["bond_break_coeff"] == pytest.approx((1 + k["s0"]) ** 2, rel=1e-15)


def _rotation(x, angle):
    """u = R x - x for a rotation by angle about z."""
    c, s = np.cos(angle), np.sin(angle)
    u = np.zeros_like(x)
    u[:, 0] = (c - 1.0) * x[:, 0] - s * x[:, 1]
    u[:, 1] = s * x[:, 0] + (c - 1.0) * x[:, 1]
    return u


def _lattice_offsets(m):
    r = np.arange(-m, m + 1)
    a, b, c = np.meshgrid(r, r, r, indexing="ij")
    off = np.stack([a.ravel(), b.ravel(), c.ravel()], axis=1).astype(np.float64)
    d2 = np.sum(off * off, axis=1)
    return off[(d2 > 0) & (d2 <= m * m + 1e-9)]


def _run_field(model, u, m, nc, K=1.0, G=0.5, influence=None):
    w = configs.lattice_sum_cube(model=model, m=m, nc=nc, field="linear", coeff=0.0, K=K, G=G,
                                 influence=influence)
    w["u"] = u(w["x"])
    s = orc.OracleSpec(w["desc"], w["x"].shape[0], w["x"], w["vol"], u=w["u"], v=w["v"], rho=w["rho"])
    s.build_bonds()
    s.compute_weighted_volume()
    s.compute_dilatation()
    s.compute_force(True)
    Phi, _ = s.compute_energy()
    return w, s, Phi


@pytest.mark.parametrize("model", ["LINEAR_PMB", "PMB"])
def test_linear_pmb_rotation(model):
    """test_force_linear_pmb's model (tstForce.hpp:1157-1167) on a rigid
    rotation: the linearized stretch of offset (a, b, c) is exactly
    (cos t - 1)(a^2 + b^2)/(a^2 + b^2 + c^2), the non-linear one 0.  W_i is the
    lattice sum of c s^2 |xi| V / 4 on interior particles; f_i = 0 by
    symmetry."""
    m, nc, K, angle = 3, 11, 1.0, 0.1
    w, s, Phi = _run_field(model, lambda x: _rotation(x, angle), m, nc, K=K)
    horizon = w["desc"].horizon
    dx = 2.0 / nc
    c = 18.0 * K / (np.pi * horizon ** 4)
    off = _lattice_offsets(m)
    a2b2 = off[:, 0] ** 2 + off[:, 1] ** 2
    xi = np.sqrt(np.sum(off * off, axis=1)) * dx
    s_lin = (np.cos(angle) - 1.0) * a2b2 / np.sum(off * off, axis=1)
    W_ref = np.sum(0.25 * c * s_lin * s_lin * xi * dx ** 3) if model == "LINEAR_PMB" else 0.0
    x, f, W = w["x"], s.get("f"), s.get("W")
    sel = interior(x, horizon, 1.1)
    assert sel.sum() == 125
    assert np.all(np.abs(f[sel]) <= 1e-12)
    assert np.all(np.abs(W[sel] - W_ref) <= 1e-12 * max(W_ref, 1e-300) + 1e-18)
    if model == "LINEAR_PMB":
        assert W_ref > 1e-6  # the linearized model sees the rotation


def test_linear_pmb_stretch():
    """Uniaxial stretch u = (e x, 0, 0): linearized stretch e a^2/(a^2+b^2+c^2)
    per offset, f_i = 0 and W_i the lattice sum on the interior; differs from
    PMB at O(e^2)."""
    m, nc, K, e = 3, 11, 1.0, 0.1
    field = lambda x: np.stack([e * x[:, 0], 0 * x[:, 0], 0 * x[:, 0]], axis=1)  # noqa: E731
    w, s, _ = _run_field("LINEAR_PMB", field, m, nc, K=K)
    horizon = w["desc"].horizon
    dx = 2.0 / nc
    c = 18.0 * K / (np.pi * horizon ** 4)
    off = _lattice_offsets(m)
    n2 = np.sum(off * off, axis=1)
    s_lin = e * off[:, 0] ** 2 / n2
    W_ref = np.sum(0.25 * c * s_lin * s_lin * np.sqrt(n2) * dx * dx ** 3)
    sel = interior(w["x"], horizon, 1.1)
    assert np.all(np.abs(s.get("f")[sel]) <= 1e-12)
    assert np.all(np.abs(s.get("W")[sel] - W_ref) <= 1e-12 * W_ref)
    _, s2, _ = _run_field("PMB", field, m, nc, K=K)
    assert np.all(np.abs(s2.get("W")[sel] - W_ref) > 1e-3 * W_ref)


def test_linear_lps_rotation():
    """test_force_linear_lps's model (tstForce.hpp:1186-1198, influence 1) on a
    rigid rotation: the dilatation uses the non-linear stretch (the weighted
    volume and dilatation passes are LPS's), so theta_i = 0 and
    W_i = sum (s_coeff / m_i) / 2 * (1/|xi|) s_lin^2 |xi|^2 V; m_i the
    lattice sum of |xi| V (influence 1/|xi|)."""
    m, nc, K, G, angle = 3, 15, 1.0, 0.5, 0.1
    w, s, _ = _run_field("LINEAR_LPS", lambda x: _rotation(x, angle), m, nc, K=K, G=G, influence=1)
    horizon = w["desc"].horizon
    dx = 2.0 / nc
    off = _lattice_offsets(m)
    n2 = np.sum(off * off, axis=1)
    xi = np.sqrt(n2) * dx
    vol = dx ** 3
    m_ref = np.sum(xi * vol)
    s_lin = (np.cos(angle) - 1.0) * (off[:, 0] ** 2 + off[:, 1] ** 2) / n2
    W_ref = np.sum(0.5 * (15.0 * G / m_ref) * (1.0 / xi) * s_lin * s_lin * xi * xi * vol)
    sel = interior(w["x"], horizon, 2.1)
    assert sel.sum() > 0
    assert np.all(np.abs(s.get("m")[sel] - m_ref) <= 1e-12 * m_ref)
    assert np.all(np.abs(s.get("theta")[sel]) <= 1e-12)
    assert np.all(np.abs(s.get("f")[sel]) <= 1e-12)
    assert np.all(np.abs(s.get("W")[sel] - W_ref) <= 1e-10 * W_ref)


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_stress_rigid_rotation_is_zero(model):
    # a rigid rotation stretches no bond: s = 0, theta = 0, so every bond's
    # force and the stress vanish (to rounding of the rotated lengths)
    w = configs.lattice_sum_cube(model=model, m=3, nc=9, field="linear", coeff=0.0, K=1.0, G=0.5)
    x = w["x"]
    ang = 0.3
    c, sn = np.cos(ang), np.sin(ang)
    u = np.zeros_like(x)
    u[:, 0] = (c - 1.0) * x[:, 0] - sn * x[:, 2]
    u[:, 2] = sn * x[:, 0] + (c - 1.0) * x[:, 2]
    s = orc.OracleSpec(w["desc"], x.shape[0], x, w["vol"], u=u, v=w["v"], rho=w["rho"])
    s.build_bonds()
    s.compute_weighted_volume()
    s.compute_dilatation()
    s.compute_force(True)
    s.compute_stress()
    sig = s.get("stress")
    assert np.max(np.abs(sig)) <= 1e-13


def test_stress_two_particle_closed_form():
    # two particles on the x axis, one bond each way: stress(0, x, x) =
    # 0.5 c s V_j (r_x / r) xi_x with xi = dx, r = dx (1 + s)
    dx, K, s0 = 0.5, 2.0, 0.01
    horizon = 1.01 * dx
    from cabanapd_amd._abi import make_desc
    desc = make_desc("PMB", horizon=horizon, K=K, dt=1e-3)
    x = np.array([[0.0, 0.0, 0.0], [dx, 0.0, 0.0]])
    u = np.array([[0.0, 0.0, 0.0], [s0 * dx, 0.0, 0.0]])
    vol = np.array([0.125, 0.25])
    s = orc.OracleSpec(desc, 2, x, vol, u=u)
    s.build_bonds()
    s.compute_force(True)
    s.compute_stress()
    sig = s.get("stress")
    c = 18.0 * K / (np.pi * horizon ** 4)
    assert abs(sig[0, 0, 0] - 0.5 * c * s0 * vol[1] * dx) <= 1e-12 * abs(sig[0, 0, 0])
    assert abs(sig[1, 0, 0] - 0.5 * c * s0 * vol[0] * dx) <= 1e-12 * abs(sig[1, 0, 0])
    assert np.count_nonzero(sig[:, ~np.eye(3, dtype=bool)]) == 0
    assert np.count_nonzero(sig[:, 1:, 1:]) == 0


def multi_material_run(model, mats, mtype, steps=20, pairs=None):
    w = configs.lattice_sum_cube(model=model, m=3, nc=9, field="linear", coeff=0.0, K=1e9, G=0.5e9,
                                 fracture=True, G0=1.0)
    x = w["x"]
    rng = np.random.default_rng(3)
    u = rng.uniform(-2e-3, 2e-3, size=x.shape) * (2.0 / 9)
    s = orc.OracleSpec(w["desc"], x.shape[0], x, w["vol"], u=u, rho=np.full(x.shape[0], 1000.0))
    s.build_bonds()
    if mats is not None:
        s.set_materials(mats, mtype)
    if pairs is not None:
        s.set_pair_models(pairs, mtype)
    w["desc"].dt = 1e-6
    s.init()
    s.step(steps)
    out = (s.get("u"), s.get("f"), s.bonds()[2])
    s.compute_energy()
    out = out + (s.get("W"), s.get("phi"))
    s.close()
    return out


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_multi_material_identical_materials_equal_single(model):
    # createMultiForceModel with two copies of one model: every pair model
    # equals the single model (averages of equal values), so the run is the
    # single-material run bit for bit
    K, G, G0 = 1e9, 0.5e9, 1.0
    n = 9 ** 3
    mtype = np.random.default_rng(1).integers(0, 2, n).astype(np.int32)
    ref = multi_material_run(model, None, None)
    got = multi_material_run(model, [(K, G, G0), (K, G, G0)], mtype)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    assert (ref[2] == 0).sum() > 0  # bonds break: the break test is exercised


@pytest.mark.parametrize("model", ["PMB", "LPS"])
def test_multi_material_label_swap(model):
    # relabelling the materials (swap the list and the types) changes nothing
    mats = [(1e9, 0.5e9, 1.0), (3e9, 1.2e9, 4.0)]
    n = 9 ** 3
    mtype = np.random.default_rng(2).integers(0, 2, n).astype(np.int32)
    a = multi_material_run(model, mats, mtype)
    b = multi_material_run(model, mats[::-1], (1 - mtype).astype(np.int32))
    for p, q in zip(a, b):
        assert np.array_equal(p, q)
    single = multi_material_run(model, None, None)
    assert not np.array_equal(a[0], single[0])
