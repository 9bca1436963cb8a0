"""Workload generators: particles, model and loading of the benchmark
specimens and of the reference's regression inputs.

Every generator returns a dict with keys
  desc (ModelDesc), x (n,3), vol (n,), rho (n,), u (n,3), v (n,3),
  nofail (n,) or None, radius (n,) or None,
  notches: (p0, v1, v2) arrays or None,
  bc: (kind, idx, value) or None, body (n,3) or None,
  prescribed: list of (field, idx, a, b) set before each force pass,
  temperature (n,) initial temperature or None,
  steps, output_every, name, dx.
Positions are in the caller order the reference's grid creation produces:
lexicographic with x fastest.
"""
import math

import numpy as np

from . import _abi
from ._abi import PDK_BC_FORCE_VALUE, make_desc


def grid_cell_centred(low, high, num_cells):
    """Particles::create on a uniform mesh, one particle per cell at the cell
    centre, vol = cell volume (CabanaPD_Particles.hpp:242-307)."""
    low = np.asarray(low, dtype=np.float64)
    high = np.asarray(high, dtype=np.float64)
    nc = np.asarray(num_cells, dtype=np.int64)
    d = (high - low) / nc
    ax = [low[k] + np.arange(nc[k]) * d[k] + 0.5 * d[k] for k in range(3)]
    zz, yy, xx = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    x = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
    vol = np.full(x.shape[0], d[0] * d[1] * d[2])
    return x, vol, d


def grid_node_centred(n_side, spacing):
    """Node-centred cube: x = (i, j, k) * spacing, i, j, k in [0, n_side)
    (custom Particles::create(x, vol), CabanaPD_Particles.hpp:310-368)."""
    a = np.arange(n_side, dtype=np.float64) * spacing
    zz, yy, xx = np.meshgrid(a, a, a, indexing="ij")
    x = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
    vol = np.full(x.shape[0], spacing ** 3)
    idx = np.stack(np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(n_side), indexing="ij"),
                   axis=-1).reshape(-1, 3)[:, ::-1]
    return x, vol, idx


def _base(desc, x, vol, rho, name, dx, steps, output_every=0):
    n = x.shape[0]
    return dict(desc=desc, x=x, vol=vol, rho=np.full(n, rho) if np.isscalar(rho) else rho,
                u=np.zeros((n, 3)), v=np.zeros((n, 3)), nofail=None, radius=None, notches=None,
                bc=None, body=None, prescribed=[], temperature=None, steps=steps, output_every=output_every,
                name=name, dx=dx)


def elastic_wave(precision="fp64"):
    """regression_test/inputs/elastic_wave.json run through
    examples/mechanics/elastic_wave.cpp: 41^3 cell-centred particles on
    [-0.5, 0.5]^3, rho 100, K 1, G 0.5, horizon 0.075 (+1e-10), influence 0,
    NoFracture, dt 0.01, int(0.6 / 0.01) = 60 steps, radial Gaussian shell
    u = 0.001 exp(-(|x| - 0.25)^2 / 0.07^2) x/|x|.

    The example asks for ForceModel<LinearLPS>, which keeps the LPS force tag
    (force_models/CabanaPD_LPS.hpp:31-32,255-272; only LinearPMB replaces it,
    force_models/CabanaPD_PMB.hpp:584), and the solver picks its force class
    by that tag (CabanaPD_Solver.hpp:96-99): the reference run, and the
    polynomial its regression test fits, is the LPS force."""
    x, vol, d = grid_cell_centred([-0.5, -0.5, -0.5], [0.5, 0.5, 0.5], (41, 41, 41))
    horizon = 0.075 + 1e-10
    desc = make_desc("LPS", precision=precision, horizon=horizon, K=1.0, G=0.5, dt=0.01, influence=0)
    steps = int(0.6 / 0.01)  # Inputs truncation (CabanaPD_Input.hpp:109)
    w = _base(desc, x, vol, 100.0, "elastic_wave", d[0], steps)
    norm = np.sqrt(np.sum(x * x, axis=1))
    amp = 0.001 * np.exp(-((norm - 0.25) ** 2) / 0.07 / 0.07)
    with np.errstate(invalid="ignore", divide="ignore"):
        comp = np.where(norm[:, None] > 0.0, x / norm[:, None], 0.0)
    w["u"] = amp[:, None] * comp
    return w


def line_profile(x, dx, profile_dim=0):
    """Particles of Region<Line> (CabanaPD_Geometry.hpp:137-168): the other two
    coordinates within +-dx/2 of 0."""
    o = [(profile_dim + 1) % 3, (profile_dim + 2) % 3]
    sel = np.ones(x.shape[0], dtype=bool)
    for k in o:
        sel &= (x[:, k] < dx / 2.0) & (x[:, k] > -dx / 2.0)
    return np.nonzero(sel)[0]


def polyfit_check(xs, us, fixture):
    """regression_test/tstElasticWave.hpp:62-93: relative error to the fitted
    polynomial and the L2 norm of the differences inside the fitted window.
    Returns (max relative error, L2 norm, points checked)."""
    c = fixture["coefficients"]
    lo, hi = fixture["x_window"]
    elo, ehi = fixture["x_exclude"]
    worst, total, k = 0.0, 0.0, 0
    for xv, uv in zip(xs, us):
        if xv > hi or xv < lo or (elo < xv < ehi):
            continue
        fit = 0.0
        for ci in c:
            fit = fit * xv + ci
        worst = max(worst, abs((uv - fit) / fit))
        total += (uv - fit) ** 2
        k += 1
    return worst, math.sqrt(total), k


def elastic_cube(n_side=41, steps=500, precision="fp64", dt=1e-7):
    """Plain FP64 PMB elastic cube: 41^3 node-centred particles, dx = 1/40,
    delta = 3.01 dx, E = 72 GPa, nu = 0.25, rho = 2440, Gaussian radial
    displacement A exp(-r^2/sigma^2) r_hat (A = 1e-4, sigma = 0.1) about the
    centre (style of examples/mechanics/elastic_wave.cpp:94-95)."""
    dx = 1.0 / (n_side - 1)
    x, vol, _ = grid_node_centred(n_side, dx)
    E, nu = 72e9, 0.25
    K = E / (3.0 * (1.0 - 2.0 * nu))
    desc = make_desc("PMB", fracture=False, precision=precision, horizon=3.01 * dx, K=K, dt=dt)
    w = _base(desc, x, vol, 2440.0, "pmb_elastic_cube_%d" % n_side, dx, steps)
    c = x - 0.5
    r = np.sqrt((c * c).sum(axis=1))
    amp = 1e-4 * np.exp(-r * r / 0.01)
    with np.errstate(invalid="ignore", divide="ignore"):
        rhat = np.where(r[:, None] > 0, c / r[:, None], 0.0)
    w["u"] = amp[:, None] * rhat
    return w


def kw_plate(num_cells=(200, 400, 18), size=(0.1, 0.2, 0.009), E=191e9, nu=0.25, rho=8000.0,
             G0=42408.0, impact_v=16.5, dt=7.5e-8, steps=1000, precision="fp64",
             horizon=None, horizon_extra=0.0, output_every=0, name=None):
    """Pre-notched plate under edge impact (Kalthoff-Winkler geometry of
    examples/mechanics/kalthoff_winkler.cpp:58-127): two notches of length
    height/2 at y = +-width/8 through the thickness, x-velocity between the
    notches on the impact edge, zero-force strip at the impact edge."""
    size = np.asarray(size, dtype=np.float64)
    low, high = -0.5 * size, 0.5 * size
    x, vol, d = grid_cell_centred(low, high, num_cells)
    dx = d[0]
    K = E / (3.0 * (1.0 - 2.0 * nu))
    if horizon is None:
        horizon = 3.01 * dx
    horizon = horizon + horizon_extra
    desc = make_desc("PMB", fracture=True, precision=precision, horizon=horizon, K=K, G0=G0, dt=dt)
    w = _base(desc, x, vol, rho, name or "pmb_kw_plate_%dx%dx%d" % tuple(num_cells), dx, steps, output_every)
    height, width, thickness = size
    y1, y2 = -width / 8.0, width / 8.0
    low_x, low_z = low[0], low[2]
    p0 = np.array([[low_x, y1, low_z], [low_x, y2, low_z]])
    v1 = np.array([[height / 2.0, 0.0, 0.0]] * 2)
    v2 = np.array([[0.0, 0.0, thickness]] * 2)
    w["notches"] = (p0, v1, v2)
    sel = (x[:, 1] > y1) & (x[:, 1] < y2) & (x[:, 0] < -0.5 * height + dx)
    w["v"][sel, 0] = impact_v
    x_bc = -0.5 * height
    lo = np.array([x_bc - dx, y1 - 0.25 * dx, -thickness])
    hi = np.array([x_bc + dx, y2 + 0.25 * dx, thickness])
    inside = np.all((x >= lo) & (x <= hi), axis=1)
    w["bc"] = (PDK_BC_FORCE_VALUE, np.nonzero(inside)[0].astype(np.int64), 0.0)
    return w


def kw_regression(precision="fp64"):
    """regression_test/inputs/kalthoff_winkler.json run through
    examples/mechanics/kalthoff_winkler.cpp: 75x151x7 cells, horizon
    0.002 (+1e-10), nu = 1/3, v0 = 16, dt = 1e-7, tf = 70 us, output every 50."""
    tf, dt = 70e-6, 1e-7
    steps = int(tf / dt)  # Inputs truncation (CabanaPD_Input.hpp:109)
    E, nu = 191e9, 1.0 / 3.0
    return kw_plate(num_cells=(75, 151, 7), size=(0.1, 0.2, 0.009), E=E, nu=nu, rho=8000.0, G0=42408.0,
                    impact_v=16.0, dt=dt, steps=steps, precision=precision, horizon=0.002,
                    horizon_extra=1e-10, output_every=50, name="kw_regression")


def crack_inclusion(precision="fp64"):
    """regression_test/inputs/crack_inclusion.json run through
    examples/mechanics/crack_inclusion.cpp: LPS fracture (influence 0) on
    80x40x4 cells of a 0.04 x 0.02 x 0.001 plate, two materials (plate
    E 72 GPa, G0 3.8; inclusion E 144 GPa, G0 38; nu 0.22, rho 2440) averaged
    across the interface (createMultiForceModel with AverageTag), inclusion
    radius 0.005 at (0.015, 0); horizon 0.001 (+1e-10); one pre-notch of
    length height/2 through the thickness at mid-height from the low-x edge;
    no-fail bands of one horizon inside the loaded rows; traction 2 MPa on the
    low/high y rows as f_y += sign(y) sigma/dy; dt 3.4e-8,
    int(22.5e-6 / 3.4e-8) = 661 steps, output every 50."""
    size = np.array([0.04, 0.02, 0.001])
    low, high = -0.5 * size, 0.5 * size
    x, vol, d = grid_cell_centred(low, high, (80, 40, 4))
    dy = d[1]
    rho0 = 2440.0
    E = (72e9, 144e9)
    nu = (0.22, 0.22)
    G0 = (3.8, 38.0)
    mats = []
    for t in range(2):
        K = E[t] / (3 * (1 - 2 * nu[t]))
        G = E[t] / (2 * (1 + nu[t]))
        mats.append((K, G, G0[t]))
    horizon = 0.001 + 1e-10
    tf, dt = 22.5e-6, 3.4e-8
    steps = int(tf / dt)  # Inputs truncation (CabanaPD_Input.hpp:109)
    desc = make_desc("LPS", fracture=True, precision=precision, horizon=horizon, K=mats[0][0], G=mats[0][1],
                     G0=mats[0][2], dt=dt, influence=0)
    w = _base(desc, x, vol, rho0, "crack_inclusion_80x40x4", d[0], steps, output_every=50)
    height, thickness = size[0], size[2]
    y_notch = 0.5 * (low[1] + high[1])
    w["notches"] = (np.array([[low[0], y_notch, low[2]]]), np.array([[height / 2.0, 0.0, 0.0]]),
                    np.array([[0.0, 0.0, thickness]]))
    # BC planes (crack_inclusion.cpp:98-105) and the no-fail bands (:122-125)
    p1_lo_y, p1_hi_y = low[1] - dy, low[1] + dy
    p2_lo_y, p2_hi_y = high[1] - dy, high[1] + dy
    w["nofail"] = ((x[:, 1] <= p1_lo_y + horizon + 1e-10) | (x[:, 1] >= p2_hi_y - horizon - 1e-10)).astype(np.int32)
    rsq = (x[:, 0] - 0.015) ** 2 + (x[:, 1] - 0.0) ** 2
    mtype = (rsq < 0.005 * 0.005).astype(np.int32)
    w["materials"] = (mats, mtype)
    # the force BC functor f(pid, 1) += b0 * sign(y) on both planes, applied
    # after the force pass every step: a constant per-particle force term
    b0 = 2e6 / dy
    inx = (x[:, 0] >= low[0]) & (x[:, 0] <= high[0]) & (x[:, 2] >= low[2]) & (x[:, 2] <= high[2])
    in1 = inx & (x[:, 1] >= p1_lo_y) & (x[:, 1] <= p1_hi_y)
    in2 = inx & (x[:, 1] >= p2_lo_y) & (x[:, 1] <= p2_hi_y)
    body = np.zeros_like(x)
    sel = in1 | in2
    body[sel, 1] = b0 * (np.abs(x[sel, 1]) / x[sel, 1])
    w["body"] = body
    return w


def dogbone_inside(x, low, high, gage=0.050, width=0.013, fillet=0.076):
    """The ASTM D638 type I specimen of examples/mechanics/dogbone_tensile_test.cpp:71-139:
    outside the gage section's width band and outside the four fillet circles
    (strict comparisons, as there)."""
    midx = 0.5 * (low[0] + high[0])
    midy = 0.5 * (low[1] + high[1])
    rsq = fillet * fillet
    xc_l, xc_r = midx - 0.5 * gage, midx + 0.5 * gage
    yc_b, yc_t = midy - 0.5 * width - fillet, midy + 0.5 * width + fillet
    keep = ~((np.abs(x[:, 0] - midx) < 0.5 * gage) & (np.abs(x[:, 1] - midy) > 0.5 * width))
    for xc, yc in ((xc_l, yc_b), (xc_r, yc_b), (xc_l, yc_t), (xc_r, yc_t)):
        keep &= ~(np.abs(x[:, 0] - xc) * np.abs(x[:, 0] - xc) + np.abs(x[:, 1] - yc) * np.abs(x[:, 1] - yc) < rsq)
    return keep


def dogbone(grip_velocity=0.1, steps=None, num_cells=(251, 29, 5)):
    """examples/mechanics/inputs/dogbone_tensile_test.json run through
    dogbone_tensile_test.cpp: PMB ElasticPerfectlyPlastic with fracture,
    E 3 GPa, nu 0.25 (K = E / (3 (1 - 2 nu))), G0 3000, sigma_y 30 MPa,
    rho 1200, horizon 0.002 (+1e-10), dt 3e-7, final time 0.015
    (int(0.015 / 3e-7) = 50000 steps); 251x29x5 cells of 0.165 x 0.019 x
    0.0032 cut to the dogbone.  Grips (inclusive boxes beyond the grip
    distance 0.115 about the centre): u = (v0 t, 0, 0) on the right grip and
    u = 0 on the left, set after each initial half step (force_update =
    false).  The example seeds the particles randomly within their cells
    (Cabana::InitRandom); here they sit at the cell centres, so the specimen
    is the same shape but not the same particle set."""
    size = np.array([0.165, 0.019, 0.0032])
    low, high = -0.5 * size, 0.5 * size
    x, vol, d = grid_cell_centred(low, high, num_cells)
    keep = dogbone_inside(x, low, high)
    x, vol = x[keep], vol[keep]
    E, nu = 3e9, 0.25
    K = E / (3 * (1 - 2 * nu))
    horizon = 0.002 + 1e-10
    dt = 3.0e-7
    if steps is None:
        steps = int(1.5e-2 / dt)
    desc = make_desc("PMB", fracture=True, horizon=horizon, K=K, G0=3000.0, dt=dt, mechanics="epp",
                     sigma_y=30e6)
    w = _base(desc, x, vol, 1200.0, "dogbone_%dx%dx%d" % tuple(num_cells), d[0], steps, output_every=200)
    midx = 0.5 * (low[0] + high[0])
    D = 0.115
    right = (x[:, 0] >= midx + 0.5 * D) & (x[:, 0] <= high[0])
    left = (x[:, 0] >= low[0]) & (x[:, 0] <= midx - 0.5 * D)
    grips = np.flatnonzero(right | left)
    a = np.zeros((grips.size, 3))
    b = np.zeros((grips.size, 3))
    b[right[grips], 0] = grip_velocity
    w["prescribed"] = [("u", grips, a, b)]
    w["right_grip"] = np.flatnonzero(right)
    return w


def thermal_deformation(steps=None):
    """examples/thermomechanics/inputs/thermal_deformation.json run through
    thermal_deformation.cpp: PMB NoFracture TemperatureDependent, E 370 GPa,
    nu 0.25, rho 3980, alpha 7.5e-6, temp0 20, horizon 0.03 (+1e-10),
    dt 7.5e-7, final time 0.01 (int(0.01 / 7.5e-7) = 13333 steps); 101x31x3
    cells of 1.0 x 0.3 x 0.03.  Temperature imposed on every particle before
    each force pass (BodyTerm with force_update = false):
    T = temp0 + 5000 (y - y_low) t."""
    size = np.array([1.0, 0.3, 0.03])
    low, high = -0.5 * size, 0.5 * size
    x, vol, d = grid_cell_centred(low, high, (101, 31, 3))
    E, nu = 370e9, 0.25
    K = E / (3 * (1 - 2 * nu))
    horizon = 0.03 + 1e-10
    dt = 7.5e-7
    if steps is None:
        steps = int(0.01 / dt)
    temp0 = 20.0
    desc = make_desc("PMB", horizon=horizon, K=K, dt=dt, thermal=True, alpha=7.5e-6, temp0=temp0)
    w = _base(desc, x, vol, 3980.0, "thermal_deformation_101x31x3", d[0], steps, output_every=100)
    n = x.shape[0]
    w["prescribed"] = [("temperature", np.arange(n), np.full(n, temp0), 5000.0 * (x[:, 1] - low[1]))]
    return w


def thermal_deformation_heat_transfer(steps=None, thermal_subcycle=None, num_cells=(49, 49, 49)):
    """examples/thermomechanics/inputs/thermal_deformation_heat_transfer.json run
    through thermal_deformation_heat_transfer.cpp: PMB NoFracture
    DynamicTemperature, E 115 GPa, nu 0.25, rho 8915, alpha 17e-6, kappa 387,
    cp 385, temp0 100 (also the initial temperature), horizon 0.00615
    (+1e-10), dt 4.2e-7, heat conduction every 45000 steps (final time 8 s);
    49^3 cells of 0.1^3.  The temperature is held at 0 on the two planes
    within one spacing of the low and high y faces (a force_update = false
    functor before each force pass)."""
    size = np.array([0.1, 0.1, 0.1])
    low, high = -0.5 * size, 0.5 * size
    x, vol, d = grid_cell_centred(low, high, num_cells)
    E, nu = 115e9, 0.25
    K = E / (3 * (1 - 2 * nu))
    horizon = 0.00615 + 1e-10
    dt = 4.2e-7
    if steps is None:
        steps = int(8.0 / dt)
    if thermal_subcycle is None:
        thermal_subcycle = 45000
    temp0 = 100.0
    desc = make_desc("PMB", horizon=horizon, K=K, dt=dt, thermal="dynamic", alpha=17e-6, temp0=temp0, kappa=387.0,
                     cp=385.0, thermal_subcycle=thermal_subcycle)
    w = _base(desc, x, vol, 8915.0, "thermal_heat_transfer_%dx%dx%d" % tuple(num_cells), d[0], steps,
              output_every=10)
    n = x.shape[0]
    w["temperature"] = np.full(n, temp0)
    dy = d[1]
    full = (x[:, 0] >= low[0]) & (x[:, 0] <= high[0]) & (x[:, 2] >= low[2]) & (x[:, 2] <= high[2])
    plane1 = full & (x[:, 1] >= high[1] - dy) & (x[:, 1] <= high[1] + dy)
    plane2 = full & (x[:, 1] >= low[1] - dy) & (x[:, 1] <= low[1] + dy)
    cold = np.flatnonzero(plane1 | plane2)
    w["prescribed"] = [("temperature", cold, np.zeros(cold.size), None)]
    return w


def kw_crack_y(x, phi):
    """Output time series value of kalthoff_winkler.cpp:134-145: max over
    particles of (phi > 0.3 ? y : 0)."""
    vals = np.where(phi > 0.3, x[:, 1], 0.0)
    return float(vals.max()) if vals.size else 0.0


def lps_cube(n_side=128, steps=500, precision="fp64", seed=44758454, dt=None, sigma=1e7, G0=3.8,
             fracture=True):
    """LPS elastic-brittle cube: n^3 node-centred particles, dx = 1/(n-1),
    delta = 3.01 dx, E = 72 GPa, nu = 0.22, rho = 2440, influence 0, seeded
    uniform perturbation |u_d| <= 1e-3 dx / sqrt(3), opposite-face traction
    +-sigma/dx on the y faces (style of examples/mechanics/crack_branching.cpp:120-132)."""
    dx = 1.0 / (n_side - 1)
    x, vol, idx = grid_node_centred(n_side, dx)
    E, nu = 72e9, 0.22
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    horizon = 3.01 * dx
    if dt is None:
        from .engine import critical_timestep
        try:
            dt = critical_timestep(dx, dx, dx, K, 2440.0, horizon, 0.85)
        except Exception:
            dt = 1.13e-6 * (127.0 * dx)
    desc = make_desc("LPS", fracture=fracture, precision=precision, horizon=horizon, K=K, G=G, G0=G0, dt=dt,
                     influence=0)
    w = _base(desc, x, vol, 2440.0, "lps_cube_%d" % n_side, dx, steps)
    rng = np.random.default_rng(seed)
    a = 1e-3 * dx / math.sqrt(3.0)
    w["u"] = rng.uniform(-a, a, size=x.shape)
    body = np.zeros_like(x)
    jy = idx[:, 1]
    body[jy == 0, 1] = -sigma / dx
    body[jy == n_side - 1, 1] = sigma / dx
    w["body"] = body
    return w


def serving_request(rng, dx=5e-4, rho=8000.0, nu=0.25, precision="mixed", n_range=(2e4, 2e6), steps_range=(200, 2000)):
    """One specimen of the serving stream: a rectangular PMB fracture plate
    with log-uniform particle count, aspect ratio U[1,4], thickness U[3,20]
    cells, E ~ U[50,210] GPa, G0 ~ logU[1, 5e4] J/m^2, 0-3 random straight
    pre-notches (placement as in examples/mechanics/random_cracks.cpp:74-114),
    an x-velocity U[1,30] m/s on the low-x edge and U[200,2000] steps.
    dx, rho and nu are not drawn (the plate values of the fracture benchmark)."""
    n_target = math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))
    aspect = rng.uniform(1.0, 4.0)
    nz = int(rng.integers(3, 21))
    nx = max(8, int(round(math.sqrt(n_target / (nz * aspect)))))
    ny = max(8, int(round(aspect * nx)))
    size = (nx * dx, ny * dx, nz * dx)
    E = rng.uniform(50e9, 210e9)
    G0 = math.exp(rng.uniform(math.log(1.0), math.log(5e4)))
    K = E / (3.0 * (1.0 - 2.0 * nu))
    horizon = 3.01 * dx
    from .engine import critical_timestep
    dt = critical_timestep(dx, dx, dx, K, rho, horizon, 0.85)
    low = -0.5 * np.asarray(size)
    high = 0.5 * np.asarray(size)
    x, vol, d = grid_cell_centred(low, high, (nx, ny, nz))
    desc = make_desc("PMB", fracture=True, precision=precision, horizon=horizon, K=K, G0=G0, dt=dt)
    steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
    w = _base(desc, x, vol, rho, "serving_%dx%dx%d" % (nx, ny, nz), dx, steps)
    nnotch = int(rng.integers(0, 4))
    if nnotch:
        thickness = size[2]
        maxl = 0.25 * min(size[0], size[1])
        minl = 0.5 * maxl
        p0, v1, v2 = [], [], []
        for _ in range(nnotch):
            rx, ry = rng.uniform(), rng.uniform()
            xc = (low[0] + maxl) + ((high[0] - maxl) - (low[0] + maxl)) * rx
            yc = (low[1] + maxl) + ((high[1] - maxl) - (low[1] + maxl)) * ry
            length = minl + (maxl - minl) * rng.uniform()
            theta = math.pi * rng.uniform()
            p0.append([xc, yc, low[2]])
            v1.append([length * math.cos(theta), length * math.sin(theta), 0.0])
            v2.append([0.0, (-1.0 + 2.0 * rng.uniform()) * thickness, thickness])
        w["notches"] = (np.array(p0), np.array(v1), np.array(v2))
    v0 = rng.uniform(1.0, 30.0)
    w["v"][x[:, 0] < low[0] + dx, 0] = v0
    return w


def dem_packing(n=1000, seed=7, model="HERTZ_JKR", polydisperse=True, dt=1.5e-7, steps=2000,
                rebuild_every=20, speed=50.0, r_mean=1e-3, sigma=0.1, clip=0.3, E=70e9, nu=0.25, rho=2600.0,
                restitution=0.5, gamma=0.05, extend_frac=0.1, gravity=True):
    """Spheres on a jittered cubic lattice (non-overlapping: each centre stays
    at least its radius inside its own cell), log-normal radii clipped to
    +-clip around r_mean, random velocities of magnitude ~speed, gravity
    f_z -= 9.8 rho (examples/dem/powder_fill.cpp:138-142)."""
    rng = np.random.default_rng(seed)
    if polydisperse:
        r = np.exp(rng.normal(math.log(r_mean), sigma, size=n))
        r = np.clip(r, (1 - clip) * r_mean, (1 + clip) * r_mean)
    else:
        r = np.full(n, r_mean)
    a = 2.05 * r.max()
    side = int(math.ceil(n ** (1.0 / 3.0)))
    cells = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij"),
                     axis=-1).reshape(-1, 3)[:n, ::-1].astype(np.float64)
    slack = (a / 2.0 - r)[:, None]
    x = (cells + 0.5) * a + rng.uniform(-1.0, 1.0, size=(n, 3)) * slack
    vol = 4.0 / 3.0 * math.pi * r ** 3
    desc = make_desc(model, radius=r_mean, radius_extend=extend_frac * r_mean, nu=nu, E=E,
                     restitution=restitution, gamma=gamma, mu_friction=0.0, dt=dt, rebuild_every=rebuild_every,
                     polydisperse=polydisperse)
    w = _base(desc, x, vol, rho, "dem_%s_%d" % (model.lower(), n), a, steps)
    w["v"] = rng.normal(0.0, speed / math.sqrt(3.0), size=(n, 3))
    w["radius"] = r if polydisperse else None
    if gravity:
        b = np.zeros((n, 3))
        b[:, 2] = -9.8 * rho
        w["body"] = b
    return w


def dem_fcc(n=1000000, seed=7, model="HERTZ_JKR", dt=1.5e-7, steps=2000, rebuild_every=20, speed=1.0,
            r_mean=1e-3, sigma=0.1, spacing=2.3, E=70e9, nu=0.25, rho=2600.0, restitution=0.5, gamma=0.05,
            extend_frac=0.1, gravity=True):
    """Dense granular box: n spheres on a face-centred cubic lattice with
    nearest-neighbour distance spacing * r_mean, radii log-normal (sigma) about
    r_mean clipped to [0.7, spacing / 2 * 0.999] r_mean so that no two spheres
    overlap (solid fraction about 0.50 for spacing 2.3), seeded random
    velocities of magnitude ~speed, gravity f_z -= 9.8 rho."""
    rng = np.random.default_rng(seed)
    r = np.exp(rng.normal(math.log(r_mean), sigma, size=n))
    r = np.clip(r, 0.7 * r_mean, 0.5 * spacing * 0.999 * r_mean)
    d = spacing * r_mean
    a = d * math.sqrt(2.0)  # cubic cell of the FCC lattice
    m = int(math.ceil((n / 4.0) ** (1.0 / 3.0)))
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    cz, cy, cx = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    corners = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1).astype(np.float64)
    x = ((corners[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a)[:n]
    vol = 4.0 / 3.0 * math.pi * r ** 3
    desc = make_desc(model, radius=r_mean, radius_extend=extend_frac * r_mean, nu=nu, E=E,
                     restitution=restitution, gamma=gamma, mu_friction=0.0, dt=dt, rebuild_every=rebuild_every,
                     polydisperse=True)
    w = _base(desc, x, vol, rho, "dem_fcc_%s_%d" % (model.lower(), n), d, steps)
    w["v"] = rng.normal(0.0, speed / math.sqrt(3.0), size=(n, 3))
    w["radius"] = r
    if gravity:
        b = np.zeros((n, 3))
        b[:, 2] = -9.8 * rho
        w["body"] = b
    box = np.prod(x.max(0) - x.min(0) + d)
    w["solid_fraction"] = float(vol.sum() / box)
    return w


def dem_packing_radii(n, seed, r_mean=1e-3, sigma=0.1, clip=0.3):
    """Radii exp(N(ln r_mean, sigma^2)) clipped to [1 - clip, 1 + clip] r_mean."""
    rng = np.random.default_rng(seed)
    r = np.exp(rng.normal(math.log(r_mean), sigma, size=n))
    return np.clip(r, (1.0 - clip) * r_mean, (1.0 + clip) * r_mean), rng


def contact_overlaps(spec, y, r):
    """Pairs (i, j) of a contact spec's list and their overlaps R_i + R_j - |y_j - y_i|."""
    off, nbr, _ = spec.bonds()
    i = np.repeat(np.arange(off.size - 1), np.diff(off))
    d = np.linalg.norm(y[nbr] - y[i], axis=1)
    return i, nbr, r[i] + r[nbr] - d


def dem_packing_walls(L, r_wall, spacing=1.8):
    """Container of frozen spheres for a cube [0, L]^3: on each face a square
    sheet of radius r_wall spheres, centre spacing spacing * r_wall (the gap at
    the centre of four is 0.27 r_wall at 1.8, so no sphere of the packing can
    pass), centres r_wall outside the face (the frozen container particles of
    examples/dem/powder_fill.cpp:63-90)."""
    k = int(math.ceil(L / (spacing * r_wall))) + 1
    g = np.linspace(0.0, L, k)
    g1, g2 = np.meshgrid(g, g, indexing="ij")
    sheets = []
    for d in range(3):
        for outer in (False, True):
            p = np.zeros((k * k, 3))
            o = [q for q in range(3) if q != d]
            p[:, o[0]] = g1.ravel()
            p[:, o[1]] = g2.ravel()
            p[:, d] = L + r_wall if outer else -r_wall
            sheets.append(p)
    return np.concatenate(sheets)


def dem_random_packing(ctx, n=1000000, fraction=0.55, seed=20261203, r_mean=1e-3, sigma=0.1, clip=0.3,
                       stage_step=0.005, relax_steps=300, gas_steps=1500, gas_speed=20.0, E=70e9, nu=0.25,
                       rho=2600.0, dt=1.5e-7, log=None):
    """A seeded random packing of n spheres in a cube at the given solid
    fraction (counted over the box that holds every sphere), radii log-normal
    about r_mean clipped to +-clip, no two spheres overlapping.

    Random sequential addition jams near 0.38, so the packing is compacted on
    this engine, inside a container of frozen sphere sheets (dem_packing_walls,
    held by a zero-force boundary condition):
      1. the spheres sit on randomly chosen sites of a face-centred cubic
         lattice whose neighbour distance exceeds twice the largest radius
         (no overlap), with a seeded jitter;
      2. the lattice is melted: gas_steps steps of undamped Hertz contact
         (restitution 1) from random directions at gas_speed;
      3. stages of affine compression of every centre and wall about the box
         centre, each raising the container's solid fraction by stage_step,
         with the velocities cleared and then relax_steps steps of damped
         Hertz contact (restitution 0.1), until the spheres' own bounding box
         holds the target fraction after the residual overlaps (a few 1e-3 of
         a radius) are removed by one common scale of the radii.
    Returns (positions, radii, achieved fraction, stages, wall positions,
    wall radius)."""
    from . import engine
    r, rng = dem_packing_radii(n, seed, r_mean, sigma, clip)
    rmax = float(r.max())
    m = int(math.ceil((n / 4.0) ** (1.0 / 3.0)))
    a = 2.0 * rmax * 1.03 * math.sqrt(2.0)  # FCC neighbour distance a / sqrt(2) = 2.06 rmax
    basis = np.array([[0.25, 0.25, 0.25], [0.75, 0.75, 0.25], [0.75, 0.25, 0.75], [0.25, 0.75, 0.75]])
    cz, cy, cx = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    corners = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1).astype(np.float64)
    sites = (corners[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    xi = sites[np.sort(rng.choice(sites.shape[0], size=n, replace=False))]
    xi = xi + rng.uniform(-0.01 * a, 0.01 * a, size=xi.shape)
    xi = xi[rng.permutation(n)]
    L = m * a
    xw = dem_packing_walls(L, r_mean)
    x = np.concatenate([xi, xw])
    rr = np.concatenate([r, np.full(xw.shape[0], r_mean)])
    N = x.shape[0]
    vol = 4.0 / 3.0 * math.pi * rr ** 3
    vol_in = float(vol[:n].sum())
    walls = np.arange(n, N)

    def spec(restitution, u, v):
        desc = make_desc("HERTZ", radius=r_mean, radius_extend=0.1 * r_mean, nu=nu, E=E, restitution=restitution,
                         dt=dt, rebuild_every=10, polydisperse=True)
        s = engine.Spec(ctx, desc, N)
        s.upload(x, vol, u=u, v=v, rho=np.full(N, rho), radius=rr)
        s.build_bonds()
        s.set_bc(PDK_BC_FORCE_VALUE, walls, 0.0)
        s.init()
        return s

    # 2. melt the lattice
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    v0 = np.zeros_like(x)
    v0[:n] = gas_speed * dirs
    s = spec(1.0, None, v0)
    s.step(gas_steps)
    u = s.get("u")
    s.close()
    if log:
        log(f"  packing: {n} spheres, {xw.shape[0]} wall spheres, lattice melted over {gas_steps} steps")
    # 3. compaction
    s = spec(0.1, u, None)
    y = x + u
    centre = np.full(3, 0.5 * L)
    side = L
    stages = 0
    final = None
    zero_v = np.zeros_like(x)
    while stages < 400:
        stages += 1
        f_box = vol_in / side ** 3
        lam = (f_box / (f_box + stage_step)) ** (1.0 / 3.0)
        side *= lam
        y = centre + lam * (y - centre)
        s.set_state(u=y - x, v=zero_v)
        s.step(relax_steps)
        y = x + s.get("u")
        yi = y[:n]
        lo = (yi - r[:, None]).min(axis=0)
        hi = (yi + r[:, None]).max(axis=0)
        frac = vol_in / float(np.prod(hi - lo))
        if log and (stages % 10 == 0 or frac >= fraction):
            log(f"  packing stage {stages}: container fraction {f_box + stage_step:.4f} bounding-box fraction "
                f"{frac:.4f}")
        if frac < fraction:
            continue
        i, j, ov = contact_overlaps(s, y, rr)
        keep = (i < n) & (j < n)
        worst = float(np.max(ov[keep] / (rr[i[keep]] + rr[j[keep]]), initial=0.0))
        scale = 1.0 - worst * (1.0 + 1e-6) if worst > 0 else 1.0
        lo = (yi - scale * r[:, None]).min(axis=0)
        hi = (yi + scale * r[:, None]).max(axis=0)
        frac = vol_in * scale ** 3 / float(np.prod(hi - lo))
        if log:
            log(f"  packing stage {stages}: worst overlap {worst:.2e} of R_i + R_j, fraction after scaling {frac:.4f}")
        if frac >= fraction:
            final = (yi.copy(), r * scale, frac, stages, y[n:].copy(), r_mean)
            break
    s.close()
    if final is None:
        raise RuntimeError("packing did not reach the solid fraction")
    return final


def dem_box(ctx, n=1000000, fraction=0.55, seed=20261203, model="HERTZ_JKR", dt=1.5e-7, steps=2000,
            rebuild_every=20, r_mean=1e-3, E=70e9, nu=0.25, rho=2600.0, restitution=0.5, gamma=0.05,
            extend_frac=0.1, log=None):
    """The granular box: the n spheres of dem_random_packing at the solid
    fraction, at rest, inside its container of frozen wall spheres (the frozen
    container particles of examples/dem/powder_fill.cpp:63-90, held by a
    zero-force boundary condition; particles n.. of the workload), Hertz-JKR
    (E 70 GPa, nu 0.25, rho 2600, restitution 0.5, surface energy 0.05 J/m^2,
    no friction), gravity f_z -= 9.8 rho on the spheres
    (examples/dem/powder_fill.cpp:138-142), contact list rebuilt every
    rebuild_every steps."""
    y, r, frac, stages, yw, rw = dem_random_packing(ctx, n=n, fraction=fraction, seed=seed, r_mean=r_mean, E=E,
                                                    nu=nu, rho=rho, log=log)
    x = np.concatenate([y, yw])
    rad = np.concatenate([r, np.full(yw.shape[0], rw)])
    N = x.shape[0]
    vol = 4.0 / 3.0 * math.pi * rad ** 3
    desc = make_desc(model, radius=r_mean, radius_extend=extend_frac * r_mean, nu=nu, E=E,
                     restitution=restitution, gamma=gamma, mu_friction=0.0, dt=dt, rebuild_every=rebuild_every,
                     polydisperse=True)
    w = _base(desc, x, vol, rho, "dem_box_%s_%d" % (model.lower(), n), 2.0 * r_mean, steps)
    w["radius"] = rad
    b = np.zeros((N, 3))
    b[:n, 2] = -9.8 * rho
    w["body"] = b
    w["bc"] = (PDK_BC_FORCE_VALUE, np.arange(n, N, dtype=np.int64), 0.0)
    w["spheres"] = n
    w["wall_spheres"] = N - n
    w["solid_fraction"] = frac
    w["packing_stages"] = stages
    return w


def lattice_sum_cube(model="PMB", m=3, nc=11, field="linear", coeff=0.1, K=1.0, G=0.5, influence=None,
                     fracture=False, G0=1000.0):
    """Known-answer lattice of unit_test/tstForce.hpp:593-661: cell-centred
    particles on [-1, 1]^3 with nc cells per side, horizon = m dx, u = s0 x
    (linear) or u_x = u11 x^2 (quadratic)."""
    x, vol, d = grid_cell_centred([-1.0, -1.0, -1.0], [1.0, 1.0, 1.0], (nc, nc, nc))
    dx = 2.0 / nc
    horizon = dx * m
    kw = dict(horizon=horizon, K=K, dt=1e-7, fracture=fracture)
    if fracture:
        kw["G0"] = G0
    if model in ("LPS", "LINEAR_LPS"):
        kw["G"] = G
    desc = make_desc(model, influence=influence, **kw)
    w = _base(desc, x, vol, 1.0, "lattice_%s_%s" % (model, field), dx, 0)
    u = np.zeros_like(x)
    if field == "linear":
        u = coeff * x
    else:
        u[:, 0] = coeff * x[:, 0] * x[:, 0]
    w["u"] = u
    return w


def colliding_blocks(n_cells=(8, 8, 8), gap_cells=3.5, speed=150.0, steps=200, dx=1e-3, K=130e9, rho=7800.0,
                     critical_stretch=0.004, contact_factor=0.9, extend_factor=0.01, precision="fp64"):
    """Two PMB fracture blocks of n_cells each, gap_cells cell spacings apart
    along x (more than the horizon, so no bond joins them), moving toward
    each other at +-speed: they interact only through the contact model on
    the PD specimen (NormalRepulsion with radius contact_factor dx / 2 and
    list skin extend_factor dx, the contact parameters of
    examples/mechanics/inputs/fragmenting_cylinder.json), and bonds at the
    impact faces break.  G0 = 9 K delta s_c^2 / 5 (the PMB relation for a
    critical stretch s_c)."""
    nx, ny, nz = n_cells
    lx = nx * dx
    xa, vola, _ = grid_cell_centred([-lx - 0.5 * gap_cells * dx, 0.0, 0.0],
                                    [-0.5 * gap_cells * dx, ny * dx, nz * dx], n_cells)
    xb, volb, _ = grid_cell_centred([0.5 * gap_cells * dx, 0.0, 0.0],
                                    [lx + 0.5 * gap_cells * dx, ny * dx, nz * dx], n_cells)
    x = np.concatenate([xa, xb])
    vol = np.concatenate([vola, volb])
    horizon = 3.01 * dx
    G0 = 9.0 * K * horizon * (critical_stretch * critical_stretch) / 5.0
    from .engine import critical_timestep
    try:
        dt = critical_timestep(dx, dx, dx, K, rho, horizon, 0.7)
    except Exception:
        dt = 1.0e-7
    desc = make_desc("PMB", fracture=True, precision=precision, horizon=horizon, K=K, G0=G0, dt=dt)
    w = _base(desc, x, vol, rho, "blocks_%dx%dx%d" % n_cells, dx, steps)
    w["v"][: xa.shape[0], 0] = speed
    w["v"][xa.shape[0]:, 0] = -speed
    w["contact"] = _abi.make_contact(contact_factor * dx / 2.0, extend_factor * dx, horizon, K)
    return w


def upload(spec, w):
    """Send a workload to an engine spec (or an oracle spec)."""
    spec.upload(w["x"], w["vol"], u=w["u"], v=w["v"], rho=w["rho"], nofail=w["nofail"], radius=w["radius"])


def setup(spec, w, init=True):
    """Build bonds, prenotch, BC/body term and Solver::init for a workload."""
    spec.build_bonds()
    if w.get("materials") is not None:
        spec.set_materials(*w["materials"])
    if w.get("pair_models") is not None:
        spec.set_pair_models(*w["pair_models"])
    if w["notches"] is not None:
        spec.prenotch(*w["notches"])
    if w["bc"] is not None:
        spec.set_bc(*w["bc"])
    if w["body"] is not None:
        spec.set_body_force(w["body"])
    if w.get("temperature") is not None:
        spec.set_field("temperature", w["temperature"])
    for field, idx, a, b in w.get("prescribed", []):
        spec.set_prescribed(field, idx, a, b)
    if w.get("contact") is not None:
        spec.set_contact(w["contact"])
    if init:
        spec.init()
