"""Lattice known answers of unit_test/tstForce.hpp, restated in numpy.

These are the reference test suite's own single-particle sums over a full
lattice neighborhood (tstForce.hpp:93-524) and its closed forms
(tstForce.hpp:779-886).  They are independent of the oracle's neighbor list
and loops: a wrong cutoff, a dropped bond term or a wrong model constant in
the oracle (or the engine) shows up as a mismatch.
"""
import math

import numpy as np

PI = math.pi


def offsets(m, dx, horizon):
    r = np.arange(-m, m + 1)
    k, j, i = np.meshgrid(r, r, r, indexing="ij")
    xi = np.stack([i.ravel() * dx, j.ravel() * dx, k.ravel() * dx], axis=1)
    mag = np.sqrt((xi * xi).sum(axis=1))
    keep = (mag > 0.0) & (mag < horizon + 1e-14)
    return xi[keep], mag[keep]


def pmb_c(K, horizon):
    return 18.0 * K / (PI * horizon ** 4)


def ref_neighbors(horizon, m):
    """computeReferenceNeighbors (tstForce.hpp:324-342)."""
    dx = horizon / m
    return offsets(m, dx, horizon)[0].shape[0]


def pmb_stress_linear(K, horizon, m, s0):
    """computeReferenceStress, LinearTag, PMB (tstForce.hpp:91-126): diagonal
    of the stress tensor, sum over the lattice neighborhood of 0.5 f_a xi_a."""
    dx = horizon / m
    vol = dx ** 3
    xiv, xi = offsets(m, dx, horizon)
    f_mag = pmb_c(K, horizon) * s0 * vol
    return [float(np.sum(0.5 * (f_mag * xiv[:, d] / xi) * xiv[:, d])) for d in range(3)]


def pmb_W_linear(K, horizon, m, s0):
    """computeReferenceStrainEnergyDensity, LinearTag, PMB (tstForce.hpp:128-155)."""
    dx = horizon / m
    vol = dx ** 3
    _, xi = offsets(m, dx, horizon)
    c = pmb_c(K, horizon)
    return float(np.sum(0.25 * c * s0 * s0 * xi * vol))


def _quad_bonds(m, dx, horizon, u11, x):
    r = np.arange(-m, m + 1)
    k, j, i = np.meshgrid(r, r, r, indexing="ij")
    xi_x, xi_y, xi_z = i.ravel() * dx, j.ravel() * dx, k.ravel() * dx
    eta_u = u11 * (2 * x * xi_x + xi_x * xi_x)
    rx = xi_x + eta_u
    rr = np.sqrt(rx * rx + xi_y * xi_y + xi_z * xi_z)
    xi = np.sqrt(xi_x * xi_x + xi_y * xi_y + xi_z * xi_z)
    keep = (xi > 0.0) & (xi < horizon + 1e-14)
    s = np.zeros_like(xi)
    s[keep] = (rr[keep] - xi[keep]) / xi[keep]
    return keep, xi_x, rx, rr, xi, s


def pmb_W_quadratic(K, horizon, m, u11, x):
    """tstForce.hpp:157-190."""
    dx = horizon / m
    vol = dx ** 3
    keep, _, _, _, xi, s = _quad_bonds(m, dx, horizon, u11, x)
    c = pmb_c(K, horizon)
    return float(np.sum(0.25 * c * s[keep] ** 2 * xi[keep] * vol))


def pmb_fx_quadratic(K, horizon, m, u11, x):
    """computeReferenceForceX, QuadraticTag, PMB (tstForce.hpp:202-236)."""
    dx = horizon / m
    vol = dx ** 3
    keep, _, rx, rr, _, s = _quad_bonds(m, dx, horizon, u11, x)
    c = pmb_c(K, horizon)
    return float(np.sum(c * s[keep] * vol * rx[keep] / rr[keep]))


def influence(xi, influence_type):
    return 1.0 / xi if influence_type == 1 else np.ones_like(xi)


def lps_weighted_volume(horizon, m, vol, influence_type):
    """tstForce.hpp:238-262."""
    dx = horizon / m
    _, xi = offsets(m, dx, horizon)
    return float(np.sum(influence(xi, influence_type) * xi * xi * vol))


def lps_theta_linear(horizon, m, s0, vol, wv, influence_type):
    """tstForce.hpp:264-287."""
    dx = horizon / m
    _, xi = offsets(m, dx, horizon)
    return float(np.sum(3.0 / wv * influence(xi, influence_type) * s0 * xi * xi * vol))


def lps_theta_quadratic(horizon, m, u11, vol, wv, x, influence_type):
    """tstForce.hpp:289-321."""
    dx = horizon / m
    keep, _, _, _, xi, s = _quad_bonds(m, dx, horizon, u11, x)
    return float(np.sum(3.0 / wv * influence(xi[keep], influence_type) * s[keep] * xi[keep] ** 2 * vol))


def lps_coeffs(K, G):
    return 3.0 * K - 5.0 * G, 15.0 * G


def lps_W_linear(K, G, horizon, m, s0, influence_type):
    """tstForce.hpp:386-422."""
    dx = horizon / m
    vol = dx ** 3
    wv = lps_weighted_volume(horizon, m, vol, influence_type)
    theta = lps_theta_linear(horizon, m, s0, vol, wv, influence_type)
    nn = ref_neighbors(horizon, m)
    tc, sc = lps_coeffs(K, G)
    _, xi = offsets(m, dx, horizon)
    infl = influence(xi, influence_type)
    w = 1.0 / nn * 0.5 * tc / 3.0 * theta * theta + 0.5 * (sc / wv) * infl * s0 * s0 * xi * xi * vol
    return float(np.sum(w))


def lps_W_quadratic(K, G, horizon, m, u11, x, influence_type):
    """tstForce.hpp:424-464."""
    dx = horizon / m
    vol = dx ** 3
    wv = lps_weighted_volume(horizon, m, vol, influence_type)
    nn = ref_neighbors(horizon, m)
    theta_i = lps_theta_quadratic(horizon, m, u11, vol, wv, x, influence_type)
    tc, sc = lps_coeffs(K, G)
    keep, _, _, _, xi, s = _quad_bonds(m, dx, horizon, u11, x)
    infl = influence(xi[keep], influence_type)
    w = 1.0 / nn * 0.5 * tc / 3.0 * theta_i * theta_i + 0.5 * (sc / wv) * infl * s[keep] ** 2 * xi[keep] ** 2 * vol
    return float(np.sum(w))


def lps_fx_quadratic(K, G, horizon, m, u11, x, influence_type):
    """tstForce.hpp:466-507."""
    dx = horizon / m
    vol = dx ** 3
    wv = lps_weighted_volume(horizon, m, vol, influence_type)
    theta_i = lps_theta_quadratic(horizon, m, u11, vol, wv, x, influence_type)
    tc, sc = lps_coeffs(K, G)
    keep, xi_x, rx, rr, xi, s = _quad_bonds(m, dx, horizon, u11, x)
    fx = 0.0
    for b in np.nonzero(keep)[0]:
        theta_j = lps_theta_quadratic(horizon, m, u11, vol, wv, x + xi_x[b], influence_type)
        inf_b = 1.0 / xi[b] if influence_type == 1 else 1.0
        coeff = (tc * theta_i / wv + tc * theta_j / wv + s[b] * (sc / wv + sc / wv)) * inf_b * xi[b] * vol
        fx += coeff * rx[b] / rr[b]
    return fx


def lps_stress_linear(K, G, horizon, m, s0, influence_type):
    """computeReferenceStress, LinearTag, LPS (tstForce.hpp:345-379): constant
    weighted volume and dilatation on the full lattice neighborhood."""
    dx = horizon / m
    vol = dx ** 3
    wv = lps_weighted_volume(horizon, m, vol, influence_type)
    theta = lps_theta_linear(horizon, m, s0, vol, wv, influence_type)
    tc, sc = lps_coeffs(K, G)
    xiv, xi = offsets(m, dx, horizon)
    f_mag = (tc * theta / wv + tc * theta / wv + s0 * (sc / wv + sc / wv)) * influence(xi, influence_type) * xi * vol
    return [float(np.sum(0.5 * (f_mag * xiv[:, d] / xi) * xiv[:, d])) for d in range(3)]
