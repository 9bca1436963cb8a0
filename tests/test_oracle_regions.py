"""Pins of the region-output restatement (oracle/regions.py) by closed forms
and independent counts: Region<Line> on the elastic-wave mesh, inclusive box
faces on an exact lattice, the cylinder's annulus counted in integers, the
Kokkos reducer identities and NaN handling, and the gated crack-extent
functor.  The Kalthoff-Winkler observable through these functions is pinned
by the reference's regression value in test_oracle_regression.py."""
import sys

import numpy as np

from cabanapd_amd import _abi, configs
from oracle import regions


def test_line_region_on_elastic_wave_mesh():
    # 41 cells per side, cell-centred on [-0.5, 0.5]: the members are the
    # particles of the middle cell index in y and z (centre within rounding of 0)
    x, _, d = configs.grid_cell_centred([-0.5] * 3, [0.5] * 3, (41, 41, 41))
    ids = regions.members(x, _abi.line_region(0, d))
    assert ids.size == 41
    k = np.rint((x + 0.5) / d - 0.5).astype(int)
    assert np.all(k[ids, 1] == 20) and np.all(k[ids, 2] == 20)
    assert np.array_equal(np.sort(k[ids, 0]), np.arange(41))
    assert np.all(np.diff(x[ids, 0]) > 0)
    assert np.array_equal(ids, configs.line_profile(x, d[0]))


def test_line_region_half_width_is_strict():
    x, _, _ = configs.grid_node_centred(9, 0.125)
    x = x - 0.5  # coordinates k/8 - 1/2, exact
    ids = regions.members(x, _abi.line_region(2, [0.25, 0.25, 0.25]))  # half width 0.125 = one spacing
    assert ids.size == 9  # the neighbouring lines at +-0.125 are excluded
    assert np.all(x[ids, 0] == 0.0) and np.all(x[ids, 1] == 0.0)


def test_box_faces_are_inclusive():
    x, _, _ = configs.grid_node_centred(9, 0.125)
    ids = regions.members(x, _abi.box_region([0.25, 0.0, 0.5], [0.625, 0.375, 0.5]))
    assert ids.size == 4 * 4 * 1
    assert np.array_equal(ids, np.sort(ids))


def test_cylinder_annulus_count():
    x, _, _ = configs.grid_node_centred(13, 1.0)
    x = x - np.array([6.0, 6.0, 0.0])
    r = _abi.cylinder_region(2.0, 5.0, 0.0, 3.0)
    ids = regions.members(x, r)
    ring = sum(1 for i in range(-6, 7) for j in range(-6, 7) if 4 <= i * i + j * j <= 25)
    assert ids.size == ring * 4
    off = _abi.cylinder_region(0.0, 1.0, 0.0, 0.0, x_center=2.0, y_center=-3.0)
    ids = regions.members(x, off)
    assert ids.size == 5
    assert set(map(tuple, x[ids, :2].tolist())) == {(2, -3), (1, -3), (3, -3), (2, -2), (2, -4)}


def test_reducer_identities_and_nan():
    assert regions.reduce([], "max") == -sys.float_info.max
    assert regions.reduce([], "min") == sys.float_info.max
    assert regions.reduce([], "sum") == 0.0
    assert regions.reduce([float("nan"), 1.0, 2.0], "max") == 2.0
    assert regions.reduce([1.0, float("nan")], "max") == 1.0
    assert regions.reduce([3.0, float("nan"), -1.0], "min") == -1.0
    assert regions.reduce([0.5, 0.25, 0.125], "sum") == 0.875


def test_gated_quantity_is_the_crack_functor():
    y = np.array([0.1, -0.2, 0.3, 0.4])
    phi = np.array([0.31, 0.9, 0.3, 0.0])
    q = regions.quantity(y, gate=phi, threshold=0.3)
    assert q.tolist() == [0.1, -0.2, 0.0, 0.0]
    assert regions.reduce(q, "max") == 0.1
    x = np.stack([np.zeros(4), y, np.zeros(4)], axis=1)
    assert regions.reduce(q, "max") == configs.kw_crack_y(x, phi)
    assert regions.quantity([2.0, 3.0], vol=[0.5, 0.25]).tolist() == [1.0, 0.75]
