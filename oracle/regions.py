"""CPU restatement of the reference's region outputs.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
series (pdk_series_*); never on a product path.

  members      Region<RectangularPrism|Cylinder|Line>::inside, CabanaPD_Geometry.hpp:54-168,
               collected in ascending particle id like ParticleSteeringVector (:170-273)
  quantity     the user functor of a time series, in the gated form of
               examples/mechanics/kalthoff_winkler.cpp:135-142 (damage > 0.3 ? y : 0)
  reduce       Kokkos::Sum/Max/Min joined over the members, OutputTimeSeries::update
               (CabanaPD_OutputProfiles.hpp:146-163); Max/Min start from the reduction
               identity (-DBL_MAX / +DBL_MAX) and replace only when strictly greater /
               smaller, so NaN never wins
"""
import sys

import numpy as np

from cabanapd_amd import _abi

DBL_MAX = sys.float_info.max


def members(x, r, ids=None):
    """Ascending caller ids of the particles of region r (an _abi.RegionDesc)
    given reference positions x[n][3]."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    if ids is not None:
        return np.unique(np.asarray(ids, dtype=np.int64))
    if r.kind == _abi.PDK_REGION_ALL:
        return np.arange(n, dtype=np.int64)
    out = []
    for p in range(n):
        x0, x1, x2 = x[p]
        if r.kind == _abi.PDK_REGION_BOX:
            ok = (r.low[0] <= x0 <= r.high[0]) and (r.low[1] <= x1 <= r.high[1]) and (r.low[2] <= x2 <= r.high[2])
        elif r.kind == _abi.PDK_REGION_CYLINDER:
            rsq = (x0 - r.x_center) * (x0 - r.x_center) + (x1 - r.y_center) * (x1 - r.y_center)
            ok = (r.r_in * r.r_in <= rsq <= r.r_out * r.r_out) and (r.low_z <= x2 <= r.high_z)
        elif r.kind == _abi.PDK_REGION_LINE:
            a = x[p, (r.dim + 1) % 3]
            b = x[p, (r.dim + 2) % 3]
            ok = (-r.half[0] < a < r.half[0]) and (-r.half[1] < b < r.half[1])
        else:
            raise ValueError("unknown region kind")
        if ok:
            out.append(p)
    return np.asarray(out, dtype=np.int64)


def quantity(value, gate=None, threshold=0.0, vol=None):
    """Per-particle value of the series' functor."""
    v = np.array(value, dtype=np.float64, copy=True)
    if gate is not None:
        g = np.asarray(gate)
        v = np.array([vi if gi > threshold else 0.0 for vi, gi in zip(v, g)])
    if vol is not None:
        v = v * np.asarray(vol)
    return v


def reduce(values, reducer):
    """Sequential join in member order."""
    if reducer == "sum":
        acc = 0.0
        for v in values:
            acc += float(v)
        return acc
    if reducer == "max":
        acc = -DBL_MAX
        for v in values:
            if v > acc:
                acc = float(v)
        return acc
    if reducer == "min":
        acc = DBL_MAX
        for v in values:
            if v < acc:
                acc = float(v)
        return acc
    raise ValueError("unknown reducer")
