"""pdk_compute_energy_noalloc (the entry point that allocates nothing: its
partial sums live in a fixed device array of the library) against
pdk_compute_energy, bit for bit: totals, per-particle energy and damage.
Two calls in a row on specimens of different sizes (the second smaller, so
stale contents of the shared array past its end would show), after which
the first specimen is queried again; a specimen beyond the fixed maximum is
refused."""
import numpy as np
import pytest

from cabanapd_amd import configs, engine
from tests.parity import ctx, say

pytestmark = pytest.mark.gpu


def _spec(w):
    s = engine.Spec(ctx(), w["desc"], w["x"].shape[0])
    configs.upload(s, w)
    configs.setup(s, w)
    return s


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _same(s):
    """Bit patterns compared, so that the NaN energies of LPS particles that
    lost every bond (the reference's 0 * inf) count as equal."""
    W0, D0 = s.compute_energy()
    f0 = (s.get("W"), s.get("phi"))
    W1, D1 = s.compute_energy_noalloc()
    f1 = (s.get("W"), s.get("phi"))
    say(f"  energy {W0!r} / {W1!r}, damage {D0!r} / {D1!r}, NaN energies {int(np.isnan(f1[0]).sum())}")
    assert np.array_equal(_bits([W0, D0]), _bits([W1, D1]))
    assert np.array_equal(_bits(f0[0]), _bits(f1[0])) and np.array_equal(_bits(f0[1]), _bits(f1[1]))
    return W1, D1


def test_noalloc_energy_equals_compute_energy():
    say("case: energy without allocation, PMB fracture plate 60x120x9 then LPS cube 20^3")
    big = configs.kw_plate(num_cells=(60, 120, 9), steps=40)
    a = _spec(big)
    a.step(40)
    first = _same(a)
    small = configs.lps_cube(n_side=20, steps=10, fracture=True)
    b = _spec(small)
    b.step(10)
    _same(b)
    assert np.array_equal(_bits(a.compute_energy_noalloc()), _bits(first))
    a.close()
    b.close()


def test_noalloc_energy_refuses_oversized_spec():
    say("case: energy without allocation, spec beyond the fixed scratch (refused before any GPU work)")
    n = 256 * 256 * 257  # 16,842,752 > 16,777,216
    desc = engine.make_desc("PMB", fracture=True, horizon=3.01 / 256, K=1e9, G0=1.0, dt=1e-9)
    s = engine.Spec(ctx(), desc, n)
    with pytest.raises(engine.PdkError, match="exceed the fixed scratch"):
        s.compute_energy_noalloc()
    s.close()
