"""The drop-in C++ API (include/CabanaPD) driven end to end on the GPU.

dropin_kalthoff_winkler is the reference's Kalthoff-Winkler example
(examples/mechanics/kalthoff_winkler.cpp) written against the drop-in headers;
it checks the reference's regression value (7.42e-2 +- 1.1e-4,
regression_test/tstKalthoffWinkler.hpp:53) on the input deck the reference
ships (tests/golden/kalthoff_winkler.json).  dropin_elastic_wave is
examples/mechanics/elastic_wave.cpp on the deck of
regression_test/tstElasticWave.hpp; its displacement profile is checked
against that test's polynomial fit and against the oracle.  dropin_force runs
the tstForce flow (Force, Neighbor, computeForce, computeEnergy) against
lattice sums.  dropin_crack_inclusion is examples/mechanics/crack_inclusion.cpp
(two materials through createMultiForceModel, a functor force boundary
condition) checking regression_test/tstCrackInclusion.hpp:55.  dropin_dogbone
(ElasticPerfectlyPlastic, displacement grips), dropin_thermal_deformation
(TemperatureDependent, imposed temperature) and dropin_thermal_heat_transfer
(DynamicTemperature, cold planes) are the reference's examples on
their decks; their force_update = false functors run on the host each step,
and the results equal the engine's prescribed-value runs bit for bit.

The drivers are built on the CPU beforehand (`make -C tests/cpp`, also done by
__graft_entry__.build()); the test builds them only when they are missing.
"""
import os
import subprocess

import pytest

from tests.parity import say

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def driver(name):
    path = os.path.join(CPP, name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", CPP, name], check=True)
    return path


def run(cmd, cwd, timeout):
    say("case: " + " ".join(os.path.basename(c) for c in cmd))
    p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    out = p.stdout + p.stderr
    say("  " + out.strip().replace("\n", "\n  ")[-2000:])
    return p.returncode, out


def test_dropin_force_known_answers(tmp_path):
    rc, out = run([driver("dropin_force")], str(tmp_path), 100)
    assert rc == 0 and "DROPIN_FORCE_OK" in out


def test_dropin_kalthoff_winkler_example(tmp_path):
    deck = os.path.join(HERE, "golden", "kalthoff_winkler.json")
    rc, out = run([driver("dropin_kalthoff_winkler"), deck, str(tmp_path / "crack_y.txt")], str(tmp_path), 110)
    assert rc == 0 and "KW_REGRESSION_OK" in out


def test_dropin_elastic_wave_example(tmp_path):
    import json

    import numpy as np

    from cabanapd_amd import configs
    from tests.parity import oracle_spec
    deck = os.path.join(HERE, "golden", "elastic_wave.json")
    prof = tmp_path / "displacement_profile.txt"
    rc, out = run([driver("dropin_elastic_wave"), deck, str(prof)], str(tmp_path), 110)
    assert rc == 0 and "EW_DONE" in out
    rows = np.array([[float(t) for t in ln.split()] for ln in prof.read_text().splitlines() if ln.strip()])
    xs, us = rows[:, 1], rows[:, 2]
    fx = json.load(open(os.path.join(HERE, "golden", "elastic_wave_polyfit.json")))
    worst, l2, k = configs.polyfit_check(xs, us, fx)
    say(f"  polyfit: worst rel {worst:.4f} l2 {l2:.3e} points {k}")
    assert k >= 20 and worst < fx["max_rel_err"] and l2 < fx["max_l2_err"]
    # the same run in the oracle; the file holds six significant digits
    w = configs.elastic_wave()
    o = oracle_spec(w)
    configs.setup(o, w)
    o.step(w["steps"])
    line = configs.line_profile(w["x"], w["dx"])
    xo, uo = w["x"][line, 0], o.get("u")[line, 0]
    assert rows.shape[0] == line.size
    a, b = np.argsort(xs), np.argsort(xo)
    assert np.allclose(xs[a], xo[b], rtol=1e-5, atol=1e-7)
    assert np.max(np.abs(us[a] - uo[b])) <= 1e-5 * np.max(np.abs(uo))


def test_dropin_crack_inclusion_example(tmp_path):
    deck = os.path.join(HERE, "golden", "crack_inclusion.json")
    rc, out = run([driver("dropin_crack_inclusion"), deck, str(tmp_path / "incl_crack_y.txt")], str(tmp_path), 110)
    assert rc == 0 and "CI_REGRESSION_OK" in out


def _deck_variant(tmp_path, name, **changes):
    import json
    deck = json.load(open(os.path.join(HERE, "golden", name)))
    for k, v in changes.items():
        deck[k]["value"] = v
    path = tmp_path / name
    path.write_text(json.dumps(deck))
    return str(path)


def test_dropin_dogbone_example(tmp_path):
    """examples/mechanics/dogbone_tensile_test.cpp (EPP, displacement grips
    through a force_update = false functor evaluated on the host each step)
    on its deck with the grips at 20 m/s for 300 steps: the final
    displacement equals the engine run of the same specimen with the grips as
    prescribed a + b t values, bit for bit."""
    import numpy as np

    from cabanapd_amd import configs
    from tests.parity import gpu_spec
    steps = 300
    deck = _deck_variant(tmp_path, "dogbone_tensile_test.json", grip_velocity=20.0, final_time=(steps + 0.5) * 3.0e-7,
                         output_frequency=100)
    out = tmp_path / "dogbone_u.bin"
    rc, log = run([driver("dropin_dogbone"), deck, str(out)], str(tmp_path), 110)
    assert rc == 0 and "DB_DONE" in log
    w = configs.dogbone(grip_velocity=20.0, steps=steps)
    u_drop = np.fromfile(out, dtype=np.float64).reshape(-1, 3)
    assert u_drop.shape == w["x"].shape
    g = gpu_spec(w)
    configs.setup(g, w)
    g.step(steps)
    u = g.get("u")
    nd = int(np.sum(u_drop != u))
    say(f"  drop-in vs engine: differing={nd} max|u|={np.max(np.abs(u)):.3e}")
    assert nd == 0
    g.close()


def test_dropin_thermal_deformation_example(tmp_path):
    """examples/thermomechanics/thermal_deformation.cpp on its deck (13333
    steps; the temperature BodyTerm evaluated on the host each step): the
    final displacement and temperature equal the engine run with the
    temperature as prescribed a + b t values, bit for bit."""
    import numpy as np

    from cabanapd_amd import configs
    from tests.parity import gpu_spec
    deck = os.path.join(HERE, "golden", "thermal_deformation.json")
    out = tmp_path / "thermal_u.bin"
    rc, log = run([driver("dropin_thermal_deformation"), deck, str(out)], str(tmp_path), 110)
    assert rc == 0 and "TD_DONE" in log
    w = configs.thermal_deformation()
    u_drop = np.fromfile(out, dtype=np.float64).reshape(-1, 3)
    t_drop = np.fromfile(str(out) + ".temp", dtype=np.float64)
    g = gpu_spec(w)
    configs.setup(g, w)
    g.step(w["steps"])
    u, T = g.get("u"), g.get("temperature")
    nd = int(np.sum(u_drop != u)) + int(np.sum(t_drop != T))
    say(f"  drop-in vs engine: differing={nd} max|u|={np.max(np.abs(u)):.3e} T max {T.max():.4f}")
    assert nd == 0
    g.close()


def test_dropin_thermal_heat_transfer_example(tmp_path):
    """examples/thermomechanics/thermal_deformation_heat_transfer.cpp on its
    deck with heat conduction every 20 steps for 100 steps (the deck's own
    cadence, 45000, needs 19M steps): the final displacement and temperature
    equal the engine run, bit for bit."""
    import numpy as np

    from cabanapd_amd import configs
    from tests.parity import gpu_spec
    steps, sub = 100, 20
    deck = _deck_variant(tmp_path, "thermal_deformation_heat_transfer.json", final_time=(steps + 0.5) * 4.2e-7,
                         thermal_subcycle_steps=sub, output_frequency=50)
    out = tmp_path / "heat_u.bin"
    rc, log = run([driver("dropin_thermal_heat_transfer"), deck, str(out)], str(tmp_path), 110)
    assert rc == 0 and "HT_DONE" in log
    w = configs.thermal_deformation_heat_transfer(steps=steps, thermal_subcycle=sub)
    u_drop = np.fromfile(out, dtype=np.float64).reshape(-1, 3)
    t_drop = np.fromfile(str(out) + ".temp", dtype=np.float64)
    g = gpu_spec(w)
    configs.setup(g, w)
    g.step(steps)
    u, T = g.get("u"), g.get("temperature")
    nd = int(np.sum(u_drop != u)) + int(np.sum(t_drop != T))
    say(f"  drop-in vs engine: differing={nd} max|u|={np.max(np.abs(u)):.3e} T in [{T.min():.4f}, {T.max():.4f}]")
    assert nd == 0
    g.close()
