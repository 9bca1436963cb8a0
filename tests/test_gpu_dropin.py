"""The drop-in C++ API (include/CabanaPD) driven end to end on the GPU.

tests/cpp/dropin_scenarios is one program of our own that exercises the
drop-in surface scenario by scenario on input decks in the reference's
format (tests/golden/*.json); it prints results, the checks live here:

* notched-impact: PMB fracture on the Kalthoff-Winkler deck, the crack-front
  Max series against the reference's regression value 7.42e-2 +- 1.1e-4
  (regression_test/tstKalthoffWinkler.hpp:53);
* pulse: a LinearLPS ForceModel (which runs the LPS force) on the elastic-wave
  deck, the displacement profile against the polynomial of
  regression_test/tstElasticWave.hpp:54-93 and against the oracle;
* inclusion: two LPS materials through createMultiForceModel(AverageTag), a
  functor traction, no-fail bands, against 5.75e-3 +- 5e-5
  (regression_test/tstCrackInclusion.hpp:55);
* grips / heated / conduction: ElasticPerfectlyPlastic, TemperatureDependent
  and DynamicTemperature specimens whose force_update = false functors run on
  the host each step; the results equal the engine's prescribed-value runs
  bit for bit;
* rebound / pulloff: DEM-only Solvers (HertzianModel, HertzianJKRModel) on the
  two-sphere decks, against the restitution (regression_test/tstHertz.hpp:135)
  and pull-off (regression_test/tstHertzJKR.hpp:145) values.

dropin_force runs the tstForce flow (Force, Neighbor, computeForce,
computeEnergy) against lattice sums.  The drivers are built on the CPU
beforehand (`make -C tests/cpp`, also done by __graft_entry__.build()); the
test builds them only when they are missing.
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


def scenario(name, deck, out, cwd, timeout=110):
    rc, log = run([driver("dropin_scenarios"), name, deck, out], cwd, timeout)
    assert rc == 0 and "SCENARIO_DONE" in log
    vals = {}
    for ln in log.splitlines():
        parts = ln.split()
        if len(parts) == 2:
            try:
                vals[parts[0]] = float(parts[1])
            except ValueError:
                pass
    return vals


def test_dropin_notched_impact_regression(tmp_path):
    deck = os.path.join(HERE, "golden", "kalthoff_winkler.json")
    vals = scenario("notched-impact", deck, str(tmp_path / "crack_y.txt"), str(tmp_path))
    say(f"  crack extent {vals['crack_extent']!r} after {vals['steps']:.0f} steps")
    assert abs(vals["crack_extent"] - 7.42e-2) <= 1.1e-4


def test_dropin_pulse_profile(tmp_path):
    import json

    import numpy as np

    from cabanapd_amd import configs
    from tests.parity import oracle_spec
    deck = os.path.join(HERE, "golden", "elastic_wave.json")
    prof = tmp_path / "displacement_profile.txt"
    scenario("pulse", deck, str(prof), str(tmp_path))
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


def test_dropin_inclusion_regression(tmp_path):
    deck = os.path.join(HERE, "golden", "crack_inclusion.json")
    vals = scenario("inclusion", deck, str(tmp_path / "incl_crack_y.txt"), str(tmp_path))
    say(f"  crack extent {vals['crack_extent']!r} after {vals['steps']:.0f} steps")
    assert abs(vals["crack_extent"] - 5.75e-3) <= 5e-5


def test_dropin_inclusion_explicit_indexing_equals_average_tag(tmp_path):
    """createMultiForceModel(particles, FullIndexing<2>{}, m0, m01, m10, m1)
    with the averaged models built by hand (ForceModel(m_a, m_b)) runs the
    inclusion deck exactly as createMultiForceModel(particles, AverageTag{},
    m0, m1): the crack-front time series files are identical."""
    deck = os.path.join(HERE, "golden", "crack_inclusion.json")
    out_a, out_b = tmp_path / "avg.txt", tmp_path / "idx.txt"
    a = scenario("inclusion", deck, str(out_a), str(tmp_path))
    b = scenario("inclusion-indexed", deck, str(out_b), str(tmp_path))
    say(f"  crack extent {a['crack_extent']!r} / {b['crack_extent']!r}")
    assert out_a.read_bytes() == out_b.read_bytes()
    assert a["crack_extent"] == b["crack_extent"]


def _deck_variant(tmp_path, name, **changes):
    import json
    deck = json.load(open(os.path.join(HERE, "golden", name)))
    for k, v in changes.items():
        deck[k]["value"] = v
    path = tmp_path / name
    path.write_text(json.dumps(deck))
    return str(path)


def test_dropin_grips_equal_engine(tmp_path):
    """The dogbone deck with the grips pulled at 20 m/s for 300 steps through
    a force_update = false displacement functor: the final displacement
    equals the engine run of the same specimen with the grips as prescribed
    a + b t values, bit for bit."""
    import numpy as np

    from cabanapd_amd import configs
    from tests.parity import gpu_spec
    steps = 300
    deck = _deck_variant(tmp_path, "dogbone_tensile_test.json", grip_velocity=20.0, final_time=(steps + 0.5) * 3.0e-7,
                         output_frequency=100)
    out = tmp_path / "dogbone_u.bin"
    vals = scenario("grips", deck, str(out), str(tmp_path))
    w = configs.dogbone(grip_velocity=20.0, steps=steps)
    u_drop = np.fromfile(out, dtype=np.float64).reshape(-1, 3)
    assert u_drop.shape == w["x"].shape
    g = gpu_spec(w)
    configs.setup(g, w)
    g.step(steps)
    u = g.get("u")
    nd = int(np.sum(u_drop != u))
    say(f"  drop-in vs engine: differing={nd} max|u|={np.max(np.abs(u)):.3e} grip force {vals['last_grip_force']:.4e}")
    assert nd == 0
    # the specimen resists the pull: the internal force on the moving grip points back (-x)
    assert vals["last_grip_force"] < 0.0
    g.close()


def test_dropin_heated_equals_engine(tmp_path):
    """The thermal-deformation deck (13333 steps; the temperature ramp set by
    a force = false BodyTerm evaluated on the host each step): the final
    displacement and temperature equal the engine run with the temperature as
    prescribed a + b t values, bit for bit."""
    import numpy as np

    from cabanapd_amd import configs
    from tests.parity import gpu_spec
    deck = os.path.join(HERE, "golden", "thermal_deformation.json")
    out = tmp_path / "thermal_u.bin"
    scenario("heated", deck, str(out), str(tmp_path))
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


def test_dropin_conduction_equals_engine(tmp_path):
    """The heat-transfer deck with conduction every 20 steps for 100 steps
    (the deck's own cadence, 45000, needs 19M steps) and both y faces held
    cold: the final displacement and temperature equal the engine run, bit
    for bit."""
    import numpy as np

    from cabanapd_amd import configs
    from tests.parity import gpu_spec
    steps, sub = 100, 20
    deck = _deck_variant(tmp_path, "thermal_deformation_heat_transfer.json", final_time=(steps + 0.5) * 4.2e-7,
                         thermal_subcycle_steps=sub, output_frequency=50)
    out = tmp_path / "heat_u.bin"
    scenario("conduction", deck, str(out), str(tmp_path))
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


def test_dropin_hertz_rebound(tmp_path):
    deck = os.path.join(HERE, "golden", "hertzian_contact.json")
    vals = scenario("rebound", deck, str(tmp_path / "unused"), str(tmp_path))
    say(f"  speed ratio {vals['speed_ratio']!r} after {vals['steps']:.0f} steps")
    assert abs(vals["speed_ratio"] - 0.5) <= 1e-3


def test_dropin_jkr_pulloff(tmp_path):
    deck = os.path.join(HERE, "golden", "hertzian_jkr_contact.json")
    vals = scenario("pulloff", deck, str(tmp_path / "unused"), str(tmp_path))
    say(f"  pull-off ratio {vals['pulloff_ratio']!r} after {vals['steps']:.0f} steps")
    assert abs(vals["pulloff_ratio"] + 1.0) <= 5e-3


def test_dropin_collide_equals_engine(tmp_path):
    """A PD force model and a NormalRepulsion contact model in one Solver
    (the four-argument constructor) on our colliding-blocks deck
    (tests/golden/colliding_blocks.json): the final displacement equals the
    engine run of the same specimen with pdk_set_contact, bit for bit."""
    import json

    import numpy as np

    from cabanapd_amd import configs
    from tests.parity import gpu_spec
    deck = os.path.join(HERE, "golden", "colliding_blocks.json")
    out = tmp_path / "blocks_u.bin"
    vals = scenario("collide", deck, str(out), str(tmp_path))
    w = configs.colliding_blocks(steps=int(vals["steps"]))
    d = json.load(open(deck))
    assert d["timestep"]["value"] == w["desc"].dt and d["horizon"]["value"] == w["desc"].horizon
    u_drop = np.fromfile(out, dtype=np.float64).reshape(-1, 3)
    assert u_drop.shape == w["x"].shape
    g = gpu_spec(w)
    configs.setup(g, w)
    g.step(w["steps"])
    u = g.get("u")
    nd = int(np.sum(u_drop != u))
    say(f"  drop-in vs engine after {w['steps']} steps: differing={nd} max|u|={np.max(np.abs(u)):.3e} "
        f"contact pairs {g.contact_pairs()}")
    assert nd == 0
    g.close()
