"""Writes tests/golden/dogbone_epp_curve.json: the grip force of the dogbone
tensile deck (examples/mechanics/inputs/dogbone_tensile_test.json, PMB
ElasticPerfectlyPlastic, grips at 0.1 m/s, 50,000 steps; particles at the
cell centres) from the oracle, every 10,000 steps.  The force is the sum over
the right grip of f_x times the particle volume, as the example's output time
series (examples/mechanics/dogbone_tensile_test.cpp:200-214).  Our own
regression fixture (the reference holds no number for this deck); run from the
repository root: python3 tests/golden/make_dogbone_curve.py"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cabanapd_amd import configs  # noqa: E402
from oracle import binding as orc  # noqa: E402


def main():
    orc.set_threads(min(os.cpu_count() or 1, 16))
    w = configs.dogbone()
    d = w["desc"]
    o = orc.OracleSpec(d, w["x"].shape[0], w["x"], w["vol"], u=w["u"], v=w["v"], rho=w["rho"])
    configs.setup(o, w)
    rg = w["right_grip"]
    vol = w["vol"]
    steps, force = [], []
    for k in range(1, 6):
        o.step(10000)
        f = o.get("f")
        steps.append(10000 * k)
        force.append(float(np.sum(f[rg, 0] * vol[rg])))
    out = {"deck": "dogbone_tensile_test.json", "grip_velocity": 0.1, "steps": steps, "grip_force_x": force,
           "yielded_bonds": int(np.count_nonzero(o.plastic())), "broken_bonds": int(np.sum(o.bonds()[2] == 0))}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "dogbone_epp_curve.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
