"""The drop-in C++ API (include/CabanaPD) driven end to end on the GPU.

dropin_kalthoff_winkler is the reference's Kalthoff-Winkler example
(examples/mechanics/kalthoff_winkler.cpp) written against the drop-in headers;
it checks the reference's regression value (7.42e-2 +- 1.1e-4,
regression_test/tstKalthoffWinkler.hpp:53) on the input deck the reference
ships (tests/golden/kalthoff_winkler.json).  dropin_force runs the tstForce
flow (Force, Neighbor, computeForce, computeEnergy) against lattice sums.

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
