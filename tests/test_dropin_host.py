"""Host-side checks of the drop-in C++ API that need no GPU:
tests/cpp/dropin_indexing (indexing functors, AverageTag model lists, the
explicit-indexing form and the engine's pair table)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def test_dropin_indexing_host():
    exe = os.path.join(CPP, "dropin_indexing")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", CPP, "dropin_indexing"], check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
