"""Drop-in header behaviour that needs no GPU: Particles::create with
Cabana::InitUniform / Cabana::InitRandom and a selection functor
(CabanaPD_Particles.hpp:241-400), through tests/cpp/dropin_particles.cpp."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def test_dropin_particles_create():
    exe = os.path.join(CPP, "dropin_particles")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", CPP, "dropin_particles"], check=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "PARTICLES_OK" in p.stdout, p.stdout + p.stderr
