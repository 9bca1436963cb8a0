"""The C-ABI library loads and exports every entry point include/pdk.h declares
(no GPU calls)."""
import ctypes
import os
import re

from cabanapd_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdk_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_header():
    lib = ctypes.CDLL(engine.LIB_PATH)
    names = declared("pdk.h")
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared("pdk.h")) == set(engine.exported_symbols())


def test_constants_match_oracle():
    from cabanapd_amd._abi import make_desc
    from oracle import binding as orc
    for d in [make_desc("PMB", fracture=True, horizon=1.505e-3, K=127.33e9, G0=42408.0),
              make_desc("LPS", fracture=True, horizon=0.0237, K=42.857e9, G=29.5e9, G0=3.8),
              make_desc("HERTZ", radius=5e-5, radius_extend=1e-8, nu=0.25, E=195.6e9, restitution=0.5),
              make_desc("HERTZ_JKR", radius=5e-5, radius_extend=1e-8, nu=0.25, E=195.6e9, restitution=1.0,
                        gamma=10.0)]:
        a = engine.model_constants(d)
        b = orc.model_constants(d)
        assert a == b


def test_critical_timestep_matches_oracle():
    from oracle import binding as orc
    args = (5e-4, 5e-4, 5e-4, 191e9 / 1.5, 8000.0, 3.01 * 5e-4, 0.85)
    assert engine.critical_timestep(*args) == orc.critical_timestep(*args)
