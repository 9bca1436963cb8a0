"""Helpers that drive the engine and the oracle through the same calls."""
import sys

import numpy as np

from cabanapd_amd import configs, engine
from oracle import binding as orc

_CTX = None


def ctx():
    global _CTX
    if _CTX is None:
        _CTX = engine.Context(0)
    return _CTX


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_spec(w, desc=None):
    s = engine.Spec(ctx(), desc or w["desc"], w["x"].shape[0])
    configs.upload(s, w)
    return s


def oracle_spec(w, desc=None):
    return orc.OracleSpec(desc or w["desc"], w["x"].shape[0], w["x"], w["vol"], u=w["u"], v=w["v"],
                          rho=w["rho"], nofail=w["nofail"], radius=w["radius"])


def pair(w, setup=True, init=True):
    g = gpu_spec(w)
    o = oracle_spec(w)
    if setup:
        configs.setup(g, w, init=init)
        configs.setup(o, w, init=init)
    return g, o


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / den)


def bonds_equal(g, o):
    go, gn, gi = g.bonds()
    oo, on, oi = o.bonds()
    return np.array_equal(go, oo) and np.array_equal(gn, on), np.array_equal(gi, oi)


def broken_set(spec):
    off, nbr, intact = spec.bonds()
    rows = np.repeat(np.arange(off.size - 1), np.diff(off))
    sel = intact == 0
    return set(zip(rows[sel].tolist(), nbr[sel].tolist()))
