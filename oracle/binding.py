"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The interface mirrors
cabanapd_amd.engine.Spec so a parity test drives both the same way.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from cabanapd_amd import _abi
from cabanapd_amd._abi import FIELD, VECTOR_FIELDS

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)
_VP = C.c_void_p

_SIGS = {
    "orc_set_threads": (None, [C.c_int32]),
    "orc_get_threads": (C.c_int32, []),
    "orc_create": (C.c_int, [C.POINTER(_abi.ModelDesc), C.c_int64, _D, _D, _D, _D, _D, _I32, _D,
                             C.POINTER(_VP)]),
    "orc_destroy": (C.c_int, [_VP]),
    "orc_model_constants": (C.c_int, [C.POINTER(_abi.ModelDesc), C.POINTER(_abi.ModelConstants)]),
    "orc_critical_timestep": (C.c_int, [C.c_double] * 7 + [_D]),
    "orc_build_neighbors": (C.c_int, [_VP]),
    "orc_bond_stats": (C.c_int, [_VP, _I64, _I32]),
    "orc_prenotch": (C.c_int, [_VP, C.c_int32, _D, _D, _D]),
    "orc_get_bonds": (C.c_int, [_VP, _I64, _I64, _U8]),
    "orc_set_bc": (C.c_int, [_VP, C.c_int32, _I64, C.c_int64, C.c_double]),
    "orc_set_body_force": (C.c_int, [_VP, _D]),
    "orc_compute_weighted_volume": (C.c_int, [_VP]),
    "orc_compute_dilatation": (C.c_int, [_VP]),
    "orc_compute_force": (C.c_int, [_VP, C.c_int32]),
    "orc_compute_energy": (C.c_int, [_VP, _D, _D]),
    "orc_compute_stress": (C.c_int, [_VP]),
    "orc_set_materials": (C.c_int, [_VP, C.c_int32, C.POINTER(_abi.Material), _I32]),
    "orc_set_pair_models": (C.c_int, [_VP, C.c_int32, C.POINTER(_abi.PairModel), _I32]),
    "orc_set_prescribed": (C.c_int, [_VP, C.c_int32, _I64, C.c_int64, _D, _D]),
    "orc_set_field": (C.c_int, [_VP, C.c_int32, _D]),
    "orc_get_plastic": (C.c_int, [_VP, _D]),
    "orc_init": (C.c_int, [_VP]),
    "orc_step": (C.c_int, [_VP, C.c_int32]),
    "orc_get": (C.c_int, [_VP, C.c_int32, _D]),
    "orc_set_state": (C.c_int, [_VP, _D, _D]),
    "orc_set_contact": (C.c_int, [_VP, C.POINTER(_abi.ContactDesc)]),
    "orc_contact_pairs": (C.c_int, [_VP, _I64]),
    "orc_cr_pow": (C.c_int, [C.c_int32, C.c_int64, _D, _D]),
    "orc_bond_pmb": (C.c_int, [C.POINTER(_abi.ModelDesc), _D, _D, C.c_double, _D, _D, _I32]),
    "orc_bond_prenotch": (C.c_int, [_D, _D, _D, _D, _D, _I32]),
    "orc_jkr_normal_force": (C.c_int, [C.POINTER(_abi.ModelDesc), C.c_double, C.c_double, C.c_double,
                                       C.c_double, _D]),
    "orc_jkr_patch_radius": (C.c_int, [C.POINTER(_abi.ModelDesc), C.c_double, _D]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, what="oracle call"):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def model_constants(desc):
    k = _abi.ModelConstants()
    _check(lib().orc_model_constants(C.byref(desc), C.byref(k)))
    return {name: getattr(k, name) for name, _ in _abi.ModelConstants._fields_}


def critical_timestep(dx, dy, dz, K, rho, horizon, safety=0.85):
    out = C.c_double()
    _check(lib().orc_critical_timestep(dx, dy, dz, K, rho, horizon, safety, C.byref(out)))
    return out.value


def cr_pow(which, x):
    """The oracle's correctly rounded contact powers: which in {"1.5", "1/3", "3", "0.5"}."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    _check(lib().orc_cr_pow({"1.5": 0, "1/3": 1, "3": 2, "0.5": 3}[which], x.size, _dp(x), _dp(out)))
    return out


def bond_pmb(desc, xi, eta, vol_j):
    f = np.zeros(3)
    w = C.c_double()
    brk = C.c_int32()
    _check(lib().orc_bond_pmb(C.byref(desc), _dp(_f64(xi)), _dp(_f64(eta)), vol_j, _dp(f), C.byref(w),
                              C.byref(brk)))
    return f, w.value, bool(brk.value)


def bond_prenotch(v1, v2, p0, xi, xj):
    keep = C.c_int32()
    _check(lib().orc_bond_prenotch(_dp(_f64(v1)), _dp(_f64(v2)), _dp(_f64(p0)), _dp(_f64(xi)), _dp(_f64(xj)),
                                   C.byref(keep)))
    return bool(keep.value)


def jkr_normal_force(desc, r, vn, vol, rho):
    out = C.c_double()
    _check(lib().orc_jkr_normal_force(C.byref(desc), r, vn, vol, rho, C.byref(out)))
    return out.value


def jkr_patch_radius(desc, delta_n):
    out = C.c_double()
    _check(lib().orc_jkr_patch_radius(C.byref(desc), delta_n, C.byref(out)))
    return out.value


class OracleSpec:
    def __init__(self, desc, n, x, vol, u=None, v=None, rho=None, nofail=None, radius=None):
        self.desc = desc
        self.n = int(n)
        n = self.n
        x = _f64(x, (n, 3))
        vol = _f64(vol, (n,))
        u = _f64(u, (n, 3))
        v = _f64(v, (n, 3))
        rho = _f64(rho, (n,))
        radius = _f64(radius, (n,))
        nf = None if nofail is None else np.ascontiguousarray(nofail, dtype=np.int32)
        self.h = _VP()
        _check(lib().orc_create(C.byref(desc), n, _dp(x), _dp(u), _dp(v), _dp(vol), _dp(rho),
                                None if nf is None else nf.ctypes.data_as(_I32), _dp(radius), C.byref(self.h)),
               "orc_create")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_bonds(self):
        _check(lib().orc_build_neighbors(self.h))

    def set_contact(self, contact):
        """A contact model on this PD specimen (an _abi.ContactDesc)."""
        _check(lib().orc_set_contact(self.h, C.byref(contact)), "orc_set_contact")

    def contact_pairs(self):
        t = C.c_int64()
        _check(lib().orc_contact_pairs(self.h, C.byref(t)))
        return t.value

    def bond_stats(self):
        t = C.c_int64()
        m = C.c_int32()
        _check(lib().orc_bond_stats(self.h, C.byref(t), C.byref(m)))
        return t.value, m.value

    def prenotch(self, p0, v1, v2):
        p0 = _f64(p0).reshape(-1, 3)
        v1 = _f64(v1).reshape(-1, 3)
        v2 = _f64(v2).reshape(-1, 3)
        _check(lib().orc_prenotch(self.h, p0.shape[0], _dp(p0), _dp(v1), _dp(v2)))

    def bonds(self):
        total, _ = self.bond_stats()
        off = np.zeros(self.n + 1, dtype=np.int64)
        nbr = np.zeros(max(total, 1), dtype=np.int64)
        intact = np.zeros(max(total, 1), dtype=np.uint8)
        _check(lib().orc_get_bonds(self.h, off.ctypes.data_as(_I64), nbr.ctypes.data_as(_I64),
                                   intact.ctypes.data_as(_U8)))
        return off, nbr[:total], intact[:total]

    def set_bc(self, kind, idx, value):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        _check(lib().orc_set_bc(self.h, kind, idx.ctypes.data_as(_I64), idx.size, value))

    def set_body_force(self, b):
        _check(lib().orc_set_body_force(self.h, _dp(_f64(b, (self.n, 3)) if b is not None else None)))

    def compute_weighted_volume(self):
        _check(lib().orc_compute_weighted_volume(self.h))

    def compute_dilatation(self):
        _check(lib().orc_compute_dilatation(self.h))

    def compute_force(self, reset=True):
        _check(lib().orc_compute_force(self.h, 1 if reset else 0))

    def compute_energy(self):
        W = C.c_double()
        D = C.c_double()
        _check(lib().orc_compute_energy(self.h, C.byref(W), C.byref(D)))
        return W.value, D.value

    def compute_stress(self):
        _check(lib().orc_compute_stress(self.h))

    def set_materials(self, mats, mtype):
        arr = _abi.materials_array(mats)
        t = np.ascontiguousarray(mtype, dtype=np.int32)
        _check(lib().orc_set_materials(self.h, len(mats), arr, t.ctypes.data_as(_I32)))

    def set_pair_models(self, pairs, mtype):
        """Explicit-indexing pair table: pairs[a * ntypes + b] (see _abi.pair_models_array)."""
        nt = int(round(len(pairs) ** 0.5))
        arr = _abi.pair_models_array(pairs)
        t = np.ascontiguousarray(mtype, dtype=np.int32)
        _check(lib().orc_set_pair_models(self.h, nt, arr, t.ctypes.data_as(_I32)))

    def set_prescribed(self, field, idx, a, b=None):
        """Values a + b t applied before each force pass (and at t = 0 in init)."""
        nc = 3 if field == "u" else 1
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        a = _f64(a, (idx.size, nc)) if idx.size else np.zeros((0, nc))
        b = _f64(b, (idx.size, nc)) if (b is not None and idx.size) else None
        _check(lib().orc_set_prescribed(self.h, FIELD[field], idx.ctypes.data_as(_I64), idx.size, _dp(a),
                                        _dp(b)))

    def set_field(self, field, values):
        _check(lib().orc_set_field(self.h, FIELD[field], _dp(_f64(values, (self.n,)))))

    def plastic(self):
        """EPP plastic stretch per bond, in the order of bonds()."""
        total, _ = self.bond_stats()
        out = np.zeros(max(total, 1), dtype=np.float64)
        _check(lib().orc_get_plastic(self.h, _dp(out)))
        return out[:total]

    def init(self):
        _check(lib().orc_init(self.h))

    def step(self, nsteps=1):
        _check(lib().orc_step(self.h, nsteps))

    def set_state(self, u=None, v=None):
        _check(lib().orc_set_state(self.h, _dp(_f64(u, (self.n, 3))), _dp(_f64(v, (self.n, 3)))))

    def get(self, field):
        shape = _abi.field_shape(field, self.n)
        out = np.empty(shape, dtype=np.float64)
        _check(lib().orc_get(self.h, FIELD[field], _dp(out)))
        return out
