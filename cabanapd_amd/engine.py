"""ctypes binding of libpdk.so (include/pdk.h).

This is a thin host-side mirror used by the tests and the benchmark; the
drop-in C++ API lives in include/CabanaPD/.  The library must exist: there is
no fallback path.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import FIELD, VECTOR_FIELDS, make_desc  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpdk.so")

_lib = None

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)
_VP = C.c_void_p

_SIGS = {
    "pdk_last_error": (C.c_char_p, []),
    "pdk_model_constants_get": (C.c_int, [C.POINTER(_abi.ModelDesc), C.POINTER(_abi.ModelConstants)]),
    "pdk_critical_timestep": (C.c_int, [C.c_double] * 7 + [_D]),
    "pdk_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(_VP)]),
    "pdk_ctx_destroy": (C.c_int, [_VP]),
    "pdk_ctx_bytes_in_use": (C.c_int, [_VP, C.POINTER(C.c_uint64)]),
    "pdk_spec_create": (C.c_int, [_VP, C.POINTER(_abi.ModelDesc), C.c_int64, C.POINTER(_VP)]),
    "pdk_spec_destroy": (C.c_int, [_VP]),
    "pdk_spec_upload": (C.c_int, [_VP, _D, _D, _D, _D, _D, _I32, _D]),
    "pdk_spec_set_state": (C.c_int, [_VP, _D, _D]),
    "pdk_spec_copy_state": (C.c_int, [_VP, _VP]),
    "pdk_spec_bytes": (C.c_int, [_VP, C.POINTER(C.c_uint64)]),
    "pdk_build_bonds": (C.c_int, [_VP]),
    "pdk_bond_stats": (C.c_int, [_VP, _I64, _I32]),
    "pdk_prenotch": (C.c_int, [_VP, C.c_int32, _D, _D, _D]),
    "pdk_download_bonds": (C.c_int, [_VP, C.c_int64, _I64, _I64, _U8]),
    "pdk_set_bc": (C.c_int, [_VP, C.c_int32, _I64, C.c_int64, C.c_double]),
    "pdk_set_body_force": (C.c_int, [_VP, _D]),
    "pdk_compute_weighted_volume": (C.c_int, [_VP]),
    "pdk_compute_dilatation": (C.c_int, [_VP]),
    "pdk_compute_force": (C.c_int, [_VP, C.c_int32]),
    "pdk_compute_energy": (C.c_int, [_VP, _D, _D]),
    "pdk_init": (C.c_int, [_VP]),
    "pdk_step": (C.c_int, [_VP, C.c_int32]),
    "pdk_steps_done": (C.c_int, [_VP, _I64]),
    "pdk_download": (C.c_int, [_VP, C.c_int32, _D]),
    "pdk_sync": (C.c_int, [_VP]),
    "pdk_set_tuning": (C.c_int, [_VP, C.c_int32]),
    "pdk_spec_layout": (C.c_int, [_VP, _I32, _I32, _I32]),
    "pdk_autotune": (C.c_int, [_VP, C.c_int32, _I32]),
    "pdk_set_graph": (C.c_int, [_VP, C.c_int32]),
    "pdk_kernel_timing": (C.c_int, [_VP, C.c_int32]),
    "pdk_kernel_time": (C.c_int, [_VP, C.c_int32, _D, _I64]),
    "pdk_kernel_name": (C.c_int, [_VP, C.c_int32, C.c_char_p, C.c_int32]),
    "pdk_submit": (C.c_int, [_VP, _VP, C.c_int32, C.POINTER(C.c_uint64)]),
    "pdk_poll": (C.c_int, [_VP, C.c_uint64, _I32]),
    "pdk_batch_progress": (C.c_int, [_VP, C.c_int32, _I32]),
    "pdk_batch_set_graph": (C.c_int, [_VP, C.c_int32]),
    "pdk_series_create": (C.c_int, [_VP, C.POINTER(_abi.RegionDesc), C.POINTER(_abi.Quantity), C.c_int32,
                                    C.c_int64, C.POINTER(_VP)]),
    "pdk_series_destroy": (C.c_int, [_VP]),
    "pdk_series_members": (C.c_int, [_VP, _I64]),
    "pdk_series_update": (C.c_int, [_VP, C.c_double]),
    "pdk_series_values": (C.c_int, [_VP, C.c_int64, _D, _D, _I64]),
    "pdk_series_gather": (C.c_int, [_VP, C.c_int64, _I64, _D, _I64]),
    "pdk_compute_energy_async": (C.c_int, [_VP]),
    "pdk_compute_energy_noalloc": (C.c_int, [_VP, _D, _D]),
    "pdk_compute_stress": (C.c_int, [_VP]),
    "pdk_set_materials": (C.c_int, [_VP, C.c_int32, C.POINTER(_abi.Material), _I32]),
    "pdk_set_pair_models": (C.c_int, [_VP, C.c_int32, C.POINTER(_abi.PairModel), _I32]),
    "pdk_set_prescribed": (C.c_int, [_VP, C.c_int32, _I64, C.c_int64, _D, _D]),
    "pdk_set_field": (C.c_int, [_VP, C.c_int32, _D]),
    "pdk_set_contact": (C.c_int, [_VP, C.POINTER(_abi.ContactDesc)]),
    "pdk_contact_pairs": (C.c_int, [_VP, _I64]),
}


class PdkError(RuntimeError):
    pass


def lib():
    """Load libpdk.so; raises if it is missing (build with __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PdkError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def _check(rc):
    if rc != 0:
        raise PdkError(lib().pdk_last_error().decode())


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def model_constants(desc):
    k = _abi.ModelConstants()
    _check(lib().pdk_model_constants_get(C.byref(desc), C.byref(k)))
    return {name: getattr(k, name) for name, _ in _abi.ModelConstants._fields_}


def critical_timestep(dx, dy, dz, K, rho, horizon, safety=0.85):
    out = C.c_double()
    _check(lib().pdk_critical_timestep(dx, dy, dz, K, rho, horizon, safety, C.byref(out)))
    return out.value


class Context:
    def __init__(self, device=0, hbm_budget=0):
        self.h = _VP()
        _check(lib().pdk_ctx_create(device, hbm_budget, C.byref(self.h)))

    def close(self):
        if self.h:
            _check(lib().pdk_ctx_destroy(self.h))
            self.h = _VP()

    def bytes_in_use(self):
        b = C.c_uint64()
        _check(lib().pdk_ctx_bytes_in_use(self.h, C.byref(b)))
        return b.value

    def submit(self, spec, nsteps):
        t = C.c_uint64()
        _check(lib().pdk_submit(self.h, spec.h, nsteps, C.byref(t)))
        return t.value

    def poll(self, ticket):
        d = C.c_int32()
        _check(lib().pdk_poll(self.h, ticket, C.byref(d)))
        return bool(d.value)

    def progress(self, max_steps=0):
        n = C.c_int32()
        _check(lib().pdk_batch_progress(self.h, max_steps, C.byref(n)))
        return n.value

    def set_graph(self, steps):
        """Replay the captured packed step for up to steps steps (0: eager launches)."""
        _check(lib().pdk_batch_set_graph(self.h, steps))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Spec:
    """One specimen on the GPU (particles, bonds, state)."""

    def __init__(self, ctx, desc, n):
        self.ctx = ctx
        self.desc = desc
        self.n = int(n)
        self.h = _VP()
        _check(lib().pdk_spec_create(ctx.h, C.byref(desc), self.n, C.byref(self.h)))

    def close(self):
        if self.h:
            _check(lib().pdk_spec_destroy(self.h))
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, x, vol, u=None, v=None, rho=None, nofail=None, radius=None):
        n = self.n
        x = _f64(x, (n, 3))
        vol = _f64(vol, (n,))
        u = _f64(u, (n, 3))
        v = _f64(v, (n, 3))
        rho = _f64(rho, (n,))
        radius = _f64(radius, (n,))
        nf = None if nofail is None else np.ascontiguousarray(nofail, dtype=np.int32)
        _check(lib().pdk_spec_upload(self.h, _dp(x), _dp(u), _dp(v), _dp(vol), _dp(rho),
                                     None if nf is None else nf.ctypes.data_as(_I32), _dp(radius)))

    def set_state(self, u=None, v=None):
        _check(lib().pdk_spec_set_state(self.h, _dp(_f64(u, (self.n, 3))), _dp(_f64(v, (self.n, 3)))))

    def copy_state_from(self, src):
        """Restart this spec from src's state between steps (pdk_spec_copy_state)."""
        _check(lib().pdk_spec_copy_state(self.h, src.h))

    def build_bonds(self):
        _check(lib().pdk_build_bonds(self.h))

    def bond_stats(self):
        t = C.c_int64()
        m = C.c_int32()
        _check(lib().pdk_bond_stats(self.h, C.byref(t), C.byref(m)))
        return t.value, m.value

    def prenotch(self, p0, v1, v2):
        p0 = _f64(p0).reshape(-1, 3)
        v1 = _f64(v1).reshape(-1, 3)
        v2 = _f64(v2).reshape(-1, 3)
        _check(lib().pdk_prenotch(self.h, p0.shape[0], _dp(p0), _dp(v1), _dp(v2)))

    def bonds(self):
        total, _ = self.bond_stats()
        off = np.zeros(self.n + 1, dtype=np.int64)
        nbr = np.zeros(max(total, 1), dtype=np.int64)
        intact = np.zeros(max(total, 1), dtype=np.uint8)
        _check(lib().pdk_download_bonds(self.h, nbr.size, off.ctypes.data_as(_I64), nbr.ctypes.data_as(_I64),
                                        intact.ctypes.data_as(_U8)))
        return off, nbr[:total], intact[:total]

    def set_bc(self, kind, idx, value):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        _check(lib().pdk_set_bc(self.h, kind, idx.ctypes.data_as(_I64), idx.size, value))

    def set_body_force(self, b):
        _check(lib().pdk_set_body_force(self.h, _dp(_f64(b, (self.n, 3)) if b is not None else None)))

    def compute_weighted_volume(self):
        _check(lib().pdk_compute_weighted_volume(self.h))

    def compute_dilatation(self):
        _check(lib().pdk_compute_dilatation(self.h))

    def compute_force(self, reset=True):
        _check(lib().pdk_compute_force(self.h, 1 if reset else 0))

    def compute_energy(self):
        W = C.c_double()
        D = C.c_double()
        _check(lib().pdk_compute_energy(self.h, C.byref(W), C.byref(D)))
        return W.value, D.value

    def init(self):
        _check(lib().pdk_init(self.h))

    def step(self, nsteps=1):
        _check(lib().pdk_step(self.h, nsteps))

    def compute_stress(self):
        _check(lib().pdk_compute_stress(self.h))

    def set_materials(self, mats, mtype):
        """createMultiForceModel(particles, AverageTag, ...): mats = [(K, G, G0), ...],
        mtype[n] the material of each particle."""
        arr = _abi.materials_array(mats)
        t = None if mtype is None else np.ascontiguousarray(mtype, dtype=np.int32)
        _check(lib().pdk_set_materials(self.h, len(mats), arr, None if t is None else t.ctypes.data_as(_I32)))

    def set_pair_models(self, pairs, mtype):
        """createMultiForceModel(particles, indexing, models...): pairs[a * ntypes + b]
        holds the constants of models[indexing(a, b)] (see _abi.pair_models_array)."""
        nt = int(round(len(pairs) ** 0.5))
        if nt * nt != len(pairs):
            raise ValueError("pairs must hold ntypes^2 entries")
        arr = _abi.pair_models_array(pairs)
        t = None if mtype is None else np.ascontiguousarray(mtype, dtype=np.int32)
        _check(lib().pdk_set_pair_models(self.h, nt, arr, None if t is None else t.ctypes.data_as(_I32)))

    def compute_energy_noalloc(self):
        """pdk_compute_energy through the library's fixed scratch (no allocation)."""
        W = C.c_double()
        D = C.c_double()
        _check(lib().pdk_compute_energy_noalloc(self.h, C.byref(W), C.byref(D)))
        return W.value, D.value

    def compute_energy_async(self):
        _check(lib().pdk_compute_energy_async(self.h))

    def set_prescribed(self, field, idx, a, b=None):
        """Values a + b t set before each force pass (and at t = 0 in init):
        field "u" (a, b [count][3]) or "temperature" (a, b [count])."""
        nc = 3 if field == "u" else 1
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        a = _f64(a, (idx.size, nc)) if idx.size else np.zeros((0, nc))
        b = _f64(b, (idx.size, nc)) if (b is not None and idx.size) else None
        _check(lib().pdk_set_prescribed(self.h, _abi.FIELD[field], idx.ctypes.data_as(_I64), idx.size, _dp(a),
                                        _dp(b)))

    def set_contact(self, contact):
        """A contact model on this PD specimen (an _abi.ContactDesc; pdk_set_contact)."""
        _check(lib().pdk_set_contact(self.h, C.byref(contact)))

    def contact_pairs(self):
        t = C.c_int64()
        _check(lib().pdk_contact_pairs(self.h, C.byref(t)))
        return t.value

    def set_field(self, field, values):
        _check(lib().pdk_set_field(self.h, _abi.FIELD[field], _dp(_f64(values, (self.n,)))))

    def series(self, region, field, component=0, reducer="max", capacity=1024, gate=None, threshold=0.0,
               weight_volume=False, ids=None):
        """Device-resident time series of field[component] reduced over a
        region (an _abi.*_region descriptor, or ids= for an explicit list)."""
        return Series(self, region, field, component, reducer, capacity, gate, threshold, weight_volume, ids)

    def steps_done(self):
        s = C.c_int64()
        _check(lib().pdk_steps_done(self.h, C.byref(s)))
        return s.value

    def get(self, field):
        shape = _abi.field_shape(field, self.n)
        out = np.empty(shape, dtype=np.float64)
        _check(lib().pdk_download(self.h, FIELD[field], _dp(out)))
        return out

    def sync(self):
        _check(lib().pdk_sync(self.h))

    def bytes(self):
        b = C.c_uint64()
        _check(lib().pdk_spec_bytes(self.h, C.byref(b)))
        return b.value

    def layout(self):
        lr = C.c_int32()
        ns = C.c_int32()
        sl = C.c_int32()
        _check(lib().pdk_spec_layout(self.h, C.byref(lr), C.byref(ns), C.byref(sl)))
        return {"lattice_rows": bool(lr.value), "stencil_entries": ns.value, "stencil_slots": bool(sl.value)}

    def set_graph(self, enable):
        _check(lib().pdk_set_graph(self.h, 1 if enable else 0))

    def set_tuning(self, unroll):
        _check(lib().pdk_set_tuning(self.h, unroll))

    def autotune(self, launches=3):
        u = C.c_int32()
        _check(lib().pdk_autotune(self.h, launches, C.byref(u)))
        return u.value

    def timing(self, enable=True):
        _check(lib().pdk_kernel_timing(self.h, 1 if enable else 0))

    def kernel_time(self, cls):
        ms = C.c_double()
        cnt = C.c_int64()
        _check(lib().pdk_kernel_time(self.h, cls, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def kernel_name(self, cls):
        buf = C.create_string_buffer(256)
        _check(lib().pdk_kernel_name(self.h, cls, buf, 256))
        return buf.value.decode()


class Series:
    """pdk_series: a region, a per-particle quantity and a reducer
    (OutputTimeSeries, CabanaPD_OutputProfiles.hpp:113-215)."""

    def __init__(self, spec, region, field, component, reducer, capacity, gate, threshold, weight_volume, ids):
        self.spec = spec
        r = region if region is not None else _abi.RegionDesc()
        self._ids = None
        if ids is not None:
            self._ids = np.ascontiguousarray(ids, dtype=np.int64)
            r.kind = _abi.PDK_REGION_INDICES
            r.ids = self._ids.ctypes.data_as(_I64)
            r.count = self._ids.size
        q = _abi.Quantity()
        q.field = FIELD[field]
        q.component = component
        q.gate_field = -1 if gate is None else FIELD[gate]
        q.gate_threshold = threshold
        q.weight_volume = 1 if weight_volume else 0
        self.h = _VP()
        _check(lib().pdk_series_create(spec.h, C.byref(r), C.byref(q), _abi.REDUCER[reducer], capacity,
                                       C.byref(self.h)))
        self._ids = None

    def close(self):
        if self.h:
            _check(lib().pdk_series_destroy(self.h))
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def members(self):
        c = C.c_int64()
        _check(lib().pdk_series_members(self.h, C.byref(c)))
        return c.value

    def update(self, time=0.0):
        _check(lib().pdk_series_update(self.h, time))

    def values(self):
        c = C.c_int64()
        _check(lib().pdk_series_values(self.h, 0, None, None, C.byref(c)))
        t = np.empty(c.value)
        v = np.empty(c.value)
        _check(lib().pdk_series_values(self.h, c.value, _dp(t), _dp(v), C.byref(c)))
        return t, v

    def gather(self):
        m = self.members()
        ids = np.empty(m, dtype=np.int64)
        v = np.empty(m)
        c = C.c_int64()
        _check(lib().pdk_series_gather(self.h, m, ids.ctypes.data_as(_I64), _dp(v), C.byref(c)))
        return ids, v
