"""ctypes mirror of include/pdk.h (types and enums only)."""
import ctypes as C

PDK_MODEL_PMB = 0
PDK_MODEL_LPS = 1
PDK_MODEL_HERTZ = 2
PDK_MODEL_HERTZ_JKR = 3
PDK_MODEL_LINEAR_PMB = 4
PDK_MODEL_LINEAR_LPS = 5

PDK_FP64 = 0
PDK_MIXED = 1

FIELD = {
    "x": 0, "u": 1, "v": 2, "f": 3, "vol": 4, "rho": 5,
    "W": 6, "phi": 7, "m": 8, "theta": 9, "stress": 10, "temperature": 11,
}
PDK_MECH_ELASTIC = 0
PDK_MECH_EPP = 1
PDK_TEMP_INDEPENDENT = 0
PDK_TEMP_DEPENDENT = 1
PDK_TEMP_DYNAMIC = 2
VECTOR_FIELDS = {"x", "u", "v", "f"}


def field_shape(field, n):
    """Host shape of a downloaded field: [n][3], [n][3][3] (stress) or [n]."""
    if field in VECTOR_FIELDS:
        return (n, 3)
    if field == "stress":
        return (n, 3, 3)
    return (n,)

PDK_BC_FORCE_VALUE = 0
PDK_BC_FORCE_UPDATE = 1

PDK_REGION_ALL, PDK_REGION_BOX, PDK_REGION_CYLINDER, PDK_REGION_LINE, PDK_REGION_INDICES = range(5)
REDUCER = {"sum": 0, "max": 1, "min": 2}


class RegionDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("dim", C.c_int32),
        ("low", C.c_double * 3), ("high", C.c_double * 3),
        ("r_in", C.c_double), ("r_out", C.c_double), ("low_z", C.c_double), ("high_z", C.c_double),
        ("x_center", C.c_double), ("y_center", C.c_double),
        ("half", C.c_double * 2),
        ("ids", C.POINTER(C.c_int64)), ("count", C.c_int64),
    ]


class Quantity(C.Structure):
    _fields_ = [
        ("field", C.c_int32), ("component", C.c_int32), ("gate_field", C.c_int32),
        ("weight_volume", C.c_int32), ("gate_threshold", C.c_double),
    ]


def box_region(low, high):
    r = RegionDesc()
    r.kind = PDK_REGION_BOX
    for d in range(3):
        r.low[d] = low[d]
        r.high[d] = high[d]
    return r


def cylinder_region(r_in, r_out, low_z, high_z, x_center=0.0, y_center=0.0):
    r = RegionDesc()
    r.kind = PDK_REGION_CYLINDER
    r.r_in, r.r_out, r.low_z, r.high_z, r.x_center, r.y_center = r_in, r_out, low_z, high_z, x_center, y_center
    return r


def line_region(dim, dx):
    """Region<Line>(profile_dim, dx): half a cell in the two other dimensions."""
    r = RegionDesc()
    r.kind = PDK_REGION_LINE
    r.dim = dim
    r.half[0] = dx[(dim + 1) % 3] / 2.0
    r.half[1] = dx[(dim + 2) % 3] / 2.0
    return r


def all_region():
    r = RegionDesc()
    r.kind = PDK_REGION_ALL
    return r


# Kernel classes of pdk_kernel_time.
KCLS_FORCE, KCLS_LPS, KCLS_ENERGY, KCLS_BUILD, KCLS_OTHER = range(5)


class ModelDesc(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("fracture", C.c_int32),
        ("influence", C.c_int32),
        ("precision", C.c_int32),
        ("horizon", C.c_double),
        ("K", C.c_double),
        ("G", C.c_double),
        ("G0", C.c_double),
        ("dt", C.c_double),
        ("neighbor_tol", C.c_double),
        ("radius", C.c_double),
        ("radius_extend", C.c_double),
        ("nu", C.c_double),
        ("E", C.c_double),
        ("restitution", C.c_double),
        ("gamma", C.c_double),
        ("mu_friction", C.c_double),
        ("rebuild_every", C.c_int32),
        ("polydisperse", C.c_int32),
        ("ordering", C.c_int32),
        ("bond_layout", C.c_int32),
        ("mechanics", C.c_int32),
        ("thermal", C.c_int32),
        ("sigma_y", C.c_double),
        ("alpha", C.c_double),
        ("temp0", C.c_double),
        ("kappa", C.c_double),
        ("cp", C.c_double),
        ("thermal_subcycle", C.c_int32),
        ("linear_microconductivity", C.c_int32),
    ]


PDK_ORDER_AUTO = 0
PDK_ORDER_MORTON = 1
PDK_ORDER_LATTICE_ROWS = 2
PDK_BONDS_AUTO = 0
PDK_BONDS_COMPACT = 1


PDK_CONTACT_NORMAL_REPULSION = 0


class ContactDesc(C.Structure):
    """pdk_contact_desc: a contact model on a PD specimen."""
    _fields_ = [("model", C.c_int32), ("reserved", C.c_int32), ("radius", C.c_double),
                ("radius_extend", C.c_double), ("horizon", C.c_double), ("K", C.c_double)]


def make_contact(radius, radius_extend, horizon, K, model=PDK_CONTACT_NORMAL_REPULSION):
    c = ContactDesc()
    c.model = model
    c.radius = radius
    c.radius_extend = radius_extend
    c.horizon = horizon
    c.K = K
    return c


class Material(C.Structure):
    _fields_ = [("K", C.c_double), ("G", C.c_double), ("G0", C.c_double)]


def materials_array(mats):
    """(K, G, G0) tuples -> a ctypes array of Material."""
    arr = (Material * len(mats))()
    for t, (K, G, G0) in enumerate(mats):
        arr[t].K, arr[t].G, arr[t].G0 = K, G, G0
    return arr


class PairModel(C.Structure):
    """pdk_pair_model: the constants of the model of one ordered type pair."""
    _fields_ = [("c", C.c_double), ("bond_break_coeff", C.c_double),
                ("theta_coeff", C.c_double * 2), ("s_coeff", C.c_double * 2)]


def pair_models_array(pairs):
    """ntypes^2 dicts (row-major over (type_i, type_j)) with keys c,
    bond_break_coeff, theta_coeff (2), s_coeff (2) -> a ctypes array."""
    arr = (PairModel * len(pairs))()
    for k, p in enumerate(pairs):
        arr[k].c = p.get("c", 0.0)
        arr[k].bond_break_coeff = p.get("bond_break_coeff", 0.0)
        for side in range(2):
            arr[k].theta_coeff[side] = p.get("theta_coeff", (0.0, 0.0))[side]
            arr[k].s_coeff[side] = p.get("s_coeff", (0.0, 0.0))[side]
    return arr


class ModelConstants(C.Structure):
    _fields_ = [(name, C.c_double) for name in (
        "c", "s0", "bond_break_coeff", "theta_coeff", "s_coeff",
        "Rs", "Es", "Gs", "beta", "coeff_h_n", "coeff_h_d", "Gamma", "a0",
        "delta_tear", "fc", "s_Y", "thermal_coeff")]


def make_desc(model="PMB", fracture=False, precision="fp64", horizon=0.0, K=0.0,
              G=0.0, G0=0.0, dt=1e-7, influence=None, neighbor_tol=1e-14,
              radius=0.0, radius_extend=0.0, nu=0.25, E=0.0, restitution=1.0,
              gamma=0.0, mu_friction=0.0, rebuild_every=0, polydisperse=False, ordering=0, bond_layout=0,
              mechanics="elastic", sigma_y=0.0, thermal=False, alpha=0.0, temp0=0.0, kappa=0.0, cp=0.0,
              thermal_subcycle=1, linear_microconductivity=False):
    """Model description with the reference's constructor defaults.

    influence defaults: PMB fracture s0 uses influence_type 1
    (CabanaPD_ForceModels.hpp:93-101), LPS uses 0
    (force_models/CabanaPD_LPS.hpp:46-47).  mechanics "epp" selects
    ElasticPerfectlyPlastic (PMB with fracture, yield stress sigma_y);
    thermal=True selects TemperatureDependent (expansion alpha, reference
    temperature temp0); thermal="dynamic" adds heat conduction (kappa, cp,
    every thermal_subcycle steps)."""
    kinds = {"PMB": PDK_MODEL_PMB, "LPS": PDK_MODEL_LPS, "LINEAR_PMB": PDK_MODEL_LINEAR_PMB,
             "LINEAR_LPS": PDK_MODEL_LINEAR_LPS,
             "HERTZ": PDK_MODEL_HERTZ, "HERTZ_JKR": PDK_MODEL_HERTZ_JKR}
    m = kinds[model.upper()]
    if influence is None:
        influence = 0 if m in (PDK_MODEL_LPS, PDK_MODEL_LINEAR_LPS) else 1
    d = ModelDesc()
    d.model = m
    d.fracture = 1 if fracture else 0
    d.influence = influence
    d.precision = {"fp64": PDK_FP64, "mixed": PDK_MIXED}[precision]
    d.horizon = horizon
    d.K = K
    d.G = G
    d.G0 = G0
    d.dt = dt
    d.neighbor_tol = neighbor_tol
    d.radius = radius
    d.radius_extend = radius_extend
    d.nu = nu
    d.E = E
    d.restitution = restitution
    d.gamma = gamma
    d.mu_friction = mu_friction
    d.rebuild_every = rebuild_every
    d.polydisperse = 1 if polydisperse else 0
    d.ordering = ordering
    d.bond_layout = bond_layout
    d.mechanics = {"elastic": PDK_MECH_ELASTIC, "epp": PDK_MECH_EPP}[mechanics.lower()]
    d.sigma_y = sigma_y
    d.thermal = (PDK_TEMP_DYNAMIC if thermal == "dynamic" else PDK_TEMP_DEPENDENT) if thermal else PDK_TEMP_INDEPENDENT
    d.kappa = kappa
    d.cp = cp
    d.thermal_subcycle = thermal_subcycle
    d.linear_microconductivity = 1 if linear_microconductivity else 0
    d.alpha = alpha
    d.temp0 = temp0
    return d
