/*
 * pdk.h -- C ABI of the MI355X-native peridynamics force/update engine.
 *
 * This is the drop-in boundary for CabanaPD's force/update hot path.  The
 * reference exposes that path only as C++ templates instantiated in user
 * translation units (no FFI).  The C++ drop-in headers under
 * include/CabanaPD/ keep the reference's class and function names and call
 * these entry points; the test suite calls them through ctypes.
 *
 * Conventions
 *   - Every entry point returns int: 0 = ok, nonzero = error; the message is
 *     available from pdk_last_error() (thread-local).  No C++ exception
 *     crosses this boundary (the reference throws std::runtime_error on the
 *     host, e.g. CabanaPD_Input.hpp:64-106; here the drop-in headers turn a
 *     nonzero status back into std::runtime_error).
 *   - Host arrays are borrowed for the duration of the call only.  Vector
 *     fields are AoS double[n][3] in the caller's particle order, exactly the
 *     reference's slice layout (AoSoA vector length 1, CabanaPD_Particles.hpp:105-126).
 *   - Device memory is owned by the opaque handles.  The engine keeps
 *     particles in its own device order (lattice rows, x fastest, for
 *     lattice specimens; Morton order otherwise) and translates on
 *     upload/download.
 *   - Calls on one spec are synchronous with respect to the host unless the
 *     name says otherwise (pdk_submit / pdk_poll), matching the reference's
 *     Kokkos::fence() after every kernel (e.g. force/CabanaPD_PMB.hpp:121-123).
 */
#ifndef PDK_H
#define PDK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- model selection (reference tags: CabanaPD_Types.hpp) --------------- */
enum pdk_model_kind {
    PDK_MODEL_PMB = 0,        /* ForceModel<PMB, Elastic, NoFracture|Fracture>  force_models/CabanaPD_PMB.hpp:149-231 */
    PDK_MODEL_LPS = 1,        /* ForceModel<LPS, Elastic, NoFracture|Fracture>  force_models/CabanaPD_LPS.hpp:151-253 */
    PDK_MODEL_HERTZ = 2,      /* HertzianModel     force_models/CabanaPD_Hertzian.hpp:23-88 */
    PDK_MODEL_HERTZ_JKR = 3,  /* HertzianJKRModel  force_models/CabanaPD_HertzianJKR.hpp:23-177 */
    PDK_MODEL_LINEAR_PMB = 4, /* ForceModel<LinearPMB, Elastic, NoFracture>  force_models/CabanaPD_PMB.hpp:575-597,
                                 Force<LinearPMB> force/CabanaPD_PMB.hpp:406-538 (linearized stretch) */
    PDK_MODEL_LINEAR_LPS = 5  /* ForceModel<LinearLPS, Elastic, NoFracture>  force_models/CabanaPD_LPS.hpp:255-273,
                                 Force<LinearLPS> force/CabanaPD_LPS.hpp:577-724 (linearized stretch in the
                                 force and energy; weighted volume and dilatation as LPS) */
};

enum pdk_precision {
    PDK_FP64 = 0,   /* plain FP64: the reference arithmetic, operation for operation */
    PDK_MIXED = 1   /* FP32 bond math, FP64 accumulation/integration, FP64 guard band for breaks */
};

enum pdk_field {
    PDK_FIELD_X = 0,      /* reference position  double[n][3] */
    PDK_FIELD_U = 1,      /* displacement        double[n][3] */
    PDK_FIELD_V = 2,      /* velocity            double[n][3] */
    PDK_FIELD_F = 3,      /* force density       double[n][3] */
    PDK_FIELD_VOL = 4,    /* volume              double[n] */
    PDK_FIELD_RHO = 5,    /* density             double[n] */
    PDK_FIELD_W = 6,      /* strain energy density double[n] */
    PDK_FIELD_PHI = 7,    /* damage              double[n] */
    PDK_FIELD_M = 8,      /* LPS weighted volume double[n] */
    PDK_FIELD_THETA = 9,  /* LPS dilatation      double[n] */
    PDK_FIELD_STRESS = 10, /* stress tensor      double[n][3][3], stress(i, a, b) at [9 i + 3 a + b];
                              the value of the last pdk_compute_stress */
    PDK_FIELD_TEMPERATURE = 11, /* temperature  double[n] (thermal models; Particles::sliceTemperature) */
    PDK_FIELD_COUNT = 12
};

enum pdk_mechanics {
    PDK_MECH_ELASTIC = 0,  /* Elastic */
    PDK_MECH_EPP = 1       /* ElasticPerfectlyPlastic: PMB with Fracture only (force_models/CabanaPD_PMB.hpp:74-147,
                              232-284); one plastic stretch per bond, zero when the bonds are built */
};

enum pdk_thermal {
    PDK_TEMP_INDEPENDENT = 0,  /* TemperatureIndependent */
    PDK_TEMP_DEPENDENT = 1,    /* TemperatureDependent: PMB / LinearPMB stretch s - alpha (T_avg - temp0)
                                  (CabanaPD_ForceModels.hpp:165-211); with Fracture the critical stretch
                                  test uses (1 + s0 + alpha (T_avg - temp0))^2 (:213-256) */
    PDK_TEMP_DYNAMIC = 2       /* DynamicTemperature: TemperatureDependent PMB plus peridynamic heat
                                  conduction with forward Euler every thermal_subcycle steps
                                  (CabanaPD_HeatTransfer.hpp:24-182, CabanaPD_ForceModels.hpp:262-310,
                                  CabanaPD_Solver.hpp:340-347) */
};

enum pdk_bc_kind {
    PDK_BC_FORCE_VALUE = 0,   /* f(pid,d) = value   CabanaPD_Boundary.hpp:77-119 */
    PDK_BC_FORCE_UPDATE = 1   /* f(pid,d) += value  CabanaPD_Boundary.hpp:121-163 */
};

/* Model/specimen description.  Derived constants (c, s0, bond_break_coeff,
 * theta/s coefficients, Hertz/JKR constants) are computed by the engine with
 * the reference's expressions (see pdk_model_constants). */
typedef struct pdk_model_desc {
    int32_t model;          /* pdk_model_kind */
    int32_t fracture;       /* 0 = NoFracture, 1 = Fracture */
    int32_t influence;      /* 0 or 1: BaseFractureModel s0 selector / LPS influence (1 => 1/|xi|) */
    int32_t precision;      /* pdk_precision */
    double horizon;         /* force horizon delta; PD neighbor cutoff = horizon + neighbor_tol */
    double K;               /* bulk modulus */
    double G;               /* shear modulus (LPS) */
    double G0;              /* fracture energy (Fracture) */
    double dt;              /* time step (VelocityVerlet, CabanaPD_Integrate.hpp:82-86) */
    double neighbor_tol;    /* Neighbor ctor tolerance, reference default 1e-14 (CabanaPD_Neighbor.hpp:46) */
    /* contact (DEM) */
    double radius;          /* ContactModel::radius (uniform) */
    double radius_extend;   /* ContactModel::radius_extend (list skin) */
    double nu;              /* Poisson's ratio */
    double E;               /* Young's modulus */
    double restitution;     /* e */
    double gamma;           /* JKR surface energy */
    double mu_friction;     /* JKR Coulomb friction */
    int32_t rebuild_every;  /* 0: rebuild when max|u-u_build| > extend (CabanaPD_Neighbor.hpp:116-119); k>0: every k steps (plus the same safety trigger) */
    int32_t polydisperse;   /* 1: per-particle radii (extension; equal radii reduce to the reference) */
    int32_t ordering;       /* particle order on the device: PDK_ORDER_AUTO / _MORTON / _LATTICE_ROWS */
    int32_t bond_layout;    /* PDK_BONDS_AUTO / PDK_BONDS_COMPACT */
    /* mechanics and thermal coupling (PMB family; zero = Elastic, TemperatureIndependent);
       both precision forms: mixed keeps the thermal strain and the plastic stretch in FP64 */
    int32_t mechanics;      /* pdk_mechanics */
    int32_t thermal;        /* pdk_thermal */
    double sigma_y;         /* yield stress (EPP): s_Y = sigma_y / (3 K) */
    double alpha;           /* thermal expansion coefficient (TemperatureDependent) */
    double temp0;           /* reference temperature (TemperatureDependent) */
    /* heat transfer (DynamicTemperature) */
    double kappa;           /* thermal conductivity */
    double cp;              /* specific heat capacity */
    int32_t thermal_subcycle;  /* heat conduction at steps k with k % thermal_subcycle == 0, with dt_th =
                                  thermal_subcycle * dt (inputs "thermal_subcycle_steps") */
    int32_t linear_microconductivity; /* 0: constant microconductivity (the reference's default);
                                         1: 4 thermal_coeff (1 - |xi| / horizon) */
} pdk_model_desc;

enum pdk_ordering {
    PDK_ORDER_AUTO = 0,         /* lattice rows when the positions form a lattice, else Morton */
    PDK_ORDER_MORTON = 1,       /* Morton order of the reference positions */
    PDK_ORDER_LATTICE_ROWS = 2  /* lattice index order, x fastest (fails if not a lattice) */
};

enum pdk_bond_layout {
    PDK_BONDS_AUTO = 0,     /* stencil slots on full lattices (neighbor ids computed, not stored), else compact */
    PDK_BONDS_COMPACT = 1   /* always the compact form: one stored neighbor id per bond */
};

/* Derived constants, for inspection and tests. */
typedef struct pdk_model_constants {
    double c, s0, bond_break_coeff, theta_coeff, s_coeff;
    double Rs, Es, Gs, beta, coeff_h_n, coeff_h_d, Gamma, a0, delta_tear, fc;
    double s_Y;  /* EPP yield stretch sigma_y / 3 / K (force_models/CabanaPD_PMB.hpp:90-96) */
    double thermal_coeff;  /* 9/2 kappa / pi / horizon^3 (CabanaPD_ForceModels.hpp:277-288) */
} pdk_model_constants;

typedef struct pdk_ctx_s* pdk_ctx;
typedef struct pdk_spec_s* pdk_spec;

/* ---- errors --------------------------------------------------------------- */
const char* pdk_last_error(void);

/* ---- derived constants ----------------------------------------------------
 * PMB: c = 18K/(pi delta^4)           force_models/CabanaPD_PMB.hpp:41-47
 * s0, bond_break_coeff                CabanaPD_ForceModels.hpp:93-102, LPS override force_models/CabanaPD_LPS.hpp:209-223
 * EPP s0 = (5 G0 / sigma_y / delta + sigma_y / K) / 6  force_models/CabanaPD_PMB.hpp:255-264
 *   (with TemperatureDependent the thermal fracture model's s0 = sqrt(5 G0 / 9 / K / delta),
 *   influence 1, force_models/CabanaPD_PMB.hpp:366-396, CabanaPD_ForceModels.hpp:228-236)
 * LPS theta/s coefficients            force_models/CabanaPD_LPS.hpp:73-83
 * Hertz / JKR constants               force_models/CabanaPD_Hertzian.hpp:41-56, force_models/CabanaPD_HertzianJKR.hpp:51-80 */
int pdk_model_constants_get(const pdk_model_desc* desc, pdk_model_constants* out);

/* Inputs::computeCriticalTimeStep lattice sum (CabanaPD_Input.hpp:217-336),
 * returns safety * dt_crit. */
int pdk_critical_timestep(double dx, double dy, double dz, double K, double rho,
                          double horizon, double safety_factor, double* dt_out);

/* ---- context ---------------------------------------------------------------
 * One context per GPU.  hbm_budget (bytes, 0 = 90% of free) bounds what specs
 * and the serving admission may allocate. */
int pdk_ctx_create(int device, uint64_t hbm_budget, pdk_ctx* ctx);
int pdk_ctx_destroy(pdk_ctx ctx);
int pdk_ctx_bytes_in_use(pdk_ctx ctx, uint64_t* bytes);

/* ---- specimen lifetime -----------------------------------------------------
 * pdk_spec_create + pdk_spec_upload replace Particles::create(x, vol)
 * (CabanaPD_Particles.hpp:310-368) followed by Particles::update with the
 * user's init functor (:429-443).  nofail and radius may be NULL. */
int pdk_spec_create(pdk_ctx ctx, const pdk_model_desc* desc, int64_t n, pdk_spec* spec);
int pdk_spec_destroy(pdk_spec spec);
int pdk_spec_upload(pdk_spec spec, const double* x, const double* u, const double* v,
                    const double* vol, const double* rho, const int32_t* nofail,
                    const double* radius);
/* Replace u, v (caller order) on an existing spec; f is reset to zero. */
int pdk_spec_set_state(pdk_spec spec, const double* u, const double* v);
/* Restart dst from src's state between steps: displacement, velocity, force,
 * intact-bond bits, LPS weighted volume / dilatation, energy, damage,
 * temperature, plastic stretch and the step count are copied device to
 * device.  No counterpart in the reference (a Solver is never forked); it lets
 * one state be advanced in both precision forms.  The specs must share the
 * context, the particle count and device order, the model and the bond
 * layout (same positions, same build); their precision may differ.  dst must
 * have its bonds built; it counts as initialised afterwards. */
int pdk_spec_copy_state(pdk_spec dst, pdk_spec src);
int pdk_spec_bytes(pdk_spec spec, uint64_t* bytes);

/* ---- bond family -------------------------------------------------------------
 * Neighbor<NoFracture|Fracture> ctor (CabanaPD_Neighbor.hpp:44-66,199-210):
 * full list of j != i with |x_j - x_i|^2 <= (horizon + tol)^2 (reference
 * positions; current positions x+u for contact models).  Built on the device
 * by cell binning into a sliced CSR in device order; every bond starts intact. */
int pdk_build_bonds(pdk_spec spec);
/* Neighbor statistics (Neighbor::getStatistics, CabanaPD_Neighbor.hpp:82-102). */
int pdk_bond_stats(pdk_spec spec, int64_t* total_bonds, int32_t* max_neighbors);
/* Prenotch<N>::create (CabanaPD_Prenotch.hpp:215-263): every bond whose
 * segment crosses one of the n plane patches (p0 + a v1 + b v2) is broken.
 * p0, v1, v2 are double[n][3]. */
int pdk_prenotch(pdk_spec spec, int32_t n, const double* p0, const double* v1, const double* v2);
/* Bonds in caller order: offsets[n+1], nbr[total] (caller ids, ascending per
 * row), intact[total] (1 = mu 1, 0 = broken). Any pointer may be NULL.
 * capacity is the length of nbr/intact; the call fails (writing nothing) when
 * the current bond count (pdk_bond_stats) exceeds it. */
int pdk_download_bonds(pdk_spec spec, int64_t capacity, int64_t* offsets, int64_t* nbr, uint8_t* intact);

/* ---- boundary conditions / body terms --------------------------------------
 * createBoundaryCondition(ForceValueBCTag|ForceUpdateBCTag, value, ...)
 * (CabanaPD_Boundary.hpp:77-176): idx are caller particle ids.  Applied to f
 * after the force pass, before the final half step (CabanaPD_Solver.hpp:365-367),
 * and once at init (:272-274).  A second call replaces the first. */
int pdk_set_bc(pdk_spec spec, int32_t kind, const int64_t* idx, int64_t n, double value);
/* Constant per-particle body force density b[n][3] (caller order) added to f
 * every step: BodyTerm with a time-independent functor (CabanaPD_BodyTerm.hpp:46-63),
 * e.g. gravity f_z -= 9.8 rho, or a face traction.  NULL clears it. */
int pdk_set_body_force(pdk_spec spec, const double* b);
/* Prescribed values where the reference applies boundary conditions and body
 * terms built with force_update = false: after the velocity-Verlet initial
 * half step and before the force pass of step k at t = k dt
 * (CabanaPD_Solver.hpp:350-352), and at t = 0 at the start of pdk_init
 * (:255-257).  field is PDK_FIELD_U (a, b double[count][3]: the three
 * displacement components, as the dogbone grips of
 * examples/mechanics/dogbone_tensile_test.cpp set them) or
 * PDK_FIELD_TEMPERATURE (a, b double[count]: the imposed temperature of
 * examples/thermomechanics/thermal_deformation.cpp).  Member idx[m] (caller
 * id, no repeats) gets a[m] + b[m] * t, evaluated in that order without
 * contraction; b may be NULL (constant values).  One set per field: a second
 * call replaces the first, count = 0 removes it.  Specs with prescribed
 * values are stepped with pdk_step (not the batch engine). */
int pdk_set_prescribed(pdk_spec spec, int32_t field, const int64_t* idx, int64_t count, const double* a,
                       const double* b);
/* Set a per-particle field from caller-order values: PDK_FIELD_TEMPERATURE
 * (thermal models; initially zero, as the reference's particle fields). */
int pdk_set_field(pdk_spec spec, int32_t field, const double* values);

/* ---- multi-material specimens ---------------------------------------------------
 * createMultiForceModel( particles, AverageTag, m_0, .., m_{nmat-1} )
 * (CabanaPD_ForceModelsMulti.hpp:338-516) for 1 <= nmat <= 4 materials that
 * share the spec's model kind, fracture form, horizon and influence (as the
 * reference requires): material t has its own K, G, G0; type[n] (caller
 * order) holds each particle's material (Particles::sliceType).  A bond
 * between materials a != b uses the averaged model: PMB c = (c_a + c_b) / 2
 * (force_models/CabanaPD_PMB.hpp:51-55); s0 = sqrt((s0_a^2 K_a + s0_b^2 K_b) /
 * (K_a + K_b)), bond_break_coeff = (1 + s0)^2 (CabanaPD_ForceModels.hpp:114-121);
 * LPS keeps each material's theta/s coefficients on its own particle's terms
 * (force_models/CabanaPD_LPS.hpp:55-69).  Both precision forms (the mixed
 * form with each pair's c, critical stretch and guard band); the force pass
 * runs on lattice-row windows where the layout has them.  Call after
 * pdk_spec_upload and before pdk_init. */
typedef struct pdk_material {
    double K, G, G0;
} pdk_material;
int pdk_set_materials(pdk_spec spec, int32_t nmat, const pdk_material* materials, const int32_t* type);

/* ---- pair models (explicit indexing) --------------------------------------------
 * createMultiForceModel( particles, indexing, m_0, .., m_{k-1} )
 * (CabanaPD_ForceModelsMulti.hpp:351-358; dispatch :228-250): the bond from a
 * particle of material type a to one of type b uses model m_{indexing(a, b)}.
 * pairs[a * ntypes + b] (1 <= ntypes <= 4) holds that model's constants as the
 * reference's model objects carry them: c (PMB forms), bond_break_coeff =
 * (1 + s0)^2 (fracture; +inf for a pair that never breaks, as
 * BaseNoFractureModel), and the LPS coefficients theta_coeff[2], s_coeff[2]
 * (force_models/CabanaPD_LPS.hpp:41-44, 117-132: [0] on particle i's terms, [1]
 * on particle j's).  type[n] (caller order) holds each particle's material.
 * This engine folds the LPS coefficients into per-particle terms, so an LPS
 * spec needs theta_coeff[0] and s_coeff[0] of every pair (a, b) to depend on a
 * only and [1] on b only, with the same per-type values (any table built from
 * per-material models and their averages is of this form); other tables are
 * rejected.  pdk_set_materials is the special case of the averaged pair table.
 * Both precision forms; call after pdk_spec_upload and before pdk_init. */
typedef struct pdk_pair_model {
    double c;
    double bond_break_coeff;
    double theta_coeff[2];
    double s_coeff[2];
} pdk_pair_model;
int pdk_set_pair_models(pdk_spec spec, int32_t ntypes, const pdk_pair_model* pairs, const int32_t* type);

/* ---- contact between the particles of a PD specimen ---------------------------
 * Solver( inputs, particles, force_model, contact_model )
 * (CabanaPD_Solver.hpp:134-151, 332-404): a second neighbor list on the current
 * positions x + u with cutoff 2 radius + radius_extend (+ the neighbor tolerance
 * at the first build, CabanaPD_Neighbor.hpp:60-64), rebuilt when the largest
 * displacement since the last build exceeds radius_extend
 * (CabanaPD_Neighbor.hpp:106-132, VelocityVerlet<Contact>,
 * CabanaPD_Integrate.hpp:147-193).  Each step the contact forces are added
 * after the PD force pass and before the force boundary conditions and the
 * final half step; Solver::init computes no contact force.  NormalRepulsion
 * (force/CabanaPD_Contact.hpp:44-80, force_models/CabanaPD_Contact.hpp:53-90):
 * for every listed pair with r < radius, f_i += 15 c ((r - radius) / delta)
 * V_j (x_j + u_j - x_i - u_i) / r with c = 18 K / (pi delta^4); FP64 in every
 * precision.  Call after pdk_build_bonds and before pdk_init; specs with
 * contact are stepped with pdk_step (not the batch engine). */
enum pdk_contact_kind {
    PDK_CONTACT_NORMAL_REPULSION = 0
};
typedef struct pdk_contact_desc {
    int32_t model;          /* pdk_contact_kind */
    int32_t reserved;
    double radius;          /* ContactModel::radius */
    double radius_extend;   /* ContactModel::radius_extend */
    double horizon;         /* NormalRepulsionModel delta */
    double K;               /* NormalRepulsionModel K */
} pdk_contact_desc;
int pdk_set_contact(pdk_spec spec, const pdk_contact_desc* contact);
/* Pairs in the contact list of a PD specimen with contact (0 without). */
int pdk_contact_pairs(pdk_spec spec, int64_t* pairs);

/* ---- force / energy (free functions, CabanaPD_Force.hpp:176-219) ---------- */
int pdk_compute_weighted_volume(pdk_spec spec);   /* Force<LPS>::computeWeightedVolume; no-op for PMB */
int pdk_compute_dilatation(pdk_spec spec);        /* Force<LPS>::computeDilatation; no-op for PMB */
int pdk_compute_force(pdk_spec spec, int32_t reset);  /* computeForce(model, force, particles, neighbor, reset) */
int pdk_compute_energy(pdk_spec spec, double* total_strain_energy, double* total_damage);  /* computeEnergy */
/* computeStress (CabanaPD_Force.hpp:221-236): reset, then per bond
 * stress(i, a, b) += f_a xi_b with f half the bond force
 * (Force<*>::computeStressFull, force/CabanaPD_PMB.hpp:157-202, 341-401, 492-536;
 * force/CabanaPD_LPS.hpp:221-267, 506-572, 677-723).  FP64 in every precision
 * (an output pass); PD models only. */
int pdk_compute_stress(pdk_spec spec);

/* ---- time stepping (Solver::init / runStep, CabanaPD_Solver.hpp:229-404) ----
 * pdk_init: Solver::init(bc) order: weighted volume (LPS NoFracture), force,
 * energy, force BC at t = 0.
 * pdk_step: nsteps x runStep(step, bc) -- velocity-Verlet fused into the
 * force pass epilogue; energy/damage are not recomputed (call
 * pdk_compute_energy at output steps, as Solver::output does). */
int pdk_init(pdk_spec spec);
int pdk_step(pdk_spec spec, int32_t nsteps);
int pdk_steps_done(pdk_spec spec, int64_t* steps);

/* Download a field in caller order (double[n][3] or double[n]). */
int pdk_download(pdk_spec spec, int32_t field, double* out);
int pdk_sync(pdk_spec spec);

/* ---- region outputs on the device ----------------------------------------------
 * A series binds a region of the specimen (its members fixed at creation,
 * from the reference positions), a per-particle quantity and a reducer.
 * pdk_series_update reduces the quantity over the members into the next slot
 * of a device-resident time series -- OutputTimeSeries::update
 * (CabanaPD_OutputProfiles.hpp:146-163) -- without copying anything to the
 * host: it is enqueued on the spec's stream and returns at once (the one
 * asynchronous call besides the serving calls).  pdk_series_values downloads
 * the series (OutputTimeSeries::print, :165-187); pdk_series_gather reads the
 * quantity of every member now, members in ascending caller id
 * (createOutputProfile, :26-62).  The members of a series are in ascending
 * caller id, the reference's steering-vector order (CabanaPD_Geometry.hpp:170-273).
 * Sum reduces in a fixed order of its own (reproducible, not the reference's
 * order: compare sums to rounding); Max/Min are exact. */
enum pdk_region_kind {
    PDK_REGION_ALL = 0,       /* every particle */
    PDK_REGION_BOX = 1,       /* Region<RectangularPrism>: low[d] <= x_d <= high[d]  CabanaPD_Geometry.hpp:54-82 */
    PDK_REGION_CYLINDER = 2,  /* Region<Cylinder>: r_in^2 <= (x-xc)^2+(y-yc)^2 <= r_out^2, low_z <= z <= high_z  :84-135 */
    PDK_REGION_LINE = 3,      /* Region<Line>: the coordinates (dim+1)%3 and (dim+2)%3 strictly within
                                 +-half[0], +-half[1]  :137-168 */
    PDK_REGION_INDICES = 4    /* the caller ids in ids[count] (a region given by a user functor,
                                 evaluated by the caller), any order, no repeats */
};
typedef struct pdk_region_desc {
    int32_t kind;             /* pdk_region_kind */
    int32_t dim;              /* LINE: profile dimension */
    double low[3], high[3];   /* BOX */
    double r_in, r_out, low_z, high_z, x_center, y_center;  /* CYLINDER */
    double half[2];           /* LINE */
    const int64_t* ids;       /* INDICES: copied at creation */
    int64_t count;            /* INDICES */
} pdk_region_desc;
enum pdk_reducer {
    PDK_REDUCE_SUM = 0,       /* Kokkos::Sum<double> */
    PDK_REDUCE_MAX = 1,       /* Kokkos::Max<double>: identity -DBL_MAX */
    PDK_REDUCE_MIN = 2        /* Kokkos::Min<double>: identity +DBL_MAX */
};
/* value(p) = field(p, component); where gate_field >= 0 it is value(p) if
 * gate_field(p) > gate_threshold, else 0 (the crack-extent functor of
 * examples/mechanics/kalthoff_winkler.cpp:135-142: damage > 0.3 ? y : 0);
 * weight_volume = 1 multiplies by the particle volume (sum of W: total strain
 * energy).  Energy and damage are the values of the last energy pass. */
typedef struct pdk_quantity {
    int32_t field;            /* pdk_field */
    int32_t component;        /* 0..2 for vector fields, 0 for scalars */
    int32_t gate_field;       /* -1 or a scalar pdk_field */
    int32_t weight_volume;    /* 0 or 1 */
    double gate_threshold;
} pdk_quantity;
typedef struct pdk_series_s* pdk_series;
int pdk_series_create(pdk_spec spec, const pdk_region_desc* region, const pdk_quantity* q, int32_t reducer,
                      int64_t capacity, pdk_series* series);
int pdk_series_destroy(pdk_series series);
int pdk_series_members(pdk_series series, int64_t* count);
int pdk_series_update(pdk_series series, double time);
int pdk_series_values(pdk_series series, int64_t capacity, double* times, double* values, int64_t* count);
int pdk_series_gather(pdk_series series, int64_t capacity, int64_t* ids, double* values, int64_t* count);
/* computeEnergy (CabanaPD_Force.hpp:200-219) without reading the totals
 * back: refreshes W and damage on the device for the series that follow. */
int pdk_compute_energy_async(pdk_spec spec);
/* No counterpart in the reference.  pdk_compute_energy for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same W, damage and totals, bit
 * for bit, with the per-block partial sums in a fixed device array of the
 * library instead of the spec's own buffer.  The call allocates nothing.
 * Limit: 16,777,216 particles per spec (a larger spec fails with an error).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (two host threads, two streams) race and must be serialised by the
 * caller. */
int pdk_compute_energy_noalloc(pdk_spec spec, double* total_strain_energy, double* total_damage);

/* ---- tuning --------------------------------------------------------------------
 * Force-pass variant: unroll = 1, 2, 4 or 8 gathers the neighbor data of that
 * many bond slots per group; unroll = 0 (stencil-slot layout only) stages
 * the neighbor data of each lattice row window in LDS instead (also for
 * several materials, or a temperature-dependent and/or ElasticPerfectlyPlastic
 * PMB; several materials together with either, and the linearized models
 * with a temperature, take the gathers).  Results do not depend on the
 * choice.  The default is 0 where available, else 4 (2 for the mixed form
 * in the compact layout).
 * pdk_autotune times the candidates on the current state (after pdk_init;
 * the probes repeat the init force pass, which is idempotent) and keeps the
 * fastest. */
int pdk_set_tuning(pdk_spec spec, int32_t unroll);
/* Device layout chosen at upload/build: lattice-row order (1) or Morton (0);
 * the number of lattice stencil entries (0 = no stencil: the mixed form
 * gathers x_j for the bond vectors); stencil_slots = 1 when the bonds are in
 * the stencil-slot form (slot q of every row is stencil entry q), 0 for the
 * compact form.  Any pointer may be NULL. */
int pdk_spec_layout(pdk_spec spec, int32_t* lattice_rows, int32_t* stencil_entries, int32_t* stencil_slots);
int pdk_autotune(pdk_spec spec, int32_t launches_per_candidate, int32_t* unroll_out);
/* enable = 1: pdk_step replays the launches of two consecutive steps as one
 * captured HIP graph (not while kernel timing is enabled).  Default 0: every
 * kernel is launched individually, which measured faster for steps of
 * 30 us and more.  Results are identical either way. */
int pdk_set_graph(pdk_spec spec, int32_t enable);

/* ---- measurement -------------------------------------------------------------
 * When enabled, every kernel launch of the spec is bracketed by HIP events on
 * the spec's stream; pdk_kernel_time returns the summed duration (ms) and the
 * launch count of one kernel class:
 *   0 = force pass (PD force + fused VV epilogue, or contact force)
 *   1 = LPS weighted-volume/dilatation pass
 *   2 = energy/damage pass
 *   3 = bond build
 *   4 = other (prologue/VV-only, BC, contact list rebuild) */
int pdk_kernel_timing(pdk_spec spec, int32_t enable);
int pdk_kernel_time(pdk_spec spec, int32_t kernel_class, double* total_ms, int64_t* launches);
/* Name of the device function behind a kernel class for the spec's current
 * configuration (what rocprofv3 reports). */
int pdk_kernel_name(pdk_spec spec, int32_t kernel_class, char* buf, int32_t buflen);

/* ---- serving: continuous batching ---------------------------------------------
 * pdk_submit queues nsteps of a spec on the context's batch engine; the batch
 * engine packs all running specs into one launch per step over their
 * concatenated slice ranges.  Requests are admitted as slots and HBM budget
 * allow; pdk_poll reports completion (*done = 1) without blocking. */
int pdk_submit(pdk_ctx ctx, pdk_spec spec, int32_t nsteps, uint64_t* ticket);
int pdk_poll(pdk_ctx ctx, uint64_t ticket, int32_t* done);
/* Advance the batch engine by up to max_steps packed steps (0 = until idle).
 * Returns the number of packed steps launched in *launched. */
int pdk_batch_progress(pdk_ctx ctx, int32_t max_steps, int32_t* launched);
/* No counterpart in the reference (a CabanaPD run is one specimen per
 * process).  The packed step's launches are captured as one HIP graph per
 * configuration of running specs and replayed for up to steps packed steps
 * at a time, never past the next completion (the launch index is then read
 * from a device counter that the graph advances); results equal the eager
 * launches bit for bit.  steps = 0: eager launches only (the default: with the
 * host queueing steps ahead, eager launches keep the GPU as busy). */
int pdk_batch_set_graph(pdk_ctx ctx, int32_t steps);

#ifdef __cplusplus
}
#endif

#endif /* PDK_H */
