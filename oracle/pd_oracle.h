/*
 * pd_oracle.h -- CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A C++17/OpenMP restatement of CabanaPD's force/update path, used by the
 * test suite as the parity checker and by bench.py as the CPU baseline.
 * Nothing in the product (libpdk.so, the drop-in headers) links or calls it.
 *
 * The functions mirror the engine's C ABI (include/pdk.h) so a parity test
 * drives both with the same sequence of calls.  Particle order is the
 * caller's; neighbor rows are ascending in caller id.
 */
#ifndef PD_ORACLE_H
#define PD_ORACLE_H

#include <stdint.h>

#include "../include/pdk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys_s* orc_sys;

void orc_set_threads(int32_t nthreads);
int32_t orc_get_threads(void);

int orc_create(const pdk_model_desc* desc, int64_t n, const double* x, const double* u,
               const double* v, const double* vol, const double* rho, const int32_t* nofail,
               const double* radius, orc_sys* out);
int orc_destroy(orc_sys s);
int orc_model_constants(const pdk_model_desc* desc, pdk_model_constants* out);
int orc_critical_timestep(double dx, double dy, double dz, double K, double rho, double horizon,
                          double safety_factor, double* dt_out);

int orc_build_neighbors(orc_sys s);
int orc_bond_stats(orc_sys s, int64_t* total, int32_t* max_neighbors);
int orc_prenotch(orc_sys s, int32_t n, const double* p0, const double* v1, const double* v2);
int orc_get_bonds(orc_sys s, int64_t* offsets, int64_t* nbr, uint8_t* intact);
int orc_set_bc(orc_sys s, int32_t kind, const int64_t* idx, int64_t n, double value);
int orc_set_body_force(orc_sys s, const double* b);

int orc_compute_weighted_volume(orc_sys s);
int orc_compute_dilatation(orc_sys s);
int orc_compute_force(orc_sys s, int32_t reset);
int orc_compute_energy(orc_sys s, double* W, double* D);
int orc_compute_stress(orc_sys s);
int orc_set_materials(orc_sys s, int32_t nmat, const pdk_material* mats, const int32_t* type);
int orc_set_pair_models(orc_sys s, int32_t ntypes, const pdk_pair_model* pairs, const int32_t* type);
int orc_set_prescribed(orc_sys s, int32_t field, const int64_t* idx, int64_t count, const double* a,
                       const double* b);
int orc_set_field(orc_sys s, int32_t field, const double* values);
int orc_get_plastic(orc_sys s, double* out);
int orc_init(orc_sys s);
int orc_step(orc_sys s, int32_t nsteps);
int orc_get(orc_sys s, int32_t field, double* out);
int orc_set_state(orc_sys s, const double* u, const double* v);
int orc_set_contact(orc_sys s, const pdk_contact_desc* contact);
int orc_contact_pairs(orc_sys s, int64_t* pairs);
/* The correctly rounded powers of the contact models, elementwise:
 * which = 0: x^1.5, 1: x^(double nearest 1/3), 2: x^3, 3: x^0.5. */
int orc_cr_pow(int32_t which, int64_t n, const double* x, double* out);

/* Single-bond kernels exposed for closed-form pins. */
int orc_bond_pmb(const pdk_model_desc* desc, const double* xi3, const double* eta3, double vol_j,
                 double* f3, double* w, int32_t* breaks);
int orc_bond_prenotch(const double* v1, const double* v2, const double* p0, const double* xi,
                      const double* xj, int32_t* keep);
int orc_jkr_normal_force(const pdk_model_desc* desc, double r, double vn, double vol, double rho,
                         double* fn);
int orc_jkr_patch_radius(const pdk_model_desc* desc, double delta_n, double* a);

#ifdef __cplusplus
}
#endif

#endif
