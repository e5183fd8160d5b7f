/*
 * ot2d_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference OVERTURNS-2D iteration, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker of
 * the HIP path.  Never linked into, loaded by or called from the product.
 * Entry points mirror include/ot2d.h with an "or_" prefix and host memory.
 */
#ifndef OT2D_ORACLE_H
#define OT2D_ORACLE_H
#include "../include/ot2d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_case or_case;

int32_t or_case_create(const ot2d_params* p, int32_t nmesh, const ot2d_mesh_desc* m, or_case** out);
int32_t or_case_destroy(or_case* c);
int32_t or_set_grid(or_case* c, int32_t im, const double* xv, const double* yv);
int32_t or_set_state(or_case* c, int32_t im, const double* q);
int32_t or_get_state(or_case* c, int32_t im, double* q);
int32_t or_set_iblank(or_case* c, int32_t im, const int32_t* ib);
int32_t or_get_iblank(or_case* c, int32_t im, int32_t* ib);
int32_t or_get_field(or_case* c, int32_t im, int32_t field, double* out);
int32_t or_set_field(or_case* c, int32_t im, int32_t field, const double* in);
int32_t or_set_istep0(or_case* c, int32_t istep0);
int32_t or_get_istep0(or_case* c, int32_t* istep0);
int32_t or_time_step(or_case* c);
int32_t or_bc(or_case* c);
int32_t or_compute_residual(or_case* c);
int32_t or_rbc(or_case* c);
int32_t or_lhs(or_case* c);
int32_t or_turbulence(or_case* c);
int32_t or_monitor(or_case* c, int32_t* mstop);
int32_t or_step(or_case* c, double* norms, int32_t* argmax);
int32_t or_iterate(or_case* c, int32_t niter, double* hist, int32_t* done);
int32_t or_forces(or_case* c, int32_t im, int32_t nsp, double out[6]);
int32_t or_set_mesh_connectivity(or_case* c, int32_t nsp, int32_t im, const int32_t* iblank, int32_t nfringe,
                                 const int32_t* imesh, const int32_t* ibc, int32_t ndonor, const int32_t* idonor,
                                 const double* frac);
int32_t or_interpolate(or_case* c);
int32_t or_sizes(const ot2d_params* p, const ot2d_mesh_desc* m, int32_t* jd, int32_t* kd,
                 int32_t* nq, int32_t* nv);
int32_t or_default_params(ot2d_params* p);

#ifdef __cplusplus
}
#endif
#endif
