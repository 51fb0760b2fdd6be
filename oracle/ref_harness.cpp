// Oracle cross-check harness: compiles the reference's own per-element kernel
// headers that need no BLAS (src/get_neighbour_q.h, src/get_bedge_q.h,
// src/set_workingQ.h, src/update_Q.h, src/calc_dt.h, src/set_ic.h,
// src/neighbour_zero.h, with src/constants/constant_util.h for FMASK and the
// physics globals) from the reference tree, and exports them with C linkage.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile (target `ref`) into
// oracle/_ref/ in a container where the reference tree is present; used by
// CPU tests to check the oracle's restatement at the reference's order N=4.
// src/euler_rhs.h and src/init_grid.h include <cblas.h>, which this image
// lacks, so they are not built here.
#include <cmath>
#include <cstdlib>
using namespace std;  // the reference's translation units have this in scope

#include "constants/constant_util.h"
#include "calc_dt.h"
#include "get_bedge_q.h"
#include "get_neighbour_q.h"
#include "neighbour_zero.h"
#include "set_ic.h"
#include "set_workingQ.h"
#include "update_Q.h"

extern "C" {
void ref_set_physics(double g, double mach, double alpha, double p, double r,
                     double u, double v, double e) {
  gam = g; bc_mach = mach; bc_alpha = alpha; bc_p = p; bc_r = r;
  bc_u = u; bc_v = v; bc_e = e;
}
void ref_get_fmask(int *out) { for (int i = 0; i < 15; i++) out[i] = FMASK[i]; }
void ref_get_neighbour_q(const int *edgeNum, const double *xL, const double *yL,
                         const double *xR, const double *yR, const double *qL,
                         const double *qR, double *exL, double *exR) {
  get_neighbour_q(edgeNum, xL, yL, xR, yR, qL, qR, exL, exR);
}
void ref_get_bedge_q(const int *type, const int *num, const double *nx,
                     const double *ny, const double *q, double *exQ) {
  get_bedge_q(type, num, nx, ny, q, exQ);
}
void ref_set_workingQ(const double *dt, const int *stage, const double *q,
                      const double *k1, const double *k2, double *w) {
  set_workingQ(dt, stage, q, k1, k2, w);
}
void ref_update_Q(const double *dt, double *q, const double *k1, const double *k2,
                  const double *k3, double *w) {
  update_Q(dt, q, k1, k2, k3, w);
}
void ref_calc_dt(const double *q, const double *fscale, double *dt1) {
  calc_dt(q, fscale, dt1);
}
void ref_set_ic(double *q, double *w) { set_ic(q, w); }
void ref_neighbour_zero(double *exQ) { neighbour_zero(exQ); }
}
