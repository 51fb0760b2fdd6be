// Tangent-linear and adjoint RHS kernels (declarations).
#pragma once
#include <hip/hip_runtime.h>

#include "dg_internal.h"

namespace dg {

struct TlmArgs {
  int ncells;
  const double *w, *wd;                 // state and its tangent [K][Np][4]
  const double *rx, *ry, *sx, *sy;      // [K][Np]
  const double *rxd, *ryd, *sxd, *syd;  // tangents (may be null)
  const double *nx, *ny, *fscale;       // [K][NF]
  const double *nxd, *nyd, *fsd;        // tangents (may be null)
  const int *fmap, *fmask;
  double *fluxd;  // scratch [K][NF][4]: tangent of the scaled Roe fluxes
  double *out;    // [K][Np][4]: tangent of the RHS
};

struct AdjArgs {
  int ncells;
  double scale;
  const double *w, *mu;                 // linearisation state and RHS cotangent
  const double *rx, *ry, *sx, *sy, *nx, *ny, *fscale;
  const int *fmap, *fmask;
  double *gmq, *gpq;                    // scratch [K][NF][4]
  double *acc;                          // += scale L'(w)^T mu
  double *g_rx, *g_ry, *g_sx, *g_sy;    // += scale (dL/dgeo)^T mu  [K][Np]
  double *g_nx, *g_ny, *g_fs;           // [K][NF]
  double *g_bcu, *g_bcv;                // [K][NF] (boundary slots only)
};

// Ld = L'(w) wd + dL/dgeo geod + dL/dbc (bcud, bcvd)
hipError_t launch_rhs_tlm(const Launch &L, const TlmArgs &a, double bcud, double bcvd);
// node_slots: [Np][2] face slots m with FMASK[m] == node (-1 when none)
hipError_t launch_rhs_adj(const Launch &L, const AdjArgs &a, const int *node_slots);
// z = a x + b y  (y may be null)
hipError_t launch_lincomb(const Launch &L, size_t n, double a, const double *x, double b,
                          const double *y, double *z);
// *dst = cfl / dt1
hipError_t launch_record_dt(const Launch &L, const unsigned long long *dt1, double cfl, double *dst);

}  // namespace dg
