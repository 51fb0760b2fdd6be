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

// State epilogue of the adjoint RHS, per-slot form (element-wise, so out may
// alias x0 and out2 may alias y0):
//   out  = e0 x0 + scale L'(w)^T mu            (x0 may be null)
//   out2 = f0 y0 + f1 x0 + f2 out              (when out2 is not null)
// (the face-pair form's epilogue is described with its fields below).  The
// reverse SSPRK3 step is three such launches with no separate
// linear-combination passes.
struct AdjArgs {
  int ncells;
  double scale;
  const double *w, *mu;                 // linearisation state and RHS cotangent
  const double *rx, *ry, *sx, *sy, *nx, *ny, *fscale;
  const int *fmap, *fmask;
  double *gmq, *gpq;                    // scratch [K][NF][4]
  const double *x0;
  double e0;
  double *out;
  const double *y0;
  double f0, f1, f2;
  double *out2;
  double *g_rx, *g_ry, *g_sx, *g_sy;    // += scale (dL/dgeo)^T mu  [K][Np], or [K] (compact)
  double *g_nx, *g_ny, *g_fs;           // [K][NF] (boundary pass: boundary slots only when compact)
  double *g_nxf, *g_nyf, *g_fsf;        // compact: face-pass cotangents per face [K][3]
  double *g_bcu, *g_bcv;                // [K][NF] (boundary slots only)
  // compact: the metric cotangents are summed over each cell's nodes and
  // the face pass's normal / fscale cotangents over each face's nodes (the
  // init_grid factors of a straight-sided cell are the same at all its
  // nodes, and a face's normal and fscale at all the face's nodes, so the
  // geometry chain needs only the sums); 0: per node, per slot
  int compact;
  int nbedges;                          // boundary edges (cell, local face)
  const int *b2c, *bnum;
  // Face-pair form (the gradient's reverse sweep; compact only; null pairs:
  // the per-slot form).  pairs [npairs][2] lists each interior face once as
  // (c * 3 + f) of its two sides, the first the lower cell; cmet [K][4]
  // holds (rx, ry, sx, sy) of each cell (its node-0 factors: a
  // straight-sided cell has the same factors at every node).  The face-pair
  // pass reads psi [K][NF][4] = -LIFT^T mu, the cotangent of the scaled
  // fluxes (the previous launch's psi_out, or launch_adj_psi for the
  // first); the cell pass runs on the matrix cores with the per-cell
  // factors and writes psi_out = -LIFT^T out.  Its state epilogue is
  //   out = em mu + scale L'(w)^T mu  (+ e0 x0 + f0 y0 when x0 is not null;
  //   x0 and y0 both given; out may alias y0; out2 unused)
  const int *pairs;
  int npairs;
  const double *cmet;
  double em;
  const double *psi;
  double *psi_out;
};

// Ld = L'(w) wd + dL/dgeo geod + dL/dbc (bcud, bcvd)
hipError_t launch_rhs_tlm(const Launch &L, const TlmArgs &a, double bcud, double bcvd);
// node_slots: [Np][2] face slots m with FMASK[m] == node (-1 when none)
hipError_t launch_rhs_adj(const Launch &L, const AdjArgs &a, const int *node_slots);
// psi [K][NF][4] = -LIFT^T mu [K][Np][4] on the matrix cores, with the
// arithmetic of the fused reverse RHS
hipError_t launch_adj_psi(const Launch &L, int ncells, const double *mu, double *psi);
// z = a x + b y  (y may be null)
hipError_t launch_lincomb(const Launch &L, size_t n, double a, const double *x, double b,
                          const double *y, double *z);

// Reverse of init_grid for the vertex coordinates: for every cell and each
// of its 6 vertex coordinates d, gv[6 c + d] = sum over the cell's factors
// of cotangent * d(factor)/d(coordinate d) (forward mode, one dual
// direction per thread).  g_J (cotangent of J at node 0) may be null.
struct GeoCot {
  const double *rx, *ry, *sx, *sy;  // [K][Np], or [K] when compact
  const double *nx, *ny, *fs;       // [K][NF]
  const double *J;                  // [K]
  // compact: rx..sy are per-cell sums, and nxf, nyf, fsf [K][3] per-face
  // sums that add to the per-slot nx, ny, fs
  int compact;
  const double *nxf, *nyf, *fsf;
};
// cell_rel [K][6]: (x0, y0, x1 - x0, y1 - y0, x2 - x0, y2 - y0) per cell; the
// factors are those of init_grid_cell_rel_t (as launch_init_grid_rel writes).
hipError_t launch_geo_vjp(const Launch &L, int ncells, const double *cell_rel, const GeoCot &g,
                          double *gv);
// rx, ry, sx, sy [K][Np], nx, ny, fs [K][NF], J [K] (node 0) and, when not
// null, cmet [K][4] (rx, ry, sx, sy of node 0) of the cells cell_rel, in the
// translation-invariant form init_grid_cell_rel_t.
hipError_t launch_init_grid_rel(const Launch &L, int ncells, const double *cell_rel, double *rx,
                                double *ry, double *sx, double *sy, double *nx, double *ny,
                                double *fs, double *J, double *cmet);
// partial[0..kSumBlocks) = block sums of x[i * stride + offset], i < n, in a
// fixed order; the caller adds the partials in index order.
constexpr int kSumBlocks = 256;
hipError_t launch_strided_sum(const Launch &L, size_t n, int stride, int offset, const double *x,
                              double *partial);
// Wall-face data of the lift functional, compact: [nwall][Nfp][7] (state at
// the face node, nx, ny, fscale) then [nwall] J.
hipError_t launch_wall_gather(const Launch &L, int nwall, const int *wcell, const int *wface, const double *q,
                              const double *nx, const double *ny, const double *fs, const double *J,
                              double *out);
// dst[idx[i]] += val[i] (indices distinct)
hipError_t launch_add_indexed(const Launch &L, int n, const int *idx, const double *val, double *dst);

}  // namespace dg
