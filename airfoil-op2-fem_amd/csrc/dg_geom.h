// init_grid of one straight-sided cell in generic arithmetic (double or
// Dual), shared by the host tangent of the geometry and the device geometry
// of the sensitivity path.  Same formulas and association order as
// init_grid_kernel (src/init_grid.h:1-106): node coordinates from the three
// vertices, xr = Dr x (etc.), J, the metric factors, and on every face node
// the outward normal and fscale = sJ / J.
#pragma once
#include "dg_dual.h"

namespace dg {

// The factors from the node coordinates X, Y [Np] (absolute, or relative to
// any fixed point: Dr and Ds annihilate constants).
template <int N, class T>
DG_HD inline void init_grid_factors_t(const double *Dr, const double *Ds, const int *fmask,
                                      const T *X, const T *Y, T *rx, T *ry, T *sx, T *sy, T *nx,
                                      T *ny, T *fs, T &Jc) {
  constexpr int Np = (N + 1) * (N + 2) / 2, Nfp = N + 1, NF = 3 * Nfp;
  T XR[Np], XS[Np], YR[Np], YS[Np], J[Np];
  for (int j = 0; j < Np; j++) {
    T a = 0.0, b = 0.0, cc = 0.0, dd = 0.0;
    for (int i = 0; i < Np; i++) {
      a += Dr[j * Np + i] * X[i];
      b += Ds[j * Np + i] * X[i];
      cc += Dr[j * Np + i] * Y[i];
      dd += Ds[j * Np + i] * Y[i];
    }
    XR[j] = a; XS[j] = b; YR[j] = cc; YS[j] = dd;
  }
  for (int i = 0; i < Np; i++) {
    J[i] = -XS[i] * YR[i] + XR[i] * YS[i];
    rx[i] = YS[i] / J[i]; sx[i] = -YR[i] / J[i]; ry[i] = -XS[i] / J[i]; sy[i] = XR[i] / J[i];
  }
  Jc = J[0];
  for (int m = 0; m < NF; m++) {
    const int f = fmask[m], face = m / Nfp;
    T a, b;
    if (face == 0) { a = YR[f]; b = -XR[f]; }
    else if (face == 1) { a = YS[f] - YR[f]; b = XR[f] - XS[f]; }
    else { a = -YS[f]; b = XS[f]; }
    T sJ = sqrt(a * a + b * b);
    nx[m] = a / sJ; ny[m] = b / sJ; fs[m] = sJ / J[f];
  }
}

// v = the three vertices (x0, y0, x1, y1, x2, y2).
template <int N, class T>
DG_HD inline void init_grid_cell_t(const double *r, const double *s, const double *Dr,
                                   const double *Ds, const int *fmask, const T v[6], T *rx,
                                   T *ry, T *sx, T *sy, T *nx, T *ny, T *fs, T &Jc) {
  constexpr int Np = (N + 1) * (N + 2) / 2;
  T X[Np], Y[Np];
  for (int d = 0; d < 2; d++) {
    T *o = d == 0 ? X : Y;
    const T a1 = 0.5 * v[2 + d], a2 = 0.5 * v[4 + d], a0 = -0.5 * v[d];
    for (int i = 0; i < Np; i++) o[i] = a1 * (1.0 + r[i]) + a2 * (1.0 + s[i]) + a0 * (s[i] + r[i]);
  }
  init_grid_factors_t<N, T>(Dr, Ds, fmask, X, Y, rx, ry, sx, sy, nx, ny, fs, Jc);
}

// The same cell from its edge vectors e = (x1 - x0, y1 - y0, x2 - x0,
// y2 - y0): node coordinates relative to vertex 0,
// x_i - x0 = (1 + r_i)/2 (x1 - x0) + (1 + s_i)/2 (x2 - x0).  Equal to
// init_grid_cell_t in exact arithmetic; its rounding is relative to the cell
// size instead of the coordinates' size, so the factors of a cell far from
// the origin (in cell sizes) are smooth functions of the vertex positions.
template <int N, class T>
DG_HD inline void init_grid_cell_rel_t(const double *r, const double *s, const double *Dr,
                                       const double *Ds, const int *fmask, const T e[4], T *rx,
                                       T *ry, T *sx, T *sy, T *nx, T *ny, T *fs, T &Jc) {
  constexpr int Np = (N + 1) * (N + 2) / 2;
  T X[Np], Y[Np];
  for (int i = 0; i < Np; i++) {
    const double a = 0.5 * (1.0 + r[i]), b = 0.5 * (1.0 + s[i]);
    X[i] = a * e[0] + b * e[2];
    Y[i] = a * e[1] + b * e[3];
  }
  init_grid_factors_t<N, T>(Dr, Ds, fmask, X, Y, rx, ry, sx, sy, nx, ny, fs, Jc);
}

}  // namespace dg
