// Boundary exterior states of types 0-3 with the free-stream velocity as an
// input, in generic arithmetic (double or Dual), and the hand-written
// reverse of the same map.  Host and device code; P is any struct with the
// members gam, bc_r, bc_p (dg::Phys).
//
// Types (src/get_bedge_q.h:24-50): 0 inflow (free stream), 1 outflow (free
// stream with the interior energy), 2 wall (reflected normal momentum); 3
// characteristic far field (no reference counterpart; the formulas of
// farfield_state).
#pragma once
#include "dg_dual.h"

namespace dg {

template <class T, class P>
DG_HD inline void boundary_t(const P &ph, int type, T bcu, T bcv, const T mQ[4], T nx, T ny, T qb[4]) {
  const double g = ph.gam;
  T bce = ph.bc_p / (g - 1.0) + 0.5 * ph.bc_r * (bcu * bcu + bcv * bcv);
  if (type == 0) {
    qb[0] = T(ph.bc_r); qb[1] = ph.bc_r * bcu; qb[2] = ph.bc_r * bcv; qb[3] = bce;
  } else if (type == 1) {
    qb[0] = T(ph.bc_r); qb[1] = ph.bc_r * bcu; qb[2] = ph.bc_r * bcv; qb[3] = mQ[3];
  } else if (type == 2) {
    T vn = nx * mQ[1] + ny * mQ[2];
    qb[0] = mQ[0]; qb[1] = mQ[1] - 2 * vn * nx; qb[2] = mQ[2] - 2 * vn * ny; qb[3] = mQ[3];
  } else {
    T rhoi = mQ[0], ui = mQ[1] / rhoi, vi = mQ[2] / rhoi;
    T pi_ = (g - 1.0) * (mQ[3] - 0.5 * rhoi * (ui * ui + vi * vi));
    T ci = sqrt(fabs(g * pi_ / rhoi));
    T vni = ui * nx + vi * ny;
    const double cinf = ::sqrt(g * ph.bc_p / ph.bc_r);
    T vninf = bcu * nx + bcv * ny;
    if (val(fabs(vni)) >= val(ci)) {
      if (val(vni) < 0.0) {
        qb[0] = T(ph.bc_r); qb[1] = ph.bc_r * bcu; qb[2] = ph.bc_r * bcv; qb[3] = bce;
      } else {
        for (int k = 0; k < 4; k++) qb[k] = mQ[k];
      }
      return;
    }
    T rp = vni + 2.0 * ci / (g - 1.0);
    T rm = vninf - 2.0 * cinf / (g - 1.0);
    T vnb = 0.5 * (rp + rm);
    T cb = 0.25 * (g - 1.0) * (rp - rm);
    T ent, ut, vt;
    if (val(vnb) > 0.0) {
      ent = pi_ / pow(rhoi, g);
      ut = ui - vni * nx; vt = vi - vni * ny;
    } else {
      ent = T(ph.bc_p / ::pow(ph.bc_r, g));
      ut = bcu - vninf * nx; vt = bcv - vninf * ny;
    }
    T rhob = pow(cb * cb / (g * ent), 1.0 / (g - 1.0));
    T pb = rhob * cb * cb / g;
    T ub = ut + vnb * nx, vb = vt + vnb * ny;
    qb[0] = rhob; qb[1] = rhob * ub; qb[2] = rhob * vb;
    qb[3] = pb / (g - 1.0) + 0.5 * rhob * (ub * ub + vb * vb);
  }
}

// Reverse of boundary_t at (bcu, bcv, mQ, nx, ny): for the cotangent bp of
// the exterior state, accumulates its cotangents of the interior state
// (bq), the normal (bnx, bny) and the free-stream velocity (bbcu, bbcv).
// The branches are those the forward takes at this point (the
// derivative of |x| at 0 is taken as +1, as in the dual arithmetic).
template <class P>
DG_HD inline void boundary_vjp(const P &ph, int type, double bcu, double bcv, const double mQ[4], double nx,
                               double ny, const double bp[4], double bq[4], double &bnx, double &bny,
                               double &bbcu, double &bbcv) {
  const double g = ph.gam, r = ph.bc_r;
  // free stream (r, r u, r v, p/(g-1) + r (u^2 + v^2)/2) at (u, v) = (bcu, bcv)
  auto freestream = [&](bool energy) {
    bbcu += r * bp[1] + (energy ? r * bcu * bp[3] : 0.0);
    bbcv += r * bp[2] + (energy ? r * bcv * bp[3] : 0.0);
  };
  if (type == 0) {
    freestream(true);
    return;
  }
  if (type == 1) {
    freestream(false);
    bq[3] += bp[3];
    return;
  }
  if (type == 2) {
    const double vn = nx * mQ[1] + ny * mQ[2];
    const double bvn = -2.0 * (nx * bp[1] + ny * bp[2]);
    bq[0] += bp[0];
    bq[1] += bp[1] + nx * bvn;
    bq[2] += bp[2] + ny * bvn;
    bq[3] += bp[3];
    bnx += -2.0 * vn * bp[1] + mQ[1] * bvn;
    bny += -2.0 * vn * bp[2] + mQ[2] * bvn;
    return;
  }
  // characteristic far field: the forward sweep
  const double rhoi = mQ[0], ui = mQ[1] / rhoi, vi = mQ[2] / rhoi;
  const double pi_ = (g - 1.0) * (mQ[3] - 0.5 * rhoi * (ui * ui + vi * vi));
  const double aa = g * pi_ / rhoi;
  const double ci = ::sqrt(::fabs(aa));
  const double vni = ui * nx + vi * ny;
  const double cinf = ::sqrt(g * ph.bc_p / ph.bc_r);
  const double vninf = bcu * nx + bcv * ny;
  if (::fabs(vni) >= ci) {
    if (vni < 0.0) {
      freestream(true);
    } else {
      for (int k = 0; k < 4; k++) bq[k] += bp[k];
    }
    return;
  }
  const double rp = vni + 2.0 * ci / (g - 1.0);
  const double rm = vninf - 2.0 * cinf / (g - 1.0);
  const double vnb = 0.5 * (rp + rm);
  const double cb = 0.25 * (g - 1.0) * (rp - rm);
  const bool inside = vnb > 0.0;
  const double ent = inside ? pi_ / ::pow(rhoi, g) : ph.bc_p / ::pow(ph.bc_r, g);
  const double ut = inside ? ui - vni * nx : bcu - vninf * nx;
  const double vt = inside ? vi - vni * ny : bcv - vninf * ny;
  const double X = cb * cb / (g * ent);
  const double rhob = ::pow(X, 1.0 / (g - 1.0));
  const double ub = ut + vnb * nx, vb = vt + vnb * ny;
  // the reverse sweep
  double brhob = bp[0] + ub * bp[1] + vb * bp[2] + 0.5 * (ub * ub + vb * vb) * bp[3];
  const double bub = rhob * bp[1] + rhob * ub * bp[3];
  const double bvb = rhob * bp[2] + rhob * vb * bp[3];
  const double bpb = bp[3] / (g - 1.0);
  // pb = rhob cb^2 / g
  brhob += cb * cb / g * bpb;
  double bcb = 2.0 * rhob * cb / g * bpb;
  // ub = ut + vnb nx, vb = vt + vnb ny
  const double but = bub, bvt = bvb, bvnb = nx * bub + ny * bvb;
  bnx += vnb * bub;
  bny += vnb * bvb;
  // rhob = X^(1/(g-1)), X = cb^2 / (g ent)
  const double bX = brhob * rhob / ((g - 1.0) * X);
  bcb += bX * 2.0 * cb / (g * ent);
  const double bent = -bX * X / ent;
  double bvni = 0.0, bvninf = 0.0, bui = 0.0, bvi = 0.0, bpi = 0.0, brhoi = 0.0;
  if (inside) {
    // ent = pi rhoi^-g, ut = ui - vni nx, vt = vi - vni ny
    bpi += bent / ::pow(rhoi, g);
    brhoi += -g * ent / rhoi * bent;
    bui += but;
    bvi += bvt;
    bvni += -nx * but - ny * bvt;
    bnx += -vni * but;
    bny += -vni * bvt;
  } else {
    // ut = bcu - vninf nx, vt = bcv - vninf ny (ent is a constant)
    bbcu += but;
    bbcv += bvt;
    bvninf += -nx * but - ny * bvt;
    bnx += -vninf * but;
    bny += -vninf * bvt;
  }
  // vnb = (rp + rm) / 2, cb = (g-1)(rp - rm) / 4
  const double brp = 0.5 * bvnb + 0.25 * (g - 1.0) * bcb;
  const double brm = 0.5 * bvnb - 0.25 * (g - 1.0) * bcb;
  // rp = vni + 2 ci / (g-1), rm = vninf - const
  bvni += brp;
  const double bci = 2.0 / (g - 1.0) * brp;
  bvninf += brm;
  // vninf = bcu nx + bcv ny
  bbcu += nx * bvninf;
  bbcv += ny * bvninf;
  bnx += bcu * bvninf;
  bny += bcv * bvninf;
  // ci = sqrt(|aa|), aa = g pi / rhoi
  const double baa = bci * (aa >= 0.0 ? 1.0 : -1.0) / (2.0 * ci);
  bpi += g / rhoi * baa;
  brhoi += -aa / rhoi * baa;
  // vni = ui nx + vi ny
  bui += nx * bvni;
  bvi += ny * bvni;
  bnx += ui * bvni;
  bny += vi * bvni;
  // pi = (g-1)(q3 - rhoi (ui^2 + vi^2) / 2)
  bq[3] += (g - 1.0) * bpi;
  brhoi += -0.5 * (g - 1.0) * (ui * ui + vi * vi) * bpi;
  bui += -(g - 1.0) * rhoi * ui * bpi;
  bvi += -(g - 1.0) * rhoi * vi * bpi;
  // ui = q1 / rhoi, vi = q2 / rhoi
  bq[1] += bui / rhoi;
  bq[2] += bvi / rhoi;
  brhoi += -(ui * bui + vi * bvi) / rhoi;
  bq[0] += brhoi;
}

}  // namespace dg
