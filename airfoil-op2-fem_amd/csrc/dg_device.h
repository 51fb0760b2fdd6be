// Device-side building blocks shared by the stage kernels: point-wise
// Euler physics (src/euler_rhs.h:7-23, 120-220; src/get_bedge_q.h:1-51;
// src/calc_dt.h:1-22), workgroup reductions and the XCD-aware tile walk.
#pragma once
#include <hip/hip_runtime.h>

#include "dg_internal.h"

namespace dg {

template <int N>
struct Ord {
  static constexpr int Np = (N + 1) * (N + 2) / 2;
  static constexpr int Nfp = N + 1;
  static constexpr int NF = 3 * (N + 1);
};

// ------------------------------------------------------------------------
// Reciprocal and square root of positive normal doubles: the hardware
// estimates (v_rcp_f64, v_rsq_f64) refined by two Newton-Raphson steps,
// which gives full double precision (within an ulp of the correctly rounded
// value) in about half the instructions of IEEE division and sqrt.  The
// inputs here are densities, sums of square roots of densities and sound
// speeds; a non-positive or non-finite input gives a non-finite result,
// which the stage kernels' state check reports.
// ------------------------------------------------------------------------

__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(y, e, y);
}

// Goldschmidt iteration on g ~ sqrt(x), h ~ 1/(2 sqrt(x)) from v_rsq_f64.
__device__ __forceinline__ double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// sqrt of a value that may also be 0 (a speed at rest) or not finite: the
// refined form for positive values, the IEEE one otherwise (sqrt(0) = 0,
// NaN stays NaN).
__device__ __forceinline__ double sqrt_nn(double x) {
  return x > 0.0 && x < __builtin_huge_val() ? sqrt_nr(x) : sqrt(x);
}

// Outward unit normal and fscale of face f of a straight-sided cell from its
// metric factors (rx, ry, sx, sy).  src/init_grid.h:80-104 forms the face
// vectors (yr, -xr), (ys - yr, xr - xs), (-ys, xs) and divides by their
// length sJ, with fscale = sJ / J; since rx = ys/J, ry = -xs/J, sx = -yr/J,
// sy = xr/J, those vectors are J (-sx, -sy), J (rx + sx, ry + sy) and
// J (-rx, -ry), so for J > 0 the unit normal is the normalised bracket and
// fscale is the bracket's length.
__device__ __forceinline__ void cell_face_geometry(double rx, double ry, double sx, double sy, int f,
                                                   double &nx, double &ny, double &fs) {
  const double a = f == 0 ? -sx : (f == 1 ? rx + sx : -rx);
  const double b = f == 0 ? -sy : (f == 1 ? ry + sy : -ry);
  fs = sqrt_nr(a * a + b * b);
  const double i = rcp_nr(fs);
  nx = a * i;
  ny = b * i;
}

// ------------------------------------------------------------------------
// Point-wise physics (src/euler_rhs.h:7-23, 120-220; src/get_bedge_q.h)
// ------------------------------------------------------------------------

__device__ __forceinline__ void euler_flux(double gam, const double q[4], double F[4],
                                           double G[4]) {
  double rho = q[0];
  const double ir = rcp_nr(rho);
  double u = q[1] * ir;
  double v = q[2] * ir;
  double p = (gam - 1) * (q[3] - 0.5 * (q[1] * u + q[2] * v));
  F[0] = q[1];
  F[1] = q[1] * u + p;
  F[2] = q[2] * u;
  F[3] = u * (q[3] + p);
  G[0] = q[2];
  G[1] = q[1] * v;
  G[2] = q[2] * v + p;
  G[3] = v * (q[3] + p);
}

// Roe flux at one face node, rotated onto the normal, rotated back, times
// fscale (src/euler_rhs.h:151-219).
__device__ __forceinline__ void roe_node(double gam, const double mQ[4], const double eQ[4],
                                         double nx, double ny, double fs, double out[4]) {
  double mR0 = mQ[0], mR1 = nx * mQ[1] + ny * mQ[2], mR2 = -ny * mQ[1] + nx * mQ[2], mR3 = mQ[3];
  double pR0 = eQ[0], pR1 = nx * eQ[1] + ny * eQ[2], pR2 = -ny * eQ[1] + nx * eQ[2], pR3 = eQ[3];
  const double imR0 = rcp_nr(mR0), ipR0 = rcp_nr(pR0);
  double mU = mR1 * imR0, mV = mR2 * imR0;
  double mP = (gam - 1) * (mR3 - 0.5 * (mR1 * mU + mR2 * mV));
  double pU = pR1 * ipR0, pV = pR2 * ipR0;
  double pP = (gam - 1) * (pR3 - 0.5 * (pR1 * pU + pR2 * pV));
  double fm0 = mR1, fm1 = mR1 * mU + mP, fm2 = mR2 * mU, fm3 = mU * (mR3 + mP);
  double fp0 = pR1, fp1 = pR1 * pU + pP, fp2 = pR2 * pU, fp3 = pU * (pR3 + pP);
  double mH = (mR3 + mP) * imR0;
  double pH = (pR3 + pP) * ipR0;
  double ms = sqrt_nr(mR0), ps = sqrt_nr(pR0);
  double rho = ms * ps;
  const double iden = rcp_nr(ms + ps);
  double u = (ms * mU + ps * pU) * iden;
  double v = (ms * mV + ps * pV) * iden;
  double H = (ms * mH + ps * pH) * iden;
  double c2 = (gam - 1) * (H - 0.5 * (u * u + v * v));
  double c = sqrt_nr(c2);
  const double ic = rcp_nr(c), ic2 = ic * ic;
  double dW1 = -0.5 * rho * (pU - mU) * ic + 0.5 * (pP - mP) * ic2;
  double dW2 = (pR0 - mR0) - (pP - mP) * ic2;
  double dW3 = rho * (pV - mV);
  double dW4 = 0.5 * rho * (pU - mU) * ic + 0.5 * (pP - mP) * ic2;
  dW1 = fabs(u - c) * dW1;
  dW2 = fabs(u) * dW2;
  dW3 = fabs(u) * dW3;
  dW4 = fabs(u + c) * dW4;
  double f0 = (fm0 + fp0) / 2.0 - (dW1 + dW2 + dW4) / 2.0;
  double f1 = (fm1 + fp1) / 2.0 - (dW1 * (u - c) + dW2 * u + dW4 * (u + c)) / 2.0;
  double f2 = (fm2 + fp2) / 2.0 - (dW1 * v + dW2 * v + dW3 + dW4 * v) / 2.0;
  double f3 = (fm3 + fp3) / 2.0 -
              (dW1 * (H - u * c) + dW2 * (u * u + v * v) + dW3 * v + dW4 * (H + u * c)) / 2.0;
  out[0] = f0 * fs;
  out[1] = (nx * f1 - ny * f2) * fs;
  out[2] = (ny * f1 + nx * f2) * fs;
  out[3] = f3 * fs;
}

// Characteristic (Riemann-invariant) far-field state; no reference
// counterpart.  Subsonic: R+ from inside, R- from the free stream, entropy
// and tangential velocity from the upwind side.  Supersonic: free stream on
// inflow, interior on outflow.
__device__ __forceinline__ void farfield_state(const Phys &ph, const double qi[4], double nx,
                                               double ny, double qb[4]) {
  const double g = ph.gam;
  double rhoi = qi[0], ui = qi[1] / rhoi, vi = qi[2] / rhoi;
  double pi_ = (g - 1.0) * (qi[3] - 0.5 * rhoi * (ui * ui + vi * vi));
  double ci = sqrt(fabs(g * pi_ / rhoi));
  double vni = ui * nx + vi * ny;
  double cinf = sqrt(g * ph.bc_p / ph.bc_r);
  double vninf = ph.bc_u * nx + ph.bc_v * ny;
  if (fabs(vni) >= ci) {
    if (vni < 0.0) {
      qb[0] = ph.bc_r; qb[1] = ph.bc_r * ph.bc_u; qb[2] = ph.bc_r * ph.bc_v; qb[3] = ph.bc_e;
    } else {
      qb[0] = qi[0]; qb[1] = qi[1]; qb[2] = qi[2]; qb[3] = qi[3];
    }
    return;
  }
  double rp = vni + 2.0 * ci / (g - 1.0);
  double rm = vninf - 2.0 * cinf / (g - 1.0);
  double vnb = 0.5 * (rp + rm);
  double cb = 0.25 * (g - 1.0) * (rp - rm);
  double ent, ut, vt;
  if (vnb > 0.0) {
    ent = pi_ / pow(rhoi, g);
    ut = ui - vni * nx; vt = vi - vni * ny;
  } else {
    ent = ph.bc_p / pow(ph.bc_r, g);
    ut = ph.bc_u - vninf * nx; vt = ph.bc_v - vninf * ny;
  }
  double rhob = pow(cb * cb / (g * ent), 1.0 / (g - 1.0));
  double pb = rhob * cb * cb / g;
  double ub = ut + vnb * nx, vb = vt + vnb * ny;
  qb[0] = rhob;
  qb[1] = rhob * ub;
  qb[2] = rhob * vb;
  qb[3] = pb / (g - 1.0) + 0.5 * rhob * (ub * ub + vb * vb);
}

// Isentropic vortex state at (x, y) and time t (vortex/src/set_ic.h:1-28,
// with the advection by uinf in both velocity components and an optional
// period in x).
__device__ __forceinline__ void vortex_state(double gam, const Vortex &vx, double t, double x, double y,
                                             double q[4]) {
  const double pi = 3.14159265358979323846;
  double dx = x - vx.x0 - vx.uinf * t;
  if (vx.period > 0.0) dx -= vx.period * floor(dx / vx.period + 0.5);
  const double dy = y - vx.y0;
  const double r2 = dx * dx + dy * dy;
  const double e1 = exp(1.0 - r2);
  const double u = vx.uinf - vx.beta * e1 * (dy / (2 * pi));
  const double v = vx.beta * e1 * (dx / (2 * pi));
  double rho = 1.0 - ((gam - 1) / (16 * gam * pi * pi)) * vx.beta * vx.beta * exp(2 * (1.0 - r2));
  rho = pow(rho, 1.0 / (gam - 1));
  const double p = pow(rho, gam);
  q[0] = rho;
  q[1] = rho * u;
  q[2] = rho * v;
  q[3] = (p / (gam - 1)) + 0.5 * rho * (u * u + v * v);
}

// Exterior state on a boundary face node (src/get_bedge_q.h:24-50 + type 3).
__device__ __forceinline__ void boundary_state(const Phys &ph, int type, const double mQ[4],
                                               double nx, double ny, double qb[4]) {
  if (type == 0) {
    qb[0] = ph.bc_r; qb[1] = ph.bc_r * ph.bc_u; qb[2] = ph.bc_r * ph.bc_v; qb[3] = ph.bc_e;
  } else if (type == 1) {
    qb[0] = ph.bc_r; qb[1] = ph.bc_r * ph.bc_u; qb[2] = ph.bc_r * ph.bc_v; qb[3] = mQ[3];
  } else if (type == 2) {
    double vn = nx * mQ[1] + ny * mQ[2];
    qb[0] = mQ[0];
    qb[1] = mQ[1] - 2 * vn * nx;
    qb[2] = mQ[2] - 2 * vn * ny;
    qb[3] = mQ[3];
  } else {
    farfield_state(ph, mQ, nx, ny, qb);
  }
}

// CFL quantity of one node (src/calc_dt.h:7-12 with (N+1)^2).
__device__ __forceinline__ double cfl_node(double gam, double n1sq, const double q[4], double fs) {
  double rho = q[0];
  const double ir = rcp_nr(rho);
  double u = q[1] * ir;
  double v = q[2] * ir;
  double p = (gam - 1.0) * (q[3] - rho * (u * u + v * v) * 0.5);
  double c = sqrt_nn(fabs(gam * p * ir));
  return n1sq * 0.5 * fs * (sqrt_nn(u * u + v * v) + c);
}

__device__ __forceinline__ unsigned long long max_bits(unsigned long long a, unsigned long long b) {
  return a > b ? a : b;
}

// *dst = max(*dst, v) for a running maximum that only grows: a coherent
// load first, and the atomic only when v is larger.  Thousands of waves
// finishing together otherwise serialise on one address; a stale load only
// costs an unneeded atomic.
__device__ __forceinline__ void atomic_max_if_larger(unsigned long long *dst, unsigned long long v) {
  const unsigned long long cur = __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (v > cur) atomicMax(dst, v);
}

// Max over a workgroup of non-negative doubles held as bit patterns (a NaN
// pattern compares above +inf, so it propagates).
template <int NT>
__device__ unsigned long long block_max_bits(unsigned long long v, unsigned long long *red) {
  for (int off = 32; off > 0; off >>= 1) v = max_bits(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  unsigned long long r = 0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < NT / 64; i++) r = max_bits(r, red[i]);
  }
  __syncthreads();
  return r;
}

// Maps a persistent workgroup's k-th visit to a tile so that the workgroups
// of one XCD (blockIdx % 8 on MI355X) walk one contiguous chunk of tiles.
__device__ __forceinline__ int xcd_tile(int k, int ntiles) {
  const int nxcd = 8;
  const int g = gridDim.x;
  if (g % nxcd != 0) return blockIdx.x + k * g;
  const int xcd = blockIdx.x % nxcd, local = blockIdx.x / nxcd, per = g / nxcd;
  const int chunk = (ntiles + nxcd - 1) / nxcd;
  const int t = local + k * per;
  if (t >= chunk) return ntiles;  // this XCD's chunk is done
  return xcd * chunk + t;
}


}  // namespace dg
