// Tangent-linear and adjoint (reverse-mode) kernels of the DG Euler RHS
// L(w) = vol(w) - LIFT * flux(w), with respect to the state, the geometric
// factors (rx, ry, sx, sy, nx, ny, fscale) and the free-stream velocity that
// the far-field boundary sees (bc_u, bc_v).
//
// Tangent:  Ld = L'(w) wd + dL/dgeo geod + dL/dbc bcd, with the point
//           functions evaluated in dual arithmetic.
// Adjoint:  out = e0 x0 + scale * L'(w)^T mu (plus an optional second
//           combination), geometry and bc accumulators += scale *
//           (dL/dgeo)^T mu, (dL/dbc)^T mu.  Gather-only in two forms.
//           Per-slot form (the C-ABI entry dg_rhs_adjoint): the face pass
//           evaluates the reverse Roe flux once per face slot and stores the
//           interior-side (gmq) and exterior-side (gpq) cotangents, the
//           boundary pass reverses the boundary states through their dual
//           Jacobian, and the cell pass applies the transposed weak
//           derivatives and the reverse Euler fluxes and gathers its own gmq
//           and the neighbours' gpq.  Face-pair form (the gradient's reverse
//           sweep): one reverse Roe flux per interior face node pair (and per
//           boundary slot, with the hand-written reverse of the boundary
//           state), then the cell pass on the matrix cores with per-cell
//           factors (adj_cell_mfma_kernel), which also forms the next
//           launch's psi = -LIFT^T out on the matrix cores.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dg_adjoint.h"
#include "dg_bc.h"
#include "dg_device.h"
#include "dg_dual.h"
#include "dg_geom.h"
#include "dg_mfma.h"

namespace dg {

// ------------------------------------------------------------------------
// Point functions in generic arithmetic (same formulas as dg_device.h)
// ------------------------------------------------------------------------

template <class T>
__device__ __forceinline__ void euler_flux_t(double gam, const T q[4], T F[4], T G[4]) {
  T rho = q[0];
  T u = q[1] / rho;
  T v = q[2] / rho;
  T p = (gam - 1) * (q[3] - 0.5 * (q[1] * u + q[2] * v));
  F[0] = q[1]; F[1] = q[1] * u + p; F[2] = q[2] * u; F[3] = u * (q[3] + p);
  G[0] = q[2]; G[1] = q[1] * v; G[2] = q[2] * v + p; G[3] = v * (q[3] + p);
}

template <class T>
__device__ __forceinline__ void roe_t(double gam, const T mQ[4], const T eQ[4], T nx, T ny, T fs,
                                      T out[4]) {
  T mR0 = mQ[0], mR1 = nx * mQ[1] + ny * mQ[2], mR2 = -ny * mQ[1] + nx * mQ[2], mR3 = mQ[3];
  T pR0 = eQ[0], pR1 = nx * eQ[1] + ny * eQ[2], pR2 = -ny * eQ[1] + nx * eQ[2], pR3 = eQ[3];
  T mU = mR1 / mR0, mV = mR2 / mR0;
  T mP = (gam - 1) * (mR3 - 0.5 * (mR1 * mU + mR2 * mV));
  T pU = pR1 / pR0, pV = pR2 / pR0;
  T pP = (gam - 1) * (pR3 - 0.5 * (pR1 * pU + pR2 * pV));
  T fm0 = mR1, fm1 = mR1 * mU + mP, fm2 = mR2 * mU, fm3 = mU * (mR3 + mP);
  T fp0 = pR1, fp1 = pR1 * pU + pP, fp2 = pR2 * pU, fp3 = pU * (pR3 + pP);
  T mH = (mR3 + mP) / mR0, pH = (pR3 + pP) / pR0;
  T ms = sqrt(mR0), ps = sqrt(pR0);
  T rho = ms * ps, den = ms + ps;
  T u = (ms * mU + ps * pU) / den, v = (ms * mV + ps * pV) / den, H = (ms * mH + ps * pH) / den;
  T c2 = (gam - 1) * (H - 0.5 * (u * u + v * v));
  T c = sqrt(c2);
  T W1 = -0.5 * rho * (pU - mU) / c + 0.5 * (pP - mP) / c2;
  T W2 = (pR0 - mR0) - (pP - mP) / c2;
  T W3 = rho * (pV - mV);
  T W4 = 0.5 * rho * (pU - mU) / c + 0.5 * (pP - mP) / c2;
  W1 = fabs(u - c) * W1;
  W2 = fabs(u) * W2;
  W3 = fabs(u) * W3;
  W4 = fabs(u + c) * W4;
  T f0 = (fm0 + fp0) / 2.0 - (W1 + W2 + W4) / 2.0;
  T f1 = (fm1 + fp1) / 2.0 - (W1 * (u - c) + W2 * u + W4 * (u + c)) / 2.0;
  T f2 = (fm2 + fp2) / 2.0 - (W1 * v + W2 * v + W3 + W4 * v) / 2.0;
  T f3 = (fm3 + fp3) / 2.0 - (W1 * (H - u * c) + W2 * (u * u + v * v) + W3 * v + W4 * (H + u * c)) / 2.0;
  out[0] = f0 * fs;
  out[1] = (nx * f1 - ny * f2) * fs;
  out[2] = (ny * f1 + nx * f2) * fs;
  out[3] = f3 * fs;
}

// ------------------------------------------------------------------------
// Reverse-mode point functions
// ------------------------------------------------------------------------

// Reverse of euler_flux: accumulates F^T bF + G^T bG into bq.
__device__ __forceinline__ void euler_flux_vjp(double gam, const double q[4], const double bF[4],
                                               const double bG[4], double bq[4]) {
  const double g1 = gam - 1.0;
  const double rho = q[0], ir = rcp_nr(rho), u = q[1] * ir, v = q[2] * ir;
  const double p = g1 * (q[3] - 0.5 * (q[1] * u + q[2] * v));
  double bu = bF[1] * q[1] + bF[2] * q[2] + bF[3] * (q[3] + p);
  double bv = bG[1] * q[1] + bG[2] * q[2] + bG[3] * (q[3] + p);
  double bp = bF[1] + bF[3] * u + bG[2] + bG[3] * v;
  double b1 = bF[0] + bF[1] * u + bG[1] * v;
  double b2 = bF[2] * u + bG[0] + bG[2] * v;
  double b3 = bF[3] * u + bG[3] * v;
  // p = g1 (q3 - 0.5 (q1 u + q2 v))
  b3 += g1 * bp;
  b1 += -0.5 * g1 * u * bp;
  bu += -0.5 * g1 * q[1] * bp;
  b2 += -0.5 * g1 * v * bp;
  bv += -0.5 * g1 * q[2] * bp;
  // u = q1/rho, v = q2/rho
  b1 += bu * ir;
  b2 += bv * ir;
  const double b0 = -(bu * u + bv * v) * ir;
  bq[0] += b0; bq[1] += b1; bq[2] += b2; bq[3] += b3;
}

// Reverse of roe_t: given the cotangent o[4] of the scaled flux, returns the
// cotangents of mQ, eQ (accumulated) and of nx, ny, fs; with R non-null also
// the unscaled flux R(mQ, eQ, n) itself (the scaled one is fs R).
__device__ __forceinline__ void roe_vjp(double gam, const double mQ[4], const double eQ[4], double nx,
                                        double ny, double fs, const double o[4], double bmQ[4],
                                        double beQ[4], double &bnx, double &bny, double &bfs,
                                        double *R = nullptr) {
  const double g1 = gam - 1.0;
  // forward sweep
  const double mR0 = mQ[0], mR1 = nx * mQ[1] + ny * mQ[2], mR2 = -ny * mQ[1] + nx * mQ[2], mR3 = mQ[3];
  const double pR0 = eQ[0], pR1 = nx * eQ[1] + ny * eQ[2], pR2 = -ny * eQ[1] + nx * eQ[2], pR3 = eQ[3];
  // reciprocals refined from the hardware estimates (dg_device.h), shared
  // by the forward recomputation and every quotient of the reverse sweep
  const double imR0 = rcp_nr(mR0), ipR0 = rcp_nr(pR0);
  const double mU = mR1 * imR0, mV = mR2 * imR0;
  const double mP = g1 * (mR3 - 0.5 * (mR1 * mU + mR2 * mV));
  const double pU = pR1 * ipR0, pV = pR2 * ipR0;
  const double pP = g1 * (pR3 - 0.5 * (pR1 * pU + pR2 * pV));
  const double fm1 = mR1 * mU + mP, fm2 = mR2 * mU, fm3 = mU * (mR3 + mP);
  const double fp1 = pR1 * pU + pP, fp2 = pR2 * pU, fp3 = pU * (pR3 + pP);
  const double mH = (mR3 + mP) * imR0, pH = (pR3 + pP) * ipR0;
  const double ms = sqrt_nr(mR0), ps = sqrt_nr(pR0);
  const double ims = rcp_nr(ms), ips = rcp_nr(ps);
  const double rho = ms * ps, den = ms + ps, iden = rcp_nr(den);
  const double u = (ms * mU + ps * pU) * iden, v = (ms * mV + ps * pV) * iden;
  const double H = (ms * mH + ps * pH) * iden;
  const double c2 = g1 * (H - 0.5 * (u * u + v * v));
  const double c = sqrt_nr(c2);
  const double ic = rcp_nr(c), ic2 = ic * ic, ic4 = ic2 * ic2;
  const double dU = pU - mU, dP = pP - mP, dV = pV - mV, dR = pR0 - mR0;
  const double W1 = -0.5 * rho * dU * ic + 0.5 * dP * ic2;
  const double W2 = dR - dP * ic2;
  const double W3 = rho * dV;
  const double W4 = 0.5 * rho * dU * ic + 0.5 * dP * ic2;
  const double a1 = fabs(u - c), a2 = fabs(u), a4 = fabs(u + c);
  const double X1 = a1 * W1, X2 = a2 * W2, X3 = a2 * W3, X4 = a4 * W4;
  const double f0 = (mR1 + pR1) / 2.0 - (X1 + X2 + X4) / 2.0;
  const double f1 = (fm1 + fp1) / 2.0 - (X1 * (u - c) + X2 * u + X4 * (u + c)) / 2.0;
  const double f2 = (fm2 + fp2) / 2.0 - (X1 * v + X2 * v + X3 + X4 * v) / 2.0;
  const double f3 = (fm3 + fp3) / 2.0 -
                    (X1 * (H - u * c) + X2 * (u * u + v * v) + X3 * v + X4 * (H + u * c)) / 2.0;
  // reverse sweep
  bfs = o[0] * f0 + o[1] * (nx * f1 - ny * f2) + o[2] * (ny * f1 + nx * f2) + o[3] * f3;
  if (R) {
    R[0] = f0; R[1] = nx * f1 - ny * f2; R[2] = ny * f1 + nx * f2; R[3] = f3;
  }
  const double b1o = o[1] * fs, b2o = o[2] * fs;
  bnx = b1o * f1 + b2o * f2;
  bny = -b1o * f2 + b2o * f1;
  const double bf0 = o[0] * fs, bf1 = nx * b1o + ny * b2o, bf2 = -ny * b1o + nx * b2o, bf3 = o[3] * fs;
  double bmR1 = 0, bpR1 = 0, bmR0 = 0, bpR0 = 0, bmR2 = 0, bpR2 = 0, bmR3 = 0, bpR3 = 0;
  double bfm1 = bf1 / 2, bfp1 = bf1 / 2, bfm2 = bf2 / 2, bfp2 = bf2 / 2, bfm3 = bf3 / 2, bfp3 = bf3 / 2;
  bmR1 += bf0 / 2; bpR1 += bf0 / 2;
  double bX1 = -bf0 / 2, bX2 = -bf0 / 2, bX3 = 0, bX4 = -bf0 / 2;
  double bu = 0, bv = 0, bc = 0, bH = 0;
  double t = -bf1 / 2;
  bX1 += t * (u - c); bX2 += t * u; bX4 += t * (u + c);
  bu += t * (X1 + X2 + X4); bc += t * (-X1 + X4);
  t = -bf2 / 2;
  bX1 += t * v; bX2 += t * v; bX3 += t; bX4 += t * v;
  bv += t * (X1 + X2 + X4);
  t = -bf3 / 2;
  bX1 += t * (H - u * c); bX2 += t * (u * u + v * v); bX3 += t * v; bX4 += t * (H + u * c);
  bH += t * (X1 + X4);
  bu += t * (-X1 * c + 2.0 * u * X2 + X4 * c);
  bc += t * (-X1 * u + X4 * u);
  bv += t * (2.0 * v * X2 + X3);
  // X = a * W
  const double bW1 = a1 * bX1, bW2 = a2 * bX2, bW3 = a2 * bX3, bW4 = a4 * bX4;
  const double ba1 = W1 * bX1, ba2 = W2 * bX2 + W3 * bX3, ba4 = W4 * bX4;
  const double s1 = (u - c) >= 0.0 ? 1.0 : -1.0, s2 = u >= 0.0 ? 1.0 : -1.0, s4 = (u + c) >= 0.0 ? 1.0 : -1.0;
  bu += ba1 * s1 + ba2 * s2 + ba4 * s4;
  bc += -ba1 * s1 + ba4 * s4;
  // W terms
  double brho = 0, bdU = 0, bdP = 0, bdV = 0, bdR = 0, bc2 = 0;
  brho += bW1 * (-0.5 * dU * ic); bdU += bW1 * (-0.5 * rho * ic);
  bc += bW1 * (0.5 * rho * dU * ic2); bdP += bW1 * (0.5 * ic2); bc2 += bW1 * (-0.5 * dP * ic4);
  bdR += bW2; bdP += -bW2 * ic2; bc2 += bW2 * dP * ic4;
  brho += bW3 * dV; bdV += bW3 * rho;
  brho += bW4 * (0.5 * dU * ic); bdU += bW4 * (0.5 * rho * ic);
  bc += bW4 * (-0.5 * rho * dU * ic2); bdP += bW4 * (0.5 * ic2); bc2 += bW4 * (-0.5 * dP * ic4);
  double bpU = bdU, bmU = -bdU, bpP = bdP, bmP = -bdP, bpV = bdV, bmV = -bdV;
  bpR0 += bdR; bmR0 -= bdR;
  // c = sqrt(c2); c2 = g1 (H - (u^2+v^2)/2)
  bc2 += bc * 0.5 * ic;
  bH += g1 * bc2; bu += -g1 * u * bc2; bv += -g1 * v * bc2;
  // Roe averages
  double bms = (bu * (mU - u) + bv * (mV - v) + bH * (mH - H)) * iden;
  double bps = (bu * (pU - u) + bv * (pV - v) + bH * (pH - H)) * iden;
  const double wm = ms * iden, wp = ps * iden;
  bmU += bu * wm; bpU += bu * wp;
  bmV += bv * wm; bpV += bv * wp;
  double bmH = bH * wm, bpH = bH * wp;
  bms += brho * ps; bps += brho * ms;
  bmR0 += bms * 0.5 * ims; bpR0 += bps * 0.5 * ips;
  // mH = (mR3 + mP)/mR0
  bmR3 += bmH * imR0; bmP += bmH * imR0; bmR0 += -bmH * mH * imR0;
  bpR3 += bpH * ipR0; bpP += bpH * ipR0; bpR0 += -bpH * pH * ipR0;
  // fluxes in the rotated frame
  bmR1 += bfm1 * mU; bmU += bfm1 * mR1 + bfm2 * mR2 + bfm3 * (mR3 + mP); bmP += bfm1 + bfm3 * mU;
  bmR2 += bfm2 * mU; bmR3 += bfm3 * mU;
  bpR1 += bfp1 * pU; bpU += bfp1 * pR1 + bfp2 * pR2 + bfp3 * (pR3 + pP); bpP += bfp1 + bfp3 * pU;
  bpR2 += bfp2 * pU; bpR3 += bfp3 * pU;
  // pressures
  bmR3 += g1 * bmP; bmR1 += -0.5 * g1 * mU * bmP; bmU += -0.5 * g1 * mR1 * bmP;
  bmR2 += -0.5 * g1 * mV * bmP; bmV += -0.5 * g1 * mR2 * bmP;
  bpR3 += g1 * bpP; bpR1 += -0.5 * g1 * pU * bpP; bpU += -0.5 * g1 * pR1 * bpP;
  bpR2 += -0.5 * g1 * pV * bpP; bpV += -0.5 * g1 * pR2 * bpP;
  // velocities
  bmR1 += bmU * imR0; bmR2 += bmV * imR0; bmR0 += -(bmU * mU + bmV * mV) * imR0;
  bpR1 += bpU * ipR0; bpR2 += bpV * ipR0; bpR0 += -(bpU * pU + bpV * pV) * ipR0;
  // rotations
  bmQ[0] += bmR0; bmQ[3] += bmR3;
  bmQ[1] += nx * bmR1 - ny * bmR2; bmQ[2] += ny * bmR1 + nx * bmR2;
  beQ[0] += bpR0; beQ[3] += bpR3;
  beQ[1] += nx * bpR1 - ny * bpR2; beQ[2] += ny * bpR1 + nx * bpR2;
  bnx += bmR1 * mQ[1] + bmR2 * mQ[2] + bpR1 * eQ[1] + bpR2 * eQ[2];
  bny += bmR1 * mQ[2] - bmR2 * mQ[1] + bpR1 * eQ[2] - bpR2 * eQ[1];
}

// ------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------

template <int N>
__global__ __launch_bounds__(256, 2) void tlm_face_kernel(TlmArgs a, Phys ph, double bcud, double bcvd) {
  constexpr int Np = Ord<N>::Np, Nfp = Ord<N>::Nfp, NF = Ord<N>::NF;
  const size_t n = (size_t)a.ncells * NF;
  for (size_t it = blockIdx.x * (size_t)blockDim.x + threadIdx.x; it < n;
       it += (size_t)gridDim.x * blockDim.x) {
    const size_t c = it / NF;
    const int m = (int)(it - c * NF);
    const int f = m / Nfp, i = m - f * Nfp;
    const int node = a.fmask[m];
    Dual mQ[4], pQ[4], fl[4];
    for (int v = 0; v < 4; v++) mQ[v] = Dual(a.w[(c * Np + node) * 4 + v], a.wd[(c * Np + node) * 4 + v]);
    Dual nx(a.nx[it], a.nxd ? a.nxd[it] : 0.0), ny(a.ny[it], a.nyd ? a.nyd[it] : 0.0);
    Dual fs(a.fscale[it], a.fsd ? a.fsd[it] : 0.0);
    const int code = a.fmap[c * 3 + f];
    if (code >= 0) {
      const size_t nb = code >> 2;
      const int nn = a.fmask[(code & 3) * Nfp + (Nfp - 1 - i)];
      for (int v = 0; v < 4; v++) pQ[v] = Dual(a.w[(nb * Np + nn) * 4 + v], a.wd[(nb * Np + nn) * 4 + v]);
    } else {
      boundary_t<Dual>(ph, -code - 1, Dual(ph.bc_u, bcud), Dual(ph.bc_v, bcvd), mQ, nx, ny, pQ);
    }
    roe_t<Dual>(ph.gam, mQ, pQ, nx, ny, fs, fl);
    for (int v = 0; v < 4; v++) a.fluxd[it * 4 + v] = fl[v].d;
  }
}

template <int N>
__global__ __launch_bounds__(256, 2) void tlm_cell_kernel(TlmArgs a, Phys ph, const double *__restrict__ ops) {
  constexpr int Np = Ord<N>::Np, NF = Ord<N>::NF;
  const double *Drw = ops, *Dsw = ops + Np * Np, *LIFT = ops + 2 * Np * Np;
  const size_t n = (size_t)a.ncells * Np;
  for (size_t it = blockIdx.x * (size_t)blockDim.x + threadIdx.x; it < n;
       it += (size_t)gridDim.x * blockDim.x) {
    const size_t c = it / Np;
    const int j = (int)(it - c * Np);
    Dual dFdr[4], dFds[4], dGdr[4], dGds[4];
    for (int i = 0; i < Np; i++) {
      Dual q[4], F[4], G[4];
      for (int v = 0; v < 4; v++) q[v] = Dual(a.w[(c * Np + i) * 4 + v], a.wd[(c * Np + i) * 4 + v]);
      euler_flux_t<Dual>(ph.gam, q, F, G);
      const double ar = Drw[j * Np + i], as = Dsw[j * Np + i];
      for (int v = 0; v < 4; v++) {
        dFdr[v] += ar * F[v]; dFds[v] += as * F[v]; dGdr[v] += ar * G[v]; dGds[v] += as * G[v];
      }
    }
    Dual rx(a.rx[it], a.rxd ? a.rxd[it] : 0.0), sx(a.sx[it], a.sxd ? a.sxd[it] : 0.0);
    Dual ry(a.ry[it], a.ryd ? a.ryd[it] : 0.0), sy(a.sy[it], a.syd ? a.syd[it] : 0.0);
    for (int v = 0; v < 4; v++) {
      Dual r = (rx * dFdr[v] + sx * dFds[v]) + (ry * dGdr[v] + sy * dGds[v]);
      double lift = 0.0;
      for (int m = 0; m < NF; m++) lift += LIFT[j * NF + m] * a.fluxd[(c * NF + m) * 4 + v];
      a.out[it * 4 + v] = r.d - lift;
    }
  }
}

constexpr int kAdjThreads = 256;
// Threads per workgroup of the face-pair pass: one wave, so that its
// per-iteration barriers hold one wave rather than four in lockstep (the
// reverse sweep on the 65,536-cell O-grid: 84.1 -> 82.6 ms against 256).
constexpr int kPairThreads = 64;

// One thread per face slot, whole cells per workgroup; the cells' mu and the
// LIFT matrix are staged in LDS for psi = -LIFT^T mu.
template <int N>
__global__ __launch_bounds__(kAdjThreads, 2) void adj_face_kernel(AdjArgs a, Phys ph, const double *__restrict__ ops) {
  constexpr int Np = Ord<N>::Np, Nfp = Ord<N>::Nfp, NF = Ord<N>::NF;
  constexpr int CPB = kAdjThreads / NF;
  __shared__ double sL[Np * NF];
  __shared__ double sMu[4][CPB * Np];
  __shared__ double sG[3][CPB * NF];  // compact: per-slot geometry cotangents of the block
  __shared__ int sFm[NF];
  for (int k = threadIdx.x; k < Np * NF; k += kAdjThreads) sL[k] = ops[2 * Np * Np + k];
  if (threadIdx.x < NF) sFm[threadIdx.x] = a.fmask[threadIdx.x];
  const int t = threadIdx.x;
  const int cl = t / NF, m = t - cl * NF;
  const int f = m / Nfp, i = m - f * Nfp;
  const int node = m < NF ? a.fmask[m] : 0;
  for (size_t base = (size_t)blockIdx.x * CPB; base < (size_t)a.ncells; base += (size_t)gridDim.x * CPB) {
    __syncthreads();
    const int nc = (int)min((size_t)CPB, (size_t)a.ncells - base);
    const bool act = cl < nc;
    const size_t c = base + cl;
    const size_t it = c * NF + m;
    // the slot's own loads (independent of the staged mu) go out first, the
    // exterior trace as soon as the face-map code is in
    int code = 0;
    double4 mq4 = make_double4(0.0, 0.0, 0.0, 0.0), pq4 = mq4;
    double nx = 0.0, ny = 0.0, fs = 0.0;
    if (act) {
      code = a.fmap[c * 3 + f];
      mq4 = *reinterpret_cast<const double4 *>(a.w + (c * Np + node) * 4);
      nx = a.nx[it]; ny = a.ny[it]; fs = a.fscale[it];
    }
    for (int k = t; k < nc * Np; k += kAdjThreads) {
      const double4 mu = *reinterpret_cast<const double4 *>(a.mu + (base * Np + k) * 4);
      sMu[0][k] = mu.x; sMu[1][k] = mu.y; sMu[2][k] = mu.z; sMu[3][k] = mu.w;
    }
    if (act && code >= 0) {
      const size_t nb = code >> 2;
      const int nn = sFm[(code & 3) * Nfp + (Nfp - 1 - i)];
      pq4 = *reinterpret_cast<const double4 *>(a.w + (nb * Np + nn) * 4);
    }
    __syncthreads();
    if (act) {
      // cotangent of the scaled flux at this slot: L = vol - LIFT flux
      double psi[4] = {0, 0, 0, 0};
      for (int j = 0; j < Np; j++) {
        const double l = sL[j * NF + m];
        for (int v = 0; v < 4; v++) psi[v] -= l * sMu[v][cl * Np + j];
      }
      double mQ[4] = {mq4.x, mq4.y, mq4.z, mq4.w}, pQ[4] = {pq4.x, pq4.y, pq4.z, pq4.w};
      if (code < 0) boundary_state(ph, -code - 1, mQ, nx, ny, pQ);  // the forward's exterior state
      double bm[4] = {0, 0, 0, 0}, bp[4] = {0, 0, 0, 0}, bnx = 0, bny = 0, bfs = 0;
      roe_vjp(ph.gam, mQ, pQ, nx, ny, fs, psi, bm, bp, bnx, bny, bfs);
      // On a boundary slot bp is the cotangent of the exterior (boundary)
      // state; it stays in gpq (no neighbour reads a boundary slot's gpq) for
      // adj_boundary_kernel, which reverses the boundary state into gmq.
      *reinterpret_cast<double4 *>(a.gmq + it * 4) = make_double4(bm[0], bm[1], bm[2], bm[3]);
      *reinterpret_cast<double4 *>(a.gpq + it * 4) = make_double4(bp[0], bp[1], bp[2], bp[3]);
      if (a.compact) {
        sG[0][t] = bnx; sG[1][t] = bny; sG[2][t] = bfs;
      } else {
        a.g_nx[it] += a.scale * bnx;
        a.g_ny[it] += a.scale * bny;
        a.g_fs[it] += a.scale * bfs;
      }
    }
    if (a.compact) {
      // one thread per face sums its Nfp slots in index order
      __syncthreads();
      if (cl < nc && i == 0) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int j = 0; j < Nfp; j++) {
          s0 += sG[0][t + j]; s1 += sG[1][t + j]; s2 += sG[2][t + j];
        }
        const size_t fc = (base + cl) * 3 + f;
        a.g_nxf[fc] += a.scale * s0;
        a.g_nyf[fc] += a.scale * s1;
        a.g_fsf[fc] += a.scale * s2;
      }
    }
  }
}

// One thread per (boundary face node, input direction): reverse of the
// boundary state through one column of its dual Jacobian (directions: the 4
// interior variables, nx, ny, bc_u, bc_v), dotted with the exterior-state
// cotangent the face pass left in gpq.  Each direction feeds a different
// accumulator (the slot's interior cotangent, the normal, the free-stream
// velocity), so no two threads write the same value.  Runs between the face
// and cell passes.
template <int N>
__global__ __launch_bounds__(256) void adj_boundary_kernel(AdjArgs a, Phys ph, const int *__restrict__ fmask) {
  constexpr int Np = Ord<N>::Np, Nfp = Ord<N>::Nfp, NF = Ord<N>::NF;
  const int n = a.nbedges * Nfp * 8;
  for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < n; it += gridDim.x * blockDim.x) {
    const int dir = it & 7, bi = it >> 3;
    const int b = bi / Nfp, i = bi - b * Nfp;
    const size_t c = a.b2c[b];
    const int f = a.bnum[b], m = f * Nfp + i;
    const int code = a.fmap[c * 3 + f];
    if (code >= 0) continue;  // not a boundary face (never for a valid mesh)
    const size_t slot = c * NF + m;
    Dual q[4], qb[4];
    for (int v = 0; v < 4; v++) q[v] = Dual(a.w[(c * Np + fmask[m]) * 4 + v], dir == v ? 1.0 : 0.0);
    const Dual dnx(a.nx[slot], dir == 4 ? 1.0 : 0.0), dny(a.ny[slot], dir == 5 ? 1.0 : 0.0);
    const Dual du(ph.bc_u, dir == 6 ? 1.0 : 0.0), dv(ph.bc_v, dir == 7 ? 1.0 : 0.0);
    boundary_t<Dual>(ph, -code - 1, du, dv, q, dnx, dny, qb);
    double s = 0.0;
    for (int v = 0; v < 4; v++) s += a.gpq[slot * 4 + v] * qb[v].d;
    if (dir < 4) a.gmq[slot * 4 + dir] += s;
    else if (dir == 4) a.g_nx[slot] += a.scale * s;
    else if (dir == 5) a.g_ny[slot] += a.scale * s;
    else if (dir == 6) a.g_bcu[slot] += a.scale * s;
    else a.g_bcv[slot] += a.scale * s;
  }
}

// Face-pair form of the face and boundary passes.
//
// The Roe flux is antisymmetric, R(p, m, -n) = -R(m, p, n) (the flux of
// src/euler_rhs.h including its dW2 (u^2 + v^2) term: swapping the sides and
// reversing the normal negates the rotated velocities, which swaps the
// 1st and 4th wave strengths with a sign change, negates the 2nd and keeps
// the 3rd, so the mass and energy components change sign and the rotated
// momentum components keep theirs, which the back-rotation with -n turns
// into a sign change).  The two sides of an interior face carry opposite
// normals, so the scaled fluxes of a face node pair, fs_o R(m, p, n) and
// fs_p R(p, m, -n) = -fs_p R(m, p, n), are reversed together: one VJP with
// the cotangent fs_o psi_o - fs_p psi_p gives both traces' cotangents
// (each written to its own slot's gmq) and the normal's, which goes to the
// first side's face (the second side's normal is its negative, so the
// geometry chain sees the same total); fs_o and fs_p get psi_o . R and
// -psi_p . R.  In exact arithmetic this is the per-slot reverse; init_grid's
// per-node rounding makes the two normals opposite only to ~1e-15 relative
// times the cell's aspect, and moves the gradient by that much.
//
// Blocks [0, nbblk): the boundary slots, one thread each: the slot's
// reverse Roe flux with the boundary state as exterior, whose cotangent the
// hand-written reverse of the boundary state (boundary_vjp) carries to the
// trace, the normal and the free-stream velocity; per-slot outputs.
// Blocks [nbblk, grid): one thread per interior face node pair, Nfp
// consecutive threads per face, the per-face geometry sums in node order.
// psi, the cotangent of the scaled fluxes (-LIFT^T mu), is read from a.psi,
// which the cell pass before (or launch_adj_psi) wrote on the matrix cores.
template <int N>
__global__ __launch_bounds__(kPairThreads, 3) void adj_pair_kernel(AdjArgs a, Phys ph,
                                                                const int *__restrict__ fmask, int nbblk) {
  constexpr int Np = Ord<N>::Np, Nfp = Ord<N>::Nfp, NF = Ord<N>::NF;
  constexpr int FPB = kPairThreads / Nfp;  // faces per workgroup visit
  __shared__ double sG[4][kPairThreads];
  __shared__ int sFm[NF];
  const int t = threadIdx.x;
  if (t < NF) sFm[t] = fmask[t];
  __syncthreads();
  if ((int)blockIdx.x < nbblk) {
    const int bi = (int)blockIdx.x * kPairThreads + t;
    if (bi >= a.nbedges * Nfp) return;
    const int b = bi / Nfp, i = bi - b * Nfp;
    const size_t c = a.b2c[b];
    const int f = a.bnum[b], m = f * Nfp + i;
    const int code = a.fmap[c * 3 + f];
    if (code >= 0) return;  // not a boundary face (never for a valid mesh)
    const size_t slot = c * NF + m;
    const double4 q4 = *reinterpret_cast<const double4 *>(a.w + (c * Np + sFm[m]) * 4);
    const double4 p4 = *reinterpret_cast<const double4 *>(a.psi + slot * 4);
    const double psi[4] = {p4.x, p4.y, p4.z, p4.w};
    const double nx = a.nx[slot], ny = a.ny[slot], fs = a.fscale[slot];
    const double mQ[4] = {q4.x, q4.y, q4.z, q4.w};
    double pQ[4];
    boundary_state(ph, -code - 1, mQ, nx, ny, pQ);  // the forward's exterior state
    double bm[4] = {0, 0, 0, 0}, bp[4] = {0, 0, 0, 0}, bnx = 0, bny = 0, bfs = 0;
    roe_vjp(ph.gam, mQ, pQ, nx, ny, fs, psi, bm, bp, bnx, bny, bfs);
    double bu = 0.0, bv = 0.0;
    boundary_vjp(ph, -code - 1, ph.bc_u, ph.bc_v, mQ, nx, ny, bp, bm, bnx, bny, bu, bv);
    *reinterpret_cast<double4 *>(a.gmq + slot * 4) = make_double4(bm[0], bm[1], bm[2], bm[3]);
    a.g_nx[slot] += a.scale * bnx;
    a.g_ny[slot] += a.scale * bny;
    a.g_fs[slot] += a.scale * bfs;
    a.g_bcu[slot] += a.scale * bu;
    a.g_bcv[slot] += a.scale * bv;
    return;
  }
  const int fl = t / Nfp, i = t - fl * Nfp;
  const int nblk = (int)gridDim.x - nbblk;
  for (size_t base = (size_t)((int)blockIdx.x - nbblk) * FPB; base < (size_t)a.npairs;
       base += (size_t)nblk * FPB) {
    const size_t e = base + fl;
    const bool act = fl < FPB && e < (size_t)a.npairs;
    int fo = 0, fp = 0;
    double gnx = 0.0, gny = 0.0, gfo = 0.0, gfp = 0.0;
    if (act) {
      fo = a.pairs[2 * e];
      fp = a.pairs[2 * e + 1];
      const int co = fo / 3, cp = fp / 3;
      const int mo = (fo - 3 * co) * Nfp + i, mp = (fp - 3 * cp) * Nfp + (Nfp - 1 - i);
      const size_t so = (size_t)co * NF + mo, sp = (size_t)cp * NF + mp;
      const double4 q4 = *reinterpret_cast<const double4 *>(a.w + ((size_t)co * Np + sFm[mo]) * 4);
      const double4 e4 = *reinterpret_cast<const double4 *>(a.w + ((size_t)cp * Np + sFm[mp]) * 4);
      const double4 po = *reinterpret_cast<const double4 *>(a.psi + so * 4);
      const double4 pp = *reinterpret_cast<const double4 *>(a.psi + sp * 4);
      const double nx = a.nx[so], ny = a.ny[so], fso = a.fscale[so], fsp = a.fscale[sp];
      const double mQ[4] = {q4.x, q4.y, q4.z, q4.w}, pQ[4] = {e4.x, e4.y, e4.z, e4.w};
      const double o[4] = {fso * po.x - fsp * pp.x, fso * po.y - fsp * pp.y, fso * po.z - fsp * pp.z,
                           fso * po.w - fsp * pp.w};
      double bm[4] = {0, 0, 0, 0}, bp[4] = {0, 0, 0, 0}, bnx = 0, bny = 0, bfs = 0, R[4];
      roe_vjp(ph.gam, mQ, pQ, nx, ny, 1.0, o, bm, bp, bnx, bny, bfs, R);
      *reinterpret_cast<double4 *>(a.gmq + so * 4) = make_double4(bm[0], bm[1], bm[2], bm[3]);
      *reinterpret_cast<double4 *>(a.gmq + sp * 4) = make_double4(bp[0], bp[1], bp[2], bp[3]);
      gnx = bnx;
      gny = bny;
      gfo = po.x * R[0] + po.y * R[1] + po.z * R[2] + po.w * R[3];
      gfp = -(pp.x * R[0] + pp.y * R[1] + pp.z * R[2] + pp.w * R[3]);
    }
    sG[0][t] = gnx; sG[1][t] = gny; sG[2][t] = gfo; sG[3][t] = gfp;
    lds_wait();
    if (act && i == 0) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (int j = 0; j < Nfp; j++) {
        s0 += sG[0][t + j]; s1 += sG[1][t + j]; s2 += sG[2][t + j]; s3 += sG[3][t + j];
      }
      // the four accumulators (two faces, three arrays: never the same
      // element) are read together, then written: one round trip
      const double v0 = a.g_nxf[fo], v1 = a.g_nyf[fo], v2 = a.g_fsf[fo], v3 = a.g_fsf[fp];
      a.g_nxf[fo] = v0 + a.scale * s0;
      a.g_nyf[fo] = v1 + a.scale * s1;
      a.g_fsf[fo] = v2 + a.scale * s2;
      a.g_fsf[fp] = v3 + a.scale * s3;
    }
    lds_wait();  // one-wave workgroups: the wave's own LDS accesses
  }
}

// ------------------------------------------------------------------------
// psi = -LIFT^T mu on the matrix cores.  A tile of 16 consecutive cells'
// records (4 groups of 4 cells) is an A operand of v_mfma_f64_16x16x4f64 as
// it lies in memory: row cell + 4 var of group g, column node k at
// ((4 g + cell) Np + k) 4 + var; with B = -LIFT [Np][NF] lane l ends with
// cell l/16 of the group, slot l%16, the 4 variables.
// ------------------------------------------------------------------------
template <int N>
struct AdjTile {
  static constexpr int Np = Ord<N>::Np, NF = Ord<N>::NF, KP = (Np + 3) / 4;
  static constexpr int RCH = kMfmaTile * Np * 2;  // 16-byte chunks of a tile's records
  static constexpr int RD = copy_to_lds_doubles<RCH>();
};

// A operand of K-step ks of group g from a record tile (zero for the
// padding columns k >= Np); ln = lane_in_loop(lane)
template <int N>
__device__ __forceinline__ double adj_opA(const double *sR, int ln, int g, int ks) {
  constexpr int Np = Ord<N>::Np;
  const int r = ln & 15, k = 4 * ks + (ln >> 4);
  return (4 * ks + 3 < Np || k < Np) ? sR[((4 * g + (r & 3)) * Np + k) * 4 + (r >> 2)] : 0.0;
}

// B operand rows of M [Np][ld] for columns j < ncol (lane l: B[4 ks + l/16][l%16]),
// times sgn
template <int N>
__device__ __forceinline__ void adj_load_b(double *b, const double *M, int ld, int ncol, double sgn, int ln) {
  constexpr int Np = Ord<N>::Np, KP = (Np + 3) / 4;
  const int j = ln & 15, k0 = ln >> 4;
#pragma unroll
  for (int ks = 0; ks < KP; ks++) {
    const int k = 4 * ks + k0;
    const double v = M[(k < Np ? k : 0) * ld + (j < ncol ? j : 0)];
    b[ks] = (k < Np && j < ncol) ? sgn * v : 0.0;
  }
}

// psi = -LIFT^T of group g's records (bP: adj_load_b of LIFT, sign -1)
template <int N>
__device__ __forceinline__ dbl4 adj_psi_group(const double *sR, int ln, int g, const double *bP) {
  constexpr int KP = AdjTile<N>::KP;
  dbl4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < KP; ks++) d = mfma16x16x4(adj_opA<N>(sR, ln, g, ks), bP[ks], d);
  return d;
}

// psi = -LIFT^T mu of every cell into psi [K][NF][4] (the first launch of a
// sweep; later launches write it as psi_out), with the arithmetic of the
// fused kernel.
template <int N>
__global__ __launch_bounds__(64) void adj_psi_mfma_kernel(int ncells, const double *__restrict__ mu,
                                                         const double *__restrict__ ops, double *__restrict__ psi) {
  constexpr int Np = Ord<N>::Np, NF = Ord<N>::NF, KP = AdjTile<N>::KP;
  constexpr int TE = kMfmaTile, RCH = AdjTile<N>::RCH, RD = AdjTile<N>::RD;
  __shared__ double sMu[RD];
  const int lane = threadIdx.x;
  for (int i = lane; i < RD; i += 64) sMu[i] = 0.0;
  const int ntiles = (ncells + TE - 1) / TE;
  for (int it_t = 0;; it_t++) {
    const int tile = xcd_tile(it_t, ntiles);
    if (tile >= ntiles) break;
    const int e0 = tile * TE, ne = min(TE, ncells - e0);
    __syncthreads();
    copy_to_lds<RCH>(mu + (size_t)e0 * Np * 4, sMu, ne * Np * 2, lane);
    const int ln = lane_in_loop(lane), jj = ln & 15, ce = ln >> 4;
    double bP[KP];
    adj_load_b<N>(bP, ops + 2 * Np * Np, NF, NF, -1.0, ln);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll
    for (int g = 0; g < 4; g++) {
      if (4 * g >= ne) continue;
      const dbl4 p = adj_psi_group<N>(sMu, ln, g, bP);
      const int e = 4 * g + ce;
      if (e < ne && jj < NF) {
        double2 *d = reinterpret_cast<double2 *>(psi + ((size_t)(e0 + e) * NF + jj) * 4);
        d[0] = make_double2(p[0], p[1]);
        d[1] = make_double2(p[2], p[3]);
      }
    }
  }
}

// Cell pass of the face-pair form on the matrix cores (per-cell metric
// factors cmet, per-cell metric cotangents).  Each wave owns tiles of 16
// consecutive cells, 4 groups of 4 cells, laid out for
// v_mfma_f64_16x16x4f64 as in the stage kernels.  The tile's mu records and
// its slot cotangents gmq (both sides of each face, written by the pair
// pass) are copied into LDS by direct-to-LDS loads; the mu record layout
// is already an A operand (adj_opA).  Per group, nu_r = Drw^T mu and
// nu_s = Dsw^T mu on the matrix cores (lane: cell, node j, the 4 variables
// in its accumulator registers); with the cell's factors phF = rx nu_r +
// sx nu_s, phG = ry nu_r + sy nu_s, and lam = (dF/dw)^T phF + (dG/dw)^T phG
// plus the slot cotangents of the node's slots; the metric cotangents
// sum_i mu_i . (Drw F)_i = sum_j F_j . nu_r_j (etc.) per cell by shuffles.
// scale lam goes to the cell's slot-cotangent rows (read by then), the
// write-back forms out = em mu + scale lam (+ e0 x0 + f0 y0) chunk by chunk,
// coalesced, with mu from the LDS copy, and leaves out in the record tile,
// from which psi_out = -LIFT^T out (the next launch's psi: its mu is this
// out) comes on the matrix cores.
// The per-cell factors are node 0's; the forward's per-node factors differ
// from them by init_grid's rounding, which moves the gradient by that much.
template <int N, bool XY>
__global__ __launch_bounds__(64, 3) void adj_cell_mfma_kernel(AdjArgs a, Phys ph, const double *__restrict__ ops,
                                                             unsigned long long fmask4) {
  constexpr int Np = Ord<N>::Np, NF = Ord<N>::NF, KP = AdjTile<N>::KP;
  constexpr int TE = kMfmaTile, G = 4;
  constexpr int RCH = AdjTile<N>::RCH, SCH = TE * NF * 2;  // 16-byte chunks of a tile's records
  constexpr int RD = AdjTile<N>::RD, SD = copy_to_lds_doubles<SCH>();
  constexpr int WB = (RCH + 63) / 64;  // write-back chunks per lane
  static_assert(Np <= 16 && NF <= 16, "one 16-column output tile per cell group");
  __shared__ double sMu[RD], sGq[SD], sMet[128], sCell[TE * 4];
  auto fmask_of = [&](int m) -> int { return (int)((fmask4 >> (4 * m)) & 15ull); };
  const int lane = threadIdx.x;
  for (int i = lane; i < RD; i += 64) sMu[i] = 0.0;
  for (int i = lane; i < SD; i += 64) sGq[i] = 0.0;
  const double *Drw = ops, *Dsw = ops + Np * Np, *LIFT = ops + 2 * Np * Np;
  const int ntiles = (a.ncells + TE - 1) / TE;
  for (int it_t = 0;; it_t++) {
    const int tile = xcd_tile(it_t, ntiles);
    if (tile >= ntiles) break;
    const int e0 = tile * TE, ne = min(TE, a.ncells - e0);
    __syncthreads();  // the previous tile is done with the LDS tiles
    // ---- the tile's loads, one round trip: mu, gmq and the factors into
    // LDS; the lane's states and the running metric cotangents into
    // registers (the epilogue's x0 and y0 chunks load after the products,
    // which would otherwise spill).  They are issued at a raised wave
    // priority, as in the stage kernels (reverse sweep on the 65,536-cell
    // O-grid 82.2 -> 81.5 ms over three alternating runs; the same in the
    // face-pair pass gained nothing more)
    __builtin_amdgcn_s_setprio(2);
    copy_to_lds<RCH>(a.mu + (size_t)e0 * Np * 4, sMu, ne * Np * 2, lane);
    copy_to_lds<SCH>(a.gmq + (size_t)e0 * NF * 4, sGq, ne * NF * 2, lane);
    copy_to_lds<TE * 2>(a.cmet + (size_t)e0 * 4, sMet, ne * 2, lane);
    const int ln = lane_in_loop(lane), jj = ln & 15, ce = ln >> 4;
    double2 wr[G][2];
#pragma unroll
    for (int g = 0; g < G; g++) {
      const int e = min(4 * g + ce, ne - 1);
      const double2 *ws = reinterpret_cast<const double2 *>(a.w + ((size_t)(e0 + e) * Np + min(jj, Np - 1)) * 4);
      wr[g][0] = ws[0]; wr[g][1] = ws[1];
    }
    double bR[KP], bS[KP];
    adj_load_b<N>(bR, Drw, Np, Np, 1.0, ln);
    adj_load_b<N>(bS, Dsw, Np, Np, 1.0, ln);
    const int nch = ne * Np * 2;
    const size_t kc = (size_t)e0 + min(ln, ne - 1);
    const double crx = a.g_rx[kc], csx = a.g_sx[kc], cry = a.g_ry[kc], csy = a.g_sy[kc];
    __builtin_amdgcn_s_setprio(0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the LDS tiles are in place
    // the face slots that read node jj (at most two)
    int ms0 = -1, ms1 = -1;
#pragma unroll
    for (int m = 0; m < NF; m++)
      if (fmask_of(m) == jj) {
        if (ms0 < 0) ms0 = m;
        else ms1 = m;
      }
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (4 * g >= ne) continue;
      const int e = 4 * g + ce;
      const bool ok = e < ne && jj < Np;
      dbl4 nr = {0.0, 0.0, 0.0, 0.0}, ns = nr;
#pragma unroll
      for (int ks = 0; ks < KP; ks++) {
        const double av = adj_opA<N>(sMu, ln, g, ks);
        nr = mfma16x16x4(av, bR[ks], nr);
        ns = mfma16x16x4(av, bS[ks], ns);
      }
      double grx = 0.0, gsx = 0.0, gry = 0.0, gsy = 0.0, lam[4] = {0, 0, 0, 0};
      if (ok) {
        const double2 *mt = reinterpret_cast<const double2 *>(sMet + e * 4);
        const double2 rxy = mt[0], sxy = mt[1];
        const double wv[4] = {wr[g][0].x, wr[g][0].y, wr[g][1].x, wr[g][1].y};
        double phF[4], phG[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
          phF[v] = rxy.x * nr[v] + sxy.x * ns[v];
          phG[v] = rxy.y * nr[v] + sxy.y * ns[v];
        }
        euler_flux_vjp(ph.gam, wv, phF, phG, lam);
        // slot cotangents of the node's slots
        if (ms0 >= 0) {
          const double2 *sl = reinterpret_cast<const double2 *>(sGq + (e * NF + ms0) * 4);
          const double2 b01 = sl[0], b23 = sl[1];
          lam[0] += b01.x; lam[1] += b01.y; lam[2] += b23.x; lam[3] += b23.y;
        }
        if (ms1 >= 0) {
          const double2 *sl = reinterpret_cast<const double2 *>(sGq + (e * NF + ms1) * 4);
          const double2 b01 = sl[0], b23 = sl[1];
          lam[0] += b01.x; lam[1] += b01.y; lam[2] += b23.x; lam[3] += b23.y;
        }
        // metric cotangents at this node: F . nu_r, F . nu_s, G . nu_r, G . nu_s
        double F[4], Gv[4];
        euler_flux(ph.gam, wv, F, Gv);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          grx += F[v] * nr[v]; gsx += F[v] * ns[v]; gry += Gv[v] * nr[v]; gsy += Gv[v] * ns[v];
        }
      }
      // per-cell sums over the cell's 16 lanes (fixed order), into the stash
#pragma unroll
      for (int d = 1; d < 16; d <<= 1) {
        grx += __shfl_xor(grx, d, 64);
        gsx += __shfl_xor(gsx, d, 64);
        gry += __shfl_xor(gry, d, 64);
        gsy += __shfl_xor(gsy, d, 64);
      }
      if (jj == 0) {
        sCell[4 * e] = grx; sCell[4 * e + 1] = gsx; sCell[4 * e + 2] = gry; sCell[4 * e + 3] = gsy;
      }
      // scale lam into the cell's slot-cotangent rows (the node reads above
      // were this group's last use of them; Np <= NF, so they hold a record)
      if (ok) {
        double2 *d = reinterpret_cast<double2 *>(sGq + (e * NF + jj) * 4);
        d[0] = make_double2(a.scale * lam[0], a.scale * lam[1]);
        d[1] = make_double2(a.scale * lam[2], a.scale * lam[3]);
      }
    }
    double2 xc[WB], yc[WB];
    if (XY) {
#pragma unroll
      for (int q = 0; q < WB; q++) {
        const size_t ch = (size_t)e0 * Np * 2 + min(ln + 64 * q, nch - 1);
        xc[q] = reinterpret_cast<const double2 *>(a.x0)[ch];
        yc[q] = reinterpret_cast<const double2 *>(a.y0)[ch];
      }
    }
    __syncthreads();
    // ---- write-back: out = em mu + scale lam (+ e0 x0 + f0 y0), coalesced;
    // out replaces mu in the record tile
#pragma unroll
    for (int q = 0; q < WB; q++) {
      const int c = ln + 64 * q;
      if (c >= nch) continue;
      const int r = c >> 1, e = r / Np, k = r - e * Np;
      const double2 l = *reinterpret_cast<const double2 *>(sGq + (e * NF + k) * 4 + 2 * (c & 1));
      double2 *mr = reinterpret_cast<double2 *>(sMu) + c;
      const double2 mu2 = *mr;
      double2 o = make_double2(a.em * mu2.x + l.x, a.em * mu2.y + l.y);
      if (XY) {
        o.x += a.e0 * xc[q].x + a.f0 * yc[q].x;
        o.y += a.e0 * xc[q].y + a.f0 * yc[q].y;
      }
      reinterpret_cast<double2 *>(a.out)[(size_t)e0 * Np * 2 + c] = o;
      *mr = o;
    }
    if (ln < ne) {
      a.g_rx[kc] = crx + a.scale * sCell[4 * ln];
      a.g_sx[kc] = csx + a.scale * sCell[4 * ln + 1];
      a.g_ry[kc] = cry + a.scale * sCell[4 * ln + 2];
      a.g_sy[kc] = csy + a.scale * sCell[4 * ln + 3];
    }
    __syncthreads();
    // ---- psi_out = -LIFT^T out of the tile's cells
    double bP[KP];
    adj_load_b<N>(bP, LIFT, NF, NF, -1.0, ln);
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (4 * g >= ne) continue;
      const dbl4 p = adj_psi_group<N>(sMu, ln, g, bP);
      const int e = 4 * g + ce;
      if (e < ne && jj < NF) {
        double2 *d = reinterpret_cast<double2 *>(a.psi_out + ((size_t)(e0 + e) * NF + jj) * 4);
        d[0] = make_double2(p[0], p[1]);
        d[1] = make_double2(p[2], p[3]);
      }
    }
  }
}

// One thread per node, whole cells per workgroup; mu, the metrics and the
// nodal Euler fluxes of the cells are staged in LDS, so each flux is
// evaluated once per node (the per-slot form's cell pass).
template <int N>
__global__ __launch_bounds__(kAdjThreads, 2) void adj_cell_kernel(AdjArgs a, Phys ph, const double *__restrict__ ops,
                                const int *__restrict__ node_slots) {
  constexpr int Np = Ord<N>::Np, Nfp = Ord<N>::Nfp, NF = Ord<N>::NF;
  constexpr int CPB = kAdjThreads / Np, NT = CPB * Np;
  __shared__ double sD[2 * Np * Np];
  __shared__ double sMu[4][NT], sMet[4][NT], sFG[8][NT];
  for (int k = threadIdx.x; k < 2 * Np * Np; k += kAdjThreads) sD[k] = ops[k];
  const double *Drw = sD, *Dsw = sD + Np * Np;
  const int t = threadIdx.x;
  const int cl = t / Np, i = t - cl * Np;
  // the face slots that read this node (at most two)
  const int ms0 = t < NT ? node_slots[2 * i] : -1, ms1 = t < NT ? node_slots[2 * i + 1] : -1;
  for (size_t base = (size_t)blockIdx.x * CPB; base < (size_t)a.ncells; base += (size_t)gridDim.x * CPB) {
    const int nc = (int)min((size_t)CPB, (size_t)a.ncells - base);
    const bool act = cl < nc;
    const size_t c = base + cl;
    const size_t it = c * Np + i;
    double wi[4];
    __syncthreads();
    if (act) {
      const double4 w4 = *reinterpret_cast<const double4 *>(a.w + it * 4);
      wi[0] = w4.x; wi[1] = w4.y; wi[2] = w4.z; wi[3] = w4.w;
      const double4 mu = *reinterpret_cast<const double4 *>(a.mu + it * 4);
      sMu[0][t] = mu.x; sMu[1][t] = mu.y; sMu[2][t] = mu.z; sMu[3][t] = mu.w;
      sMet[0][t] = a.rx[it]; sMet[1][t] = a.sx[it]; sMet[2][t] = a.ry[it]; sMet[3][t] = a.sy[it];
      double F[4], G[4];
      euler_flux(ph.gam, wi, F, G);
      for (int v = 0; v < 4; v++) { sFG[v][t] = F[v]; sFG[4 + v][t] = G[v]; }
    }
    __syncthreads();
    double grx = 0, gsx = 0, gry = 0, gsy = 0;
    double o[4] = {0, 0, 0, 0};
    if (act) {
      const int c0 = cl * Np;
      // transposed weak derivatives with the metric terms
      double phF[4] = {0, 0, 0, 0}, phG[4] = {0, 0, 0, 0};
      for (int j = 0; j < Np; j++) {
        const double cr = Drw[j * Np + i], cs = Dsw[j * Np + i];
        const double aF = cr * sMet[0][c0 + j] + cs * sMet[1][c0 + j];
        const double aG = cr * sMet[2][c0 + j] + cs * sMet[3][c0 + j];
        for (int v = 0; v < 4; v++) {
          const double mu = sMu[v][c0 + j];
          phF[v] += aF * mu;
          phG[v] += aG * mu;
        }
      }
      double lam[4] = {0, 0, 0, 0};
      euler_flux_vjp(ph.gam, wi, phF, phG, lam);
      // face contributions that read this node: its own slots' interior
      // side (gmq) and the facing slots' exterior side (gpq)
      const int msl[2] = {ms0, ms1};
      for (int s = 0; s < 2; s++) {
        const int m = msl[s];
        if (m < 0) continue;
        const double4 gm = *reinterpret_cast<const double4 *>(a.gmq + (c * NF + m) * 4);
        lam[0] += gm.x; lam[1] += gm.y; lam[2] += gm.z; lam[3] += gm.w;
        const int code = a.fmap[c * 3 + m / Nfp];
        if (code >= 0) {
          const size_t nb = code >> 2;
          const int mp = (code & 3) * Nfp + (Nfp - 1 - m % Nfp);
          const double4 gp = *reinterpret_cast<const double4 *>(a.gpq + (nb * NF + mp) * 4);
          lam[0] += gp.x; lam[1] += gp.y; lam[2] += gp.z; lam[3] += gp.w;
        }
      }
      const double4 x4 = a.x0 ? *reinterpret_cast<const double4 *>(a.x0 + it * 4) : make_double4(0, 0, 0, 0);
      const double xv[4] = {x4.x, x4.y, x4.z, x4.w};
      for (int v = 0; v < 4; v++) o[v] = (a.x0 ? a.e0 * xv[v] : 0.0) + a.scale * lam[v];
      *reinterpret_cast<double4 *>(a.out + it * 4) = make_double4(o[0], o[1], o[2], o[3]);
      if (a.out2) {
        const double4 y4 = *reinterpret_cast<const double4 *>(a.y0 + it * 4);
        const double yv[4] = {y4.x, y4.y, y4.z, y4.w};
        double o2[4];
        for (int v = 0; v < 4; v++) o2[v] = a.f0 * yv[v] + (a.x0 ? a.f1 * xv[v] : 0.0) + a.f2 * o[v];
        *reinterpret_cast<double4 *>(a.out2 + it * 4) = make_double4(o2[0], o2[1], o2[2], o2[3]);
      }
      // metric cotangents: d/drx[i] = mu[i] . (Drw F)[i], etc.
      double dFdr[4] = {0, 0, 0, 0}, dFds[4] = {0, 0, 0, 0}, dGdr[4] = {0, 0, 0, 0}, dGds[4] = {0, 0, 0, 0};
      for (int k = 0; k < Np; k++) {
        const double ar = Drw[i * Np + k], as = Dsw[i * Np + k];
        for (int v = 0; v < 4; v++) {
          const double F = sFG[v][c0 + k], G = sFG[4 + v][c0 + k];
          dFdr[v] += ar * F; dFds[v] += as * F; dGdr[v] += ar * G; dGds[v] += as * G;
        }
      }
      for (int v = 0; v < 4; v++) {
        const double mu = sMu[v][t];
        grx += mu * dFdr[v]; gsx += mu * dFds[v]; gry += mu * dGdr[v]; gsy += mu * dGds[v];
      }
      if (!a.compact) {
        a.g_rx[it] += a.scale * grx;
        a.g_sx[it] += a.scale * gsx;
        a.g_ry[it] += a.scale * gry;
        a.g_sy[it] += a.scale * gsy;
      }
    }
    if (a.compact) {
      // per-cell sums: each node's four metric cotangents go through LDS
      // (the metric tiles are no longer read), one thread per cell adds
      // them in node order
      __syncthreads();
      if (t < NT) {
        sMet[0][t] = grx; sMet[1][t] = gsx; sMet[2][t] = gry; sMet[3][t] = gsy;
      }
      __syncthreads();
      if (act && i == 0) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int j = 0; j < Np; j++) {
          s0 += sMet[0][t + j]; s1 += sMet[1][t + j]; s2 += sMet[2][t + j]; s3 += sMet[3][t + j];
        }
        a.g_rx[c] += a.scale * s0;
        a.g_sx[c] += a.scale * s1;
        a.g_ry[c] += a.scale * s2;
        a.g_sy[c] += a.scale * s3;
      }
    }
  }
}

__global__ void lincomb_kernel(size_t n, double a, const double *__restrict__ x, double b,
                               const double *__restrict__ y, double *z) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    z[i] = a * x[i] + (y ? b * y[i] : 0.0);
}

// Geometry of the sensitivity path from the per-cell records cell_rel
// [K][6] = (x0, y0, x1 - x0, y1 - y0, x2 - x0, y2 - y0): the factors of
// init_grid_cell_rel_t (translation-invariant rounding), J at node 0.
template <int N>
__global__ __launch_bounds__(128) void init_grid_rel_kernel(int ncells, const double *__restrict__ ops,
                                                            const int *__restrict__ fmask,
                                                            const double *__restrict__ cell_rel,
                                                            double *rx, double *ry, double *sx,
                                                            double *sy, double *nx, double *ny,
                                                            double *fs, double *J, double *cmet) {
  constexpr int Np = Ord<N>::Np, NF = Ord<N>::NF;
  const double *Dr = ops + 2 * Np * Np + Np * NF;
  const double *Ds = Dr + Np * Np;
  const double *r = Ds + Np * Np;
  const double *s = r + Np;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < (size_t)ncells;
       c += (size_t)gridDim.x * blockDim.x) {
    const double e[4] = {cell_rel[6 * c + 2], cell_rel[6 * c + 3], cell_rel[6 * c + 4],
                         cell_rel[6 * c + 5]};
    double a[Np], b[Np], cc[Np], dd[Np], p[NF], q[NF], f[NF], Jc;
    init_grid_cell_rel_t<N, double>(r, s, Dr, Ds, fmask, e, a, b, cc, dd, p, q, f, Jc);
    for (int i = 0; i < Np; i++) {
      rx[c * Np + i] = a[i]; ry[c * Np + i] = b[i]; sx[c * Np + i] = cc[i]; sy[c * Np + i] = dd[i];
    }
    for (int m = 0; m < NF; m++) { nx[c * NF + m] = p[m]; ny[c * NF + m] = q[m]; fs[c * NF + m] = f[m]; }
    J[c] = Jc;
    if (cmet) {
      cmet[4 * c] = a[0]; cmet[4 * c + 1] = b[0]; cmet[4 * c + 2] = cc[0]; cmet[4 * c + 3] = dd[0];
    }
  }
}

// One thread per cell and absolute vertex coordinate d = 2 t + (0: x, 1: y):
// vertex t moves e (t = 1: edge 1, t = 2: edge 2, t = 0: both, negatively).
template <int N>
__global__ __launch_bounds__(128) void geo_vjp_kernel(int ncells, const double *__restrict__ ops,
                                                      const int *__restrict__ fmask,
                                                      const double *__restrict__ cell_rel, GeoCot g,
                                                      double *gv) {
  constexpr int Np = Ord<N>::Np, NF = Ord<N>::NF;
  const double *Dr = ops + 2 * Np * Np + Np * NF;
  const double *Ds = Dr + Np * Np;
  const double *r = Ds + Np * Np;
  const double *s = r + Np;
  const size_t n = (size_t)ncells * 6;
  for (size_t it = blockIdx.x * (size_t)blockDim.x + threadIdx.x; it < n;
       it += (size_t)gridDim.x * blockDim.x) {
    const size_t c = it / 6;
    const int d = (int)(it - c * 6), t = d >> 1, comp = d & 1;
    Dual e[4];
    for (int k = 0; k < 4; k++) {
      const int et = 1 + (k >> 1);  // the vertex whose edge vector e[k] belongs to
      const double seed = (k & 1) != comp ? 0.0 : (t == 0 ? -1.0 : (t == et ? 1.0 : 0.0));
      e[k] = Dual(cell_rel[6 * c + 2 + k], seed);
    }
    Dual rx[Np], ry[Np], sx[Np], sy[Np], nx[NF], ny[NF], fs[NF], Jc;
    init_grid_cell_rel_t<N, Dual>(r, s, Dr, Ds, fmask, e, rx, ry, sx, sy, nx, ny, fs, Jc);
    double acc = g.J ? g.J[c] * Jc.d : 0.0;
    if (g.compact) {
      // the factors' derivatives are the same at every node of the cell
      // (every node of a face) in exact arithmetic: node 0's (the face's
      // first node's) stand for them
      acc += g.rx[c] * rx[0].d + g.ry[c] * ry[0].d + g.sx[c] * sx[0].d + g.sy[c] * sy[0].d;
      for (int f = 0; f < 3; f++) {
        const size_t k = c * 3 + f;
        const int m = f * (NF / 3);
        acc += g.nxf[k] * nx[m].d + g.nyf[k] * ny[m].d + g.fsf[k] * fs[m].d;
      }
    } else {
      for (int i = 0; i < Np; i++) {
        const size_t k = c * Np + i;
        acc += g.rx[k] * rx[i].d + g.ry[k] * ry[i].d + g.sx[k] * sx[i].d + g.sy[k] * sy[i].d;
      }
    }
    for (int m = 0; m < NF; m++) {
      const size_t k = c * NF + m;
      acc += g.nx[k] * nx[m].d + g.ny[k] * ny[m].d + g.fs[k] * fs[m].d;
    }
    gv[it] = acc;
  }
}

// Compact copy of what the lift functional reads on the wall faces: per
// (wall face w, face node i) the state at the node and nx, ny, fscale of the
// slot, [nwall][Nfp][7]; then J of each wall cell, [nwall].
__global__ void wall_gather_kernel(int nwall, int Np, int Nfp, const int *__restrict__ wcell,
                                   const int *__restrict__ wface, const int *__restrict__ fmask,
                                   const double *__restrict__ q, const double *__restrict__ nx,
                                   const double *__restrict__ ny, const double *__restrict__ fs,
                                   const double *__restrict__ J, double *out) {
  const int NF = 3 * Nfp, n = nwall * Nfp;
  for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < n; it += gridDim.x * blockDim.x) {
    const int w = it / Nfp, i = it - w * Nfp;
    const size_t k = wcell[w];
    const int m = wface[w] * Nfp + i;
    const size_t node = k * Np + fmask[m];
    double *o = out + (size_t)it * 7;
    for (int v = 0; v < 4; v++) o[v] = q[node * 4 + v];
    o[4] = nx[k * NF + m]; o[5] = ny[k * NF + m]; o[6] = fs[k * NF + m];
    if (i == 0) out[(size_t)n * 7 + w] = J[k];
  }
}

// partial[b] = sum of x[i * stride + offset] over block b's contiguous
// chunk of i, in a fixed order (deterministic for a fixed grid).
__global__ __launch_bounds__(256) void strided_sum_kernel(size_t n, int stride, int offset,
                                                          const double *__restrict__ x, double *partial) {
  __shared__ double red[256];
  const size_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const size_t i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  double acc = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += 256) acc += x[i * stride + offset];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void add_indexed_kernel(int n, const int *__restrict__ idx, const double *__restrict__ val,
                                   double *dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[idx[i]] += val[i];
}

static int grid_of(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  return g < 1 ? 1 : (int)g;
}

template <int N>
static hipError_t tlm_n(const Launch &L, const TlmArgs &a, double bcud, double bcvd) {
  const size_t nf = (size_t)a.ncells * Ord<N>::NF, nc = (size_t)a.ncells * Ord<N>::Np;
  hipLaunchKernelGGL(tlm_face_kernel<N>, dim3(grid_of(nf)), dim3(256), 0, L.stream, a, L.ph, bcud, bcvd);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(tlm_cell_kernel<N>, dim3(grid_of(nc)), dim3(256), 0, L.stream, a, L.ph, L.ops);
  return hipGetLastError();
}

static int cell_blocks(size_t ncells, int cells_per_block) {
  size_t g = (ncells + cells_per_block - 1) / cells_per_block;
  if (g > 65536) g = 65536;
  return g < 1 ? 1 : (int)g;
}

template <int N>
static hipError_t adj_n(const Launch &L, const AdjArgs &a, const int *node_slots) {
  hipError_t e = hipSuccess;
  if (a.pairs) {
    // face-pair form: the face-pair pass (boundary blocks first, so that
    // their longer threads start early; psi read from a.psi), then the cell
    // pass on the matrix cores, which writes psi_out
    if (!a.cmet || !a.psi || !a.psi_out || a.out2 || (a.x0 && !a.y0)) return hipErrorInvalidValue;
    const int nbblk = (a.nbedges * Ord<N>::Nfp + kPairThreads - 1) / kPairThreads;
    const int fpb = kPairThreads / Ord<N>::Nfp;
    int npb = (a.npairs + fpb - 1) / fpb;
    // The pair workgroups (a grid-stride walk over the pairs) are capped at
    // 2/3 of the pass's resident capacity: a grid larger than the device
    // holds keeps workgroups pending, and those take every slot that frees
    // up, so the forward recomputation launched beside the sweep on the
    // second stream hardly runs until the pass drains.  Measured on the
    // 65,536-cell O-grid (one-wave workgroups): reverse sweep 86.6 ms at 1/3
    // of the capacity, 82.7 ms at 2/3, 83.6 ms at the full capacity (and
    // 86.8 ms uncapped with four-wave workgroups).
    static int pair_per_cu = -1;
    if (pair_per_cu < 0 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pair_per_cu, adj_pair_kernel<N>, kPairThreads, 0) != hipSuccess ||
         pair_per_cu < 1))
      pair_per_cu = 1;
    if (L.cus > 0) npb = std::min(npb, std::max(1, 2 * L.cus * pair_per_cu / 3));
    if (nbblk + npb > 0)
      hipLaunchKernelGGL(adj_pair_kernel<N>, dim3(nbblk + npb), dim3(kPairThreads), 0, L.stream, a, L.ph, L.fmask,
                         nbblk);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // one-wave workgroups, one 16-cell tile each (the XCD-chunked walk gives
    // each XCD a contiguous range of tiles)
    int grid = (a.ncells + kMfmaTile - 1) / kMfmaTile;
    if (L.max_blocks > 0 && grid > L.max_blocks) grid = L.max_blocks;
    if (grid >= 8) grid += (8 - grid % 8) % 8;
    if (a.x0)
      hipLaunchKernelGGL((adj_cell_mfma_kernel<N, true>), dim3(grid), dim3(64), 0, L.stream, a, L.ph, L.ops,
                         (unsigned long long)L.fmask4);
    else
      hipLaunchKernelGGL((adj_cell_mfma_kernel<N, false>), dim3(grid), dim3(64), 0, L.stream, a, L.ph, L.ops,
                         (unsigned long long)L.fmask4);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(adj_face_kernel<N>, dim3(cell_blocks(a.ncells, kAdjThreads / Ord<N>::NF)),
                     dim3(kAdjThreads), 0, L.stream, a, L.ph, L.ops);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.nbedges > 0) {
    const int nb = (a.nbedges * Ord<N>::Nfp * 8 + 255) / 256;
    hipLaunchKernelGGL(adj_boundary_kernel<N>, dim3(nb), dim3(256), 0, L.stream, a, L.ph, L.fmask);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(adj_cell_kernel<N>, dim3(cell_blocks(a.ncells, kAdjThreads / Ord<N>::Np)),
                     dim3(kAdjThreads), 0, L.stream, a, L.ph, L.ops, node_slots);
  return hipGetLastError();
}

hipError_t launch_rhs_tlm(const Launch &L, const TlmArgs &a, double bcud, double bcvd) {
  if (a.ncells == 0) return hipSuccess;
  switch (L.N) {
    case 1: return tlm_n<1>(L, a, bcud, bcvd);
    case 2: return tlm_n<2>(L, a, bcud, bcvd);
    case 3: return tlm_n<3>(L, a, bcud, bcvd);
    case 4: return tlm_n<4>(L, a, bcud, bcvd);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_rhs_adj(const Launch &L, const AdjArgs &a, const int *node_slots) {
  if (a.ncells == 0) return hipSuccess;
  switch (L.N) {
    case 1: return adj_n<1>(L, a, node_slots);
    case 2: return adj_n<2>(L, a, node_slots);
    case 3: return adj_n<3>(L, a, node_slots);
    case 4: return adj_n<4>(L, a, node_slots);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_adj_psi(const Launch &L, int ncells, const double *mu, double *psi) {
  if (ncells == 0) return hipSuccess;
  int grid = (ncells + kMfmaTile - 1) / kMfmaTile;
  if (grid >= 8) grid += (8 - grid % 8) % 8;
  switch (L.N) {
    case 1: hipLaunchKernelGGL(adj_psi_mfma_kernel<1>, dim3(grid), dim3(64), 0, L.stream, ncells, mu, L.ops, psi); break;
    case 2: hipLaunchKernelGGL(adj_psi_mfma_kernel<2>, dim3(grid), dim3(64), 0, L.stream, ncells, mu, L.ops, psi); break;
    case 3: hipLaunchKernelGGL(adj_psi_mfma_kernel<3>, dim3(grid), dim3(64), 0, L.stream, ncells, mu, L.ops, psi); break;
    case 4: hipLaunchKernelGGL(adj_psi_mfma_kernel<4>, dim3(grid), dim3(64), 0, L.stream, ncells, mu, L.ops, psi); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_lincomb(const Launch &L, size_t n, double a, const double *x, double b,
                          const double *y, double *z) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(lincomb_kernel, dim3(grid_of(n)), dim3(256), 0, L.stream, n, a, x, b, y, z);
  return hipGetLastError();
}

hipError_t launch_geo_vjp(const Launch &L, int ncells, const double *cell_rel, const GeoCot &g,
                          double *gv) {
  if (ncells == 0) return hipSuccess;
  size_t nb = ((size_t)ncells * 6 + 127) / 128;
  const int grid = nb > 65536 ? 65536 : (int)nb;
  switch (L.N) {
    case 1: hipLaunchKernelGGL(geo_vjp_kernel<1>, dim3(grid), dim3(128), 0, L.stream, ncells, L.ops, L.fmask, cell_rel, g, gv); break;
    case 2: hipLaunchKernelGGL(geo_vjp_kernel<2>, dim3(grid), dim3(128), 0, L.stream, ncells, L.ops, L.fmask, cell_rel, g, gv); break;
    case 3: hipLaunchKernelGGL(geo_vjp_kernel<3>, dim3(grid), dim3(128), 0, L.stream, ncells, L.ops, L.fmask, cell_rel, g, gv); break;
    case 4: hipLaunchKernelGGL(geo_vjp_kernel<4>, dim3(grid), dim3(128), 0, L.stream, ncells, L.ops, L.fmask, cell_rel, g, gv); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_init_grid_rel(const Launch &L, int ncells, const double *cell_rel, double *rx,
                                double *ry, double *sx, double *sy, double *nx, double *ny,
                                double *fs, double *J, double *cmet) {
  if (ncells == 0) return hipSuccess;
  size_t nb = ((size_t)ncells + 127) / 128;
  const int grid = nb > 65536 ? 65536 : (int)nb;
#define DG_IGR(n) hipLaunchKernelGGL(init_grid_rel_kernel<n>, dim3(grid), dim3(128), 0, L.stream, ncells, \
                                     L.ops, L.fmask, cell_rel, rx, ry, sx, sy, nx, ny, fs, J, cmet)
  switch (L.N) {
    case 1: DG_IGR(1); break;
    case 2: DG_IGR(2); break;
    case 3: DG_IGR(3); break;
    case 4: DG_IGR(4); break;
    default: return hipErrorInvalidValue;
  }
#undef DG_IGR
  return hipGetLastError();
}

hipError_t launch_strided_sum(const Launch &L, size_t n, int stride, int offset, const double *x,
                              double *partial) {
  hipLaunchKernelGGL(strided_sum_kernel, dim3(kSumBlocks), dim3(256), 0, L.stream, n, stride, offset, x,
                     partial);
  return hipGetLastError();
}

hipError_t launch_wall_gather(const Launch &L, int nwall, const int *wcell, const int *wface, const double *q,
                              const double *nx, const double *ny, const double *fs, const double *J,
                              double *out) {
  if (nwall == 0) return hipSuccess;
  const int n = nwall * L.Nfp;
  hipLaunchKernelGGL(wall_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, L.stream, nwall, L.Np, L.Nfp,
                     wcell, wface, L.fmask, q, nx, ny, fs, J, out);
  return hipGetLastError();
}

hipError_t launch_add_indexed(const Launch &L, int n, const int *idx, const double *val, double *dst) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(add_indexed_kernel, dim3((n + 255) / 256), dim3(256), 0, L.stream, n, idx, val, dst);
  return hipGetLastError();
}

}  // namespace dg
