/*
 * Oracle: CPU restatement of the reference DG Euler kernels (N-generic).
 *
 * TEST INFRASTRUCTURE ONLY -- see dg_oracle.h.  Compiled with
 * -ffp-contract=off so that every multiply and add rounds separately, as the
 * reference's plain C++ loops do without FMA contraction.
 *
 * Deviations from the reference, each deliberate:
 *  - the hardcoded order N=4 (15/5/60 literals) is generalised to c->N;
 *  - calc_dt's factor (4+1)^2 becomes (N+1)^2 (as vortex/src/calc_dt.h:12);
 *  - dead flux evaluations (src/euler_rhs.h:250-278, 131-149) are skipped:
 *    their results are never read;
 *  - get_bedge_q gains type 3, a characteristic (Riemann-invariant) far
 *    field, which the reference does not have.
 */
#include "dg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXNP 64
#define MAXNF 32

static double wtime(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

/* netlib dgemv, row-major NoTrans == column-major Trans:
 * temp = sum_i A[j][i] x[i] in ascending i; y[j] = beta*y[j] + alpha*temp. */
static void naive_dgemv(int m, int n, double alpha, const double *A, const double *x,
                        int incx, double beta, double *y, int incy) {
  if (beta == 0.0) {
    for (int j = 0; j < m; j++) y[j * incy] = 0.0;
  } else if (beta != 1.0) {
    for (int j = 0; j < m; j++) y[j * incy] = beta * y[j * incy];
  }
  for (int j = 0; j < m; j++) {
    double temp = 0.0;
    for (int i = 0; i < n; i++) temp = temp + A[j * n + i] * x[i * incx];
    y[j * incy] = y[j * incy] + alpha * temp;
  }
}

/* src/euler_rhs.h:7-23 */
void or_euler_flux(const or_consts *c, const double *q, double *F, double *G,
                   double *rho, double *u, double *v, double *p) {
  const double gam = c->gam;
  *rho = q[0];
  *u = q[1] / (*rho);
  *v = q[2] / (*rho);
  *p = (gam - 1) * (q[3] - 0.5 * (q[1] * (*u) + q[2] * (*v)));
  F[0] = q[1];
  F[1] = q[1] * (*u) + (*p);
  F[2] = q[2] * (*u);
  F[3] = (*u) * (q[3] + (*p));
  G[0] = q[2];
  G[1] = q[1] * (*v);
  G[2] = q[2] * (*v) + (*p);
  G[3] = (*v) * (q[3] + (*p));
}

/* src/euler_rhs.h:120-220.  Roe flux without entropy fix, evaluated in the
 * frame rotated onto the face normal, rotated back and scaled by fscale. */
void or_roe(const or_consts *c, double *flux, const double *nx, const double *ny,
            const double *fscale, const double *q, const double *pQ) {
  const double gam = c->gam;
  const int nf = 3 * c->Nfp;
  for (int i = 0; i < nf; i++) {
    const double *mQ = q + 4 * c->FMASK[i];
    const double *eQ = pQ + 4 * i;
    double mRoeQ[4], pRoeQ[4];
    mRoeQ[0] = mQ[0];
    mRoeQ[1] = nx[i] * mQ[1] + ny[i] * mQ[2];
    mRoeQ[2] = -ny[i] * mQ[1] + nx[i] * mQ[2];
    mRoeQ[3] = mQ[3];
    pRoeQ[0] = eQ[0];
    pRoeQ[1] = nx[i] * eQ[1] + ny[i] * eQ[2];
    pRoeQ[2] = -ny[i] * eQ[1] + nx[i] * eQ[2];
    pRoeQ[3] = eQ[3];

    double fxMQ[4], fyMQ[4], mRoeRho, mRoeU, mRoeV, mRoeP;
    or_euler_flux(c, mRoeQ, fxMQ, fyMQ, &mRoeRho, &mRoeU, &mRoeV, &mRoeP);
    double fxPQ[4], fyPQ[4], pRoeRho, pRoeU, pRoeV, pRoeP;
    or_euler_flux(c, pRoeQ, fxPQ, fyPQ, &pRoeRho, &pRoeU, &pRoeV, &pRoeP);

    double mH = (mRoeQ[3] + mRoeP) / mRoeRho;
    double pH = (pRoeQ[3] + pRoeP) / pRoeRho;
    double mRoeRhoSqrt = sqrt(mRoeRho);
    double pRoeRhoSqrt = sqrt(pRoeRho);
    double rho = mRoeRhoSqrt * pRoeRhoSqrt;
    double u = (mRoeRhoSqrt * mRoeU + pRoeRhoSqrt * pRoeU) / (mRoeRhoSqrt + pRoeRhoSqrt);
    double v = (mRoeRhoSqrt * mRoeV + pRoeRhoSqrt * pRoeV) / (mRoeRhoSqrt + pRoeRhoSqrt);
    double H = (mRoeRhoSqrt * mH + pRoeRhoSqrt * pH) / (mRoeRhoSqrt + pRoeRhoSqrt);
    double c2 = (gam - 1) * (H - 0.5 * (u * u + v * v));
    double cs = sqrt(c2);

    double dW1 = -0.5 * rho * (pRoeU - mRoeU) / cs + 0.5 * (pRoeP - mRoeP) / c2;
    double dW2 = (pRoeRho - mRoeRho) - (pRoeP - mRoeP) / c2;
    double dW3 = rho * (pRoeV - mRoeV);
    double dW4 = 0.5 * rho * (pRoeU - mRoeU) / cs + 0.5 * (pRoeP - mRoeP) / c2;
    dW1 = fabs(u - cs) * dW1;
    dW2 = fabs(u) * dW2;
    dW3 = fabs(u) * dW3;
    dW4 = fabs(u + cs) * dW4;

    double fx[4];
    fx[0] = (fxMQ[0] + fxPQ[0]) / 2.0;
    fx[1] = (fxMQ[1] + fxPQ[1]) / 2.0;
    fx[2] = (fxMQ[2] + fxPQ[2]) / 2.0;
    fx[3] = (fxMQ[3] + fxPQ[3]) / 2.0;
    fx[0] = fx[0] - (dW1 + dW2 + dW4) / 2.0;
    fx[1] = fx[1] - (dW1 * (u - cs) + dW2 * u + dW4 * (u + cs)) / 2.0;
    fx[2] = fx[2] - (dW1 * v + dW2 * v + dW3 + dW4 * v) / 2.0;
    fx[3] = fx[3] - (dW1 * (H - u * cs) + dW2 * (u * u + v * v) + dW3 * v + dW4 * (H + u * cs)) / 2.0;

    flux[i * 4] = fx[0];
    flux[i * 4 + 1] = nx[i] * fx[1] - ny[i] * fx[2];
    flux[i * 4 + 2] = ny[i] * fx[1] + nx[i] * fx[2];
    flux[i * 4 + 3] = fx[3];
    flux[i * 4] *= fscale[i];
    flux[i * 4 + 1] *= fscale[i];
    flux[i * 4 + 2] *= fscale[i];
    flux[i * 4 + 3] *= fscale[i];
  }
}

/* src/euler_rhs.h:222-291 */
void or_euler_rhs(const or_consts *c, const double *q, double *exteriorQ,
                  const double *rx, const double *ry, const double *sx,
                  const double *sy, const double *fscale, const double *nx,
                  const double *ny, double *qRHS) {
  const int Np = c->Np, nf = 3 * c->Nfp;
  double F[4 * MAXNP], G[4 * MAXNP];
  for (int i = 0; i < Np; i++) {
    double rho, u, v, p;
    or_euler_flux(c, &q[i * 4], &F[i * 4], &G[i * 4], &rho, &u, &v, &p);
  }
  for (int i = 0; i < 4; i++) {
    double dFdr[MAXNP], dFds[MAXNP], dGdr[MAXNP], dGds[MAXNP];
    naive_dgemv(Np, Np, 1.0, c->Drw, &F[i], 4, 0.0, dFdr, 1);
    naive_dgemv(Np, Np, 1.0, c->Dsw, &F[i], 4, 0.0, dFds, 1);
    naive_dgemv(Np, Np, 1.0, c->Drw, &G[i], 4, 0.0, dGdr, 1);
    naive_dgemv(Np, Np, 1.0, c->Dsw, &G[i], 4, 0.0, dGds, 1);
    for (int j = 0; j < Np; j++)
      qRHS[i + j * 4] = (rx[j] * dFdr[j] + sx[j] * dFds[j]) + (ry[j] * dGdr[j] + sy[j] * dGds[j]);
  }
  double flux[4 * 3 * MAXNF];
  or_roe(c, flux, nx, ny, fscale, q, exteriorQ);
  for (int i = 0; i < 4; i++)
    naive_dgemv(Np, nf, -1.0, c->LIFT, &flux[i], 4, 1.0, qRHS + i, 4);
  for (int i = 0; i < 4 * nf; i++) exteriorQ[i] = 0.0;
}

/* src/get_neighbour_q.h:1-89.  The orientation test compares the first
 * vertex of each cell's local edge; on counter-clockwise cells sharing an
 * edge (and across periodic seams, where unwrapped coordinates differ by a
 * period) it always yields reverse = true. */
void or_get_neighbour_q(const or_consts *c, const int *edgeNum, const double *xL,
                        const double *yL, const double *xR, const double *yR,
                        const double *qL, const double *qR, double *exteriorQL,
                        double *exteriorQR) {
  const int Nfp = c->Nfp;
  int edgeL = edgeNum[0];
  int edgeR = edgeNum[1];
  int reverse = !(xL[edgeL] == xR[edgeR] && yL[edgeL] == yR[edgeR]);

  int exInd = edgeL * 4 * Nfp;
  const int *fmask = c->FMASK + edgeR * Nfp;
  for (int i = 0; i < Nfp; i++) {
    int rInd = reverse ? 4 * fmask[Nfp - i - 1] : 4 * fmask[i];
    exteriorQL[exInd + 4 * i] += qR[rInd];
    exteriorQL[exInd + 4 * i + 1] += qR[rInd + 1];
    exteriorQL[exInd + 4 * i + 2] += qR[rInd + 2];
    exteriorQL[exInd + 4 * i + 3] += qR[rInd + 3];
  }
  exInd = edgeR * 4 * Nfp;
  fmask = c->FMASK + edgeL * Nfp;
  for (int i = 0; i < Nfp; i++) {
    int lInd = reverse ? 4 * fmask[Nfp - i - 1] : 4 * fmask[i];
    exteriorQR[exInd + 4 * i] += qL[lInd];
    exteriorQR[exInd + 4 * i + 1] += qL[lInd + 1];
    exteriorQR[exInd + 4 * i + 2] += qL[lInd + 2];
    exteriorQR[exInd + 4 * i + 3] += qL[lInd + 3];
  }
}

/* Characteristic far-field state from the interior state qi and the
 * free-stream state, for the outward unit normal (nx, ny): Riemann
 * invariants R+ = vn_i + 2c_i/(g-1) (from inside) and R- = vn_inf -
 * 2c_inf/(g-1) (from outside) when subsonic; entropy and tangential
 * velocity from the upwind side; interior or free stream when supersonic. */
static void farfield_state(const or_consts *c, const double *qi, double nx, double ny,
                           double *qb) {
  const double g = c->gam;
  double rhoi = qi[0], ui = qi[1] / rhoi, vi = qi[2] / rhoi;
  double pi_ = (g - 1.0) * (qi[3] - 0.5 * rhoi * (ui * ui + vi * vi));
  double ci = sqrt(fabs(g * pi_ / rhoi));
  double vni = ui * nx + vi * ny;
  double rinf = c->bc_r, uinf = c->bc_u, vinf = c->bc_v, pinf = c->bc_p;
  double cinf = sqrt(g * pinf / rinf);
  double vninf = uinf * nx + vinf * ny;
  if (fabs(vni) >= ci) {
    if (vni < 0.0) { /* supersonic inflow */
      qb[0] = rinf; qb[1] = rinf * uinf; qb[2] = rinf * vinf; qb[3] = c->bc_e;
    } else {         /* supersonic outflow */
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
    ent = pinf / pow(rinf, g);
    ut = uinf - vninf * nx; vt = vinf - vninf * ny;
  }
  double rhob = pow(cb * cb / (g * ent), 1.0 / (g - 1.0));
  double pb = rhob * cb * cb / g;
  double ub = ut + vnb * nx, vb = vt + vnb * ny;
  qb[0] = rhob;
  qb[1] = rhob * ub;
  qb[2] = rhob * vb;
  qb[3] = pb / (g - 1.0) + 0.5 * rhob * (ub * ub + vb * vb);
}

/* src/get_bedge_q.h:1-51: 0 inflow (free stream), 1 outflow (free-stream
 * density and momentum, interior energy), 2 slip wall (reflect the normal
 * momentum); 3 characteristic far field (no reference counterpart). */
void or_get_bedge_q(const or_consts *c, const int *bedge_type, const int *bedgeNum,
                    const double *nx, const double *ny, const double *q,
                    double *exteriorQ) {
  const int Nfp = c->Nfp;
  int exInd = *bedgeNum * 4 * Nfp;
  int nInd = *bedgeNum * Nfp;
  const int *fmask = c->FMASK + *bedgeNum * Nfp;
  const double bc_r = c->bc_r, bc_u = c->bc_u, bc_v = c->bc_v, bc_e = c->bc_e;
  if (*bedge_type == 0) {
    for (int i = 0; i < Nfp; i++) {
      exteriorQ[exInd + i * 4] += bc_r;
      exteriorQ[exInd + i * 4 + 1] += bc_r * bc_u;
      exteriorQ[exInd + i * 4 + 2] += bc_r * bc_v;
      exteriorQ[exInd + i * 4 + 3] += bc_e;
    }
  } else if (*bedge_type == 1) {
    for (int i = 0; i < Nfp; i++) {
      int qInd = fmask[i] * 4;
      exteriorQ[exInd + i * 4] += bc_r;
      exteriorQ[exInd + i * 4 + 1] += bc_r * bc_u;
      exteriorQ[exInd + i * 4 + 2] += bc_r * bc_v;
      exteriorQ[exInd + i * 4 + 3] += q[qInd + 3];
    }
  } else if (*bedge_type == 2) {
    for (int i = 0; i < Nfp; i++) {
      int qInd = fmask[i] * 4;
      exteriorQ[exInd + i * 4] += q[qInd];
      exteriorQ[exInd + i * 4 + 1] += q[qInd + 1] - 2 * (nx[nInd + i] * q[qInd + 1] + ny[nInd + i] * q[qInd + 2]) * nx[nInd + i];
      exteriorQ[exInd + i * 4 + 2] += q[qInd + 2] - 2 * (nx[nInd + i] * q[qInd + 1] + ny[nInd + i] * q[qInd + 2]) * ny[nInd + i];
      exteriorQ[exInd + i * 4 + 3] += q[qInd + 3];
    }
  } else {
    for (int i = 0; i < Nfp; i++) {
      int qInd = fmask[i] * 4;
      double qb[4];
      farfield_state(c, q + qInd, nx[nInd + i], ny[nInd + i], qb);
      exteriorQ[exInd + i * 4] += qb[0];
      exteriorQ[exInd + i * 4 + 1] += qb[1];
      exteriorQ[exInd + i * 4 + 2] += qb[2];
      exteriorQ[exInd + i * 4 + 3] += qb[3];
    }
  }
}

/* vortex/src/get_bedge_q.h:1-40 (the vortex formula as or_vortex_state's) */
void or_get_bedge_q_exact(const or_consts *c, const int *bedgeNum, const double *x,
                          const double *y, double t, const double *exact, double *exteriorQ) {
  const int Nfp = c->Nfp;
  int exInd = *bedgeNum * 4 * Nfp;
  const int *fmask = c->FMASK + *bedgeNum * Nfp;
  for (int i = 0; i < Nfp; i++) {
    double qb[4];
    or_vortex_state(c->gam, exact[0], exact[1], exact[2], exact[3], t, x[fmask[i]], y[fmask[i]],
                    exact[4], qb);
    exteriorQ[exInd + i * 4] += qb[0];
    exteriorQ[exInd + i * 4 + 1] += qb[1];
    exteriorQ[exInd + i * 4 + 2] += qb[2];
    exteriorQ[exInd + i * 4 + 3] += qb[3];
  }
}

/* src/set_workingQ.h:1-12 */
void or_set_workingQ(const or_consts *c, const double *dt, const int *stage,
                     const double *q, const double *k1, const double *k2,
                     double *workingQ) {
  const int n = 4 * c->Np;
  if (*stage == 0) {
    for (int i = 0; i < n; i++) workingQ[i] = q[i] + (*dt) * k1[i];
  } else {
    for (int i = 0; i < n; i++) workingQ[i] = q[i] + (*dt) * (k1[i] / 4.0 + k2[i] / 4.0);
  }
}

/* src/update_Q.h:1-12 */
void or_update_Q(const or_consts *c, const double *dt, double *q, const double *rk1,
                 const double *rk2, const double *rk3, double *workingQ) {
  const int n = 4 * c->Np;
  for (int i = 0; i < n; i++) {
    q[i] = q[i] + (*dt) * (rk1[i] / 6.0 + rk2[i] / 6.0 + 2.0 * rk3[i] / 3.0);
    workingQ[i] = q[i];
  }
}

/* src/calc_dt.h:1-22, with (N+1)^2.  The quirk fscale[FMASK[i]] (a node
 * index used on a face-slot array) is reproduced: it changes dt. */
void or_calc_dt(const or_consts *c, const double *q, const double *fscale, double *dt1) {
  const double gam = c->gam;
  const int nf = 3 * c->Nfp;
  double dt1_arr[3 * MAXNF];
  for (int i = 0; i < nf; i++) {
    int f = c->FMASK[i];
    double rho = q[f * 4];
    double u = q[f * 4 + 1] / rho;
    double v = q[f * 4 + 2] / rho;
    double p = (gam - 1.0) * (q[f * 4 + 3] - rho * (u * u + v * v) * 0.5);
    double cs = sqrt(fabs(gam * p / rho));
    dt1_arr[i] = ((double)((c->N + 1) * (c->N + 1))) * 0.5 * fscale[f] * (sqrt(u * u + v * v) + cs);
  }
  double mx = *dt1;
  for (int i = 0; i < nf; i++)
    if (dt1_arr[i] > mx) mx = dt1_arr[i];
  *dt1 = mx;
}

/* src/neighbour_zero.h:1-5 */
void or_neighbour_zero(const or_consts *c, double *exQ) {
  for (int i = 0; i < 4 * 3 * c->Nfp; i++) exQ[i] = 0.0;
}

/* src/init_grid.h:1-106 (dcopy/daxpy/dscal/dgemv restated as netlib loops) */
void or_init_grid(const or_consts *c, const double *n0, const double *n1,
                  const double *n2, double *nodeX, double *nodeY, double *x,
                  double *y, double *xr, double *yr, double *xs, double *ys,
                  double *rx, double *ry, double *sx, double *sy, double *nx,
                  double *ny, double *fscale) {
  const int Np = c->Np, Nfp = c->Nfp;
  const double *r = c->r, *s = c->s;
  double temp[MAXNP], J[MAXNP];
  nodeX[0] = n0[0]; nodeX[1] = n1[0]; nodeX[2] = n2[0];
  nodeY[0] = n0[1]; nodeY[1] = n1[1]; nodeY[2] = n2[1];
  for (int d = 0; d < 2; d++) {
    double *o = d == 0 ? x : y;
    for (int i = 0; i < Np; i++) o[i] = 1.0;                 /* dcopy(ones) */
    for (int i = 0; i < Np; i++) o[i] = o[i] + 1.0 * r[i];   /* daxpy(1, r) */
    double a = 0.5 * n1[d];
    for (int i = 0; i < Np; i++) o[i] = a * o[i];            /* dscal */
    for (int i = 0; i < Np; i++) temp[i] = 1.0;
    for (int i = 0; i < Np; i++) temp[i] = temp[i] + 1.0 * s[i];
    a = 0.5 * n2[d];
    for (int i = 0; i < Np; i++) o[i] = o[i] + a * temp[i];
    for (int i = 0; i < Np; i++) temp[i] = s[i];
    for (int i = 0; i < Np; i++) temp[i] = temp[i] + 1.0 * r[i];
    a = -0.5 * n0[d];
    for (int i = 0; i < Np; i++) o[i] = o[i] + a * temp[i];
  }
  naive_dgemv(Np, Np, 1.0, c->Dr, x, 1, 0.0, xr, 1);
  naive_dgemv(Np, Np, 1.0, c->Ds, x, 1, 0.0, xs, 1);
  naive_dgemv(Np, Np, 1.0, c->Dr, y, 1, 0.0, yr, 1);
  naive_dgemv(Np, Np, 1.0, c->Ds, y, 1, 0.0, ys, 1);
  for (int i = 0; i < Np; i++) J[i] = -xs[i] * yr[i] + xr[i] * ys[i];
  for (int i = 0; i < Np; i++) {
    rx[i] = ys[i] / J[i];
    sx[i] = -yr[i] / J[i];
    ry[i] = -xs[i] / J[i];
    sy[i] = xr[i] / J[i];
  }
  const int *F = c->FMASK;
  for (int i = 0; i < Nfp; i++) {
    nx[i] = yr[F[i]];
    ny[i] = -xr[F[i]];
  }
  for (int i = 0; i < Nfp; i++) {
    nx[Nfp + i] = ys[F[Nfp + i]] - yr[F[Nfp + i]];
    ny[Nfp + i] = xr[F[Nfp + i]] - xs[F[Nfp + i]];
  }
  for (int i = 0; i < Nfp; i++) {
    nx[2 * Nfp + i] = -ys[F[2 * Nfp + i]];
    ny[2 * Nfp + i] = xs[F[2 * Nfp + i]];
  }
  for (int i = 0; i < 3 * Nfp; i++) {
    double sJ = sqrt(nx[i] * nx[i] + ny[i] * ny[i]);
    nx[i] = nx[i] / sJ;
    ny[i] = ny[i] / sJ;
    fscale[i] = sJ / J[F[i]];
  }
}

/* src/set_ic.h:1-12 */
void or_set_ic(const or_consts *c, double *q, double *workingQ) {
  for (int i = 0; i < c->Np; i++) {
    q[i * 4] = c->bc_r;
    q[i * 4 + 1] = c->bc_r * c->bc_u;
    q[i * 4 + 2] = c->bc_r * c->bc_v;
    q[i * 4 + 3] = c->bc_e;
    for (int v = 0; v < 4; v++) workingQ[i * 4 + v] = q[i * 4 + v];
  }
}

/* ------------------------------------------------------------------ */

void or_init_grid_all(const or_consts *c, int ncells, const double *cell_xy,
                      double *nodeX, double *nodeY, double *x, double *y,
                      double *xr, double *yr, double *xs, double *ys, double *rx,
                      double *ry, double *sx, double *sy, double *nx, double *ny,
                      double *fscale) {
  const int Np = c->Np, nf = 3 * c->Nfp;
  for (int k = 0; k < ncells; k++) {
    const double *v = cell_xy + 6 * k;
    or_init_grid(c, v, v + 2, v + 4, nodeX + 3 * k, nodeY + 3 * k, x + Np * k,
                 y + Np * k, xr + Np * k, yr + Np * k, xs + Np * k, ys + Np * k,
                 rx + Np * k, ry + Np * k, sx + Np * k, sy + Np * k, nx + nf * k,
                 ny + nf * k, fscale + nf * k);
  }
}

double or_calc_dt_all(const or_consts *c, const or_mesh *m, const double *q,
                      double dt1_in, int nthreads) {
  const int Np = c->Np, nf = 3 * c->Nfp;
  double result = dt1_in;
  if (nthreads <= 1) {
    for (int k = 0; k < m->ncells; k++) or_calc_dt(c, q + 4 * Np * k, m->fscale + nf * k, &result);
    return result;
  }
  /* per-thread partials, combined afterwards (src/openmp/calc_dt_kernel.cpp:43-72) */
  double partial[256];
  int nt = nthreads > 256 ? 256 : nthreads;
  for (int t = 0; t < nt; t++) partial[t] = dt1_in;
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
    int nth = omp_get_num_threads();
#else
    int t = 0, nth = 1;
#endif
    long lo = (long)m->ncells * t / nth, hi = (long)m->ncells * (t + 1) / nth;
    double acc = dt1_in;
    for (long k = lo; k < hi; k++) or_calc_dt(c, q + 4 * Np * k, m->fscale + nf * k, &acc);
    partial[t] = acc;
  }
  for (int t = 0; t < nt; t++)
    if (partial[t] > result) result = partial[t];
  return result;
}

static void edge_body(const or_consts *c, const or_mesh *m, or_state *st, int e) {
  const int Np = c->Np, nf = 3 * c->Nfp;
  int L = m->edge2cells[2 * e], R = m->edge2cells[2 * e + 1];
  or_get_neighbour_q(c, m->edgeNum + 2 * e, m->nodeX + 3 * L, m->nodeY + 3 * L,
                     m->nodeX + 3 * R, m->nodeY + 3 * R, st->workingQ + 4 * Np * L,
                     st->workingQ + 4 * Np * R, st->exteriorQ + 4 * nf * L,
                     st->exteriorQ + 4 * nf * R);
}

void or_ssprk3_step_timed(const or_consts *c, const or_mesh *m, or_state *st,
                          double t, double dt, int nthreads, const int *cptr, const int *clist,
                          int ncol, double *kt) {
  const double frac[3] = {0.0, 1.0, 0.5};
  const int Np = c->Np, nf = 3 * c->Nfp;
  const int K = m->ncells;
  double *rk[3] = {st->rk1, st->rk2, st->rk3};
  int nt = nthreads < 1 ? 1 : nthreads;
  for (int j = 0; j < 3; j++) {
    double t0 = wtime();
    if (nt == 1 || cptr == 0) {
      for (int e = 0; e < m->nedges; e++) edge_body(c, m, st, e);
    } else {
      for (int col = 0; col < ncol; col++) {
        int lo = cptr[col], hi = cptr[col + 1];
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int idx = lo; idx < hi; idx++) edge_body(c, m, st, clist[idx]);
      }
    }
    double t1 = wtime();
    for (int b = 0; b < m->nbedges; b++) {
      int k = m->bedge2cells[b];
      if (m->bedge_type[b] == 4)
        or_get_bedge_q_exact(c, m->bedgeNum + b, m->x + Np * k, m->y + Np * k, t + frac[j] * dt,
                             m->exact, st->exteriorQ + 4 * nf * k);
      else
        or_get_bedge_q(c, m->bedge_type + b, m->bedgeNum + b, m->nx + nf * k, m->ny + nf * k,
                       st->workingQ + 4 * Np * k, st->exteriorQ + 4 * nf * k);
    }
    double t2 = wtime();
    double *out = rk[j];
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int k = 0; k < K; k++)
      or_euler_rhs(c, st->workingQ + 4 * Np * k, st->exteriorQ + 4 * nf * k,
                   m->rx + Np * k, m->ry + Np * k, m->sx + Np * k, m->sy + Np * k,
                   m->fscale + nf * k, m->nx + nf * k, m->ny + nf * k, out + 4 * Np * k);
    double t3 = wtime();
    if (j != 2) {
#pragma omp parallel for schedule(static) num_threads(nt)
      for (int k = 0; k < K; k++)
        or_set_workingQ(c, &dt, &j, st->q + 4 * Np * k, st->rk1 + 4 * Np * k,
                        st->rk2 + 4 * Np * k, st->workingQ + 4 * Np * k);
    }
    double t4 = wtime();
    if (kt) { kt[0] += t1 - t0; kt[1] += t2 - t1; kt[2] += t3 - t2; kt[3] += t4 - t3; }
  }
  double t5 = wtime();
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int k = 0; k < K; k++)
    or_update_Q(c, &dt, st->q + 4 * Np * k, st->rk1 + 4 * Np * k, st->rk2 + 4 * Np * k,
                st->rk3 + 4 * Np * k, st->workingQ + 4 * Np * k);
  if (kt) kt[4] += wtime() - t5;
}

void or_ssprk3_step(const or_consts *c, const or_mesh *m, or_state *st, double dt,
                    int nthreads, const int *cptr, const int *clist, int ncol) {
  or_ssprk3_step_timed(c, m, st, 0.0, dt, nthreads, cptr, clist, ncol, 0);
}

int or_colour_edges(int ncells, int nedges, const int *edge2cells, int *colour_ptr,
                    int *colour_list) {
  unsigned char *used = (unsigned char *)calloc((size_t)ncells, 1);
  unsigned char *col = (unsigned char *)malloc((size_t)nedges + 1);
  int ncol = 0;
  for (int e = 0; e < nedges; e++) {
    int L = edge2cells[2 * e], R = edge2cells[2 * e + 1];
    unsigned mask = used[L] | used[R];
    int k = 0;
    while (mask & (1u << k)) k++;
    if (k >= 8) { free(used); free(col); return -1; }
    col[e] = (unsigned char)k;
    used[L] |= (unsigned char)(1u << k);
    used[R] |= (unsigned char)(1u << k);
    if (k + 1 > ncol) ncol = k + 1;
  }
  int pos = 0;
  for (int k = 0; k < ncol; k++) {
    colour_ptr[k] = pos;
    for (int e = 0; e < nedges; e++)
      if (col[e] == k) colour_list[pos++] = e;
  }
  colour_ptr[ncol] = pos;
  free(used);
  free(col);
  return ncol;
}

double or_run(const or_consts *c, const or_mesh *m, or_state *st, int nsteps,
              double cfl, int nthreads, double *dt_out, double *ktime) {
  int *cptr = 0, *clist = 0, ncol = 0;
  if (nthreads > 1) {
    cptr = (int *)malloc(9 * sizeof(int));
    clist = (int *)malloc(((size_t)m->nedges + 1) * sizeof(int));
    ncol = or_colour_edges(m->ncells, m->nedges, m->edge2cells, cptr, clist);
  }
  double t6 = wtime();
  double dt1 = or_calc_dt_all(c, m, st->q, 0.0, nthreads);
  if (ktime) ktime[5] += wtime() - t6;
  double dt = cfl / dt1;
  double t = 0.0;
  for (int i = 0; i < nsteps; i++) {
    or_ssprk3_step_timed(c, m, st, t, dt, nthreads, cptr, clist, ncol, ktime);
    t += dt;
    t6 = wtime();
    dt1 = or_calc_dt_all(c, m, st->q, 0.0, nthreads);
    if (ktime) ktime[5] += wtime() - t6;
    dt = cfl / dt1;
  }
  free(cptr);
  free(clist);
  if (dt_out) *dt_out = dt;
  return t;
}

double or_run_dt(const or_consts *c, const or_mesh *m, or_state *st, int nsteps, double dt,
                 int nthreads) {
  int *cptr = 0, *clist = 0, ncol = 0;
  if (nthreads > 1) {
    cptr = (int *)malloc(9 * sizeof(int));
    clist = (int *)malloc(((size_t)m->nedges + 1) * sizeof(int));
    ncol = or_colour_edges(m->ncells, m->nedges, m->edge2cells, cptr, clist);
  }
  double t = 0.0;
  for (int i = 0; i < nsteps; i++) {
    or_ssprk3_step_timed(c, m, st, t, dt, nthreads, cptr, clist, ncol, 0);
    t += dt;
  }
  free(cptr);
  free(clist);
  return t;
}

/* ---- edge builder (tools/vtk2CGNS.cpp:112-199 contract) ---- */

typedef struct { int a, b, cell, j; } half_edge;

static int he_cmp(const void *x, const void *y) {
  const half_edge *p = (const half_edge *)x, *q = (const half_edge *)y;
  if (p->a != q->a) return p->a < q->a ? -1 : 1;
  if (p->b != q->b) return p->b < q->b ? -1 : 1;
  if (p->cell != q->cell) return p->cell < q->cell ? -1 : 1;
  return p->j < q->j ? -1 : (p->j > q->j);
}

int or_build_edges(int ncells, const int *cells, int *nedges, int *edge2cells,
                   int *edgeNum, int *edge2nodes, int *nbedges, int *bedge2cells,
                   int *bedgeNum, int *bedge2nodes) {
  size_t nh = (size_t)ncells * 3;
  half_edge *h = (half_edge *)malloc(nh * sizeof(half_edge));
  for (int k = 0; k < ncells; k++)
    for (int j = 0; j < 3; j++) {
      int p1 = cells[3 * k + j], p2 = cells[3 * k + (j + 1) % 3];
      half_edge *e = &h[3 * (size_t)k + j];
      e->a = p1 < p2 ? p1 : p2;
      e->b = p1 < p2 ? p2 : p1;
      e->cell = k;
      e->j = j;
    }
  qsort(h, nh, sizeof(half_edge), he_cmp);
  int ne = 0, nb = 0;
  size_t i = 0;
  while (i < nh) {
    size_t n = 1;
    while (i + n < nh && h[i + n].a == h[i].a && h[i + n].b == h[i].b) n++;
    if (n > 2 || h[i].a == h[i].b) { free(h); return -1; }
    if (n == 2) {
      if (h[i].cell == h[i + 1].cell) { free(h); return -2; }
      edge2cells[2 * ne] = h[i].cell;
      edge2cells[2 * ne + 1] = h[i + 1].cell;
      edgeNum[2 * ne] = h[i].j;
      edgeNum[2 * ne + 1] = h[i + 1].j;
      if (edge2nodes) { edge2nodes[2 * ne] = h[i].a; edge2nodes[2 * ne + 1] = h[i].b; }
      ne++;
    } else {
      bedge2cells[nb] = h[i].cell;
      bedgeNum[nb] = h[i].j;
      if (bedge2nodes) { bedge2nodes[2 * nb] = h[i].a; bedge2nodes[2 * nb + 1] = h[i].b; }
      nb++;
    }
    i += n;
  }
  free(h);
  *nedges = ne;
  *nbedges = nb;
  return 0;
}

/* ---- isentropic vortex (vortex/src/set_ic.h:1-28, corrected) ---- */

void or_vortex_state(double gam, double uinf, double beta, double x0, double y0,
                     double t, double x, double y, double period_x, double *q4) {
  const double pi = 3.14159265358979323846;
  double dx = x - x0 - uinf * t;
  if (period_x > 0.0) dx -= period_x * floor(dx / period_x + 0.5);
  double dy = y - y0;
  double r2 = dx * dx + dy * dy;
  double e1 = exp(1.0 - r2);
  double u = uinf - beta * e1 * (dy / (2 * pi));
  double v = beta * e1 * (dx / (2 * pi));
  double rho = 1.0 - ((gam - 1) / (16 * gam * pi * pi)) * beta * beta * exp(2 * (1.0 - r2));
  rho = pow(rho, 1.0 / (gam - 1));
  double p = pow(rho, gam);
  q4[0] = rho;
  q4[1] = rho * u;
  q4[2] = rho * v;
  q4[3] = (p / (gam - 1)) + 0.5 * rho * (u * u + v * v);
}

void or_vortex_l2(const or_consts *c, int ncells, const double *x, const double *y,
                  const double *J, const double *mass, const double *q, double gam,
                  double uinf, double beta, double x0, double y0, double t,
                  double period_x, double *err4) {
  const int Np = c->Np;
  double acc[4] = {0, 0, 0, 0};
  for (int k = 0; k < ncells; k++) {
    double e[4][MAXNP];
    for (int i = 0; i < Np; i++) {
      double ex[4];
      or_vortex_state(gam, uinf, beta, x0, y0, t, x[Np * k + i], y[Np * k + i], period_x, ex);
      for (int v = 0; v < 4; v++) e[v][i] = q[4 * (Np * k + i) + v] - ex[v];
    }
    for (int v = 0; v < 4; v++) {
      double s = 0.0;
      for (int i = 0; i < Np; i++) {
        double row = 0.0;
        for (int j = 0; j < Np; j++) row += mass[i * Np + j] * e[v][j];
        s += e[v][i] * row;
      }
      acc[v] += J[k] * s;
    }
  }
  for (int v = 0; v < 4; v++) err4[v] = sqrt(acc[v]);
}

void or_wall_forces(const or_consts *c, int nwall, const int *wcell, const int *wface,
                    const double *q, const double *nx, const double *ny,
                    const double *fscale, const double *J, const double *facew,
                    double *fxy) {
  const int Np = c->Np, Nfp = c->Nfp, NF = 3 * c->Nfp;
  double fx = 0.0, fy = 0.0;
  for (int b = 0; b < nwall; b++) {
    long k = wcell[b];
    int f = wface[b];
    double bx = 0.0, by = 0.0;
    for (int i = 0; i < Nfp; i++) {
      int m = f * Nfp + i;
      const double *qq = q + (k * Np + c->FMASK[m]) * 4;
      double rho = qq[0], u = qq[1] / rho, v = qq[2] / rho;
      double p = (c->gam - 1.0) * (qq[3] - 0.5 * rho * (u * u + v * v));
      double sJ = fscale[k * NF + m] * J[k];
      bx += facew[i] * sJ * p * nx[k * NF + m];
      by += facew[i] * sJ * p * ny[k * NF + m];
    }
    fx += bx;
    fy += by;
  }
  fxy[0] = fx;
  fxy[1] = fy;
}

void or_wall_cp(const or_consts *c, int nwall, const int *wcell, const int *wface,
                const double *q, double qdyn, double *cp) {
  const int Np = c->Np, Nfp = c->Nfp;
  for (int b = 0; b < nwall; b++) {
    long k = wcell[b];
    for (int i = 0; i < Nfp; i++) {
      const double *qq = q + (k * Np + c->FMASK[wface[b] * Nfp + i]) * 4;
      double rho = qq[0], u = qq[1] / rho, v = qq[2] / rho;
      double p = (c->gam - 1.0) * (qq[3] - 0.5 * rho * (u * u + v * v));
      cp[b * Nfp + i] = (p - c->bc_p) / qdyn;
    }
  }
}
