// HIP kernels for gfx950 (MI355X): the drop-in per-set loops of the DG
// Euler path (face traces, boundary states, SSPRK3 stage updates, CFL
// reduction), the set-up loops (init_grid, initial states) and the wall
// forces.  FP64 throughout.  The fused stage kernel (Roe fluxes, volume and
// LIFT products on the matrix cores, stage epilogues) is in dg_stage_kernel.h.
#include <hip/hip_runtime.h>

#include "dg_device.h"
#include "dg_internal.h"

namespace dg {

// ------------------------------------------------------------------------
// Stage kernel: the matrix-core kernel of dg_stage_kernel.h
// ------------------------------------------------------------------------

hipError_t launch_rhs(const Launch &L, const StageArgs &a, bool gather, int epi) {
  if (!gather && epi != EPI_RHS) return hipErrorInvalidValue;
  return launch_rhs_mfma(L, a, gather, epi);
}

// ------------------------------------------------------------------------
// Face-trace loops of the drop-in API
// ------------------------------------------------------------------------

// src/get_neighbour_q.h: exteriorQ[c][f][i] += workingQ[nb][FMASK[nf][Nfp-1-i]].
// One lane per (cell, face node), gather-only.
__global__ void get_neighbour_q_kernel(int ncells, int Np, int Nfp, const int *__restrict__ fmap,
                                       const int *__restrict__ fmask, const double *__restrict__ w,
                                       double *__restrict__ ext) {
  const int NF = 3 * Nfp;
  const size_t n = (size_t)ncells * NF;
  for (size_t it = blockIdx.x * (size_t)blockDim.x + threadIdx.x; it < n;
       it += (size_t)gridDim.x * blockDim.x) {
    const size_t c = it / NF;
    const int m = (int)(it - c * NF);
    const int f = m / Nfp, i = m - f * Nfp;
    const int code = fmap[c * 3 + f];
    if (code < 0) continue;
    const int nb = code >> 2, nf = code & 3;
    const int node = fmask[nf * Nfp + (Nfp - 1 - i)];
    const double2 *src = reinterpret_cast<const double2 *>(w + ((size_t)nb * Np + node) * 4);
    double2 *dst = reinterpret_cast<double2 *>(ext + it * 4);
    double2 a0 = src[0], a1 = src[1], d0 = dst[0], d1 = dst[1];
    dst[0] = make_double2(d0.x + a0.x, d0.y + a0.y);
    dst[1] = make_double2(d1.x + a1.x, d1.y + a1.y);
  }
}

// src/get_bedge_q.h: one lane per (boundary edge, face node).
__global__ void get_bedge_q_kernel(int nbedges, int Np, int Nfp, const int *__restrict__ b2c,
                                   const int *__restrict__ bnum, const int *__restrict__ btype,
                                   const int *__restrict__ fmask, Phys ph,
                                   const double *__restrict__ nx, const double *__restrict__ ny,
                                   const double *__restrict__ w, double *__restrict__ ext) {
  const int NF = 3 * Nfp;
  const size_t n = (size_t)nbedges * Nfp;
  for (size_t it = blockIdx.x * (size_t)blockDim.x + threadIdx.x; it < n;
       it += (size_t)gridDim.x * blockDim.x) {
    const size_t b = it / Nfp;
    const int i = (int)(it - b * Nfp);
    const size_t c = b2c[b];
    const int f = bnum[b];
    const int m = f * Nfp + i;
    if (btype[b] == kBcExact) continue;  // exact states: exact_bc_kernel
    const double *mq = w + ((size_t)c * Np + fmask[m]) * 4;
    double mQ[4] = {mq[0], mq[1], mq[2], mq[3]}, qb[4];
    boundary_state(ph, btype[b], mQ, nx[c * NF + m], ny[c * NF + m], qb);
    double *d = ext + (c * NF + m) * 4;
    d[0] += qb[0]; d[1] += qb[1]; d[2] += qb[2]; d[3] += qb[3];
  }
}

// vortex/src/get_bedge_q.h:1-40: one lane per (boundary edge, face node);
// the exact vortex state at the node's (x, y) and time t, added to or stored
// in the face slot.  The time of a fused stage is read on the device.
__global__ void exact_bc_kernel(int nb, int Np, int Nfp, const int *__restrict__ b2c,
                                const int *__restrict__ bnum, const int *__restrict__ btype,
                                const int *__restrict__ fmask, double gam, const double *__restrict__ x,
                                const double *__restrict__ y, Vortex vx, double t_host,
                                const double *__restrict__ t_dev, const unsigned long long *__restrict__ dt1,
                                double cfl, const double *__restrict__ dt_fixed, double frac, int accumulate,
                                double *__restrict__ ext) {
  const int NF = 3 * Nfp;
  double t = t_host;
  if (t_dev) t = t_dev[0] + frac * (dt_fixed ? dt_fixed[0] : cfl / __longlong_as_double((long long)dt1[0]));
  const size_t n = (size_t)nb * Nfp;
  for (size_t it = blockIdx.x * (size_t)blockDim.x + threadIdx.x; it < n;
       it += (size_t)gridDim.x * blockDim.x) {
    const size_t b = it / Nfp;
    if (btype && btype[b] != kBcExact) continue;
    const int i = (int)(it - b * Nfp);
    const size_t c = b2c[b];
    const int m = bnum[b] * Nfp + i;
    const size_t node = c * Np + fmask[m];
    double qb[4];
    vortex_state(gam, vx, t, x[node], y[node], qb);
    double *d = ext + (c * NF + m) * 4;
    if (accumulate) {
      d[0] += qb[0]; d[1] += qb[1]; d[2] += qb[2]; d[3] += qb[3];
    } else {
      d[0] = qb[0]; d[1] = qb[1]; d[2] = qb[2]; d[3] = qb[3];
    }
  }
}

// src/set_workingQ.h
__global__ void set_workingQ_kernel(size_t n, double dt, int stage, const double *__restrict__ q,
                                    const double *__restrict__ k1, const double *__restrict__ k2,
                                    double *__restrict__ w) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    if (stage == 0)
      w[i] = q[i] + dt * k1[i];
    else
      w[i] = q[i] + dt * (k1[i] / 4.0 + k2[i] / 4.0);
  }
}

// src/update_Q.h
__global__ void update_Q_kernel(size_t n, double dt, double *__restrict__ q,
                                const double *__restrict__ k1, const double *__restrict__ k2,
                                const double *__restrict__ k3, double *__restrict__ w) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double v = q[i] + dt * (k1[i] / 6.0 + k2[i] / 6.0 + 2.0 * k3[i] / 3.0);
    q[i] = v;
    w[i] = v;
  }
}

// src/calc_dt.h: per-cell max over face slots, then a device max.
__global__ __launch_bounds__(256) void calc_dt_kernel(int ncells, int N, int Np, int Nfp,
                                                      const int *__restrict__ fmask, double gam,
                                                      const double *__restrict__ q,
                                                      const double *__restrict__ fscale,
                                                      unsigned long long *out) {
  __shared__ unsigned long long red[4];
  const int NF = 3 * Nfp;
  const double n1sq = (double)((N + 1) * (N + 1));
  unsigned long long best = 0ull;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < (size_t)ncells;
       c += (size_t)gridDim.x * blockDim.x) {
    for (int i = 0; i < NF; i++) {
      const int f = fmask[i];
      const double *qq = q + (c * Np + f) * 4;
      double qv[4] = {qq[0], qq[1], qq[2], qq[3]};
      double v = cfl_node(gam, n1sq, qv, fscale[c * NF + f]) + 0.0;
      best = max_bits(best, (unsigned long long)__double_as_longlong(v));
    }
  }
  unsigned long long r = block_max_bits<256>(best, red);
  if (threadIdx.x == 0) atomic_max_if_larger(out, r);
}

// src/init_grid.h: one lane per cell.  Every product and sum is rounded on
// its own (no multiply-add contraction) in the order of the reference's BLAS
// calls (dcopy/daxpy/dscal for the node coordinates, a dot product per row
// for dgemv), and division and square root are the correctly rounded IEEE
// operations, so the factors are the same bits as a sequential host build
// with contraction off.  The metrics of a cell whose coordinates are many
// cell sizes from the origin are differences of nearly equal numbers, and a
// contracted multiply-add moves them by ~1e-15 x (coordinate / cell size);
// over many steps that difference would show in the solution.
__global__ void init_grid_kernel(int ncells, int Np, int Nfp, const double *__restrict__ ops,
                                 const int *__restrict__ fmask, const double *__restrict__ cell_xy,
                                 double *nodeX, double *nodeY, double *x, double *y, double *xr,
                                 double *yr, double *xs, double *ys, double *rx, double *ry,
                                 double *sx, double *sy, double *nx, double *ny, double *fscale,
                                 double *Jc) {
#pragma clang fp contract(off)
  const int NF = 3 * Nfp;
  const double *Dr = ops + 2 * Np * Np + Np * NF;
  const double *Ds = Dr + Np * Np;
  const double *r = Ds + Np * Np;
  const double *s = r + Np;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < (size_t)ncells;
       c += (size_t)gridDim.x * blockDim.x) {
    const double *v = cell_xy + 6 * c;
    nodeX[3 * c] = v[0]; nodeX[3 * c + 1] = v[2]; nodeX[3 * c + 2] = v[4];
    nodeY[3 * c] = v[1]; nodeY[3 * c + 1] = v[3]; nodeY[3 * c + 2] = v[5];
    double *X = x + c * Np, *Y = y + c * Np;
    for (int d = 0; d < 2; d++) {
      double *o = d == 0 ? X : Y;
      const double a1 = 0.5 * v[2 + d], a2 = 0.5 * v[4 + d], a0 = -0.5 * v[d];
      for (int i = 0; i < Np; i++) {
        double t = a1 * (1.0 + r[i]);
        t = t + a2 * (1.0 + s[i]);
        o[i] = t + a0 * (s[i] + r[i]);
      }
    }
    double *XR = xr + c * Np, *XS = xs + c * Np, *YR = yr + c * Np, *YS = ys + c * Np;
    for (int j = 0; j < Np; j++) {
      double a = 0, b = 0, cc = 0, dd = 0;
      for (int i = 0; i < Np; i++) {
        a += Dr[j * Np + i] * X[i];
        b += Ds[j * Np + i] * X[i];
        cc += Dr[j * Np + i] * Y[i];
        dd += Ds[j * Np + i] * Y[i];
      }
      XR[j] = a; XS[j] = b; YR[j] = cc; YS[j] = dd;
    }
    double J0 = 0.0;
    for (int i = 0; i < Np; i++) {
      double J = -XS[i] * YR[i] + XR[i] * YS[i];
      if (i == 0) J0 = J;
      rx[c * Np + i] = YS[i] / J;
      sx[c * Np + i] = -YR[i] / J;
      ry[c * Np + i] = -XS[i] / J;
      sy[c * Np + i] = XR[i] / J;
    }
    if (Jc) Jc[c] = J0;
    for (int m = 0; m < NF; m++) {
      const int f = fmask[m];
      const int face = m / Nfp;
      double nxv, nyv;
      if (face == 0) { nxv = YR[f]; nyv = -XR[f]; }
      else if (face == 1) { nxv = YS[f] - YR[f]; nyv = XR[f] - XS[f]; }
      else { nxv = -YS[f]; nyv = XS[f]; }
      double sJ = sqrt(nxv * nxv + nyv * nyv);
      double J = -XS[f] * YR[f] + XR[f] * YS[f];
      nx[c * NF + m] = nxv / sJ;
      ny[c * NF + m] = nyv / sJ;
      fscale[c * NF + m] = sJ / J;
    }
  }
}

// src/set_ic.h
__global__ void set_ic_kernel(size_t nnodes, Phys ph, double *q, double *w) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nnodes;
       i += (size_t)gridDim.x * blockDim.x) {
    double v[4] = {ph.bc_r, ph.bc_r * ph.bc_u, ph.bc_r * ph.bc_v, ph.bc_e};
    for (int k = 0; k < 4; k++) { q[4 * i + k] = v[k]; w[4 * i + k] = v[k]; }
  }
}

// Isentropic vortex at time t (vortex/src/set_ic.h:1-28, with free-stream
// speed uinf).
__global__ void set_vortex_kernel(size_t nnodes, double gam, const double *__restrict__ x,
                                  const double *__restrict__ y, double uinf, double beta,
                                  double x0, double y0, double period_x, double t, double *q,
                                  double *w) {
  Vortex vx;
  vx.uinf = uinf; vx.beta = beta; vx.x0 = x0; vx.y0 = y0; vx.period = period_x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nnodes;
       i += (size_t)gridDim.x * blockDim.x) {
    double qv[4];
    vortex_state(gam, vx, t, x[i], y[i], qv);
    for (int k = 0; k < 4; k++) { q[4 * i + k] = qv[k]; w[4 * i + k] = qv[k]; }
  }
}

// Cells per pass of l2_error_kernel: one thread per node, then one per
// (cell, variable), within 256 threads.
__host__ __device__ inline int l2_cells_per_pass(int Np) { return min(256 / Np, 64); }

// Error against the exact vortex (vortex/src/l2_error.h:1-29).  Each
// workgroup takes a contiguous run of cells, l2_cells_per_pass per pass: one
// thread per node forms e = q - exact and writes e^2 to error (when set),
// then one thread per (cell, variable) adds sum_i e_i^2 and, when the metric
// factors are given, J e^T M e with J = xr ys - xs yr at node 0.  The
// workgroup's two sums per variable go to partial[blockIdx][0..3] and
// [4..7]; the grid depends on ncells only, so the summation order is fixed.
__global__ __launch_bounds__(256) void l2_error_kernel(int ncells, int Np, double gam, Vortex vx, double t,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       const double *__restrict__ q,
                                                       const double *__restrict__ xr,
                                                       const double *__restrict__ yr,
                                                       const double *__restrict__ xs,
                                                       const double *__restrict__ ys,
                                                       const double *__restrict__ mass, double *error,
                                                       double *partial) {
  __shared__ double sE[256 * 4];
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const int cb = l2_cells_per_pass(Np);
  const int npass = (ncells + cb - 1) / cb;
  const int per = (npass + gridDim.x - 1) / gridDim.x;
  const int p0 = blockIdx.x * per, p1 = min(npass, p0 + per);
  const bool weighted = xr != nullptr;
  double acc_n = 0.0, acc_m = 0.0;  // thread (cell tid / 4, variable tid % 4)
  for (int p = p0; p < p1; p++) {
    const int c0 = p * cb, nc = min(cb, ncells - c0);
    if (tid < nc * Np) {
      const size_t i = (size_t)c0 * Np + tid;
      double ex[4];
      vortex_state(gam, vx, t, x[i], y[i], ex);
      const double2 qa = *reinterpret_cast<const double2 *>(q + 4 * i);
      const double2 qb = *reinterpret_cast<const double2 *>(q + 4 * i + 2);
      const double e0 = qa.x - ex[0], e1 = qa.y - ex[1], e2 = qb.x - ex[2], e3 = qb.y - ex[3];
      sE[4 * tid] = e0; sE[4 * tid + 1] = e1; sE[4 * tid + 2] = e2; sE[4 * tid + 3] = e3;
      if (error) {
        *reinterpret_cast<double2 *>(error + 4 * i) = make_double2(e0 * e0, e1 * e1);
        *reinterpret_cast<double2 *>(error + 4 * i + 2) = make_double2(e2 * e2, e3 * e3);
      }
    }
    __syncthreads();
    if (tid < 4 * nc) {
      const int cl = tid >> 2, v = tid & 3;
      const double *e = sE + 4 * cl * Np + v;
      double sn = 0.0, sm = 0.0;
      for (int a = 0; a < Np; a++) {
        const double ea = e[4 * a];
        sn += ea * ea;
        if (weighted) {
          double row = 0.0;
          for (int b = 0; b < Np; b++) row += mass[a * Np + b] * e[4 * b];
          sm += ea * row;
        }
      }
      acc_n += sn;
      if (weighted) {
        const size_t n0 = (size_t)(c0 + cl) * Np;
        acc_m += (xr[n0] * ys[n0] - xs[n0] * yr[n0]) * sm;
      }
    }
    __syncthreads();
  }
  // tree over the threads of one variable: every stride is a multiple of 4
  for (int k = 0; k < 2; k++) {
    red[tid] = k == 0 ? acc_n : acc_m;
    __syncthreads();
    for (int w = 128; w >= 4; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    if (tid < 4) partial[8 * blockIdx.x + 4 * k + tid] = red[tid];
    __syncthreads();
  }
}

__global__ void zero_kernel(size_t n, double *p) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

__global__ void copy_kernel(size_t n, const double *__restrict__ s, double *__restrict__ d) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

// Surface pressure force of each wall face: integral of p n over the face
// with the 1D face mass matrix as quadrature (weights facew[i] = row sums,
// face parameter length 2, so sJ = fscale * J is half the face length).
__global__ void wall_forces_kernel(int nwall, int Np, int Nfp, const int *__restrict__ fmask,
                                   double gam, const int *__restrict__ wcell,
                                   const int *__restrict__ wface, const double *__restrict__ q,
                                   const double *__restrict__ nx, const double *__restrict__ ny,
                                   const double *__restrict__ fscale, const double *__restrict__ J,
                                   const double *__restrict__ facew, double pinf, double qdyn,
                                   double *out, double *cp) {
  const int NF = 3 * Nfp;
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nwall; b += gridDim.x * blockDim.x) {
    const size_t c = wcell[b];
    const int f = wface[b];
    double fx = 0.0, fy = 0.0;
    for (int i = 0; i < Nfp; i++) {
      const int m = f * Nfp + i;
      const double *qq = q + (c * Np + fmask[m]) * 4;
      const double rho = qq[0], u = qq[1] / rho, v = qq[2] / rho;
      const double p = (gam - 1.0) * (qq[3] - 0.5 * rho * (u * u + v * v));
      const double sJ = fscale[c * NF + m] * J[c];
      fx += facew[i] * sJ * p * nx[c * NF + m];
      fy += facew[i] * sJ * p * ny[c * NF + m];
      cp[(size_t)b * Nfp + i] = (p - pinf) / qdyn;
    }
    out[2 * (size_t)b] = fx;
    out[2 * (size_t)b + 1] = fy;
  }
}

hipError_t launch_wall_forces(const Launch &L, int nwall, const int *wall_cell, const int *wall_face,
                              const double *q, const double *nx, const double *ny,
                              const double *fscale, const double *J, const double *facew,
                              double pinf, double qdyn, double *out, double *cp) {
  if (nwall == 0) return hipSuccess;
  hipLaunchKernelGGL(wall_forces_kernel, dim3((nwall + 127) / 128), dim3(128), 0, L.stream, nwall,
                     L.Np, L.Nfp, L.fmask, L.ph.gam, wall_cell, wall_face, q, nx, ny, fscale, J,
                     facew, pinf, qdyn, out, cp);
  return hipGetLastError();
}

static int grid_for(size_t n, int bs, int cap) {
  size_t g = (n + bs - 1) / bs;
  if (g > (size_t)cap) g = cap;
  return g < 1 ? 1 : (int)g;
}

hipError_t launch_get_neighbour_q(const Launch &L, int ncells, const int *fmap, const double *w,
                                  double *ext) {
  size_t n = (size_t)ncells * L.NF;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(get_neighbour_q_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, L.stream,
                     ncells, L.Np, L.Nfp, fmap, L.fmask, w, ext);
  return hipGetLastError();
}

hipError_t launch_get_bedge_q(const Launch &L, int nbedges, const int *b2c, const int *bnum,
                              const int *btype, const double *nx, const double *ny,
                              const double *w, double *ext) {
  size_t n = (size_t)nbedges * L.Nfp;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(get_bedge_q_kernel, dim3(grid_for(n, 256, 4096)), dim3(256), 0, L.stream,
                     nbedges, L.Np, L.Nfp, b2c, bnum, btype, L.fmask, L.ph, nx, ny, w, ext);
  return hipGetLastError();
}

hipError_t launch_set_workingQ(const Launch &L, int ncells, double dt, int stage, const double *q,
                               const double *k1, const double *k2, double *w) {
  size_t n = (size_t)ncells * L.Np * 4;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(set_workingQ_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, L.stream, n,
                     dt, stage, q, k1, k2, w);
  return hipGetLastError();
}

hipError_t launch_update_Q(const Launch &L, int ncells, double dt, double *q, const double *k1,
                           const double *k2, const double *k3, double *w) {
  size_t n = (size_t)ncells * L.Np * 4;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(update_Q_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, L.stream, n, dt,
                     q, k1, k2, k3, w);
  return hipGetLastError();
}

hipError_t launch_calc_dt(const Launch &L, int ncells, const double *q, const double *fscale,
                          unsigned long long *d_out) {
  if (ncells == 0) return hipSuccess;
  hipLaunchKernelGGL(calc_dt_kernel, dim3(grid_for(ncells, 256, 2048)), dim3(256), 0, L.stream,
                     ncells, L.N, L.Np, L.Nfp, L.fmask, L.ph.gam, q, fscale, d_out);
  return hipGetLastError();
}

hipError_t launch_init_grid(const Launch &L, int ncells, const double *cell_xy, double *nodeX,
                            double *nodeY, double *x, double *y, double *xr, double *yr,
                            double *xs, double *ys, double *rx, double *ry, double *sx,
                            double *sy, double *nx, double *ny, double *fscale, double *J) {
  if (ncells == 0) return hipSuccess;
  hipLaunchKernelGGL(init_grid_kernel, dim3(grid_for(ncells, 128, 8192)), dim3(128), 0, L.stream,
                     ncells, L.Np, L.Nfp, L.ops, L.fmask, cell_xy, nodeX, nodeY, x, y, xr, yr, xs,
                     ys, rx, ry, sx, sy, nx, ny, fscale, J);
  return hipGetLastError();
}

// Per-cell metric record (rx, ry, sx, sy) of straight-sided cells: the
// factors init_grid wrote at the cell's node 0 (equal at every node up to
// rounding).
__global__ void cell_metrics_kernel(int ncells, int Np, const double *__restrict__ rx,
                                    const double *__restrict__ ry, const double *__restrict__ sx,
                                    const double *__restrict__ sy, double *cmet) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < (size_t)ncells;
       c += (size_t)gridDim.x * blockDim.x) {
    const size_t n = c * Np;
    double2 *o = reinterpret_cast<double2 *>(cmet + 4 * c);
    o[0] = make_double2(rx[n], ry[n]);
    o[1] = make_double2(sx[n], sy[n]);
  }
}

hipError_t launch_cell_metrics(const Launch &L, int ncells, const double *rx, const double *ry,
                               const double *sx, const double *sy, double *cmet) {
  if (ncells == 0) return hipSuccess;
  hipLaunchKernelGGL(cell_metrics_kernel, dim3(grid_for(ncells, 256, 8192)), dim3(256), 0, L.stream,
                     ncells, L.Np, rx, ry, sx, sy, cmet);
  return hipGetLastError();
}

hipError_t launch_set_ic(const Launch &L, int ncells, double *q, double *w) {
  size_t n = (size_t)ncells * L.Np;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(set_ic_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, L.stream, n, L.ph,
                     q, w);
  return hipGetLastError();
}

hipError_t launch_exact_bc(const Launch &L, int nb, const int *b2c, const int *bnum, const int *btype,
                           const double *x, const double *y, const Vortex &v, double t_host,
                           const double *t_dev, const unsigned long long *dt1, double cfl,
                           const double *dt_fixed, double frac, bool accumulate, double *ext) {
  const size_t n = (size_t)nb * L.Nfp;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(exact_bc_kernel, dim3(grid_for(n, 128, 4096)), dim3(128), 0, L.stream, nb, L.Np, L.Nfp,
                     b2c, bnum, btype, L.fmask, L.ph.gam, x, y, v, t_host, t_dev, dt1, cfl, dt_fixed, frac,
                     accumulate ? 1 : 0, ext);
  return hipGetLastError();
}

hipError_t launch_set_vortex(const Launch &L, int ncells, const double *x, const double *y,
                             double uinf, double beta, double x0, double y0, double period_x,
                             double t, double *q, double *w) {
  size_t n = (size_t)ncells * L.Np;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(set_vortex_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, L.stream, n,
                     L.ph.gam, x, y, uinf, beta, x0, y0, period_x, t, q, w);
  return hipGetLastError();
}

int l2_error_blocks(int Np, int ncells) {
  const int cb = l2_cells_per_pass(Np);
  const int npass = (ncells + cb - 1) / cb;
  return npass < kL2Blocks ? npass : kL2Blocks;
}

hipError_t launch_l2_error(const Launch &L, int ncells, const Vortex &v, double t, const double *x,
                           const double *y, const double *q, const double *xr, const double *yr,
                           const double *xs, const double *ys, const double *mass, double *error,
                           double *partial) {
  const int g = l2_error_blocks(L.Np, ncells);
  if (g == 0) return hipSuccess;
  hipLaunchKernelGGL(l2_error_kernel, dim3(g), dim3(256), 0, L.stream, ncells, L.Np, L.ph.gam, v, t, x, y, q,
                     xr, yr, xs, ys, mass, error, partial);
  return hipGetLastError();
}

hipError_t launch_zero(const Launch &L, size_t n, double *p) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, L.stream, n, p);
  return hipGetLastError();
}

hipError_t launch_copy(const Launch &L, size_t n, const double *src, double *dst) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, L.stream, n, src,
                     dst);
  return hipGetLastError();
}

// out[i] = the 4-double node record rec[i] of q (zeros for rec[i] < 0).
__global__ void gather_records_kernel(int n, const long long *__restrict__ rec, const double *__restrict__ q,
                                      double *out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const long long r = rec[i];
    for (int v = 0; v < 4; v++) out[4 * (size_t)i + v] = r >= 0 ? q[4 * r + v] : 0.0;
  }
}

hipError_t launch_gather_records(const Launch &L, int n, const long long *rec, const double *q, double *out) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_records_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, L.stream, n, rec, q,
                     out);
  return hipGetLastError();
}

}  // namespace dg
