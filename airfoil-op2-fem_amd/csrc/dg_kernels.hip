// HIP kernels for gfx950 (MI355X): nodal DG Euler right-hand side, SSPRK3
// stage updates, CFL reduction and set-up loops.  FP64 throughout.
//
// Work decomposition of the stage kernel (rhs_kernel):
//   * a workgroup owns a tile of TE consecutive cells; the tile's state
//     records ([TE][Np][4] doubles, contiguous in the op_dat layout) are
//     staged in LDS with coalesced 16-byte loads;
//   * face phase: one lane per (cell, face node) gathers the neighbour trace
//     straight from the neighbour's state record through the face map
//     (gather-only, owner computes: no atomics, no colouring), or builds the
//     boundary state, and evaluates the Roe flux into LDS;
//   * volume phase: one lane per (cell, node) evaluates the Euler fluxes F, G
//     into LDS;
//   * combine phase: one lane per (cell, node) applies the weak derivative
//     rows of Drw, Dsw, the metric terms and the LIFT row, then runs the
//     epilogue (write qRHS, or an SSPRK3 stage update, or the final update
//     fused with the CFL max-reduction).
// The grid is persistent (a few workgroups per CU) and walks tiles in
// XCD-contiguous chunks so that a cell and its neighbours are usually
// processed by workgroups of the same XCD, whose L2 then holds the shared
// face traces.
#include <hip/hip_runtime.h>

#include "dg_device.h"
#include "dg_internal.h"

namespace dg {

// ------------------------------------------------------------------------
// Stage kernel
// ------------------------------------------------------------------------

template <int N, int TE, int NT, bool GATHER, int EPI>
__global__ __launch_bounds__(NT) void rhs_kernel(StageArgs a, const double *__restrict__ ops,
                                                 const int *__restrict__ fmask, Phys ph,
                                                 unsigned long long facenodes) {
  constexpr int Np = Ord<N>::Np, Nfp = Ord<N>::Nfp, NF = Ord<N>::NF;
  constexpr int SW = Np * 4;          // LDS stride of a state record
  constexpr int SFG = Np * 8 + 2;     // padded so cells fall in different banks
  constexpr int SFL = NF * 4 + 2;
  constexpr int NOPS = 2 * Np * Np + Np * NF;
  __shared__ double sW[TE * SW];
  __shared__ double sFG[TE * SFG];
  __shared__ double sFl[TE * SFL];
  __shared__ double sOp[NOPS];
  __shared__ int sFm[NF];
  __shared__ unsigned long long sRed[NT / 64];

  const int tid = threadIdx.x;
  for (int i = tid; i < NOPS; i += NT) sOp[i] = ops[i];
  if (tid < NF) sFm[tid] = fmask[tid];
  const double *sDrw = sOp, *sDsw = sOp + Np * Np, *sLIFT = sOp + 2 * Np * Np;

  double dt = 0.0;
  if (EPI != EPI_RHS) dt = a.dt_fixed ? *a.dt_fixed : a.cfl / __longlong_as_double((long long)*a.dt1_cur);
  if (EPI == EPI_STAGE0 && blockIdx.x == 0 && tid == 0) *a.dt1_next = 0ull;
  unsigned long long cfl_bits = 0ull;
  bool bad = false;
  const double n1sq = (double)((N + 1) * (N + 1));

  const int ntiles = (a.ncells + TE - 1) / TE;
  for (int k = 0;; k++) {
    const int tile = xcd_tile(k, ntiles);
    if (tile >= ntiles) break;
    const int e0 = tile * TE;
    const int ne = min(TE, a.ncells - e0);
    __syncthreads();  // LDS reuse across tiles (and sOp/sFm on the first pass)
    {
      const double2 *src = reinterpret_cast<const double2 *>(a.w + (size_t)e0 * SW);
      double2 *dst = reinterpret_cast<double2 *>(sW);
      for (int i = tid; i < ne * SW / 2; i += NT) dst[i] = src[i];
    }
    __syncthreads();

    // face phase: Roe flux per (cell, face node)
    for (int it = tid; it < ne * NF; it += NT) {
      const int e = it / NF, m = it - e * NF;
      const int f = m / Nfp, i = m - f * Nfp;
      const size_t gm = (size_t)(e0 + e) * NF + m;
      const double nxv = a.nx[gm], nyv = a.ny[gm], fs = a.fscale[gm];
      double mQ[4], pQ[4];
      const double *ms = &sW[e * SW + sFm[m] * 4];
      mQ[0] = ms[0]; mQ[1] = ms[1]; mQ[2] = ms[2]; mQ[3] = ms[3];
      if (GATHER) {
        const int code = a.fmap[(size_t)(e0 + e) * 3 + f];
        if (code >= 0) {
          const int nb = code >> 2, nf = code & 3;
          const int node = sFm[nf * Nfp + (Nfp - 1 - i)];
          const double2 *src = reinterpret_cast<const double2 *>(a.w + ((size_t)nb * Np + node) * 4);
          double2 p01 = src[0], p23 = src[1];
          pQ[0] = p01.x; pQ[1] = p01.y; pQ[2] = p23.x; pQ[3] = p23.y;
        } else {
          boundary_state(ph, -code - 1, mQ, nxv, nyv, pQ);
        }
      } else {
        const double2 *src = reinterpret_cast<const double2 *>(a.ext + gm * 4);
        double2 p01 = src[0], p23 = src[1];
        pQ[0] = p01.x; pQ[1] = p01.y; pQ[2] = p23.x; pQ[3] = p23.y;
      }
      double fl[4];
      roe_node(ph.gam, mQ, pQ, nxv, nyv, fs, fl);
      double *d = &sFl[e * SFL + m * 4];
      d[0] = fl[0]; d[1] = fl[1]; d[2] = fl[2]; d[3] = fl[3];
    }
    // volume phase: F, G per (cell, node)
    for (int it = tid; it < ne * Np; it += NT) {
      const int e = it / Np, i = it - e * Np;
      double qv[4], F[4], G[4];
      const double *s = &sW[e * SW + i * 4];
      qv[0] = s[0]; qv[1] = s[1]; qv[2] = s[2]; qv[3] = s[3];
      euler_flux(ph.gam, qv, F, G);
      double *d = &sFG[e * SFG + i * 8];
      d[0] = F[0]; d[1] = F[1]; d[2] = F[2]; d[3] = F[3];
      d[4] = G[0]; d[5] = G[1]; d[6] = G[2]; d[7] = G[3];
    }
    __syncthreads();
    if (!GATHER) {  // drop-in euler_rhs zeroes exteriorQ on exit (OP_RW)
      double2 *z = reinterpret_cast<double2 *>(a.ext + (size_t)e0 * NF * 4);
      for (int i = tid; i < ne * NF * 2; i += NT) z[i] = make_double2(0.0, 0.0);
    }

    // combine phase
    for (int it = tid; it < ne * Np; it += NT) {
      const int e = it / Np, j = it - e * Np;
      const size_t gn = (size_t)(e0 + e) * Np + j;
      double dFdr[4] = {0, 0, 0, 0}, dFds[4] = {0, 0, 0, 0};
      double dGdr[4] = {0, 0, 0, 0}, dGds[4] = {0, 0, 0, 0};
      const double *fg = &sFG[e * SFG];
#pragma unroll
      for (int i = 0; i < Np; i++) {
        const double ar = sDrw[j * Np + i], as = sDsw[j * Np + i];
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const double Fv = fg[i * 8 + v], Gv = fg[i * 8 + 4 + v];
          dFdr[v] += ar * Fv;
          dFds[v] += as * Fv;
          dGdr[v] += ar * Gv;
          dGds[v] += as * Gv;
        }
      }
      const double rxv = a.rx[gn], sxv = a.sx[gn], ryv = a.ry[gn], syv = a.sy[gn];
      double res[4], lift[4] = {0, 0, 0, 0};
      const double *fl = &sFl[e * SFL];
#pragma unroll
      for (int m = 0; m < NF; m++) {
        const double l = sLIFT[j * NF + m];
#pragma unroll
        for (int v = 0; v < 4; v++) lift[v] += l * fl[m * 4 + v];
      }
#pragma unroll
      for (int v = 0; v < 4; v++)
        res[v] = ((rxv * dFdr[v] + sxv * dFds[v]) + (ryv * dGdr[v] + syv * dGds[v])) - lift[v];

      double o[4];
      const double *wv = &sW[e * SW + j * 4];
      if (EPI == EPI_RHS) {
#pragma unroll
        for (int v = 0; v < 4; v++) o[v] = res[v];
      } else if (EPI == EPI_STAGE0) {
#pragma unroll
        for (int v = 0; v < 4; v++) o[v] = wv[v] + dt * res[v];
      } else {
        const double2 *qs = reinterpret_cast<const double2 *>(a.q + gn * 4);
        double2 q01 = qs[0], q23 = qs[1];
        double qv[4] = {q01.x, q01.y, q23.x, q23.y};
        if (EPI == EPI_STAGE1) {
#pragma unroll
          for (int v = 0; v < 4; v++) o[v] = 0.75 * qv[v] + 0.25 * (wv[v] + dt * res[v]);
        } else {
#pragma unroll
          for (int v = 0; v < 4; v++) o[v] = (qv[v] + 2.0 * (wv[v] + dt * res[v])) * (1.0 / 3.0);
        }
      }
      double2 *dst = reinterpret_cast<double2 *>(a.out + gn * 4);
      dst[0] = make_double2(o[0], o[1]);
      dst[1] = make_double2(o[2], o[3]);
      if (EPI == EPI_STAGE2) {
        if ((facenodes >> j) & 1ull) {
          double c = cfl_node(ph.gam, n1sq, o, a.fscale[(size_t)(e0 + e) * NF + j]) + 0.0;
          unsigned long long b = (unsigned long long)__double_as_longlong(c);
          cfl_bits = max_bits(cfl_bits, b);
        }
        bad |= !(isfinite(o[0]) && isfinite(o[1]) && isfinite(o[2]) && isfinite(o[3])) ||
               !(o[0] > 0.0);
      }
    }
  }
  if (EPI == EPI_STAGE2) {
    unsigned long long r = block_max_bits<NT>(cfl_bits, sRed);
    if (tid == 0) atomic_max_if_larger(a.dt1_next, r);
    if (bad) atomicOr(a.flag, 1u);
    if (blockIdx.x == 0 && tid == 0) *a.t_dev += dt;
  }
}

template <int N>
struct TileCfg {
  static constexpr int TE = 32;
  static constexpr int NT0 = Ord<N>::NF * TE;
  static constexpr int NT = NT0 <= 256 ? 256 : (NT0 <= 384 ? 384 : (NT0 <= 512 ? 512 : 1024));
};

template <int N, bool GATHER, int EPI>
static hipError_t launch_rhs_t(const Launch &L, const StageArgs &a) {
  constexpr int TE = TileCfg<N>::TE, NT = TileCfg<N>::NT;
  const int ntiles = (a.ncells + TE - 1) / TE;
  int grid = L.cus > 0 ? 2 * L.cus : 512;
  if (L.max_blocks > 0 && grid > L.max_blocks) grid = L.max_blocks;
  if (grid > ntiles) grid = ntiles;
  if (grid >= 8) grid -= grid % 8;
  if (grid < 1) return hipSuccess;
  hipLaunchKernelGGL((rhs_kernel<N, TE, NT, GATHER, EPI>), dim3(grid), dim3(NT), 0, L.stream, a,
                     L.ops, L.fmask, L.ph, (unsigned long long)L.facenodes);
  return hipGetLastError();
}

template <int N>
static hipError_t launch_rhs_n(const Launch &L, const StageArgs &a, bool gather, int epi) {
  if (!gather) return launch_rhs_t<N, false, EPI_RHS>(L, a);
  switch (epi) {
    case EPI_RHS: return launch_rhs_t<N, true, EPI_RHS>(L, a);
    case EPI_STAGE0: return launch_rhs_t<N, true, EPI_STAGE0>(L, a);
    case EPI_STAGE1: return launch_rhs_t<N, true, EPI_STAGE1>(L, a);
    case EPI_STAGE2: return launch_rhs_t<N, true, EPI_STAGE2>(L, a);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_rhs(const Launch &L, const StageArgs &a, bool gather, int epi) {
  if (!gather && epi != EPI_RHS) return hipErrorInvalidValue;
  if (L.variant == 0) return launch_rhs_mfma(L, a, gather, epi);
  switch (L.N) {
    case 1: return launch_rhs_n<1>(L, a, gather, epi);
    case 2: return launch_rhs_n<2>(L, a, gather, epi);
    case 3: return launch_rhs_n<3>(L, a, gather, epi);
    case 4: return launch_rhs_n<4>(L, a, gather, epi);
  }
  return hipErrorInvalidValue;
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
    const double *mq = w + ((size_t)c * Np + fmask[m]) * 4;
    double mQ[4] = {mq[0], mq[1], mq[2], mq[3]}, qb[4];
    boundary_state(ph, btype[b], mQ, nx[c * NF + m], ny[c * NF + m], qb);
    double *d = ext + (c * NF + m) * 4;
    d[0] += qb[0]; d[1] += qb[1]; d[2] += qb[2]; d[3] += qb[3];
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

// Isentropic vortex (vortex/src/set_ic.h:1-28, with free-stream speed uinf).
__global__ void set_vortex_kernel(size_t nnodes, double gam, const double *__restrict__ x,
                                  const double *__restrict__ y, double uinf, double beta,
                                  double x0, double y0, double period_x, double *q, double *w) {
  const double pi = 3.14159265358979323846;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nnodes;
       i += (size_t)gridDim.x * blockDim.x) {
    double dx = x[i] - x0;
    if (period_x > 0.0) dx -= period_x * floor(dx / period_x + 0.5);
    double dy = y[i] - y0;
    double r2 = dx * dx + dy * dy;
    double e1 = exp(1.0 - r2);
    double u = uinf - beta * e1 * (dy / (2 * pi));
    double v = beta * e1 * (dx / (2 * pi));
    double rho = 1.0 - ((gam - 1) / (16 * gam * pi * pi)) * beta * beta * exp(2 * (1.0 - r2));
    rho = pow(rho, 1.0 / (gam - 1));
    double p = pow(rho, gam);
    double qv[4] = {rho, rho * u, rho * v, (p / (gam - 1)) + 0.5 * rho * (u * u + v * v)};
    for (int k = 0; k < 4; k++) { q[4 * i + k] = qv[k]; w[4 * i + k] = qv[k]; }
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

hipError_t launch_set_vortex(const Launch &L, int ncells, const double *x, const double *y,
                             double uinf, double beta, double x0, double y0, double period_x,
                             double *q, double *w) {
  size_t n = (size_t)ncells * L.Np;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(set_vortex_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, L.stream, n,
                     L.ph.gam, x, y, uinf, beta, x0, y0, period_x, q, w);
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
