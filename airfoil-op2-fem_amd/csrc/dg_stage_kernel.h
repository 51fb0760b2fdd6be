// Stage kernel with the dense contractions on the FP64 matrix cores of gfx950.
//
// Each wave (64 lanes) owns tiles of 16 consecutive cells, split into 4
// groups of 4 cells; a workgroup is 1, 2 or 4 such waves.  For one group the
// volume term is four 16x16 products on v_mfma_f64_16x16x4f64,
//
//     D[cell + 4*var][node j] = sum_k  A[cell + 4*var][k] * B[k][j],
//
// with A = F or G of the group's 4 cells (rows: 4 cells x 4 variables,
// columns: the Np source nodes padded to a multiple of 4) and B = Drw^T or
// Dsw^T (padded to 16 output nodes).  B is the same for every cell, so each
// lane keeps its B entries in registers for the whole kernel; A is staged in
// LDS already in the operand layout (the padding columns are not stored: the
// lanes that would read them supply zeros), so every MFMA reads one
// contiguous 512-byte row.  The FP64 accumulator layout is row-interleaved: lane l holds
// D[l/16 + 4*r][l%16], r = 0..3.  Rows are therefore ordered (cell + 4*var),
// so lane l holds cell l/16 of the group, node l%16, variables r = 0..3: the
// metric terms are then lane-local, the LIFT product (A = Roe fluxes,
// B = -LIFT^T) lands in the same registers, and the epilogue stores each
// lane's 4 variables of one node as one 32-byte record (the op_dat layout),
// so stores are coalesced.
//
// The Euler fluxes (one lane per (cell, node)) and the Roe fluxes (one lane
// per (cell, face node), neighbour traces gathered from the neighbour's
// state record through the face map) run on the vector ALUs; at N=3 a
// 16-cell tile is 192 Roe items (three full passes of the wave) and 160
// Euler items.  A tile runs: Roe fluxes -> LIFT products (the accumulators
// start from the surface term) -> Euler fluxes -> volume products ->
// epilogue.  The kernel is bound by memory latency more than by bandwidth or
// arithmetic, so the order of loads is chosen to shorten each tile's chain
// of dependent memory round trips within the register budget of three waves
// per SIMD: the next tile's face-map codes are loaded one tile ahead (during
// the volume products), a Roe item's own trace, exterior trace and geometry
// go out together (a boundary slot reads its own record as a stand-in
// exterior), one item ahead of its flux, and each group's
// epilogue loads (metric factors, w, q) go out before its products (the
// stage-1 kernel's metric factors of all four groups before the first).
// Running the surface term first keeps the accumulators out of the Roe
// phase, the phase with the most live values.
//
// Compact geometry (CG): on straight-sided cells the metric factors are the
// same at every node, so the kernel reads one (rx, ry, sx, sy) record per
// cell, forms the contravariant fluxes rx F + ry G and sx F + sy G at each
// node, and the volume term is Drw (rx F + ry G) + Dsw (sx F + sy G): two
// products per group instead of four, one accumulator, and 32 bytes of
// geometry per cell instead of the 608 of the per-node factors, normals and
// fscale (which it derives per face, cell_face_geometry).
#pragma once
#include <hip/hip_runtime.h>

#include "dg_device.h"
#include "dg_internal.h"
#include "dg_mfma.h"

namespace dg {

// LDS per wave (doubles): the F and G tiles [4 groups][Np columns][16 rows]
// (the face-flux tile [4][4 KF columns][16 rows] and the copy of the tile's
// state records reuse them).  A column holds its 16 rows in a slot of kColStride doubles and an F or G
// group starts kGroupSkew doubles past the end of the previous one.  The
// flux writes put the nodes of one cell (same row, consecutive columns) on
// neighbouring lanes; with 16-double columns those all fall in one bank pair
// of the 32 four-byte banks (a 10- or 12-way conflict on every store).  An
// 18-double slot moves consecutive columns two banks apart and the skew
// separates cells of different groups, while the operand reads (16 rows of
// one column per 16 lanes) stay within one slot.  The face-flux tile uses
// 17-double slots and no skew (consecutive columns one bank pair apart, at
// most 2-way conflicts between neighbouring cells), so that it and the copy
// of the tile's state records fit beside each other within the F and G
// tiles' footprint (at N=3 a wave needs 11,648 bytes: twelve one-wave
// workgroups per compute unit, the register limit, take 137 KiB of its 160).
constexpr int kColStride = 18, kGroupSkew = 2, kFaceColStride = 17;
template <int N>
constexpr int mfma_fg_group() {  // doubles per group in the F and G tiles
  return kColStride * Ord<N>::Np + kGroupSkew;
}
template <int N>
constexpr int mfma_fl_group() {  // doubles per group in the face-flux tile
  return kFaceColStride * 4 * ((Ord<N>::NF + 3) / 4);
}
// The tile's state records (16 cells x Np x 4) are copied into LDS by the
// direct-to-LDS loads issued at the start of the tile, past the face-flux
// tile (which the face phase writes meanwhile); the Euler-flux phase reads
// them there, then writes the F and G tiles over the same memory.
template <int N>
constexpr int mfma_wtile_off() {
  return 4 * mfma_fl_group<N>();
}
template <int N>
constexpr int mfma_wtile_doubles() {  // whole KiB per copy instruction (copy_to_lds)
  return 128 * ((kMfmaTile * Ord<N>::Np * 2 + 63) / 64);
}
template <int N>
constexpr int mfma_tile_doubles() {
  constexpr int fg = 2 * 4 * mfma_fg_group<N>(), fw = mfma_wtile_off<N>() + mfma_wtile_doubles<N>();
  return fg > fw ? fg : fw;
}
// The stage-2 kernel's CFL quantity reads fscale in its epilogue (from the
// fscale array, or formed from the cell's metric record on the per-cell
// geometry) rather than from a table its face phase would leave in LDS: with
// that table (12.6 KB per wave) one-wave workgroups ran the stage-2 kernel
// ~40% slower than at 11.4 KB (591 vs 429 us on the 724x724 vortex), as if
// fewer of them were resident than the occupancy query reports.
template <int N, int EPI, bool CG>
constexpr int mfma_lds_doubles() {
  return mfma_tile_doubles<N>();
}

// A workgroup is 1, 2 or 4 waves (16, 32 or 64 cells per workgroup visit);
// each wave owns one 16-cell tile of the workgroup's block of tiles and its
// own LDS tiles, so the waves of a workgroup never wait for each other (a
// phase waits only for its own wave's LDS accesses, lds_wait); the walk
// over blocks of tiles is uniform across the workgroup, and waves whose
// tile lies past the end only skip the work.
template <int N, bool GATHER, int EPI, bool CG>
__global__ __launch_bounds__(256, 3) void rhs_mfma_kernel(StageArgs a, const double *__restrict__ ops,
                                                       unsigned long long fmask4, Phys ph0,
                                                       unsigned long long facenodes, int walk) {
  constexpr int Np = Ord<N>::Np, Nfp = Ord<N>::Nfp, NF = Ord<N>::NF;
  constexpr int KP = (Np + 3) / 4;  // K-steps over source nodes
  constexpr int KF = (NF + 3) / 4;  // K-steps over face nodes
  constexpr int TE = kMfmaTile, G = 4;
  constexpr int NIT = (TE * Np + 63) / 64;  // Euler-flux items per lane
  constexpr int NRT = (TE * NF + 63) / 64;  // Roe-flux items per lane
  constexpr int GS = mfma_fg_group<N>();    // doubles per group in the F and G tiles
  constexpr int GL = mfma_fl_group<N>();    // doubles per group in the face-flux tile
  constexpr int CS = kColStride, CSL = kFaceColStride;
  constexpr bool kItemsInLoop = EPI == EPI_STAGE1 || EPI == EPI_STAGE2 || N >= 4;  // see lane_in_loop
  static_assert(Np <= 16, "one 16-row output tile per cell group");
  // per wave: F tiles [G][Np][CS] then G tiles [G][Np][CS] (column k of a
  // group holds its rows cell + 4*var; the padding columns k >= Np of the
  // last K-step are not stored, the lanes supply them as zeros); the
  // face-flux tiles [G][4 KF][CS] use the same memory before the Euler fluxes.
  // No static LDS: a one-wave workgroup's dynamic share is sized so that 12
  // of them fit a compute unit's 160 KiB in 512-byte allocation granules;
  // FMASK comes packed in a kernel argument instead.
  extern __shared__ double lds_dyn[];
  auto fmask_of = [&](int m) -> int { return (int)((fmask4 >> (4 * m)) & 15ull); };

  const int lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // uniform in the wave
  double *sA = lds_dyn + wave * mfma_lds_doubles<N, EPI, CG>();
  double *sW = sA + mfma_wtile_off<N>();  // the tile's state records (direct-to-LDS copy)
  for (int i = lane; i < mfma_tile_doubles<N>(); i += 64) sA[i] = 0.0;

  // B operands: lane l holds B[k = 4*ks + l/16][j = l%16]
  const int jj = lane & 15, kk = lane >> 4;
  const double *Drw = ops, *Dsw = ops + Np * Np, *LIFT = ops + 2 * Np * Np;
  double bR[KP], bS[KP], bL[KF];
#pragma unroll
  for (int ks = 0; ks < KP; ks++) {
    const int k = 4 * ks + kk;
    const bool ok = jj < Np && k < Np;
    bR[ks] = ok ? Drw[jj * Np + k] : 0.0;
    bS[ks] = ok ? Dsw[jj * Np + k] : 0.0;
  }
#pragma unroll
  for (int ks = 0; ks < KF; ks++) {
    const int m = 4 * ks + kk;
    bL[ks] = (jj < Np && m < NF) ? -LIFT[jj * NF + m] : 0.0;
  }

  const bool multi = a.ncases > 0;
  if (EPI == EPI_STAGE0 && blockIdx.x == 0 && wave == 0) {
    for (int c = lane; c < a.ncases; c += 64) a.dt1_next[c] = 0ull;
    if (a.dt_rec && lane == 0) a.dt_rec[0] = a.cfl / __longlong_as_double((long long)a.dt1_cur[0]);
  }
  if (EPI == EPI_STAGE2 && blockIdx.x == 0 && wave == 0)
    for (int c = lane; c < a.ncases; c += 64)
      a.t_dev[c] += a.dt_fixed ? a.dt_fixed[c] : a.cfl / __longlong_as_double((long long)a.dt1_cur[c]);
  unsigned long long cfl_bits = 0ull;
  int cfl_case = -1;
  bool bad = false;
  const double n1sq = (double)((N + 1) * (N + 1));
  double *sF = sA, *sG = sA + G * GS, *sL = sA;
  const int ce = lane >> 4;  // cell within a group owned by this lane's outputs
  // A operand of K-step ks of group g from an F or G tile: lane l supplies
  // column k = 4 ks + l/16, row l%16, and zero for the padding k >= Np
  auto opA = [&](const double *t, int g, int ks) -> double {
    return (4 * ks + 3 < Np || 4 * ks + kk < Np) ? t[g * GS + CS * (4 * ks + kk) + jj] : 0.0;
  };

  const int ntiles = multi ? a.tile_off[a.ncases] : (a.ncells + TE - 1) / TE;
  const int nblk = (ntiles + wpb - 1) / wpb;  // blocks of wpb tiles
  // The it_t-th tile of this wave: 0 = past the walk (uniform across the
  // workgroup), 1 = a tile of cells [e0, e0 + ne) of case cs, 2 = past the
  // last tile (no work).
  auto locate = [&](int it_t, int &e0_, int &ne_, int &cs_) -> int {
    const int blk = walk == 0 ? xcd_tile(it_t, nblk) : (int)blockIdx.x + it_t * (int)gridDim.x;
    e0_ = 0; ne_ = 0; cs_ = 0;
    if (blk >= nblk) return 0;
    const int tile = blk * wpb + wave;
    if (tile >= ntiles) return 2;
    if (multi) {
      int lo = 0, hi = a.ncases - 1;  // case owning this tile
      if (a.tile_case) {
        lo = a.tile_case[tile];
      } else {
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (a.tile_off[mid] <= tile) lo = mid; else hi = mid - 1;
        }
      }
      cs_ = lo;
      e0_ = a.cell_off[cs_] + (tile - a.tile_off[cs_]) * TE;
      ne_ = min(TE, a.cell_off[cs_ + 1] - e0_);
    } else {
      e0_ = tile * TE;
      ne_ = min(TE, a.ncells - e0_);
    }
    return 1;
  };
  // Face-map codes of a tile's Roe items (lane + 64 j -> (cell, face node)).
  auto load_codes = [&](int e0_, int ne_, int *c) {
    const int lr = lane_in_loop(lane);
#pragma unroll
    for (int j = 0; j < NRT; j++) {
      const int it = lr + 64 * j;
      c[j] = (GATHER && it < ne_ * NF) ? a.fmap[(size_t)(e0_ + it / NF) * 3 + (it % NF) / Nfp] : 0;
    }
  };
  int e0, ne, cs;
  int st = locate(0, e0, ne, cs);
  int code[NRT];
  load_codes(e0, ne, code);
  // The face-map codes are read here and at the end of each tile (empty asm
  // statements that use them): the compiler's wait insertion then knows at
  // the top of the tile loop that no code load is in flight.  Otherwise it
  // placed waits for all loads in flight between the tile's direct-to-LDS
  // copy instructions, two extra memory round trips at the start of every
  // tile.
#pragma unroll
  for (int j = 0; j < NRT; j++) asm volatile("" : : "v"(code[j]));
  // A case's physics and dt are loaded when a tile of another case comes,
  // not per tile: a per-tile dt load, needed at once, made every tile wait
  // for the previous tile's stores (one count for loads and stores).  With
  // the codes' waits above, 724x724 vortex 1.163-1.166 -> 1.141-1.145 ms per
  // step over three alternating runs (the codes' reads alone: no change).
  Phys ph = ph0;
  double dt = 0.0;
  int ph_case = -1;  // case whose physics and dt ph and dt hold
  for (int it_tile = 0; st != 0; it_tile++) {
    const bool active = st == 1;
    if (active && multi && cs != ph_case) {
      ph = a.phys[cs];
      if (EPI != EPI_RHS)
        dt = a.dt_fixed ? a.dt_fixed[cs] : a.cfl / __longlong_as_double((long long)a.dt1_cur[cs]);
      ph_case = cs;
      // used here, so that no load of them is in flight past this branch
      asm volatile("" : : "v"(ph.gam), "v"(ph.bc_r), "v"(ph.bc_u), "v"(ph.bc_v), "v"(ph.bc_e), "v"(ph.bc_p),
                   "v"(dt));
    }
    if (EPI == EPI_STAGE2 && active && cs != cfl_case) {
      // flush the running CFL max of the previous case (wave-uniform branch)
      if (cfl_case >= 0) {
        unsigned long long r = cfl_bits;
        for (int off = 32; off > 0; off >>= 1) r = max_bits(r, __shfl_xor(r, off, 64));
        if (lane == 0) atomic_max_if_larger(a.dt1_next + cfl_case, r);
      }
      cfl_case = cs;
      cfl_bits = 0ull;
    }
    // The previous tile's volume products are done with the LDS tiles.  A
    // wave's LDS tiles are its own, so its phases need only its own LDS
    // accesses complete, not a workgroup barrier (whose release fence would
    // also wait for the previous tile's stores).
    lds_wait();

    // the tile's state records go to LDS while the face phase runs; the
    // Euler-flux phase reads them there instead of waiting on its own loads.
    // Each stretch of code that issues a tile's loads (this copy and the
    // first Roe item, each next Roe item, each group's epilogue loads) runs
    // at a raised wave priority, so that a wave about to put loads in flight
    // is picked before waves with arithmetic ready: more of the CU's 12
    // waves keep loads outstanding (724x724 vortex: 1.147-1.160 -> 1.140-
    // 1.147 ms per step over four alternating runs on two boxes; the copy
    // alone or the epilogue alone measured no gain, priority 3 the same as 2)
    __builtin_amdgcn_s_setprio(2);
    // (issued unseen by the compiler: its waits in the face phase stay exact)
    copy_to_lds_unseen<TE * Np * 2>(a.w + (size_t)e0 * Np * 4, sW, active ? ne * Np * 2 : 1, lane);

    // ---- Roe fluxes per (cell, face node), written in the A-operand layout
    {
      const int nrt = ne * NF;
      const int lr = kItemsInLoop ? lane_in_loop(lane) : lane;
      // one item's loads: own trace, exterior trace and geometry go out
      // together, before any branch; a boundary slot reads its own record as
      // a stand-in exterior, which boundary_state then replaces; an
      // exact-state slot reads the state exact_bc_kernel stored in its
      // exteriorQ slot for this stage; the drop-in euler_rhs (GATHER false)
      // reads exteriorQ.  An item past the tile's end re-reads item 0's data
      // (its face-map code is 0), which nobody uses.
      struct RoeIn {
        double2 m01, m23, p01, p23;
        double g[4];  // nx, ny, fscale; or the cell's rx, ry, sx, sy (CG)
      };
      auto load_item = [&](int j, RoeIn &r) {
        const int it0 = lr + 64 * j;
        const int it = it0 < nrt ? it0 : 0;
        const int e = it / NF, m = it - e * NF;
        const int i = m - (m / Nfp) * Nfp;
        const size_t cell = (size_t)(e0 + e);
        const size_t gm = cell * NF + m;
        if constexpr (CG) {
          const double2 *g = reinterpret_cast<const double2 *>(a.cmet + cell * 4);
          const double2 rxy = g[0], sxy = g[1];
          r.g[0] = rxy.x; r.g[1] = rxy.y; r.g[2] = sxy.x; r.g[3] = sxy.y;
        } else {
          r.g[0] = a.nx[gm]; r.g[1] = a.ny[gm]; r.g[2] = a.fscale[gm];
        }
        const size_t own = cell * Np + fmask_of(m);
        const bool interior = GATHER && code[j] >= 0;
        const bool exact = GATHER && code[j] == -(1 + kBcExact);
        const size_t nrec =
            interior ? (size_t)(code[j] >> 2) * Np + fmask_of((code[j] & 3) * Nfp + (Nfp - 1 - i)) : own;
        const double2 *ms = reinterpret_cast<const double2 *>(a.w + own * 4);
        const double2 *ps = (GATHER && !exact) ? reinterpret_cast<const double2 *>(a.w + nrec * 4)
                                               : reinterpret_cast<const double2 *>(a.ext + gm * 4);
        r.m01 = ms[0]; r.m23 = ms[1];
        r.p01 = ps[0]; r.p23 = ps[1];
      };
      // one item ahead: item j + 1's loads go out before item j's flux
      // (724x724 vortex: stage 0 373 -> 367 us, stages 1 and 2 1-3%, the
      // per-cell geometry step 1.10 -> 1.03 ms)
      RoeIn cur, nxt;
      load_item(0, cur);
      __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int j = 0; j < NRT; j++) {
        if (j + 1 < NRT) {
          __builtin_amdgcn_s_setprio(2);
          load_item(j + 1, nxt);
          __builtin_amdgcn_s_setprio(0);
        }
        const int it = lr + 64 * j;
        if (it < nrt) {
          const int e = it / NF, m = it - e * NF;
          double nxv, nyv, fs;
          if constexpr (CG) {
            cell_face_geometry(cur.g[0], cur.g[1], cur.g[2], cur.g[3], m / Nfp, nxv, nyv, fs);
          } else {
            nxv = cur.g[0]; nyv = cur.g[1]; fs = cur.g[2];
          }
          double mQ[4] = {cur.m01.x, cur.m01.y, cur.m23.x, cur.m23.y};
          double pQ[4] = {cur.p01.x, cur.p01.y, cur.p23.x, cur.p23.y};
          if (!GATHER) {
            double2 *z = reinterpret_cast<double2 *>(a.ext + ((size_t)(e0 + e) * NF + m) * 4);
            z[0] = make_double2(0.0, 0.0);  // OP_RW: euler_rhs zeroes exteriorQ
            z[1] = make_double2(0.0, 0.0);
          } else if (code[j] < 0 && code[j] != -(1 + kBcExact)) {
            boundary_state(ph, -code[j] - 1, mQ, nxv, nyv, pQ);
          }
          double fl[4];
          roe_node(ph.gam, mQ, pQ, nxv, nyv, fs, fl);
          const int off = (e >> 2) * GL + CSL * m + (e & 3);
#pragma unroll
          for (int v = 0; v < 4; v++) sL[off + 4 * v] = fl[v];
        }
        if (j + 1 < NRT) cur = nxt;
      }
      if (NF < 4 * KF) {  // zero padding columns (orders whose 3(N+1) is not a multiple of 4)
        constexpr int PADF = 4 * KF - NF;
        for (int it = lane; it < ne * PADF; it += 64) {
          const int e = it / PADF, m = NF + (it - e * PADF);
          const int off = (e >> 2) * GL + CSL * m + (e & 3);
#pragma unroll
          for (int v = 0; v < 4; v++) sL[off + 4 * v] = 0.0;
        }
      }
    }
    lds_wait();

    // ---- LIFT products: the accumulators start from the surface term
    dbl4 acc[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      acc[g] = dbl4{0.0, 0.0, 0.0, 0.0};
      if (4 * g >= ne) continue;
#pragma unroll
      for (int ks = 0; ks < KF; ks++) acc[g] = mfma16x16x4(sL[g * GL + CSL * (4 * ks + kk) + jj], bL[ks], acc[g]);
    }
    lds_wait();  // the Euler fluxes overwrite the face-flux tiles

    // ---- Euler fluxes per (cell, node), written in the A-operand layout.
    // The padding columns are never read (opA supplies their zeros), so a
    // value the face phase left in this memory cannot reach the products (a
    // failed case must not reach the next tile, which may belong to another
    // case).
    {
      const int nit = ne * Np;
      const int lr = kItemsInLoop ? lane_in_loop(lane) : lane;
      // The copied state records: every item's record is read (into
      // registers) before any F or G store, which reuse this memory; a
      // wave's LDS instructions execute in order.
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      double2 wq[NIT][2];
#pragma unroll
      for (int j = 0; j < NIT; j++) {
        const int it = lr + 64 * j;
        const double2 *src = reinterpret_cast<const double2 *>(sW + (it < nit ? it : 0) * 4);
        wq[j][0] = src[0];
        wq[j][1] = src[1];
      }
#pragma unroll
      for (int j = 0; j < NIT; j++) {
        const int it = lr + 64 * j;
        if (it >= nit) continue;
        const int e = it / Np, k = it - e * Np;
        double F[4], Gv[4];
        double qv[4] = {wq[j][0].x, wq[j][0].y, wq[j][1].x, wq[j][1].y};
        euler_flux(ph.gam, qv, F, Gv);
        if constexpr (CG) {  // contravariant fluxes
          const double2 *mt = reinterpret_cast<const double2 *>(a.cmet + (size_t)(e0 + e) * 4);
          const double2 rxy = mt[0], sxy = mt[1];
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const double f = F[v], g = Gv[v];
            F[v] = rxy.x * f + rxy.y * g;
            Gv[v] = sxy.x * f + sxy.y * g;
          }
        }
        // A row = cell + 4*var, A column = source node k
        const int off = (e >> 2) * GS + CS * k + (e & 3);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          sF[off + 4 * v] = F[v];
          sG[off + 4 * v] = Gv[v];
        }
      }
    }
    lds_wait();

    // the next tile's face-map codes load while this tile's volume products run
    int e0n, nen, csn;
    const int stn = locate(it_tile + 1, e0n, nen, csn);
    int ncode[NRT];
    load_codes(e0n, nen, ncode);

    // The stage-1 kernel loads the per-node metric factors of all four groups
    // here, so that only w and q are left to each group's epilogue (724x724
    // vortex: stage 1 401-405 -> 387-396 us; the same in stage 0, or rx and
    // sx alone in stage 2, measured 1.5% and 5% slower, all four in stage 2
    // spill).  At N=4 the held factors spill in stage 1 too (582 -> 566 us
    // without them), so there the epilogue loads them.
    constexpr bool kMetricsAhead = !CG && EPI == EPI_STAGE1 && N <= 3;
    double hrx[G], hsx[G], hry[G], hsy[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      hrx[g] = hsx[g] = hry[g] = hsy[g] = 0.0;
      if (kMetricsAhead) {
        const int e = 4 * g + ce;
        const bool ok = e < ne && jj < Np;
        const size_t gn = (size_t)(e0 + (ok ? e : 0)) * Np + (ok ? jj : 0);
        if (4 * g < ne) { hrx[g] = a.rx[gn]; hsx[g] = a.sx[gn]; hry[g] = a.ry[gn]; hsy[g] = a.sy[gn]; }
      }
    }
    // ---- volume products, metric terms and the epilogue, one 4-cell group at a time
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (4 * g >= ne) continue;
      const int e = 4 * g + ce;
      const bool ok = e < ne && jj < Np;
      const size_t gn = (size_t)(e0 + (ok ? e : 0)) * Np + (ok ? jj : 0);
      // This group's global loads (metric factors unless loaded above, stage
      // input w, base state q) go out before its products, so their latency
      // overlaps the matrix work.  (Loading w after the products, which kept 8 fewer
      // registers live across them, put a dependent round trip into every
      // group's epilogue: stage 2 measured 432 -> 411 us with it moved up.)
      double rxv = 0.0, sxv = 0.0, ryv = 0.0, syv = 0.0;
      double2 w01 = make_double2(0.0, 0.0), w23 = w01, q01 = w01, q23 = w01;
      __builtin_amdgcn_s_setprio(2);
      // Every lane issues them: a lane past the group's cells or nodes reads
      // the tile's node 0 (its values are not used).  The conditional form
      // zeroed the destination registers first, and the compiler made each
      // group wait there for loads still in flight (stage 1 388-391 ->
      // 377-382 us with unconditional loads; stages 0 and 2 unchanged).
      {
        if constexpr (kMetricsAhead) {
          rxv = hrx[g]; sxv = hsx[g]; ryv = hry[g]; syv = hsy[g];
        } else if constexpr (!CG) {
          rxv = a.rx[gn]; sxv = a.sx[gn]; ryv = a.ry[gn]; syv = a.sy[gn];
        }
        if (EPI != EPI_RHS) {
          const double2 *ws = reinterpret_cast<const double2 *>(a.w + gn * 4);
          w01 = ws[0]; w23 = ws[1];
        }
        if (EPI == EPI_STAGE1 || EPI == EPI_STAGE2) {
          const double2 *qs = reinterpret_cast<const double2 *>(a.q + gn * 4);
          q01 = qs[0]; q23 = qs[1];
        }
      }
      __builtin_amdgcn_s_setprio(0);
      dbl4 c = acc[g];
      if constexpr (CG) {
#pragma unroll
        for (int ks = 0; ks < KP; ks++) {
          c = mfma16x16x4(opA(sF, g, ks), bR[ks], c);
          c = mfma16x16x4(opA(sG, g, ks), bS[ks], c);
        }
      } else {
        // F's two products with rx, sx, then G's with ry, sy: two live
        // accumulators instead of four
        dbl4 d1 = {0, 0, 0, 0}, d2 = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KP; ks++) {
          const double aF = opA(sF, g, ks);
          d1 = mfma16x16x4(aF, bR[ks], d1);
          d2 = mfma16x16x4(aF, bS[ks], d2);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) c[r] += rxv * d1[r] + sxv * d2[r];
        d1 = dbl4{0, 0, 0, 0};
        d2 = dbl4{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KP; ks++) {
          const double aG = opA(sG, g, ks);
          d1 = mfma16x16x4(aG, bR[ks], d1);
          d2 = mfma16x16x4(aG, bS[ks], d2);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) c[r] += ryv * d1[r] + syv * d2[r];
      }
      if (!ok) continue;
      double o[4];
      if (EPI == EPI_RHS) {
#pragma unroll
        for (int v = 0; v < 4; v++) o[v] = c[v];
      } else {
        const double wv[4] = {w01.x, w01.y, w23.x, w23.y};
        if (EPI == EPI_STAGE0) {
#pragma unroll
          for (int v = 0; v < 4; v++) o[v] = wv[v] + dt * c[v];
        } else {
          const double qv[4] = {q01.x, q01.y, q23.x, q23.y};
          if (EPI == EPI_STAGE1) {
#pragma unroll
            for (int v = 0; v < 4; v++) o[v] = 0.75 * qv[v] + 0.25 * (wv[v] + dt * c[v]);
          } else {
#pragma unroll
            for (int v = 0; v < 4; v++) o[v] = (qv[v] + 2.0 * (wv[v] + dt * c[v])) * (1.0 / 3.0);
          }
        }
      }
      double2 *dst = reinterpret_cast<double2 *>(a.out + gn * 4);
      dst[0] = make_double2(o[0], o[1]);
      dst[1] = make_double2(o[2], o[3]);
      if (EPI == EPI_STAGE2) {
        if ((facenodes >> jj) & 1ull) {
          // src/calc_dt.h:11 reads fscale at face-node index FMASK[i], i.e.
          // at this node's index jj read as a face-node slot (jj < Np <= NF);
          // on the per-cell geometry formed as the face phase formed it
          double fsq;
          if constexpr (CG) {
            const double2 *mt = reinterpret_cast<const double2 *>(a.cmet + (size_t)(e0 + e) * 4);
            const double2 rxy = mt[0], sxy = mt[1];
            double nxq, nyq;
            cell_face_geometry(rxy.x, rxy.y, sxy.x, sxy.y, jj / Nfp, nxq, nyq, fsq);
          } else {
            fsq = a.fscale[(size_t)(e0 + e) * NF + jj];
          }
          const double cv = cfl_node(ph.gam, n1sq, o, fsq) + 0.0;
          cfl_bits = max_bits(cfl_bits, (unsigned long long)__double_as_longlong(cv));
        }
        bad |= !(isfinite(o[0]) && isfinite(o[1]) && isfinite(o[2]) && isfinite(o[3])) ||
               !(o[0] > 0.0);
      }
    }
    e0 = e0n; ne = nen; cs = csn; st = stn;
#pragma unroll
    for (int j = 0; j < NRT; j++) {
      asm volatile("" : : "v"(ncode[j]));  // see the loop's head
      code[j] = ncode[j];
    }
  }
  if (EPI == EPI_STAGE2) {
    if (cfl_case >= 0) {
      unsigned long long r = cfl_bits;
      for (int off = 32; off > 0; off >>= 1) r = max_bits(r, __shfl_xor(r, off, 64));
      if (lane == 0) atomic_max_if_larger(a.dt1_next + cfl_case, r);
    }
    if (__any(bad) && lane == 0) atomicOr(a.flag, 1u);
  }
}

// The stage-0 instantiations, one per order and geometry mode: compiled in
// dg_stage0_mfma.hip, declared where the launchers are.
#define DG_STAGE0_INSTANCES(X) \
  X(1, false) X(1, true) X(2, false) X(2, true) X(3, false) X(3, true) X(4, false) X(4, true)
#define DG_STAGE0_KERNEL(N, CG)                                                                    \
  __global__ void rhs_mfma_kernel<N, true, EPI_STAGE0, CG>(StageArgs, const double *,             \
                                                           unsigned long long, Phys,               \
                                                           unsigned long long, int);

}  // namespace dg
