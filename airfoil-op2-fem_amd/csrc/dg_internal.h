// Internal declarations shared by the C-ABI layer and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dg {

// Physics constants (op_decl_const scalars of src/airfoil.cpp:183-190).
struct Phys {
  double gam, bc_r, bc_u, bc_v, bc_e, bc_p;
};

// Boundary type whose exterior state is the exact isentropic vortex at the
// face node and stage time (vortex/src/get_bedge_q.h:1-40).
constexpr int kBcExact = 4;

// Isentropic vortex (vortex/src/set_ic.h:1-28 with free-stream speed uinf,
// advected with it, periodic in x with period `period` when it is > 0):
// centre (x0, y0), strength beta.
struct Vortex {
  double uinf = 0.0, beta = 5.0, x0 = 5.0, y0 = 0.0, period = 0.0;
};

// Layout of the per-context operator buffer on the device (doubles):
//   Drw[Np*Np] | Dsw[Np*Np] | LIFT[Np*NF] | Dr[Np*Np] | Ds[Np*Np] | r[Np] | s[Np]
inline int ops_size(int Np, int NF) { return 4 * Np * Np + Np * NF + 2 * Np; }

// Everything a launch needs from the context.
struct Launch {
  int N, Np, Nfp, NF;
  hipStream_t stream;
  const double *ops;   // device operator buffer
  const int *fmask;    // device FMASK [NF]
  Phys ph;
  uint64_t facenodes;  // bit j set when node j lies on a face
  uint64_t fmask4;     // FMASK[m] in bits 4m..4m+3 (entries < 16, NF <= 15 for N <= 4)
  int cus;             // compute units of the device
  int max_blocks;      // optional cap on persistent grids (0: none)
  // matrix-core kernel launch shape (the autotuner picks per mesh; one-wave
  // workgroups with the XCD-chunked walk are within 1% of the best shape on
  // the 724x724 vortex and the best on 65k-cell meshes)
  int wpb = 1;         // waves (16-cell tiles) per workgroup, 1 / 2 / 4
  int walk = 0;        // 0 XCD-chunked block walk, 1 round-robin walk
};

// Arguments of the stage kernels.
struct StageArgs {
  const double *w;        // stage input state [K][Np][4]
  double *ext;            // exteriorQ (drop-in euler_rhs only)
  const double *rx, *ry, *sx, *sy;   // [K][Np]
  const double *fscale, *nx, *ny;    // [K][NF]
  // Per-cell metric factors of straight-sided cells, [K][4] = (rx, ry, sx,
  // sy), or null.  When set, the matrix-core kernel reads these instead of
  // the per-node factors and forms each face's normal and fscale from them
  // (see cell_face_geometry); the per-node arrays are then not read.
  const double *cmet;
  const int *fmap;        // [K][3] face map
  double *out;            // qRHS, or the next state
  const double *q;        // SSPRK3 base state (stages 1, 2)
  const unsigned long long *dt1_cur;  // device dt1 per case (bits of a non-negative double)
  unsigned long long *dt1_next;       // reduction target per case (stage 2), zeroed in stage 0
  double cfl;
  double *t_dev;          // simulated time per case, += dt in stage 2
  unsigned int *flag;     // bit 0: non-finite value seen
  int ncells;
  // Case table of the fused solver (ncases >= 1): case c owns cells
  // [cell_off[c], cell_off[c+1]) and tiles [tile_off[c], tile_off[c+1]) of the
  // concatenated arrays, with its own physics constants and dt.  ncases == 0
  // (the per-loop entry points): one case of ncells cells, physics from the
  // launch, no dt.
  int ncases;
  const int *cell_off, *tile_off;
  const int *tile_case;   // case of each tile (tile_off[ncases] entries), or null: binary search of tile_off
  const Phys *phys;
  // Frozen time step (per case): when non-null the stage updates use
  // dt_fixed[case] instead of cfl / dt1 (sensitivity runs replay a recorded
  // dt sequence).
  const double *dt_fixed;
  // Stage 0 of a recorded run: workgroup 0 stores the step's time step
  // cfl / dt1_cur[0] here (case 0; sensitivity recording), or null.
  double *dt_rec;
  // Host-side count of 16-cell tiles of a case table (0: unknown); lets the
  // launcher size the persistent grid so that every wave gets the same
  // number of tiles.
  int ntiles_hint;
};

// Cells per tile of the matrix-core stage kernel (one wave per tile).
constexpr int kMfmaTile = 16;

// Epilogue kinds of the stage kernel.
enum Epi { EPI_RHS = 0, EPI_STAGE0 = 1, EPI_STAGE1 = 2, EPI_STAGE2 = 3 };

hipError_t launch_rhs(const Launch &L, const StageArgs &a, bool gather, int epi);
// Exact vortex states on boundary edges (vortex/src/get_bedge_q.h:1-40):
// for boundary edges b in [0, nb) whose type is kBcExact (every edge when
// btype is null), at each face node m of (b2c[b], bnum[b]) the vortex state
// at (x, y) of node FMASK[m] and time t goes to ext[(cell * NF + m) * 4]:
// added (accumulate) or stored.  t = t_host when t_dev is null, else
// t_dev[0] + frac * dt with dt = dt_fixed[0] or cfl / dt1[0] (the stage time
// of a fused step, read on the device).
hipError_t launch_exact_bc(const Launch &L, int nb, const int *b2c, const int *bnum, const int *btype,
                           const double *x, const double *y, const Vortex &v, double t_host,
                           const double *t_dev, const unsigned long long *dt1, double cfl,
                           const double *dt_fixed, double frac, bool accumulate, double *ext);
hipError_t launch_rhs_mfma(const Launch &L, const StageArgs &a, bool gather, int epi);
// Errors against the exact vortex v at time t (vortex/src/l2_error.h:1-29):
// error[i][k] = (q[i][k] - exact)^2 when error is set; partial[8 g + k] and
// partial[8 g + 4 + k] (k < 4) get workgroup g's sum of those squares and,
// when xr is set, of J_c e^T mass e over its cells (J_c = xr ys - xs yr at
// the cell's node 0).  l2_error_blocks(Np, ncells) <= kL2Blocks workgroups.
constexpr int kL2Blocks = 256;
int l2_error_blocks(int Np, int ncells);
hipError_t launch_l2_error(const Launch &L, int ncells, const Vortex &v, double t, const double *x,
                           const double *y, const double *q, const double *xr, const double *yr,
                           const double *xs, const double *ys, const double *mass, double *error,
                           double *partial);
hipError_t launch_get_neighbour_q(const Launch &L, int ncells, const int *fmap,
                                  const double *w, double *ext);
hipError_t launch_get_bedge_q(const Launch &L, int nbedges, const int *b2c,
                              const int *bnum, const int *btype, const double *nx,
                              const double *ny, const double *w, double *ext);
hipError_t launch_set_workingQ(const Launch &L, int ncells, double dt, int stage,
                               const double *q, const double *k1, const double *k2,
                               double *w);
hipError_t launch_update_Q(const Launch &L, int ncells, double dt, double *q,
                           const double *k1, const double *k2, const double *k3,
                           double *w);
hipError_t launch_calc_dt(const Launch &L, int ncells, const double *q,
                          const double *fscale, unsigned long long *d_out);
hipError_t launch_init_grid(const Launch &L, int ncells, const double *cell_xy,
                            double *nodeX, double *nodeY, double *x, double *y,
                            double *xr, double *yr, double *xs, double *ys,
                            double *rx, double *ry, double *sx, double *sy,
                            double *nx, double *ny, double *fscale, double *J);
// cmet[c] = (rx, ry, sx, sy) at node 0 of cell c, [K][4]
hipError_t launch_cell_metrics(const Launch &L, int ncells, const double *rx, const double *ry,
                               const double *sx, const double *sy, double *cmet);
hipError_t launch_set_ic(const Launch &L, int ncells, double *q, double *w);
hipError_t launch_set_vortex(const Launch &L, int ncells, const double *x,
                             const double *y, double uinf, double beta, double x0,
                             double y0, double period_x, double t, double *q, double *w);
hipError_t launch_zero(const Launch &L, size_t n, double *p);
hipError_t launch_copy(const Launch &L, size_t n, const double *src, double *dst);
// out[i][0..4) = q node record rec[i] (zeros where rec[i] < 0)
hipError_t launch_gather_records(const Launch &L, int n, const long long *rec, const double *q, double *out);
// Pressure forces on the listed wall faces: out[2*b] = (fx, fy) of face b;
// cp[b*Nfp + i] = (p - p_inf) / qdyn at its nodes.
hipError_t launch_wall_forces(const Launch &L, int nwall, const int *wall_cell,
                              const int *wall_face, const double *q, const double *nx,
                              const double *ny, const double *fscale, const double *J,
                              const double *facew, double pinf, double qdyn, double *out,
                              double *cp);

}  // namespace dg
