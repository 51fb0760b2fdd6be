/*
 * dg_euler.h -- C ABI of the MI355X (gfx950) nodal DG Euler solver.
 *
 * Drop-in for the OP2 par_loop backends of the reference's DG Euler hot
 * path.  The reference driver (src/airfoil.cpp:252-334) calls, per SSPRK3
 * stage, op_par_loop_get_neighbour_q, op_par_loop_get_bedge_q,
 * op_par_loop_euler_rhs and op_par_loop_set_workingQ, then per step
 * op_par_loop_update_Q and op_par_loop_calc_dt (declared at
 * src/airfoil_op.cpp:17-98).  Each dg_<kernel> entry point below performs the
 * same set-level loop on device memory in the reference's op_dat layouts:
 *
 *   q, workingQ, rk1..3, qRHS   [ncells][Np][4]      (variables fastest)
 *   x, y, xr, .., rx, ry, sx, sy [ncells][Np]
 *   nx, ny, fscale               [ncells][3*Nfp]
 *   exteriorQ                    [ncells][3][Nfp][4]
 *   nodeX, nodeY                 [ncells][3]
 *   edge2cells, edgeNum          [nedges][2]          int32
 *   bedge2cells, bedgeNum, bedge_type [nbedges]       int32
 *
 * with Np = (N+1)(N+2)/2 and Nfp = N+1 for the context's order N (the
 * reference hardcodes N=4).  Pointers named d_* are device pointers owned by
 * the caller; the library allocates only what a context or solver owns.
 *
 * Error handling: every int-returning function returns DG_OK (0) or a
 * negative DG_ERR_* code and never exits; dg_last_error(ctx) gives the text
 * (the reference aborts instead: cutilSafeCall, cg_error_exit, exit(1) in
 * op_decl_const_char, src/cuda/airfoil_kernels.cu:33-107).
 *
 * Execution: work is enqueued on the context's stream (dg_set_stream,
 * default the null stream) and returns without synchronising, except
 * functions that hand a host scalar back (dg_calc_dt, dg_solver_* getters),
 * which synchronise that stream.  One context per host thread.
 */
#ifndef DG_EULER_H
#define DG_EULER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DG_OK = 0,
  DG_ERR_ARG = -1,       /* bad argument (null pointer, size, order, name) */
  DG_ERR_STATE = -2,     /* call out of order (e.g. no mesh declared) */
  DG_ERR_HIP = -3,       /* HIP runtime error */
  DG_ERR_NAME = -4,      /* unknown constant name */
  DG_ERR_MESH = -5,      /* invalid mesh connectivity */
  DG_ERR_NONFINITE = -6  /* non-finite or non-positive state detected */
};

/* Boundary-condition types of bedge_type: 0-2 of src/load_mesh.cpp:147-172;
 * 3 a characteristic far field (no reference counterpart); 4 the exact
 * isentropic-vortex state at the face node and stage time, the boundary of
 * the reference's vortex case (vortex/src/get_bedge_q.h:1-40). */
enum { DG_BC_INFLOW = 0, DG_BC_OUTFLOW = 1, DG_BC_WALL = 2, DG_BC_FARFIELD = 3, DG_BC_EXACT = 4 };

typedef struct dg_ctx dg_ctx;
typedef struct dg_mesh dg_mesh;
typedef struct dg_solver dg_solver;

/* ------------------------------------------------------------------ */
/* Context and constants                                              */
/* ------------------------------------------------------------------ */

/* Creates a context for polynomial order N in [1,4] on the current HIP
 * device.  The operator tables are generated for N (Hesthaven & Warburton
 * construction) and uploaded; dg_set_const may replace them.  Replaces
 * op_init (src/airfoil.cpp:67). */
int dg_ctx_create(int order, dg_ctx **out);
void dg_ctx_destroy(dg_ctx *ctx);
/* Last error text of ctx; with ctx == NULL, of the calling thread's last
 * failed call that had no context (creation, mesh generation). */
const char *dg_last_error(const dg_ctx *ctx);
int dg_order(const dg_ctx *ctx, int *N, int *Np, int *Nfp);
/* Cap on the persistent grid of the stage kernels (no reference
 * counterpart): max_blocks > 0 limits the workgroups of a launch, 0 (the
 * default) launches as many as the compute units hold resident. */
int dg_set_max_blocks(dg_ctx *ctx, int max_blocks);
/* Stream for all later work of ctx (a hipStream_t; NULL = null stream). */
int dg_set_stream(dg_ctx *ctx, void *hip_stream);

/* Host-only operator generator (no device needed): fills the order-N tables
 * that the reference declares as literals in src/constants/constant_*.h.  Any
 * output pointer may be NULL.  Sizes: r, s [Np]; Dr, Ds, Drw, Dsw, mass [Np*Np];
 * LIFT [Np*3*Nfp]; FMASK [3*Nfp]; face_mass [Nfp*Nfp]. */
int dg_generate_operators(int order, double *r, double *s, double *Dr, double *Ds,
                          double *Drw, double *Dsw, double *LIFT, int *FMASK,
                          double *mass, double *face_mass);

/* Replaces op_decl_const2 -> op_decl_const_char (src/airfoil_op.cpp:283-299,
 * src/cuda/airfoil_kernels.cu:33-107).  Names: gam, bc_mach, bc_alpha, bc_p,
 * bc_r, bc_u, bc_v, bc_e (dim 1, double); ones, r, s (Np); Dr, Ds, Drw, Dsw
 * (Np*Np); LIFT (Np*3*Nfp); FMASK (3*Nfp, int); and the vortex case's
 * vortex_x0, vortex_y0, vortex_beta (dim 1, vortex/src/vortex.cpp:171-174).
 * dim must match.  Unknown names return DG_ERR_NAME instead of exiting. */
int dg_set_const(dg_ctx *ctx, const char *name, int dim, const void *host_data);
int dg_get_const(const dg_ctx *ctx, const char *name, int dim, void *host_out);
/* Free-stream set-up of src/airfoil.cpp:79-99: bc_u, bc_v from Mach and the
 * angle of attack (degrees), bc_e = p/(gam-1) + rho(u^2+v^2)/2. */
int dg_set_freestream(dg_ctx *ctx, double gam, double mach, double alpha_deg,
                      double p, double rho);

/* Declares the mesh sets and maps (op_decl_set/op_decl_map of
 * src/airfoil.cpp:134-146) from host arrays.  cell2nodes ([ncells][3],
 * counter-clockwise) may be NULL; when given, neighbour orientation is
 * checked.  Builds the owner-computes face map on the device.  A face not
 * covered exactly once by the edge lists returns DG_ERR_MESH. */
int dg_decl_mesh(dg_ctx *ctx, int ncells, const int *cell2nodes, int nedges,
                 const int *edge2cells, const int *edgeNum, int nbedges,
                 const int *bedge2cells, const int *bedgeNum, const int *bedge_type);
/* The connectivity checks of dg_decl_mesh on the host alone (no device, no
 * context): every (cell, local edge) covered exactly once by the edge and
 * boundary-edge lists, indices and boundary types in range, and, with
 * cell2nodes, neighbours traversing their shared edge in opposite
 * directions (a clockwise cell among counter-clockwise ones fails).
 * DG_ERR_MESH with the reason in dg_last_error(NULL). */
int dg_mesh_check(int ncells, const int *cell2nodes, int nedges, const int *edge2cells,
                  const int *edgeNum, int nbedges, const int *bedge2cells,
                  const int *bedgeNum, const int *bedge_type);

/* ------------------------------------------------------------------ */
/* The reference's per-set loops (drop-ins for op_par_loop_<kernel>)   */
/* ------------------------------------------------------------------ */

/* src/init_grid.h:1-106 over cells.  d_cell_xy is [ncells][3][2]: the three
 * vertex coordinates of each cell (node_coords through cell2nodes). */
int dg_init_grid(dg_ctx *ctx, const double *d_cell_xy, double *d_nodeX,
                 double *d_nodeY, double *d_x, double *d_y, double *d_xr,
                 double *d_yr, double *d_xs, double *d_ys, double *d_rx,
                 double *d_ry, double *d_sx, double *d_sy, double *d_nx,
                 double *d_ny, double *d_fscale);
/* src/set_ic.h:1-12 over cells (uniform free stream). */
int dg_set_ic(dg_ctx *ctx, double *d_q, double *d_workingQ);
/* The vortex case's set_ic (vortex/src/set_ic.h:1-28, called at
 * vortex/src/vortex.cpp:208-213) over cells: q = workingQ = the isentropic
 * vortex of dg_set_exact_state at the nodes' coordinates (d_x, d_y:
 * init_grid's [ncells][Np]) and time t. */
int dg_set_ic_vortex(dg_ctx *ctx, double t, const double *d_x, const double *d_y, double *d_q,
                     double *d_workingQ);
/* src/neighbour_zero.h:1-5 over cells. */
int dg_neighbour_zero(dg_ctx *ctx, double *d_exteriorQ);
/* src/get_neighbour_q.h:1-89 over edges: exteriorQ += neighbour face traces
 * (OP_INC).  Gather-only: every face slot is written by the thread that owns
 * it, so no atomics or colouring are needed and results are deterministic.
 * nodeX/nodeY are accepted for signature parity; orientation comes from the
 * declared topology (always reversed on counter-clockwise meshes). */
int dg_get_neighbour_q(dg_ctx *ctx, const double *d_nodeX, const double *d_nodeY,
                       const double *d_workingQ, double *d_exteriorQ);
/* src/get_bedge_q.h:1-51 over boundary edges (OP_INC), types 0-3 (edges of
 * type 4 are left to dg_get_bedge_q_exact). */
int dg_get_bedge_q(dg_ctx *ctx, const double *d_nx, const double *d_ny,
                   const double *d_workingQ, double *d_exteriorQ);
/* The vortex case's boundary loop (vortex/src/get_bedge_q.h:1-40) over every
 * boundary edge (OP_INC): exteriorQ += the exact isentropic-vortex state at
 * the face nodes' coordinates (d_x, d_y: init_grid's [ncells][Np]) and time
 * t, the stage time t + {0, 1, 1/2} dt of vortex/src/vortex.cpp:228,248-258.
 * The vortex is dg_set_exact_state's (u_inf advects it in both components;
 * the reference's own formula hardcodes u_inf = 1 and drops the -t in v). */
int dg_get_bedge_q_exact(dg_ctx *ctx, const double *d_x, const double *d_y, double t,
                         double *d_exteriorQ);
/* Isentropic vortex of the exact-state boundaries (type 4) and
 * dg_get_bedge_q_exact: free-stream speed uinf in x, strength beta, centre
 * (x0, y0) at t = 0, periodic in x with period_x (0: not periodic).
 * Defaults: 0, 5, 5, 0, 0; dg_solver_set_vortex sets the same values. */
int dg_set_exact_state(dg_ctx *ctx, double uinf, double beta, double x0, double y0,
                       double period_x);
/* The vortex case's error loop (vortex/src/l2_error.h:1-29, called at
 * vortex/src/vortex.cpp:314-328) over cells: d_error ([ncells][Np][4],
 * OP_WRITE; may be NULL) = (q - exact)^2 per node and variable, and
 * err[k] += the sum of those squares over every node (the four OP_INC
 * globals r, ru, rv, ener).  The exact state is dg_set_exact_state's vortex
 * at time t and the node coordinates d_x, d_y (init_grid's); the reference's
 * own formula hardcodes u_inf = 1 and drops the -t in v.  The sum is formed
 * in a fixed order that depends on ncells only: repeated calls give the same
 * bits.  Synchronises the context's stream. */
int dg_l2_error(dg_ctx *ctx, const double *d_x, const double *d_y, const double *d_q, double t,
                double *d_error, double err[4]);
/* Mass-weighted L2 norm of the same errors (no reference counterpart; the
 * norm a convergence study needs): err[k] = sqrt(sum_cells J e^T M e) with
 * M the reference-element mass matrix and J = xr ys - xs yr at the cell's
 * node 0 (exact for straight-sided cells).  err is overwritten. */
int dg_l2_error_mass(dg_ctx *ctx, const double *d_x, const double *d_y, const double *d_xr,
                     const double *d_yr, const double *d_xs, const double *d_ys,
                     const double *d_q, double t, double err[4]);
/* src/euler_rhs.h:222-291 over cells: volume term, Roe surface flux, LIFT;
 * zeroes exteriorQ on exit (OP_RW) and writes qRHS (OP_WRITE). */
int dg_euler_rhs(dg_ctx *ctx, const double *d_workingQ, double *d_exteriorQ,
                 const double *d_rx, const double *d_ry, const double *d_sx,
                 const double *d_sy, const double *d_fscale, const double *d_nx,
                 const double *d_ny, double *d_qRHS);
/* src/set_workingQ.h:1-12 over cells; stage 0 or 1. */
int dg_set_workingQ(dg_ctx *ctx, double dt, int stage, const double *d_q,
                    const double *d_k1, const double *d_k2, double *d_workingQ);
/* src/update_Q.h:1-12 over cells. */
int dg_update_Q(dg_ctx *ctx, double dt, double *d_q, const double *d_rk1,
                const double *d_rk2, const double *d_rk3, double *d_workingQ);
/* src/calc_dt.h:1-22 over cells with OP_MAX: *dt1 = max(*dt1, cell values).
 * Synchronises the stream. */
int dg_calc_dt(dg_ctx *ctx, const double *d_q, const double *d_fscale, double *dt1);

/* ------------------------------------------------------------------ */
/* Fused kernels (no reference counterpart; same mathematics)          */
/* ------------------------------------------------------------------ */

/* get_neighbour_q + get_bedge_q + euler_rhs in one pass: the neighbour and
 * boundary traces are gathered straight from workingQ, exteriorQ is never
 * formed.  Writes qRHS. */
int dg_rhs_fused(dg_ctx *ctx, const double *d_workingQ, const double *d_rx,
                 const double *d_ry, const double *d_sx, const double *d_sy,
                 const double *d_fscale, const double *d_nx, const double *d_ny,
                 double *d_qRHS);

/* ------------------------------------------------------------------ */
/* Mesh generation (host only)                                        */
/* ------------------------------------------------------------------ */

/* Seeded jittered periodic grid: nx*ny quads over [x0,x0+Lx]x[y0,y0+Ly],
 * each split SW->NE into two CCW triangles; vertices off the seams jittered
 * by U(-jitter*h, jitter*h) from splitmix64(seed). */
int dg_mesh_periodic(int nx, int ny, double x0, double y0, double Lx, double Ly,
                     uint64_t seed, double jitter, dg_mesh **out);
/* Seeded jittered nx*ny quad grid over [x0,x0+Lx]x[y0,y0+Ly], not periodic,
 * quads split SW->NE into two CCW triangles, vertices off the boundary
 * jittered as dg_mesh_periodic's; every boundary edge gets type bc_type
 * (DG_BC_EXACT for the vortex case with exact-state boundaries). */
int dg_mesh_rect(int nx, int ny, double x0, double y0, double Lx, double Ly, uint64_t seed,
                 double jitter, int bc_type, dg_mesh **out);
/* O-grid around NACA0012 (naca0012.geo:14 profile); wall on the profile,
 * characteristic far field on the outer circle. */
int dg_mesh_ogrid(int ncirc, int nrad, double stretch, double rfar, dg_mesh **out);
/* Legacy-VTK ASCII triangle mesh with tools/vtk2CGNS.cpp edge numbering and
 * src/load_mesh.cpp:147-172 boundary classification. */
int dg_mesh_read_vtk(const char *path, dg_mesh **out);
void dg_mesh_destroy(dg_mesh *m);
int dg_mesh_sizes(const dg_mesh *m, int *nnodes, int *ncells, int *nedges, int *nbedges);
/* Copies one field to host memory: "coords" (double [nnodes][2]), "cells"
 * (int [ncells][3]), "cell_xy" (double [ncells][3][2]), "edge2cells",
 * "edgeNum", "edge2nodes" (int [nedges][2]), "bedge2cells", "bedgeNum",
 * "bedge_type" (int [nbedges]), "bedge2nodes" (int [nbedges][2]). */
int dg_mesh_get(const dg_mesh *m, const char *field, void *host_out);
/* Replaces the boundary class of every boundary edge (types: int [nbedges],
 * each a DG_BC_* value; nbedges must be the mesh's count).  The vortex case
 * classifies nothing: its loader reads no boundary types
 * (vortex/src/load_mesh.cpp:122-145) and every boundary edge gets the exact
 * state (vortex/src/vortex.cpp:250-260), i.e. DG_BC_EXACT for all of them
 * after dg_mesh_read_vtk.  Solvers and dg_decl_mesh copy the classes when
 * they are created or called, so a change reaches only those created later.
 * On an error the classes are left unchanged. */
int dg_mesh_set_bedge_types(dg_mesh *m, int nbedges, const int *types);

/* ------------------------------------------------------------------ */
/* Solver: the SSPRK3 driver of src/airfoil.cpp:204-334 on device      */
/* ------------------------------------------------------------------ */

/* Declares the mesh on ctx, allocates the op_dats (q, workingQ, rk1..3,
 * exteriorQ, geometry) on the device and runs init_grid and set_ic. */
int dg_solver_create(dg_ctx *ctx, const dg_mesh *mesh, dg_solver **out);
void dg_solver_destroy(dg_solver *s);
/* Isentropic vortex initial state (vortex/src/set_ic.h:1-28 with the
 * free-stream speed uinf and the advection of the centre): centre (x0,y0),
 * strength beta, gamma from the context, periodic in x with period_x (0 =
 * not periodic). */
int dg_solver_set_vortex(dg_solver *s, double uinf, double beta, double x0,
                         double y0, double period_x);
int dg_solver_set_q(dg_solver *s, const double *host_q);
int dg_solver_get_q(dg_solver *s, double *host_q);
/* Device pointer of an op_dat owned by the solver: "q", "workingQ", "rk1",
 * "rk2", "rk3", "exteriorQ", "x", "y", "rx", "ry", "sx", "sy", "nx", "ny",
 * "fscale", "nodeX", "nodeY", "J" ([ncells], cell Jacobian). */
int dg_solver_dat(dg_solver *s, const char *name, void **d_ptr);

enum {
  DG_MODE_FUSED = 0,     /* fused stage kernels, dt kept on the device */
  DG_MODE_REFERENCE = 1  /* the reference's 9 loops per step, dt on the host */
};
/* Runs nsteps SSPRK3 steps with dt = cfl/dt1 recomputed after every step
 * (src/airfoil.cpp:231-237,322-331; the reference uses cfl = 1).  The first
 * call computes dt from the current q.  Asynchronous in DG_MODE_FUSED. */
int dg_solver_run(dg_solver *s, int nsteps, double cfl, int mode);
/* As dg_solver_run in DG_MODE_FUSED, with two steps (one per parity of the
 * device dt slots) captured once into a hipGraph and replayed; an odd step
 * count runs its last step directly.  The graph is rebuilt when cfl
 * changes.  Results are identical to dg_solver_run. */
int dg_solver_run_graph(dg_solver *s, int nsteps, double cfl);
/* Simulated time and the dt the next step will use (synchronises). */
int dg_solver_time(dg_solver *s, double *t, double *dt);
/* The solver's stream (the context's stream, or a private non-blocking
 * stream when the context uses the null stream, which cannot be captured). */
int dg_solver_stream(dg_solver *s, void **hip_stream);

/* ------------------------------------------------------------------ */
/* Batched sweep (no reference counterpart)                           */
/* ------------------------------------------------------------------ */

/* ncases independent cases of any sizes in one solver: the meshes are
 * concatenated into one set of op_dats and every fused stage is one launch
 * of the persistent matrix-core kernel over a tile table covering all
 * cases.  Case k has free stream (gam, p, rho of the context, Mach mach[k],
 * angle alpha_deg[k]; both NULL: the context's free stream for every case)
 * and its own dt and time.  A case's results are bit-identical to the same
 * case run alone in a single-case solver. */
int dg_solver_create_batch(dg_ctx *ctx, int ncases, const dg_mesh *const *meshes,
                           const double *mach, const double *alpha_deg, dg_solver **out);
int dg_solver_ncases(const dg_solver *s, int *ncases);
int dg_solver_case_set_q(dg_solver *s, int k, const double *host_q);
int dg_solver_case_get_q(dg_solver *s, int k, double *host_q);
int dg_solver_case_time(dg_solver *s, int k, double *t, double *dt);
int dg_solver_case_forces(dg_solver *s, int k, double chord, double *cl, double *cd,
                          double *fxy);
/* Resets time to 0 and forces dt to be recomputed from q on the next run. */
int dg_solver_reset_time(dg_solver *s);
/* Returns DG_ERR_NONFINITE if any step so far produced a non-finite value
 * in the CFL reduction (fused mode), DG_OK otherwise (synchronises). */
int dg_solver_check(dg_solver *s);
/* Pressure force on the wall faces (bc 2), integrated with the 1D face mass
 * matrix as quadrature, as lift and drag coefficients relative to the free
 * stream of the context (angle bc_alpha, dynamic pressure rho|V|^2/2) and
 * the chord; fxy (may be NULL) gets the force vector.  Synchronises.  No
 * reference counterpart (the reference writes only the solution). */
int dg_solver_forces(dg_solver *s, double chord, double *cl, double *cd, double *fxy);
/* Surface pressure coefficient (p - p_inf)/(rho|V|^2/2) at the Nfp nodes of
 * each wall face, in wall-face order, with the nodes' x coordinate.  With
 * x and cp NULL only *nwall is returned. */
int dg_solver_wall_cp(dg_solver *s, int *nwall, double *x, double *cp);
/* Replaces save_solution (src/save_solution.cpp:13-47) without CGNS: vertex
 * values of case k (0 for a single-case solver), each vertex taking the
 * state at the corner node (0, N, Np-1) of the last cell containing it;
 * *nnodes receives the vertex count (arrays may be NULL to query it).
 * dg_solver_write_vtk writes the case's mesh and the four fields (Density,
 * Pressure, VelocityX, VelocityY) as a legacy-VTK ASCII file. */
int dg_solver_vertex_fields(dg_solver *s, int k, int *nnodes, double *rho, double *vx, double *vy,
                            double *p);
int dg_solver_write_vtk(dg_solver *s, int k, const char *path);
/* Kernel timing with HIP events on the context stream: enable, then read the
 * average duration (ms) and launch count per kernel name ("rhs_stage0",
 * "rhs_stage1", "rhs_stage2", "calc_dt", "euler_rhs", ...). */
int dg_solver_profile(dg_solver *s, int enable);
int dg_solver_kernel_time(dg_solver *s, const char *kernel, double *avg_ms, int *count);

/* Launch shape of the matrix-core stage kernel (no reference counterpart):
 * wpb = waves (16-cell tiles) per workgroup, 1, 2 or 4 (16 / 32 / 64
 * elements per workgroup); walk = 0 XCD-chunked (the workgroups of one XCD
 * walk one contiguous range of tiles), 1 round-robin; max_blocks caps the
 * persistent grid (0: occupancy-limited).  The shape never changes results.
 * Default: wpb 1, walk 0, max_blocks 0. */
int dg_solver_set_launch(dg_solver *s, int wpb, int walk, int max_blocks);
int dg_solver_launch_config(const dg_solver *s, int *wpb, int *walk, int *max_blocks);
/* Geometry the fused stage kernels read (no reference counterpart).
 * DG_GEOM_NODE (default): the per-node rx, ry, sx, sy and per face-node nx,
 * ny, fscale of src/init_grid.h, as src/euler_rhs.h and the per-loop entries
 * take them (608 bytes per cell at N=3).  DG_GEOM_CELL: one (rx, ry, sx, sy)
 * record per cell (the factors at node 0), with each face's normal and
 * fscale formed from it and the volume term taken as
 * Drw (rx F + ry G) + Dsw (sx F + sy G): 32 bytes of geometry per cell.
 * Straight-sided cells have the same factors at every node in exact
 * arithmetic, but init_grid's per-node values (Dr and Ds applied to node
 * coordinates) scatter by about 1e-15 x (coordinate size / cell size)
 * relative, 1e-12 and more on fine or thin cells, and DG_GEOM_CELL does not
 * carry that scatter: its results leave the reference's by that much (about
 * 2e-12 relative in rho v after 100 steps of the 64x64 vortex).  The
 * geometry factors are formed when the solver is created. */
#define DG_GEOM_NODE 0
#define DG_GEOM_CELL 1
int dg_solver_set_geometry(dg_solver *s, int mode);
/* Autotuner: times trial_steps fused steps (CFL cfl) of the solver's own
 * cases for each of 18 shapes (wpb 1/2/4 x walk 0/1 x 4/8/12 resident waves
 * per CU), restores the state, time and dt it started from and keeps the
 * fastest.  table (may be NULL) receives [18][4] rows (wpb, walk,
 * waves_per_cu, ms_per_step); *ntable the row count. */
int dg_solver_autotune(dg_solver *s, int trial_steps, double cfl, double *table, int *ntable);

/* ------------------------------------------------------------------ */
/* Tangent-linear and adjoint RHS (no reference counterpart)          */
/* ------------------------------------------------------------------ */

/* d_out = L'(w) wd + dL/dgeo geod + dL/d(bc_u, bc_v) (bcud, bcvd) on the
 * declared mesh, L the RHS of dg_rhs_fused.  Geometry tangents may be NULL
 * (zero).  d_scratch holds ncells*3*Nfp*4 doubles. */
int dg_rhs_tangent(dg_ctx *ctx, const double *d_w, const double *d_wd, const double *d_rx,
                   const double *d_ry, const double *d_sx, const double *d_sy,
                   const double *d_fscale, const double *d_nx, const double *d_ny,
                   const double *d_rxd, const double *d_ryd, const double *d_sxd,
                   const double *d_syd, const double *d_fsd, const double *d_nxd,
                   const double *d_nyd, double bcud, double bcvd, double *d_scratch,
                   double *d_out);
/* d_acc += scale L'(w)^T mu; the geometry cotangents d_g* ([ncells][Np] for
 * rx..sy, [ncells][3*Nfp] for fscale, nx, ny, bcu, bcv) += scale (dL/dgeo)^T
 * mu.  Gather-only, deterministic.  d_scratch holds 2*ncells*3*Nfp*4 doubles.
 * d_w, d_mu, d_acc and d_scratch must be 32-byte aligned (DG_ERR_ARG). */
int dg_rhs_adjoint(dg_ctx *ctx, const double *d_w, const double *d_mu, double scale,
                   const double *d_rx, const double *d_ry, const double *d_sx,
                   const double *d_sy, const double *d_fscale, const double *d_nx,
                   const double *d_ny, double *d_scratch, double *d_acc, double *d_grx,
                   double *d_gry, double *d_gsx, double *d_gsy, double *d_gfs, double *d_gnx,
                   double *d_gny, double *d_gbcu, double *d_gbcv);

/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): dg_rhs_adjoint without d_scratch, with the same kernels and
 * results.  The slot cotangents live in a static device array of
 * DG_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
 * process ends, so 2 * ncells * 3 * Nfp * 4 doubles must fit: at most
 * 87,381 cells at N=3 (69,905 at N=4); a larger mesh returns DG_ERR_ARG.
 * The call allocates nothing and only launches kernels on the context's
 * stream.  Every call on one GPU shares that GPU's array, which keeps the
 * last call's contents (nothing needs clearing: every slot is written
 * before it is read), so calls that overlap on one GPU, from other streams
 * or threads, race.  d_w, d_mu and d_acc must be 32-byte aligned. */
#define DG_NOALLOC_SCRATCH_BYTES (64u << 20)
int dg_rhs_adjoint_noalloc(dg_ctx *ctx, const double *d_w, const double *d_mu, double scale,
                           const double *d_rx, const double *d_ry, const double *d_sx,
                           const double *d_sy, const double *d_fscale, const double *d_nx,
                           const double *d_ny, double *d_acc, double *d_grx, double *d_gry,
                           double *d_gsx, double *d_gsy, double *d_gfs, double *d_gnx,
                           double *d_gny, double *d_gbcu, double *d_gbcv);

/* ------------------------------------------------------------------ */
/* Lift sensitivities of the airfoil case (no reference counterpart)  */
/* ------------------------------------------------------------------ */

typedef struct dg_sens dg_sens;
/* O-grid around NACA0012 (as dg_mesh_ogrid) whose surface carries nbu
 * Hicks-Henne bumps on the upper and nbl on the lower surface (peaks at
 * chord positions i/(n+1), width exponent 3; a positive amplitude moves the
 * surface up), free stream (Mach, alpha in degrees) of the context's gamma,
 * p and rho.  The objective is C_L after the recorded number of steps. */
int dg_sens_create(dg_ctx *ctx, int ncirc, int nrad, double stretch, double rfar, int nbu,
                   int nbl, double mach, double alpha_deg, dg_sens **out);
void dg_sens_destroy(dg_sens *s);
/* Sets the bump amplitudes (NULL: all zero), runs nsteps SSPRK3 steps with
 * dt = cfl/dt1 and records the dt sequence, which every later evaluation of
 * this object replays (dt frozen).  Returns C_L. */
int dg_sens_record(dg_sens *s, const double *bumps, int nsteps, double cfl, double *cl);
/* Sets the angle of attack (degrees) of later record / tangent / gradient
 * calls; the recorded dt sequence is kept until the next record. */
int dg_sens_set_alpha(dg_sens *s, double alpha_deg);
/* C_L with the recorded dt sequence at another angle / shape (for finite
 * differences); the recorded point is restored afterwards. */
int dg_sens_eval(dg_sens *s, double alpha_deg, const double *bumps, double *cl);
/* Tangent-linear run at the recorded point: dC_L for the direction
 * (dalpha [deg], dbumps[nbu+nbl]). */
int dg_sens_tangent(dg_sens *s, double dalpha, const double *dbumps, double *cl, double *dcl);
/* Discrete adjoint at the recorded point with checkpoints every ckpt_every
 * steps: dC_L/dalpha (per degree) and dC_L/dbumps[nbu+nbl]. */
int dg_sens_gradient(dg_sens *s, int ckpt_every, double *cl, double *dcl_dalpha,
                     double *dcl_dbumps);
/* Wall-clock times (ms) of the last record, gradient and tangent calls. */
int dg_sens_timing(const dg_sens *s, double *forward_ms, double *gradient_ms, double *tangent_ms);
/* Parts of the last gradient call (ms): the checkpointed forward run, the
 * reverse sweep (segment recomputation included) and the host work (C_L
 * cotangent, free-stream and geometry chains). */
int dg_sens_gradient_breakdown(const dg_sens *s, double *forward_ms, double *reverse_ms,
                               double *host_ms);
/* The recorded dt sequence (n must equal the recorded step count). */
int dg_sens_dts(const dg_sens *s, int n, double *dts);

#ifdef __cplusplus
}
#endif
#endif /* DG_EULER_H */
