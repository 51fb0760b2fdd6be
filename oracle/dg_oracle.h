/*
 * Oracle: CPU restatement of the reference DG Euler kernels (N-generic).
 *
 * TEST INFRASTRUCTURE ONLY.  The shipped library never links, loads or calls
 * this code; tests/, __graft_entry__.smoke() and the CPU-baseline leg of
 * bench.py use it as the checker (and as the timed CPU baseline).
 *
 * Each function restates one reference per-element kernel with the same
 * arithmetic and operation order, generalised from the hardcoded order N=4
 * (Np=15, Nfp=5) to any order.  BLAS calls in the reference
 * (cblas_dgemv/dcopy/daxpy/dscal, OpenBLAS, version unpinned) are restated as
 * the naive netlib loops.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int N, Np, Nfp;
  const double *r, *s, *Dr, *Ds, *Drw, *Dsw, *LIFT;
  const int *FMASK;
  double gam, bc_mach, bc_alpha, bc_p, bc_r, bc_u, bc_v, bc_e;
} or_consts;

/* src/euler_rhs.h:7-23 */
void or_euler_flux(const or_consts *c, const double *q, double *F, double *G,
                   double *rho, double *u, double *v, double *p);
/* src/euler_rhs.h:120-220 */
void or_roe(const or_consts *c, double *flux, const double *nx, const double *ny,
            const double *fscale, const double *q, const double *pQ);
/* src/euler_rhs.h:222-291 */
void or_euler_rhs(const or_consts *c, const double *q, double *exteriorQ,
                  const double *rx, const double *ry, const double *sx,
                  const double *sy, const double *fscale, const double *nx,
                  const double *ny, double *qRHS);
/* src/get_neighbour_q.h:1-89 */
void or_get_neighbour_q(const or_consts *c, const int *edgeNum, const double *xL,
                        const double *yL, const double *xR, const double *yR,
                        const double *qL, const double *qR, double *exteriorQL,
                        double *exteriorQR);
/* src/get_bedge_q.h:1-51 (types 0,1,2) plus type 3, a characteristic far field */
void or_get_bedge_q(const or_consts *c, const int *bedge_type, const int *bedgeNum,
                    const double *nx, const double *ny, const double *q,
                    double *exteriorQ);
/* vortex/src/get_bedge_q.h:1-40 for one boundary edge: exteriorQ += the exact
 * isentropic-vortex state (or_vortex_state, exact = {uinf, beta, x0, y0,
 * period_x}) at the face nodes' coordinates x, y [Np] and time t. */
void or_get_bedge_q_exact(const or_consts *c, const int *bedgeNum, const double *x,
                          const double *y, double t, const double *exact, double *exteriorQ);
/* src/set_workingQ.h:1-12 */
void or_set_workingQ(const or_consts *c, const double *dt, const int *stage,
                     const double *q, const double *k1, const double *k2,
                     double *workingQ);
/* src/update_Q.h:1-12 */
void or_update_Q(const or_consts *c, const double *dt, double *q, const double *rk1,
                 const double *rk2, const double *rk3, double *workingQ);
/* src/calc_dt.h:1-22 */
void or_calc_dt(const or_consts *c, const double *q, const double *fscale, double *dt1);
/* src/neighbour_zero.h:1-5 */
void or_neighbour_zero(const or_consts *c, double *exQ);
/* src/init_grid.h:1-106 */
void or_init_grid(const or_consts *c, const double *n0, const double *n1,
                  const double *n2, double *nodeX, double *nodeY, double *x,
                  double *y, double *xr, double *yr, double *xs, double *ys,
                  double *rx, double *ry, double *sx, double *sy, double *nx,
                  double *ny, double *fscale);
/* src/set_ic.h:1-12 */
void or_set_ic(const or_consts *c, double *q, double *workingQ);

/* ---- set-level loops (the op_par_loop bodies of the seq backend) ---- */

typedef struct {
  int ncells, nedges, nbedges;
  const int *edge2cells;   /* [E][2] */
  const int *edgeNum;      /* [E][2] */
  const int *bedge2cells;  /* [B]    */
  const int *bedgeNum;     /* [B]    */
  const int *bedge_type;   /* [B]    */
  const double *nodeX, *nodeY;                 /* [K][3]      */
  const double *rx, *ry, *sx, *sy;             /* [K][Np]     */
  const double *nx, *ny, *fscale;              /* [K][3*Nfp]  */
  /* boundary type 4 (exact isentropic-vortex state): node coordinates
   * [K][Np] and the vortex {uinf, beta, x0, y0, period_x}; may be NULL when
   * no boundary edge has type 4 */
  const double *x, *y;
  const double *exact;
} or_mesh;

typedef struct {
  double *q, *workingQ, *exteriorQ, *rk1, *rk2, *rk3;
} or_state;

/* init_grid over all cells: cell_xy is [K][3][2] vertex coordinates. */
void or_init_grid_all(const or_consts *c, int ncells, const double *cell_xy,
                      double *nodeX, double *nodeY, double *x, double *y,
                      double *xr, double *yr, double *xs, double *ys, double *rx,
                      double *ry, double *sx, double *sy, double *nx, double *ny,
                      double *fscale);

/* calc_dt over all cells with OP_MAX semantics: returns max(dt1_in, local). */
double or_calc_dt_all(const or_consts *c, const or_mesh *m, const double *q,
                      double dt1_in, int nthreads);

/* One SSPRK3 stage-sequence step, as src/airfoil.cpp:253-317, starting at
 * time 0 (or_run passes the running time to the exact-state boundaries,
 * which evaluate at t + {0, 1, 1/2} dt, vortex/src/vortex.cpp:228,248-258).
 * dt is the step size for this step.  nthreads<=1 runs the seq loop order; >1 runs the
 * openmp structure (contiguous chunks for direct loops, coloured edge loops,
 * per-thread partials for the max). */
void or_ssprk3_step(const or_consts *c, const or_mesh *m, or_state *st, double dt,
                    int nthreads, const int *edge_colour_ptr, const int *edge_colour_list,
                    int ncolours);

/* Runs nsteps steps, recomputing dt = cfl / dt1 after every step as
 * src/airfoil.cpp:231-237,322-331.  Returns the final time; *dt_out gets the
 * dt that the next step would use.  Per-kernel wall times (s) accumulate into
 * ktime[6] = {get_neighbour_q, get_bedge_q, euler_rhs, set_workingQ,
 * update_Q, calc_dt} when non-NULL. */
double or_run(const or_consts *c, const or_mesh *m, or_state *st, int nsteps,
              double cfl, int nthreads, double *dt_out, double *ktime);

/* nsteps steps of the fixed size dt from time 0 (the step sequence of or_run
 * without the CFL recomputation); returns the final time. */
double or_run_dt(const or_consts *c, const or_mesh *m, or_state *st, int nsteps, double dt,
                 int nthreads);

/* Greedy colouring of edges so that no two edges of one colour share a cell
 * (the race-freedom contract of OP2's plan colouring).  colour_ptr has
 * room for ncolours+1 entries (<= 8), colour_list for nedges. */
int or_colour_edges(int ncells, int nedges, const int *edge2cells, int *colour_ptr,
                    int *colour_list);

/* Edge builder with the contract of tools/vtk2CGNS.cpp:112-199: local edge
 * j of a CCW triangle joins vertices (j, j+1 mod 3); edges are emitted in
 * ascending (min node, max node) order; cells[0]/num[0] is the first cell
 * (lowest index) that holds the edge.  Returns the number of interior edges
 * in *nedges and boundary edges in *nbedges.  Output arrays must have room
 * for 3*ncells entries (edge2cells/edgeNum: 2 per edge). */
int or_build_edges(int ncells, const int *cells, int *nedges, int *edge2cells,
                   int *edgeNum, int *edge2nodes, int *nbedges, int *bedge2cells,
                   int *bedgeNum, int *bedge2nodes);

/* Isentropic vortex (vortex/src/set_ic.h:1-28, with the free-stream speed
 * uinf and the advection term in both velocity components). */
void or_vortex_state(double gam, double uinf, double beta, double x0, double y0,
                     double t, double x, double y, double period_x, double *q4);

/* Mass-weighted L2 norm of (q - exact vortex) per variable over the mesh:
 * sqrt(sum_k J_k e^T M e). */
void or_vortex_l2(const or_consts *c, int ncells, const double *x, const double *y,
                  const double *J, const double *mass, const double *q, double gam,
                  double uinf, double beta, double x0, double y0, double t,
                  double period_x, double *err4);

/* Pressure force on the faces listed in (wcell, wface): sum over faces and
 * face nodes of facew[i] * fscale * J * p * (nx, ny), faces in list order.
 * No reference counterpart (surface C_L of the airfoil case). */
void or_wall_forces(const or_consts *c, int nwall, const int *wcell, const int *wface,
                    const double *q, const double *nx, const double *ny,
                    const double *fscale, const double *J, const double *facew,
                    double *fxy);

/* Pressure coefficient (p - bc_p) / qdyn at the Nfp nodes of each face listed
 * in (wcell, wface), p from the state at the face node (FMASK), faces in list
 * order: cp[b * Nfp + i].  No reference counterpart (surface C_p). */
void or_wall_cp(const or_consts *c, int nwall, const int *wcell, const int *wface,
                const double *q, double qdyn, double *cp);

#ifdef __cplusplus
}
#endif
#endif
