// The isentropic-vortex driver in the source form of vortex/src/vortex.cpp,
// built against include/op2_dg.h: its declarations (no boundary classes,
// the vortex constants) and its loop nest (vortex/src/vortex.cpp:195-328):
// init_grid, set_ic at t = 0, one calc_dt with dt = 0.1 / dt1 kept for the
// whole run, per stage neighbour_zero, get_neighbour_q with the edge's node
// coordinates, the exact-state get_bedge_q at t + {0, 1, 1/2} dt, euler_rhs
// and set_workingQ, update_Q per step, and l2_error at the end.  The mesh is
// a legacy-VTK triangle file read by the C ABI (this image has no CGNS) and
// the operator tables are generated for the requested order; the seq kernel
// headers are not included (the interface does not evaluate the kernel
// identifier).
//
//   vortex_seq_form_driver <mesh.vtk> <order> <iter> <out.bin>
//
// Writes q (double [ncells][Np][4]), then t, dt and the four l2_error sums.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "../../include/dg_euler.h"
#include "../../include/op2_dg.h"

static double gam, dt;
static double vortex_x0 = 5.0, vortex_y0 = 0.0, vortex_beta = 5.0;
// the constant tables, named as vortex/src/constants/*.h names them
static double *ones, *r, *s, *Dr, *Ds, *Drw, *Dsw, *LIFT;
static int *FMASK;

int main(int argc, char **argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s <mesh.vtk> <order> <iter> <out.bin>\n", argv[0]);
    return 2;
  }
  const int N = std::atoi(argv[2]), iter = std::atoi(argv[3]);
  const char *out = argv[4];
  const int Np = (N + 1) * (N + 2) / 2, NF = 3 * (N + 1);

  dg_mesh *mesh = nullptr;
  if (dg_mesh_read_vtk(argv[1], &mesh) != DG_OK) {
    std::fprintf(stderr, "mesh: %s\n", dg_last_error(nullptr));
    return 1;
  }
  int numNodes, numCells, numEdges, numBoundaryEdges;
  dg_mesh_sizes(mesh, &numNodes, &numCells, &numEdges, &numBoundaryEdges);
  std::vector<double> coords(2 * (size_t)numNodes);
  std::vector<int> cgnsCells(3 * (size_t)numCells), edge2node(2 * (size_t)numEdges), edge2cell(2 * (size_t)numEdges),
      edgeNum(2 * (size_t)numEdges), bedge2node(2 * (size_t)numBoundaryEdges), bedge2cell(numBoundaryEdges),
      bedgeNum(numBoundaryEdges);
  dg_mesh_get(mesh, "coords", coords.data());
  dg_mesh_get(mesh, "cells", cgnsCells.data());
  dg_mesh_get(mesh, "edge2nodes", edge2node.data());
  dg_mesh_get(mesh, "edge2cells", edge2cell.data());
  dg_mesh_get(mesh, "edgeNum", edgeNum.data());
  dg_mesh_get(mesh, "bedge2nodes", bedge2node.data());
  dg_mesh_get(mesh, "bedge2cells", bedge2cell.data());
  dg_mesh_get(mesh, "bedgeNum", bedgeNum.data());

  op_init(argc, argv, 2);
  gam = 1.4;

  // operator tables for order N (the reference ships the N=4 literals)
  std::vector<double> r_(Np), s_(Np), Dr_(Np * Np), Ds_(Np * Np), Drw_(Np * Np), Dsw_(Np * Np),
      LIFT_(Np * NF), mass(Np * Np), fmass((N + 1) * (N + 1)), ones_(Np, 1.0);
  std::vector<int> FMASK_(NF);
  ones = ones_.data(); r = r_.data(); s = s_.data(); Dr = Dr_.data(); Ds = Ds_.data();
  Drw = Drw_.data(); Dsw = Dsw_.data(); LIFT = LIFT_.data(); FMASK = FMASK_.data();
  if (dg_generate_operators(N, r, s, Dr, Ds, Drw, Dsw, LIFT, FMASK, mass.data(), fmass.data()) != DG_OK) {
    std::fprintf(stderr, "operators: %s\n", dg_last_error(nullptr));
    return 1;
  }

  const size_t K = numCells;
  std::vector<double> nodeX(3 * K), nodeY(3 * K), x(Np * K), y(Np * K), xr(Np * K), yr(Np * K), xs(Np * K),
      ys(Np * K), rx(Np * K), ry(Np * K), sx(Np * K), sy(Np * K), nx(NF * K), ny(NF * K), fscale(NF * K),
      q(4 * Np * K), error(4 * Np * K), workingQ(4 * Np * K), exteriorQ(4 * NF * K), rk1(4 * Np * K),
      rk2(4 * Np * K), rk3(4 * Np * K);

  op_set nodes = op_decl_set(numNodes, "nodes");
  op_set cells = op_decl_set(numCells, "cells");
  op_set edges = op_decl_set(numEdges, "edges");
  op_set bedges = op_decl_set(numBoundaryEdges, "bedges");

  op_map cell2nodes = op_decl_map(cells, nodes, 3, cgnsCells.data(), "cell2nodes");
  op_map edge2nodes = op_decl_map(edges, nodes, 2, edge2node.data(), "edge2nodes");
  op_map edge2cells = op_decl_map(edges, cells, 2, edge2cell.data(), "edge2cells");
  op_map bedge2nodes = op_decl_map(bedges, nodes, 2, bedge2node.data(), "bedge2nodes");
  op_map bedge2cells = op_decl_map(bedges, cells, 1, bedge2cell.data(), "bedge2cells");

  op_dat node_coords = op_decl_dat(nodes, 2, "double", coords.data(), "node_coords");
  op_dat d_nodeX = op_decl_dat(cells, 3, "double", nodeX.data(), "nodeX");
  op_dat d_nodeY = op_decl_dat(cells, 3, "double", nodeY.data(), "nodeY");
  op_dat d_x = op_decl_dat(cells, Np, "double", x.data(), "x");
  op_dat d_y = op_decl_dat(cells, Np, "double", y.data(), "y");
  op_dat d_xr = op_decl_dat(cells, Np, "double", xr.data(), "xr");
  op_dat d_yr = op_decl_dat(cells, Np, "double", yr.data(), "yr");
  op_dat d_xs = op_decl_dat(cells, Np, "double", xs.data(), "xs");
  op_dat d_ys = op_decl_dat(cells, Np, "double", ys.data(), "ys");
  op_dat d_rx = op_decl_dat(cells, Np, "double", rx.data(), "rx");
  op_dat d_ry = op_decl_dat(cells, Np, "double", ry.data(), "ry");
  op_dat d_sx = op_decl_dat(cells, Np, "double", sx.data(), "sx");
  op_dat d_sy = op_decl_dat(cells, Np, "double", sy.data(), "sy");
  op_dat d_nx = op_decl_dat(cells, NF, "double", nx.data(), "nx");
  op_dat d_ny = op_decl_dat(cells, NF, "double", ny.data(), "ny");
  op_dat d_fscale = op_decl_dat(cells, NF, "double", fscale.data(), "fscale");
  op_dat d_q = op_decl_dat(cells, 4 * Np, "double", q.data(), "q");
  op_dat d_error = op_decl_dat(cells, 4 * Np, "double", error.data(), "error");
  op_dat d_workingQ = op_decl_dat(cells, 4 * Np, "double", workingQ.data(), "workingQ");
  op_dat rk[3];
  rk[0] = op_decl_dat(cells, 4 * Np, "double", rk1.data(), "rk1");
  rk[1] = op_decl_dat(cells, 4 * Np, "double", rk2.data(), "rk2");
  rk[2] = op_decl_dat(cells, 4 * Np, "double", rk3.data(), "rk3");
  op_dat d_exteriorQ = op_decl_dat(cells, 4 * NF, "double", exteriorQ.data(), "exteriorQ");
  op_dat d_edgeNum = op_decl_dat(edges, 2, "int", edgeNum.data(), "edgeNum");
  op_dat d_bedgeNum = op_decl_dat(bedges, 1, "int", bedgeNum.data(), "bedgeNum");

  op_decl_const(1, "double", &gam);
  op_decl_const(1, "double", &vortex_x0);
  op_decl_const(1, "double", &vortex_y0);
  op_decl_const(1, "double", &vortex_beta);
  op_decl_const(Np, "double", ones);
  op_decl_const(Np, "double", r);
  op_decl_const(Np, "double", s);
  op_decl_const(Np * Np, "double", Dr);
  op_decl_const(Np * Np, "double", Ds);
  op_decl_const(Np * Np, "double", Drw);
  op_decl_const(Np * Np, "double", Dsw);
  op_decl_const(NF, "int", FMASK);
  op_decl_const(Np * NF, "double", LIFT);

  double t = 0.0;
  op_par_loop(init_grid, "init_grid", cells,
                        op_arg_dat(node_coords, 0, cell2nodes, 2, "double", OP_READ),
                        op_arg_dat(node_coords, 1, cell2nodes, 2, "double", OP_READ),
                        op_arg_dat(node_coords, 2, cell2nodes, 2, "double", OP_READ),
                        op_arg_dat(d_nodeX, -1, OP_ID, 3, "double", OP_WRITE),
                        op_arg_dat(d_nodeY, -1, OP_ID, 3, "double", OP_WRITE),
                        op_arg_dat(d_x, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_y, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_xr, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_yr, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_xs, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_ys, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_rx, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_ry, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_sx, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_sy, -1, OP_ID, Np, "double", OP_WRITE),
                        op_arg_dat(d_nx, -1, OP_ID, NF, "double", OP_WRITE),
                        op_arg_dat(d_ny, -1, OP_ID, NF, "double", OP_WRITE),
                        op_arg_dat(d_fscale, -1, OP_ID, NF, "double", OP_WRITE));

  double ic_t = 0.0;
  op_par_loop(set_ic, "set_ic", cells, op_arg_gbl(&ic_t, 1, "double", OP_READ),
                     op_arg_dat(d_x, -1, OP_ID, Np, "double", OP_READ),
                     op_arg_dat(d_y, -1, OP_ID, Np, "double", OP_READ),
                     op_arg_dat(d_q, -1, OP_ID, 4 * Np, "double", OP_WRITE),
                     op_arg_dat(d_workingQ, -1, OP_ID, 4 * Np, "double", OP_WRITE));

  double dt1 = 0.0;
  op_par_loop(calc_dt, "calc_dt", cells, op_arg_dat(d_q, -1, OP_ID, 4 * Np, "double", OP_READ),
                      op_arg_dat(d_fscale, -1, OP_ID, NF, "double", OP_READ),
                      op_arg_gbl(&dt1, 1, "double", OP_MAX));
  dt = 1.0 / dt1;
  dt *= 0.1;

  const double rk_t_frac[3] = {0.0, 1.0, 0.5};
  for (int i = 0; i < iter; i++) {
    for (int j = 0; j < 3; j++) {
      op_par_loop(neighbour_zero, "neighbour_zero", cells,
                                 op_arg_dat(d_exteriorQ, -1, OP_ID, 4 * NF, "double", OP_WRITE));
      op_par_loop(get_neighbour_q, "get_neighbour_q", edges,
                                  op_arg_dat(d_edgeNum, -1, OP_ID, 2, "int", OP_READ),
                                  op_arg_dat(node_coords, 0, edge2nodes, 2, "double", OP_READ),
                                  op_arg_dat(node_coords, 1, edge2nodes, 2, "double", OP_READ),
                                  op_arg_dat(d_nodeX, 0, edge2cells, 3, "double", OP_READ),
                                  op_arg_dat(d_nodeY, 0, edge2cells, 3, "double", OP_READ),
                                  op_arg_dat(d_nodeX, 1, edge2cells, 3, "double", OP_READ),
                                  op_arg_dat(d_nodeY, 1, edge2cells, 3, "double", OP_READ),
                                  op_arg_dat(d_workingQ, 0, edge2cells, 4 * Np, "double", OP_READ),
                                  op_arg_dat(d_workingQ, 1, edge2cells, 4 * Np, "double", OP_READ),
                                  op_arg_dat(d_exteriorQ, 0, edge2cells, 4 * NF, "double", OP_INC),
                                  op_arg_dat(d_exteriorQ, 1, edge2cells, 4 * NF, "double", OP_INC));
      double t1 = t + rk_t_frac[j] * dt;
      op_par_loop(get_bedge_q, "get_bedge_q", bedges, op_arg_dat(d_bedgeNum, -1, OP_ID, 1, "int", OP_READ),
                              op_arg_dat(node_coords, 0, bedge2nodes, 2, "double", OP_READ),
                              op_arg_dat(node_coords, 1, bedge2nodes, 2, "double", OP_READ),
                              op_arg_dat(d_nodeX, 0, bedge2cells, 3, "double", OP_READ),
                              op_arg_dat(d_nodeY, 0, bedge2cells, 3, "double", OP_READ),
                              op_arg_dat(d_x, 0, bedge2cells, Np, "double", OP_READ),
                              op_arg_dat(d_y, 0, bedge2cells, Np, "double", OP_READ),
                              op_arg_gbl(&t1, 1, "double", OP_READ),
                              op_arg_dat(d_exteriorQ, 0, bedge2cells, 4 * NF, "double", OP_INC));
      op_par_loop(euler_rhs, "euler_rhs", cells, op_arg_dat(d_workingQ, -1, OP_ID, 4 * Np, "double", OP_READ),
                            op_arg_dat(d_exteriorQ, -1, OP_ID, 4 * NF, "double", OP_READ),
                            op_arg_dat(d_rx, -1, OP_ID, Np, "double", OP_READ),
                            op_arg_dat(d_ry, -1, OP_ID, Np, "double", OP_READ),
                            op_arg_dat(d_sx, -1, OP_ID, Np, "double", OP_READ),
                            op_arg_dat(d_sy, -1, OP_ID, Np, "double", OP_READ),
                            op_arg_dat(d_fscale, -1, OP_ID, NF, "double", OP_READ),
                            op_arg_dat(d_nx, -1, OP_ID, NF, "double", OP_READ),
                            op_arg_dat(d_ny, -1, OP_ID, NF, "double", OP_READ),
                            op_arg_dat(rk[j], -1, OP_ID, 4 * Np, "double", OP_WRITE));
      if (j != 2)
        op_par_loop(set_workingQ, "set_workingQ", cells, op_arg_gbl(&dt, 1, "double", OP_READ),
                                 op_arg_gbl(&j, 1, "int", OP_READ),
                                 op_arg_dat(d_q, -1, OP_ID, 4 * Np, "double", OP_READ),
                                 op_arg_dat(rk[0], -1, OP_ID, 4 * Np, "double", OP_READ),
                                 op_arg_dat(rk[1], -1, OP_ID, 4 * Np, "double", OP_READ),
                                 op_arg_dat(d_workingQ, -1, OP_ID, 4 * Np, "double", OP_WRITE));
    }
    op_par_loop(update_Q, "update_Q", cells, op_arg_gbl(&dt, 1, "double", OP_READ),
                         op_arg_dat(d_q, -1, OP_ID, 4 * Np, "double", OP_RW),
                         op_arg_dat(rk[0], -1, OP_ID, 4 * Np, "double", OP_READ),
                         op_arg_dat(rk[1], -1, OP_ID, 4 * Np, "double", OP_READ),
                         op_arg_dat(rk[2], -1, OP_ID, 4 * Np, "double", OP_READ),
                         op_arg_dat(d_workingQ, -1, OP_ID, 4 * Np, "double", OP_WRITE));
    t += dt;
  }

  double r_error = 0.0, ru_error = 0.0, rv_error = 0.0, ener_error = 0.0;
  op_par_loop(l2_error, "l2_error", cells, op_arg_dat(d_x, -1, OP_ID, Np, "double", OP_READ),
                       op_arg_dat(d_y, -1, OP_ID, Np, "double", OP_READ),
                       op_arg_dat(d_q, -1, OP_ID, 4 * Np, "double", OP_READ),
                       op_arg_gbl(&t, 1, "double", OP_READ),
                       op_arg_gbl(&r_error, 1, "double", OP_INC),
                       op_arg_gbl(&ru_error, 1, "double", OP_INC),
                       op_arg_gbl(&rv_error, 1, "double", OP_INC),
                       op_arg_gbl(&ener_error, 1, "double", OP_INC),
                       op_arg_dat(d_error, -1, OP_ID, 4 * Np, "double", OP_WRITE));

  std::vector<double> sol((size_t)4 * Np * op_get_size(cells));
  op_fetch_data(d_q, sol.data());
  FILE *f = std::fopen(out, "wb");
  if (!f) return 1;
  const double tail[6] = {t, dt, r_error, ru_error, rv_error, ener_error};
  std::fwrite(sol.data(), sizeof(double), sol.size(), f);
  std::fwrite(tail, sizeof(double), 6, f);
  std::fclose(f);
  op_timing_output();
  std::cout << "t: " << t << std::endl;
  std::cout << "Rho L2-Error: " << r_error << std::endl;
  op_exit();
  dg_mesh_destroy(mesh);
  return 0;
}
