// Mesh generation and connectivity for the DG solver (host C++).
//
// The reference reads a CGNS mesh produced offline by tools/vtk2CGNS.cpp
// (src/load_mesh.cpp:36-192).  Neither CGNS nor Gmsh exists here, so the
// library generates its meshes itself and restates the connectivity contract
// of tools/vtk2CGNS.cpp:112-199:
//   * triangles are counter-clockwise;
//   * local edge j of a triangle joins vertices (j, j+1 mod 3);
//   * interior edges carry (cell, local edge) of both sides, the lower cell
//     index first; boundary edges carry one (cell, local edge);
//   * edges are listed in ascending (min node, max node) order.
// On top of the edge lists it builds the owner-computes face map used by the
// GPU kernels: for every (cell, local face) the neighbour (cell, face) packed
// as 4*cell+face, or -(1+bc_type) on a boundary face.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dg {

struct Mesh {
  int nnodes = 0, ncells = 0;
  std::vector<double> coords;    // [nnodes][2]
  std::vector<int> cells;        // [ncells][3] counter-clockwise, 0-based
  std::vector<double> cell_xy;   // [ncells][3][2] vertex coordinates per cell
                                 // (unwrapped across periodic seams)
  std::vector<int> node_tag;     // per node: 0 interior, 1 wall ring, 2 far-field ring
  std::vector<int> node_k;       // O-grids: circumferential index of the node's radial line
  std::vector<double> node_w;    // O-grids: d(node y) / d(surface y offset of its line)
  int nedges = 0, nbedges = 0;
  std::vector<int> edge2nodes, edge2cells, edgeNum;      // [nedges][2]
  std::vector<int> bedge2nodes;                          // [nbedges][2]
  std::vector<int> bedge2cells, bedgeNum, bedge_type;    // [nbedges]
};

uint64_t splitmix64(uint64_t &state);
double uniform01(uint64_t &state);

// Periodic [x0,x0+Lx] x [y0,y0+Ly] grid of nx*ny quads, each split by its
// SW->NE diagonal into two counter-clockwise triangles (2*nx*ny cells).
// Every vertex off the seams i=0 and j=0 is moved by U(-jitter*h, jitter*h)
// in x and in y, drawn from splitmix64(seed): two draws per node in
// row-major node order (x first), for every node.
bool gen_periodic(int nx, int ny, double x0, double y0, double Lx, double Ly,
                  uint64_t seed, double jitter, Mesh &m, std::string &err);

// O-grid around the closed-trailing-edge NACA0012 profile of
// naca0012.geo:14: ncirc cells around, nrad cells out to a circle of radius
// rfar centred at (0.5, 0), radial spacing growing geometrically by `stretch`.
// Inner ring edges are walls (bc 2), outer ring edges far field (bc 3).
bool gen_ogrid(int ncirc, int nrad, double stretch, double rfar, Mesh &m,
               std::string &err);

// Jittered nx x ny quad grid over [x0, x0+Lx] x [y0, y0+Ly] (not periodic),
// each quad split SW->NE into two CCW triangles; vertices off the boundary
// jittered by U(-jitter*h, jitter*h) from splitmix64(seed); every boundary
// edge gets type bc_type.
bool gen_rect(int nx, int ny, double x0, double y0, double Lx, double Ly, uint64_t seed, double jitter,
              int bc_type, Mesh &m, std::string &err);

// The same O-grid with the surface points of line k moved by surf_dy[k] in
// y; interior nodes of line k move by (1 - rho_j) surf_dy[k] (rho_j the
// radial blending parameter), the far-field ring stays fixed.
bool gen_ogrid_shaped(int ncirc, int nrad, double stretch, double rfar,
                      const std::vector<double> *surf_dy, Mesh &m, std::string &err);
// Chord position of the surface point of radial line kc, and its side:
// +1 upper surface, -1 lower, 0 leading or trailing edge.
double ogrid_surface_x(int kc, int ncirc);
int ogrid_surface_side(int kc, int ncirc);
// Hicks-Henne bump sin(pi x^(ln 0.5 / ln xpk))^width on (0, 1), peak 1 at xpk.
double hicks_henne(double x, double xpk, double width);

// NACA0012 half thickness at chord position x in [0,1] (naca0012.geo:14).
double naca0012_thickness(double x);

// Builds edge lists from m.cells with the tools/vtk2CGNS.cpp contract.
// Boundary edges get bedge_type from node_tag (both nodes tagged 1 -> 2 wall,
// else 3 far field) when node_tag is present, else 0.
bool build_edges(Mesh &m, std::string &err);

// Face map: fmap[3*cell+face] = 4*nbr_cell+nbr_face for interior faces,
// -(1+type) for boundary faces.  Checks that every face slot is covered
// exactly once and, when cells is non-null, that neighbours traverse their
// shared edge in opposite directions (counter-clockwise cells).
bool build_face_map(int ncells, int nedges, const int *edge2cells, const int *edgeNum,
                    int nbedges, const int *bedge2cells, const int *bedgeNum,
                    const int *bedge_type, const int *cells, std::vector<int> &fmap,
                    std::string &err);

// Legacy-VTK ASCII unstructured grid reader: keeps triangles (cell type 5),
// flips clockwise ones to counter-clockwise as tools/vtk2CGNS.cpp:112-129,
// classifies boundary edges as src/load_mesh.cpp:147-172 (0 inflow,
// 1 outflow, 2 wall).  Returns false on parse errors or unclassified edges.
bool read_vtk(const std::string &path, Mesh &m, std::string &err);

}  // namespace dg
