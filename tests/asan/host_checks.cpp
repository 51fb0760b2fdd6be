// Host-only checks of the library's C++ mesh and operator code, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/asan/Makefile): the
// operator generator at every order it accepts, the periodic and O-grid
// generators (shaped too), the edge and face-map builders, and the
// legacy-VTK reader on a good file and on faulty ones.  Any sanitizer report
// aborts the run; the expected failures must be reported, not crash.
//   host_checks <dir for scratch files> [mesh.vtk ...]
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../../airfoil-op2-fem_amd/csrc/dg_mesh.h"
#include "../../airfoil-op2-fem_amd/csrc/dg_ops.h"

static int failures = 0;
#define CHECK(cond, what)                                 \
  do {                                                    \
    if (!(cond)) {                                        \
      std::fprintf(stderr, "FAILED: %s\n", (what));       \
      failures++;                                         \
    }                                                     \
  } while (0)

static void check_mesh(const dg::Mesh &m, const char *what) {
  std::vector<int> fmap;
  std::string err;
  bool ok = dg::build_face_map(m.ncells, m.nedges, m.edge2cells.data(), m.edgeNum.data(), m.nbedges,
                               m.bedge2cells.data(), m.bedgeNum.data(), m.bedge_type.data(), m.cells.data(),
                               fmap, err);
  CHECK(ok, what);
  for (int c = 0; c < m.ncells; c++) {
    const double *v = &m.cell_xy[6 * (size_t)c];
    const double a = (v[2] - v[0]) * (v[5] - v[1]) - (v[3] - v[1]) * (v[4] - v[0]);
    if (!(a > 0.0)) { CHECK(false, what); break; }
  }
}

static void write(const std::string &path, const std::string &text) {
  std::ofstream f(path);
  f << text;
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  for (int n = 1; n <= 8; n++) {
    dg::Operators o;
    const char *e = nullptr;
    const bool ok = dg::build_operators(n, o, &e);
    CHECK(ok, "build_operators");
    if (!ok) continue;
    double worst = 0.0;  // Dr and Ds annihilate constants
    for (int i = 0; i < o.Np; i++) {
      double a = 0.0, b = 0.0;
      for (int j = 0; j < o.Np; j++) { a += o.Dr[i * o.Np + j]; b += o.Ds[i * o.Np + j]; }
      worst = std::fmax(worst, std::fmax(std::fabs(a), std::fabs(b)));
    }
    CHECK(worst < 1e-10, "Dr 1 = 0");
  }
  {
    dg::Operators o;
    const char *e = nullptr;
    CHECK(!dg::build_operators(0, o, &e) && e, "order 0 rejected");
  }
  std::string err;
  dg::Mesh m;
  CHECK(dg::gen_periodic(17, 9, 0.0, -5.0, 10.0, 10.0, 7, 0.1, m, err), "gen_periodic");
  check_mesh(m, "periodic mesh");
  CHECK(!dg::gen_periodic(2, 2, 0.0, 0.0, 1.0, 1.0, 0, 0.0, m, err), "tiny periodic grid rejected");
  CHECK(dg::gen_ogrid(40, 12, 1.1, 20.0, m, err), "gen_ogrid");
  check_mesh(m, "O-grid");
  std::vector<double> dy(40, 0.0);
  for (int k = 0; k < 40; k++) dy[k] = 1e-3 * dg::hicks_henne(dg::ogrid_surface_x(k, 40), 0.4, 3.0);
  CHECK(dg::gen_ogrid_shaped(40, 12, 1.1, 20.0, &dy, m, err), "gen_ogrid_shaped");
  check_mesh(m, "shaped O-grid");
  std::vector<double> short_dy(3, 0.0);
  CHECK(!dg::gen_ogrid_shaped(40, 12, 1.1, 20.0, &short_dy, m, err), "short offsets rejected");

  // face-map faults: a duplicated edge and an uncovered face
  CHECK(dg::gen_ogrid(24, 6, 1.2, 10.0, m, err), "gen_ogrid small");
  {
    std::vector<int> e2c = m.edge2cells, en = m.edgeNum, fmap;
    e2c.push_back(e2c[0]); e2c.push_back(e2c[1]);
    en.push_back(en[0]); en.push_back(en[1]);
    CHECK(!dg::build_face_map(m.ncells, m.nedges + 1, e2c.data(), en.data(), m.nbedges, m.bedge2cells.data(),
                              m.bedgeNum.data(), m.bedge_type.data(), m.cells.data(), fmap, err),
          "duplicate edge rejected");
    CHECK(!dg::build_face_map(m.ncells, m.nedges, m.edge2cells.data(), m.edgeNum.data(), m.nbedges - 1,
                              m.bedge2cells.data() + 1, m.bedgeNum.data() + 1, m.bedge_type.data() + 1,
                              m.cells.data(), fmap, err),
          "uncovered face rejected");
    std::vector<int> cw = m.cells;
    std::swap(cw[3 * 5 + 1], cw[3 * 5 + 2]);
    CHECK(!dg::build_face_map(m.ncells, m.nedges, m.edge2cells.data(), m.edgeNum.data(), m.nbedges,
                              m.bedge2cells.data(), m.bedgeNum.data(), m.bedge_type.data(), cw.data(), fmap, err),
          "clockwise cell rejected");
  }

  // VTK files given on the command line must read and check
  for (int i = 2; i < argc; i++) {
    CHECK(dg::read_vtk(argv[i], m, err), argv[i]);
    check_mesh(m, argv[i]);
  }
  // faulty VTK files must be reported
  const std::string hdr = "# vtk DataFile Version 2.0\nx\nASCII\nDATASET UNSTRUCTURED_GRID\n";
  const std::string types1 = "CELL_TYPES 1\n5\n";
  const struct { const char *name, *body; } bad[] = {
      {"range.vtk", "POINTS 3 double\n0 0 0\n1 0 0\n0 1 0\nCELLS 1 4\n3 0 1 9\n"},
      {"neg.vtk", "POINTS 3 double\n0 0 0\n1 0 0\n0 1 0\nCELLS 1 4\n3 0 -1 2\n"},
      {"flat.vtk", "POINTS 3 double\n0 0 0\n1 0 0\n2 0 0\nCELLS 1 4\n3 0 1 2\n"},
      {"trunc.vtk", "POINTS 4 double\n0 0 0\n1 0\n"},
      {"count.vtk", "POINTS -5 double\n"},
      {"bigcell.vtk", "POINTS 3 double\n0 0 0\n1 0 0\n0 1 0\nCELLS 1 4\n1000000 0 1 2\n"},
      {"slant.vtk", "POINTS 3 double\n0 0 0\n3 0.5 0\n0 2 0\nCELLS 1 4\n3 0 1 2\n"},
  };
  for (const auto &b : bad) {
    const std::string path = dir + "/" + b.name;
    write(path, hdr + b.body + (std::string(b.name) == "trunc.vtk" || std::string(b.name) == "count.vtk" ? "" : types1));
    CHECK(!dg::read_vtk(path, m, err) && !err.empty(), b.name);
  }
  CHECK(!dg::read_vtk(dir + "/does-not-exist.vtk", m, err), "missing file");
  std::printf("%s\n", failures ? "host checks FAILED" : "host checks ok");
  return failures ? 1 : 0;
}
