// Mesh generation and connectivity (host C++).
#include "dg_mesh.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>

namespace dg {

uint64_t splitmix64(uint64_t &state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform01(uint64_t &state) {
  return (double)(splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

bool gen_periodic(int nx, int ny, double x0, double y0, double Lx, double Ly,
                  uint64_t seed, double jitter, Mesh &m, std::string &err) {
  if (nx < 2 || ny < 2) {
    err = "periodic grid needs at least 2x2 quads";
    return false;
  }
  if ((long long)nx * ny * 2 > (1ll << 29)) {
    err = "periodic grid too large";
    return false;
  }
  const double hx = Lx / nx, hy = Ly / ny;
  m = Mesh();
  m.nnodes = nx * ny;
  m.coords.resize(2 * (size_t)m.nnodes);
  uint64_t st = seed;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double dx = (2.0 * uniform01(st) - 1.0) * jitter * hx;
      double dy = (2.0 * uniform01(st) - 1.0) * jitter * hy;
      size_t n = (size_t)j * nx + i;
      bool seam = (i == 0) || (j == 0);
      m.coords[2 * n] = x0 + i * hx + (seam ? 0.0 : dx);
      m.coords[2 * n + 1] = y0 + j * hy + (seam ? 0.0 : dy);
    }
  m.ncells = 2 * nx * ny;
  m.cells.resize(3 * (size_t)m.ncells);
  m.cell_xy.resize(6 * (size_t)m.ncells);
  auto node = [&](int i, int j) { return (j % ny) * nx + (i % nx); };
  auto put = [&](size_t k, int i, int j, int slot) {
    int n = node(i, j);
    m.cells[3 * k + slot] = n;
    m.cell_xy[6 * k + 2 * slot] = m.coords[2 * (size_t)n] + (i >= nx ? Lx : 0.0);
    m.cell_xy[6 * k + 2 * slot + 1] = m.coords[2 * (size_t)n + 1] + (j >= ny ? Ly : 0.0);
  };
  size_t k = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      // SW, SE, NE and SW, NE, NW: both counter-clockwise
      put(k, i, j, 0); put(k, i + 1, j, 1); put(k, i + 1, j + 1, 2); k++;
      put(k, i, j, 0); put(k, i + 1, j + 1, 1); put(k, i, j + 1, 2); k++;
    }
  return build_edges(m, err);
}

bool gen_rect(int nx, int ny, double x0, double y0, double Lx, double Ly, uint64_t seed, double jitter,
              int bc_type, Mesh &m, std::string &err) {
  if (nx < 1 || ny < 1) {
    err = "rectangle grid needs at least 1x1 quads";
    return false;
  }
  if ((long long)nx * ny * 2 > (1ll << 29)) {
    err = "rectangle grid too large";
    return false;
  }
  if (bc_type < 0 || bc_type > 4) {
    err = "boundary type must be in [0, 4]";
    return false;
  }
  const double hx = Lx / nx, hy = Ly / ny;
  const int nx1 = nx + 1, ny1 = ny + 1;
  m = Mesh();
  m.nnodes = nx1 * ny1;
  m.coords.resize(2 * (size_t)m.nnodes);
  uint64_t st = seed;
  for (int j = 0; j < ny1; j++)
    for (int i = 0; i < nx1; i++) {
      double dx = (2.0 * uniform01(st) - 1.0) * jitter * hx;
      double dy = (2.0 * uniform01(st) - 1.0) * jitter * hy;
      const bool edge = i == 0 || j == 0 || i == nx || j == ny;
      size_t n = (size_t)j * nx1 + i;
      m.coords[2 * n] = x0 + i * hx + (edge ? 0.0 : dx);
      m.coords[2 * n + 1] = y0 + j * hy + (edge ? 0.0 : dy);
    }
  m.ncells = 2 * nx * ny;
  m.cells.resize(3 * (size_t)m.ncells);
  m.cell_xy.resize(6 * (size_t)m.ncells);
  auto put = [&](size_t k, int i, int j, int slot) {
    const int n = j * nx1 + i;
    m.cells[3 * k + slot] = n;
    m.cell_xy[6 * k + 2 * slot] = m.coords[2 * (size_t)n];
    m.cell_xy[6 * k + 2 * slot + 1] = m.coords[2 * (size_t)n + 1];
  };
  size_t k = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      put(k, i, j, 0); put(k, i + 1, j, 1); put(k, i + 1, j + 1, 2); k++;
      put(k, i, j, 0); put(k, i + 1, j + 1, 1); put(k, i, j + 1, 2); k++;
    }
  if (!build_edges(m, err)) return false;
  for (int &t : m.bedge_type) t = bc_type;
  return true;
}

double naca0012_thickness(double x) {
  if (x <= 0.0) return 0.0;
  return 0.594689181 * (0.298222773 * std::sqrt(x) - 0.127125232 * x -
                        0.357907906 * x * x + 0.291984971 * x * x * x -
                        0.105174696 * x * x * x * x);
}

double hicks_henne(double x, double xpk, double width) {
  if (x <= 0.0 || x >= 1.0) return 0.0;
  const double e = std::log(0.5) / std::log(xpk);
  return std::pow(std::sin(M_PI * std::pow(x, e)), width);
}

double ogrid_surface_x(int kc, int ncirc) {
  return 0.5 + 0.5 * std::cos(2.0 * M_PI * kc / ncirc);
}

int ogrid_surface_side(int kc, int ncirc) {
  if (kc == 0 || 2 * kc == ncirc) return 0;
  return 2 * kc < ncirc ? 1 : -1;
}

bool gen_ogrid(int ncirc, int nrad, double stretch, double rfar, Mesh &m,
               std::string &err) {
  return gen_ogrid_shaped(ncirc, nrad, stretch, rfar, nullptr, m, err);
}

bool gen_ogrid_shaped(int ncirc, int nrad, double stretch, double rfar,
                      const std::vector<double> *surf_dy, Mesh &m, std::string &err) {
  if (ncirc < 8 || nrad < 1 || stretch <= 0.0 || rfar <= 1.0) {
    err = "invalid O-grid parameters";
    return false;
  }
  if ((long long)ncirc * nrad * 2 > (1ll << 29)) {
    err = "O-grid too large";
    return false;
  }
  m = Mesh();
  const int nr1 = nrad + 1;
  m.nnodes = ncirc * nr1;
  m.coords.resize(2 * (size_t)m.nnodes);
  m.node_tag.assign(m.nnodes, 0);
  m.node_k.assign(m.nnodes, 0);
  m.node_w.assign(m.nnodes, 0.0);
  if (surf_dy && (int)surf_dy->size() != ncirc) {
    err = "surface offsets need one value per circumferential node";
    return false;
  }
  // radial blending parameter: geometric growth by `stretch` per cell
  std::vector<double> rho(nr1);
  double total = 0.0, h = 1.0;
  for (int j = 0; j < nrad; j++) { total += h; h *= stretch; }
  double acc = 0.0;
  h = 1.0;
  rho[0] = 0.0;
  for (int j = 1; j <= nrad; j++) { acc += h; h *= stretch; rho[j] = acc / total; }
  rho[nrad] = 1.0;
  for (int kc = 0; kc < ncirc; kc++) {
    double phi = 2.0 * M_PI * kc / ncirc;
    double xs = 0.5 + 0.5 * std::cos(phi);
    double ys = naca0012_thickness(xs);
    if (phi > M_PI) ys = -ys;
    if (kc == 0 || 2 * kc == ncirc) ys = 0.0;
    if (surf_dy) ys += (*surf_dy)[kc];
    double xf = 0.5 + rfar * std::cos(phi), yf = rfar * std::sin(phi);
    for (int j = 0; j <= nrad; j++) {
      size_t n = (size_t)j * ncirc + kc;
      m.coords[2 * n] = xs + (xf - xs) * rho[j];
      m.coords[2 * n + 1] = ys + (yf - ys) * rho[j];
      m.node_tag[n] = j == 0 ? 1 : (j == nrad ? 2 : 0);
      m.node_k[n] = kc;
      m.node_w[n] = 1.0 - rho[j];  // d(node y)/d(surface y offset)
    }
  }
  m.ncells = 2 * ncirc * nrad;
  m.cells.resize(3 * (size_t)m.ncells);
  m.cell_xy.resize(6 * (size_t)m.ncells);
  auto node = [&](int kc, int j) { return j * ncirc + (kc % ncirc); };
  size_t k = 0;
  auto put = [&](int a, int b, int c) {
    int v[3] = {a, b, c};
    for (int s = 0; s < 3; s++) {
      m.cells[3 * k + s] = v[s];
      m.cell_xy[6 * k + 2 * s] = m.coords[2 * (size_t)v[s]];
      m.cell_xy[6 * k + 2 * s + 1] = m.coords[2 * (size_t)v[s] + 1];
    }
    k++;
  };
  for (int j = 0; j < nrad; j++)
    for (int kc = 0; kc < ncirc; kc++) {
      // (k,j),(k,j+1),(k+1,j+1) and (k,j),(k+1,j+1),(k+1,j): counter-clockwise
      put(node(kc, j), node(kc, j + 1), node(kc + 1, j + 1));
      put(node(kc, j), node(kc + 1, j + 1), node(kc + 1, j));
    }
  for (size_t c = 0; c < (size_t)m.ncells; c++) {
    const double *v = &m.cell_xy[6 * c];
    double a = (v[2] - v[0]) * (v[5] - v[1]) - (v[3] - v[1]) * (v[4] - v[0]);
    if (!(a > 0.0)) {
      err = "O-grid produced a degenerate or clockwise cell";
      return false;
    }
  }
  return build_edges(m, err);
}

namespace {
struct HalfEdge {
  int a, b, cell, j;
};
}  // namespace

bool build_edges(Mesh &m, std::string &err) {
  const size_t nh = 3 * (size_t)m.ncells;
  std::vector<HalfEdge> h(nh);
  for (int c = 0; c < m.ncells; c++)
    for (int j = 0; j < 3; j++) {
      int p1 = m.cells[3 * (size_t)c + j], p2 = m.cells[3 * (size_t)c + (j + 1) % 3];
      h[3 * (size_t)c + j] = {std::min(p1, p2), std::max(p1, p2), c, j};
    }
  std::sort(h.begin(), h.end(), [](const HalfEdge &x, const HalfEdge &y) {
    if (x.a != y.a) return x.a < y.a;
    if (x.b != y.b) return x.b < y.b;
    if (x.cell != y.cell) return x.cell < y.cell;
    return x.j < y.j;
  });
  m.edge2nodes.clear(); m.edge2cells.clear(); m.edgeNum.clear();
  m.bedge2nodes.clear(); m.bedge2cells.clear(); m.bedgeNum.clear(); m.bedge_type.clear();
  size_t i = 0;
  while (i < nh) {
    size_t n = 1;
    while (i + n < nh && h[i + n].a == h[i].a && h[i + n].b == h[i].b) n++;
    if (h[i].a == h[i].b) { err = "edge with identical points"; return false; }
    if (n > 2) { err = "edge shared by more than two cells"; return false; }
    if (n == 2) {
      if (h[i].cell == h[i + 1].cell) { err = "edge with identical cells"; return false; }
      m.edge2nodes.push_back(h[i].a); m.edge2nodes.push_back(h[i].b);
      m.edge2cells.push_back(h[i].cell); m.edge2cells.push_back(h[i + 1].cell);
      m.edgeNum.push_back(h[i].j); m.edgeNum.push_back(h[i + 1].j);
    } else {
      m.bedge2nodes.push_back(h[i].a); m.bedge2nodes.push_back(h[i].b);
      m.bedge2cells.push_back(h[i].cell);
      m.bedgeNum.push_back(h[i].j);
      int type = 0;
      if (!m.node_tag.empty())
        type = (m.node_tag[h[i].a] == 1 && m.node_tag[h[i].b] == 1) ? 2 : 3;
      m.bedge_type.push_back(type);
    }
    i += n;
  }
  m.nedges = (int)m.edge2cells.size() / 2;
  m.nbedges = (int)m.bedge2cells.size();
  return true;
}

bool build_face_map(int ncells, int nedges, const int *edge2cells, const int *edgeNum,
                    int nbedges, const int *bedge2cells, const int *bedgeNum,
                    const int *bedge_type, const int *cells, std::vector<int> &fmap,
                    std::string &err) {
  const int kUnset = -1000;
  fmap.assign(3 * (size_t)ncells, kUnset);
  for (int e = 0; e < nedges; e++) {
    int L = edge2cells[2 * e], R = edge2cells[2 * e + 1];
    int fL = edgeNum[2 * e], fR = edgeNum[2 * e + 1];
    if (L < 0 || L >= ncells || R < 0 || R >= ncells || fL < 0 || fL > 2 || fR < 0 || fR > 2) {
      err = "edge " + std::to_string(e) + " has an out-of-range cell or local edge";
      return false;
    }
    if (fmap[3 * (size_t)L + fL] != kUnset || fmap[3 * (size_t)R + fR] != kUnset) {
      err = "face slot assigned twice (edge " + std::to_string(e) + ")";
      return false;
    }
    if (cells) {
      int a0 = cells[3 * (size_t)L + fL], a1 = cells[3 * (size_t)L + (fL + 1) % 3];
      int b0 = cells[3 * (size_t)R + fR], b1 = cells[3 * (size_t)R + (fR + 1) % 3];
      if (!(a0 == b1 && a1 == b0)) {
        err = "edge " + std::to_string(e) +
              ": neighbouring cells do not traverse the shared edge in opposite "
              "directions (cells must be counter-clockwise)";
        return false;
      }
    }
    fmap[3 * (size_t)L + fL] = 4 * R + fR;
    fmap[3 * (size_t)R + fR] = 4 * L + fL;
  }
  for (int b = 0; b < nbedges; b++) {
    int c = bedge2cells[b], f = bedgeNum[b], t = bedge_type ? bedge_type[b] : 0;
    if (c < 0 || c >= ncells || f < 0 || f > 2 || t < 0 || t > 4) {
      err = "boundary edge " + std::to_string(b) + " has an out-of-range cell, edge or type";
      return false;
    }
    if (fmap[3 * (size_t)c + f] != kUnset) {
      err = "face slot assigned twice (boundary edge " + std::to_string(b) + ")";
      return false;
    }
    fmap[3 * (size_t)c + f] = -(1 + t);
  }
  for (size_t s = 0; s < fmap.size(); s++)
    if (fmap[s] == kUnset) {
      err = "face of cell " + std::to_string(s / 3) + " (local edge " +
            std::to_string(s % 3) + ") has neither a neighbour nor a boundary edge";
      return false;
    }
  return true;
}

bool read_vtk(const std::string &path, Mesh &m, std::string &err) {
  std::ifstream in(path);
  if (!in) { err = "cannot open " + path; return false; }
  m = Mesh();
  std::string tok;
  std::vector<double> pts;
  std::vector<std::vector<int>> cellv;
  std::vector<int> types;
  while (in >> tok) {
    if (tok == "POINTS") {
      long n = -1; std::string ty;
      in >> n >> ty;
      if (!in || n < 0 || n > (1l << 30)) { err = "bad POINTS header in " + path; return false; }
      pts.resize(3 * n);
      for (long i = 0; i < 3 * n; i++) in >> pts[i];
    } else if (tok == "CELLS") {
      long n = -1, sz = -1;
      in >> n >> sz;
      if (!in || n < 0 || n > (1l << 29)) { err = "bad CELLS header in " + path; return false; }
      cellv.resize(n);
      for (long i = 0; i < n && in; i++) {
        int cnt = -1; in >> cnt;
        if (!in || cnt < 1 || cnt > 64) { err = "bad cell record in " + path; return false; }
        cellv[i].resize(cnt);
        for (int j = 0; j < cnt; j++) in >> cellv[i][j];
      }
    } else if (tok == "CELL_TYPES") {
      long n = -1; in >> n;
      if (!in || n < 0 || n > (1l << 29)) { err = "bad CELL_TYPES header in " + path; return false; }
      types.resize(n);
      for (long i = 0; i < n; i++) in >> types[i];
    }
    if (in.fail()) { err = "truncated or malformed legacy VTK file " + path; return false; }
  }
  if (pts.empty() || cellv.empty() || types.size() != cellv.size()) {
    err = "malformed legacy VTK file " + path;
    return false;
  }
  m.nnodes = (int)(pts.size() / 3);
  m.coords.resize(2 * (size_t)m.nnodes);
  for (int i = 0; i < m.nnodes; i++) {
    m.coords[2 * i] = pts[3 * i];
    m.coords[2 * i + 1] = pts[3 * i + 1];
  }
  for (size_t c = 0; c < cellv.size(); c++) {
    if (types[c] != 5 || cellv[c].size() != 3) continue;
    int p1 = cellv[c][0], p2 = cellv[c][1], p3 = cellv[c][2];
    for (int p : {p1, p2, p3})
      if (p < 0 || p >= m.nnodes) {
        err = "cell " + std::to_string(c) + " names point " + std::to_string(p) + " of " +
              std::to_string(m.nnodes);
        return false;
      }
    const double *x = m.coords.data();
    // tools/vtk2CGNS.cpp:114-125: a clockwise triangle is turned round
    double val = (x[2 * p2] - x[2 * p1]) * (x[2 * p3 + 1] - x[2 * p2 + 1]) -
                 (x[2 * p2 + 1] - x[2 * p1 + 1]) * (x[2 * p3] - x[2 * p2]);
    if (val < 0) std::swap(p1, p3);
    if (!(val != 0.0)) {
      err = "cell " + std::to_string(c) + " has zero area";
      return false;
    }
    m.cells.push_back(p1); m.cells.push_back(p2); m.cells.push_back(p3);
  }
  m.ncells = (int)(m.cells.size() / 3);
  m.cell_xy.resize(6 * (size_t)m.ncells);
  for (size_t c = 0; c < (size_t)m.ncells; c++)
    for (int s = 0; s < 3; s++) {
      int n = m.cells[3 * c + s];
      m.cell_xy[6 * c + 2 * s] = m.coords[2 * (size_t)n];
      m.cell_xy[6 * c + 2 * s + 1] = m.coords[2 * (size_t)n + 1];
    }
  if (!build_edges(m, err)) return false;
  for (int b = 0; b < m.nbedges; b++) {
    int n1 = m.bedge2nodes[2 * b], n2 = m.bedge2nodes[2 * b + 1];
    double x1 = m.coords[2 * n1], y1 = m.coords[2 * n1 + 1];
    double x2 = m.coords[2 * n2], y2 = m.coords[2 * n2 + 1];
    int t;
    if (x1 > -1.0 && x1 < 2.0 && y1 > -1.0 && y1 < 1.0 && x2 > -1.0 && x2 < 2.0 &&
        y2 > -1.0 && y2 < 1.0)
      t = 2;
    else if (y1 == y2)
      t = 2;
    else if (x1 == x2)
      t = x1 < 0.0 ? 0 : 1;
    else {
      err = "unclassified boundary edge " + std::to_string(b);
      return false;
    }
    m.bedge_type[b] = t;
  }
  return true;
}

}  // namespace dg
