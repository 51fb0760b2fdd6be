// OP2-compatible host interface (include/op2_dg.h) over the dg_euler.h C
// ABI: sets, maps and datasets with device mirrors, name-dispatched
// constants, and the nine par_loop host stubs of src/airfoil_op.cpp:17-98
// mapped onto the drop-in entry points.  Host code only.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dg_euler.h"
#include "../../include/op2_dg.h"

struct op_set_core {
  int size = 0;
  std::string name;
};

struct op_map_core {
  op_set from = nullptr, to = nullptr;
  int dim = 0;
  std::vector<int> data;
  std::string name;
};

struct op_dat_core {
  op_set set = nullptr;
  int dim = 0, type_size = 0;
  std::string type, name;
  char *host = nullptr;   // the user's array (OP2 adopts it)
  void *dev = nullptr;    // device mirror
  size_t bytes() const { return (size_t)set->size * dim * type_size; }
};

namespace {

struct Timing {
  double ms = 0.0;
  int calls = 0;
};

struct State {
  std::vector<std::unique_ptr<op_set_core>> sets;
  std::vector<std::unique_ptr<op_map_core>> maps;
  std::vector<std::unique_ptr<op_dat_core>> dats;
  struct Const {
    int dim, type_size;
    std::string type;
    std::vector<char> bytes;
  };
  std::map<std::string, Const> consts;  // stashed until the context exists
  dg_ctx *ctx = nullptr;
  // connectivity the context's face map was built from
  op_set cells = nullptr;
  op_map c2n = nullptr, e2c = nullptr, b2c = nullptr;
  op_dat enum_ = nullptr, bnum = nullptr, btype = nullptr;
  bool mesh = false;
  double *cell_xy = nullptr;
  std::map<std::string, Timing> timing;
};

State &st() {
  static State s;
  return s;
}

[[noreturn]] void die(const std::string &msg) { throw std::runtime_error("op2_dg: " + msg); }

void check(int rc, const char *what) {
  if (rc != DG_OK) {
    const char *e = dg_last_error(st().ctx);
    die(std::string(what) + ": " + (e ? e : "error"));
  }
}

void hcheck(hipError_t e, const char *what) {
  if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

// Order from the stashed constants (r has Np entries) or, failing that,
// from a state-like dat (4 * Np doubles per cell).
int infer_order() {
  State &s = st();
  int np = 0;
  auto it = s.consts.find("r");
  if (it != s.consts.end()) np = it->second.dim;
  for (int n = 1; n <= 6; n++)
    if ((n + 1) * (n + 2) / 2 == np) return n;
  die("cannot infer the order: declare the constant r (Np values) before the first par_loop");
}

void ensure_ctx() {
  State &s = st();
  if (s.ctx) return;
  const int N = infer_order();
  check(dg_ctx_create(N, &s.ctx), "op_par_loop: context");
  // scalars first (they derive nothing), then the tables
  for (auto &kv : s.consts)
    check(dg_set_const(s.ctx, kv.first.c_str(), kv.second.dim, kv.second.bytes.data()),
          ("op_decl_const2(\"" + kv.first + "\")").c_str());
}

op_dat find_int_dat(op_set set, int dim, const char *name) {
  for (auto &d : st().dats)
    if (d->set == set && d->dim == dim && d->type_size == (int)sizeof(int) && d->name == name) return d.get();
  return nullptr;
}

// Declares the mesh to the context from the registered connectivity.
void declare_mesh(op_set cells) {
  State &s = st();
  ensure_ctx();
  if (!s.e2c)
    for (auto &m : s.maps)
      if (m->to == cells && m->dim == 2) { s.e2c = m.get(); break; }
  if (s.e2c && !s.enum_)
    for (auto &d : s.dats)
      if (d->set == s.e2c->from && d->dim == 2 && d->type_size == (int)sizeof(int)) { s.enum_ = d.get(); break; }
  if (!s.b2c)
    for (auto &m : s.maps)
      if (m->to == cells && m->dim == 1) { s.b2c = m.get(); break; }
  if (s.b2c && !s.bnum) s.bnum = find_int_dat(s.b2c->from, 1, "bedgeNum");
  if (s.b2c && !s.btype) s.btype = find_int_dat(s.b2c->from, 1, "bedge_type");
  if (!s.c2n)
    for (auto &m : s.maps)
      if (m->from == cells && m->dim == 3) { s.c2n = m.get(); break; }
  const int ne = s.e2c ? s.e2c->from->size : 0;
  const int nb = s.b2c ? s.b2c->from->size : 0;
  if (ne > 0 && !s.enum_) die("no int dat of dimension 2 on the edge set (edgeNum)");
  if (nb > 0 && !s.bnum) die("boundary edges need an int dat named bedgeNum");
  // Without a bedge_type dat every boundary edge carries the exact vortex
  // state: the vortex case declares no boundary classes
  // (vortex/src/vortex.cpp:123-168) and applies the exact state on every
  // boundary edge (:250-260).
  std::vector<int> exact;
  if (nb > 0 && !s.btype) exact.assign((size_t)nb, DG_BC_EXACT);
  check(dg_decl_mesh(s.ctx, cells->size, s.c2n ? s.c2n->data.data() : nullptr, ne,
                     ne ? s.e2c->data.data() : nullptr, ne ? (const int *)s.enum_->host : nullptr, nb,
                     nb ? s.b2c->data.data() : nullptr, nb ? (const int *)s.bnum->host : nullptr,
                     nb ? (s.btype ? (const int *)s.btype->host : exact.data()) : nullptr),
        "op_par_loop: mesh");
  s.cells = cells;
  s.mesh = true;
}

void ensure_mesh(op_set cells) {
  State &s = st();
  if (!s.mesh || s.cells != cells) declare_mesh(cells);
}

double *dptr(const op_arg &a, op_set set, int dim, const char *loop, int k) {
  if (!a.dat) die(std::string(loop) + ": argument " + std::to_string(k) + " is not a dataset");
  if (a.dat->set != set) die(std::string(loop) + ": argument " + std::to_string(k) + " (" + a.dat->name + ") is on the wrong set");
  if (a.dat->dim != dim || a.dat->type_size != (int)sizeof(double))
    die(std::string(loop) + ": argument " + std::to_string(k) + " (" + a.dat->name + ") has the wrong dimension or type");
  return (double *)a.dat->dev;
}

void need_direct(const op_arg &a, const char *loop, int k) {
  if (a.map != OP_ID) die(std::string(loop) + ": argument " + std::to_string(k) + " must be direct (OP_ID)");
}

void need_mapped(const op_arg &a, op_set from, int idx, const char *loop, int k) {
  if (!a.map || a.map->from != from || a.idx != idx)
    die(std::string(loop) + ": argument " + std::to_string(k) + " must be mapped from " + from->name + " with index " +
        std::to_string(idx));
}

void sync_and_time(const char *name, std::chrono::steady_clock::time_point t0) {
  hcheck(hipDeviceSynchronize(), name);
  Timing &t = st().timing[name];
  t.ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  t.calls++;
}

int Np() {
  int n, np, nfp;
  check(dg_order(st().ctx, &n, &np, &nfp), "dg_order");
  return np;
}

}  // namespace

void op_init(int, char **, int) {}

void op_exit() {
  State &s = st();
  for (auto &d : s.dats) (void)hipFree(d->dev);
  (void)hipFree(s.cell_xy);
  if (s.ctx) dg_ctx_destroy(s.ctx);
  s = State();
}

op_set op_decl_set(int size, char const *name) {
  if (size < 0) die("op_decl_set: negative size");
  auto p = std::make_unique<op_set_core>();
  p->size = size;
  p->name = name ? name : "";
  st().sets.push_back(std::move(p));
  return st().sets.back().get();
}

op_map op_decl_map(op_set from, op_set to, int dim, int *imap, char const *name) {
  if (!from || !to || dim < 1 || (!imap && from->size > 0)) die("op_decl_map: bad arguments");
  auto p = std::make_unique<op_map_core>();
  p->from = from; p->to = to; p->dim = dim;
  p->data.assign(imap, imap + (size_t)from->size * dim);
  for (int v : p->data)
    if (v < 0 || v >= to->size) die(std::string("op_decl_map: ") + (name ? name : "") + " has an index out of range");
  p->name = name ? name : "";
  st().maps.push_back(std::move(p));
  return st().maps.back().get();
}

op_dat op_decl_dat_char(op_set set, int dim, char const *type, int type_size, char *data, char const *name) {
  if (!set || dim < 1 || type_size < 1) die("op_decl_dat: bad arguments");
  auto p = std::make_unique<op_dat_core>();
  p->set = set; p->dim = dim; p->type_size = type_size;
  p->type = type ? type : ""; p->name = name ? name : "";
  p->host = data;
  const size_t b = p->bytes();
  hcheck(hipMalloc(&p->dev, b ? b : 8), "op_decl_dat: device mirror");
  if (b && data) hcheck(hipMemcpy(p->dev, data, b, hipMemcpyHostToDevice), "op_decl_dat: upload");
  else if (b) hcheck(hipMemset(p->dev, 0, b), "op_decl_dat: zero");
  st().dats.push_back(std::move(p));
  return st().dats.back().get();
}

void op_decl_const_char(int dim, char const *type, int type_size, char *data, char const *name) {
  if (!name || !data || dim < 1) die("op_decl_const2: bad arguments");
  State &s = st();
  State::Const c{dim, type_size, type ? type : "", std::vector<char>(data, data + (size_t)dim * type_size)};
  s.consts[name] = c;
  if (s.ctx) check(dg_set_const(s.ctx, name, dim, data), (std::string("op_decl_const2(\"") + name + "\")").c_str());
}

op_arg op_arg_dat(op_dat dat, int idx, op_map map, int dim, char const *type, op_access acc) {
  if (!dat) die("op_arg_dat: null dataset");
  if (dim != dat->dim) die("op_arg_dat: dimension of " + dat->name + " does not match its declaration");
  if (map && (idx < 0 || idx >= map->dim)) die("op_arg_dat: map index out of range for " + dat->name);
  return op_arg{dat, map, idx, dim, type, acc, nullptr};
}

op_arg op_arg_gbl_char(char *data, int dim, const char *type, int, op_access acc) {
  if (!data) die("op_arg_gbl: null pointer");
  return op_arg{nullptr, nullptr, -1, dim, type, acc, data};
}

void op_timers(double *cpu, double *wall) {
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  if (cpu) *cpu = t;
  if (wall) *wall = t;
}

int op_get_size(op_set set) { return set ? set->size : 0; }

void op_fetch_data_char(op_dat dat, char *dest) {
  if (!dat || !dest) die("op_fetch_data: bad arguments");
  hcheck(hipDeviceSynchronize(), "op_fetch_data");
  if (dat->bytes()) hcheck(hipMemcpy(dest, dat->dev, dat->bytes(), hipMemcpyDeviceToHost), "op_fetch_data");
}

void op_timing_output() {
  std::printf("  %-20s %8s %12s\n", "loop", "calls", "time (ms)");
  for (auto &kv : st().timing) std::printf("  %-20s %8d %12.3f\n", kv.first.c_str(), kv.second.calls, kv.second.ms);
}

// --------------------------------------------------------------------------
// par_loop host stubs
// --------------------------------------------------------------------------

void op_par_loop_init_grid(char const *name, op_set cells, op_arg a0, op_arg a1, op_arg a2, op_arg a3,
                           op_arg a4, op_arg a5, op_arg a6, op_arg a7, op_arg a8, op_arg a9, op_arg a10,
                           op_arg a11, op_arg a12, op_arg a13, op_arg a14, op_arg a15, op_arg a16,
                           op_arg a17) {
  auto t0 = std::chrono::steady_clock::now();
  State &s = st();
  const op_arg v[3] = {a0, a1, a2};
  for (int k = 0; k < 3; k++) {
    need_mapped(v[k], cells, k, "init_grid", k);
    if (v[k].map != v[0].map || v[k].dat != v[0].dat || v[k].dat->dim != 2) die("init_grid: vertex arguments");
  }
  s.c2n = v[0].map;
  ensure_mesh(cells);
  const int K = cells->size, np = Np(), nf = 3 * (int)std::lround((std::sqrt(8.0 * np + 1.0) - 1.0) / 2.0);
  // per-cell vertex coordinates through cell2nodes (host side: the coordinates
  // never change on the device)
  std::vector<double> cxy((size_t)K * 6);
  const double *xy = (const double *)a0.dat->host;
  for (int c = 0; c < K; c++)
    for (int k = 0; k < 3; k++) {
      const int n = a0.map->data[(size_t)c * 3 + k];
      cxy[(size_t)c * 6 + 2 * k] = xy[2 * (size_t)n];
      cxy[(size_t)c * 6 + 2 * k + 1] = xy[2 * (size_t)n + 1];
    }
  (void)hipFree(s.cell_xy);
  s.cell_xy = nullptr;
  hcheck(hipMalloc(&s.cell_xy, sizeof(double) * cxy.size() + 8), "init_grid");
  hcheck(hipMemcpy(s.cell_xy, cxy.data(), sizeof(double) * cxy.size(), hipMemcpyHostToDevice), "init_grid");
  const op_arg o[15] = {a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13, a14, a15, a16, a17};
  double *p[15];
  for (int k = 0; k < 15; k++) {
    need_direct(o[k], "init_grid", 3 + k);
    p[k] = dptr(o[k], cells, k < 2 ? 3 : (k < 12 ? np : nf), "init_grid", 3 + k);
  }
  check(dg_init_grid(s.ctx, s.cell_xy, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11],
                     p[12], p[13], p[14]),
        name);
  sync_and_time(name, t0);
}

void op_par_loop_set_ic(char const *name, op_set cells, op_arg a0, op_arg a1) {
  auto t0 = std::chrono::steady_clock::now();
  ensure_mesh(cells);
  const int np = Np();
  need_direct(a0, "set_ic", 0); need_direct(a1, "set_ic", 1);
  check(dg_set_ic(st().ctx, dptr(a0, cells, 4 * np, "set_ic", 0), dptr(a1, cells, 4 * np, "set_ic", 1)), name);
  sync_and_time(name, t0);
}

void op_par_loop_neighbour_zero(char const *name, op_set cells, op_arg a0) {
  auto t0 = std::chrono::steady_clock::now();
  ensure_mesh(cells);
  need_direct(a0, "neighbour_zero", 0);
  check(dg_neighbour_zero(st().ctx, (double *)a0.dat->dev), name);
  sync_and_time(name, t0);
}

void op_par_loop_calc_dt(char const *name, op_set cells, op_arg a0, op_arg a1, op_arg a2) {
  auto t0 = std::chrono::steady_clock::now();
  ensure_mesh(cells);
  const int np = Np();
  need_direct(a0, "calc_dt", 0); need_direct(a1, "calc_dt", 1);
  if (!a2.gbl || a2.acc != OP_MAX || a2.dim != 1) die("calc_dt: argument 2 must be a global OP_MAX double");
  const int nf = a1.dat->dim;
  check(dg_calc_dt(st().ctx, dptr(a0, cells, 4 * np, "calc_dt", 0), dptr(a1, cells, nf, "calc_dt", 1),
                   (double *)a2.gbl),
        name);
  sync_and_time(name, t0);
}

void op_par_loop_get_neighbour_q(char const *name, op_set edges, op_arg a0, op_arg a1, op_arg a2, op_arg a3,
                                 op_arg a4, op_arg a5, op_arg a6, op_arg a7, op_arg a8) {
  auto t0 = std::chrono::steady_clock::now();
  State &s = st();
  need_direct(a0, "get_neighbour_q", 0);
  // nodeX[0], nodeY[0], nodeX[1], nodeY[1], workingQ[0], workingQ[1],
  // exteriorQ[0], exteriorQ[1] through edge2cells (src/airfoil.cpp:268-276)
  const op_arg m[8] = {a1, a2, a3, a4, a5, a6, a7, a8};
  const int side[8] = {0, 0, 1, 1, 0, 1, 0, 1};
  for (int k = 0; k < 8; k++) {
    need_mapped(m[k], edges, side[k], "get_neighbour_q", k + 1);
    if (m[k].map != a1.map) die("get_neighbour_q: every mapped argument must use the same edge map");
  }
  if (a1.dat != a3.dat || a2.dat != a4.dat || a5.dat != a6.dat || a7.dat != a8.dat)
    die("get_neighbour_q: the two sides of an edge must read the same datasets");
  op_set cells = a1.map->to;
  if (s.e2c != a1.map || s.enum_ != a0.dat) {
    s.e2c = a1.map;
    s.enum_ = a0.dat;
    s.mesh = false;
  }
  ensure_mesh(cells);
  const int np = Np(), nf = a7.dat->dim / 4;
  check(dg_get_neighbour_q(s.ctx, dptr(a1, cells, 3, name, 1), dptr(a2, cells, 3, name, 2),
                           dptr(a5, cells, 4 * np, name, 5), dptr(a7, cells, 4 * nf, name, 7)),
        name);
  sync_and_time(name, t0);
}

void op_par_loop_get_bedge_q(char const *name, op_set bedges, op_arg a0, op_arg a1, op_arg a2, op_arg a3,
                             op_arg a4, op_arg a5) {
  auto t0 = std::chrono::steady_clock::now();
  State &s = st();
  need_direct(a0, "get_bedge_q", 0); need_direct(a1, "get_bedge_q", 1);
  const op_arg m[4] = {a2, a3, a4, a5};
  for (int k = 0; k < 4; k++) need_mapped(m[k], bedges, 0, "get_bedge_q", k + 2);
  op_set cells = a2.map->to;
  if (s.b2c != a2.map || s.btype != a0.dat || s.bnum != a1.dat) {
    s.b2c = a2.map;
    s.btype = a0.dat;
    s.bnum = a1.dat;
    s.mesh = false;
  }
  ensure_mesh(cells);
  const int np = Np(), nf = a2.dat->dim;
  check(dg_get_bedge_q(s.ctx, dptr(a2, cells, nf, name, 2), dptr(a3, cells, nf, name, 3),
                       dptr(a4, cells, 4 * np, name, 4), dptr(a5, cells, 4 * nf, name, 5)),
        name);
  sync_and_time(name, t0);
}

void op_par_loop_euler_rhs(char const *name, op_set cells, op_arg a0, op_arg a1, op_arg a2, op_arg a3, op_arg a4,
                           op_arg a5, op_arg a6, op_arg a7, op_arg a8, op_arg a9) {
  auto t0 = std::chrono::steady_clock::now();
  ensure_mesh(cells);
  const op_arg a[10] = {a0, a1, a2, a3, a4, a5, a6, a7, a8, a9};
  for (int k = 0; k < 10; k++) need_direct(a[k], "euler_rhs", k);
  const int np = Np(), nf = a6.dat->dim;
  check(dg_euler_rhs(st().ctx, dptr(a0, cells, 4 * np, name, 0), dptr(a1, cells, 4 * nf, name, 1),
                     dptr(a2, cells, np, name, 2), dptr(a3, cells, np, name, 3), dptr(a4, cells, np, name, 4),
                     dptr(a5, cells, np, name, 5), dptr(a6, cells, nf, name, 6), dptr(a7, cells, nf, name, 7),
                     dptr(a8, cells, nf, name, 8), dptr(a9, cells, 4 * np, name, 9)),
        name);
  sync_and_time(name, t0);
}

void op_par_loop_set_workingQ(char const *name, op_set cells, op_arg a0, op_arg a1, op_arg a2, op_arg a3,
                              op_arg a4, op_arg a5) {
  auto t0 = std::chrono::steady_clock::now();
  ensure_mesh(cells);
  if (!a0.gbl || !a1.gbl) die("set_workingQ: arguments 0 and 1 must be globals (dt, stage)");
  const int np = Np();
  const op_arg a[4] = {a2, a3, a4, a5};
  for (int k = 0; k < 4; k++) need_direct(a[k], "set_workingQ", k + 2);
  check(dg_set_workingQ(st().ctx, *(const double *)a0.gbl, *(const int *)a1.gbl, dptr(a2, cells, 4 * np, name, 2),
                        dptr(a3, cells, 4 * np, name, 3), dptr(a4, cells, 4 * np, name, 4),
                        dptr(a5, cells, 4 * np, name, 5)),
        name);
  sync_and_time(name, t0);
}

void op_par_loop_update_Q(char const *name, op_set cells, op_arg a0, op_arg a1, op_arg a2, op_arg a3, op_arg a4,
                          op_arg a5) {
  auto t0 = std::chrono::steady_clock::now();
  ensure_mesh(cells);
  if (!a0.gbl) die("update_Q: argument 0 must be a global (dt)");
  const int np = Np();
  const op_arg a[5] = {a1, a2, a3, a4, a5};
  for (int k = 0; k < 5; k++) need_direct(a[k], "update_Q", k + 1);
  check(dg_update_Q(st().ctx, *(const double *)a0.gbl, dptr(a1, cells, 4 * np, name, 1),
                    dptr(a2, cells, 4 * np, name, 2), dptr(a3, cells, 4 * np, name, 3),
                    dptr(a4, cells, 4 * np, name, 4), dptr(a5, cells, 4 * np, name, 5)),
        name);
  sync_and_time(name, t0);
}

// --------------------------------------------------------------------------
// The vortex case's loop forms (vortex/src/vortex.cpp:195-328), reached
// through op_par_loop by name and argument count
// --------------------------------------------------------------------------
namespace {

// The exact vortex of the vortex case: the declared vortex_x0, vortex_y0,
// vortex_beta (defaults vortex/src/constants/constant_util.h:6-8) and the
// free-stream speed 1 that vortex/src/set_ic.h:8 hardcodes.
void set_vortex_exact_state() {
  State &s = st();
  auto val = [&](const char *n, double d) {
    auto it = s.consts.find(n);
    if (it == s.consts.end()) return d;
    if (it->second.type_size != (int)sizeof(double) || it->second.dim != 1)
      die(std::string("op_decl_const: ") + n + " must be one double");
    double v;
    std::memcpy(&v, it->second.bytes.data(), sizeof(double));
    return v;
  };
  check(dg_set_exact_state(s.ctx, 1.0, val("vortex_beta", 5.0), val("vortex_x0", 5.0), val("vortex_y0", 0.0), 0.0),
        "vortex exact state");
}

double gbl_double(const op_arg &a, const char *loop, int k) {
  if (!a.gbl || a.dim != 1) die(std::string(loop) + ": argument " + std::to_string(k) + " must be a global double");
  return *(const double *)a.gbl;
}

// set_ic(t, x, y, q, workingQ) (vortex/src/vortex.cpp:208-213)
void vortex_set_ic(char const *name, op_set cells, const op_arg *a) {
  auto t0 = std::chrono::steady_clock::now();
  ensure_mesh(cells);
  set_vortex_exact_state();
  const int np = Np();
  const double t = gbl_double(a[0], "set_ic", 0);
  for (int k = 1; k < 5; k++) need_direct(a[k], "set_ic", k);
  check(dg_set_ic_vortex(st().ctx, t, dptr(a[1], cells, np, name, 1), dptr(a[2], cells, np, name, 2),
                         dptr(a[3], cells, 4 * np, name, 3), dptr(a[4], cells, 4 * np, name, 4)),
        name);
  sync_and_time(name, t0);
}

// get_neighbour_q(edgeNum, node_coords[0..1] via edge2nodes, then the
// airfoil form's eight arguments) (vortex/src/vortex.cpp:236-247): the node
// coordinates only serve the reference's orientation test, which the face
// map replaces
void vortex_get_neighbour_q(char const *name, op_set edges, const op_arg *a) {
  for (int k = 1; k < 3; k++) {
    if (!a[k].map || a[k].map->from != edges || a[k].idx != k - 1 || a[k].dat != a[1].dat || a[k].map != a[1].map)
      die(std::string("get_neighbour_q: argument ") + std::to_string(k) + " must be the node coordinates through the "
          "edge-to-node map with index " + std::to_string(k - 1));
  }
  op_par_loop_get_neighbour_q(name, edges, a[0], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10]);
}

// get_bedge_q(bedgeNum, node_coords[0..1] via bedge2nodes, nodeX, nodeY, x,
// y via bedge2cells, t, exteriorQ) (vortex/src/vortex.cpp:250-260): the exact
// state at time t on every boundary edge
void vortex_get_bedge_q(char const *name, op_set bedges, const op_arg *a) {
  auto t0 = std::chrono::steady_clock::now();
  State &s = st();
  need_direct(a[0], "get_bedge_q", 0);
  for (int k = 1; k < 3; k++)
    if (!a[k].map || a[k].map->from != bedges || a[k].idx != k - 1)
      die(std::string("get_bedge_q: argument ") + std::to_string(k) +
          " must be the node coordinates through the boundary-edge-to-node map");
  for (int k : {3, 4, 5, 6, 8}) need_mapped(a[k], bedges, 0, "get_bedge_q", k);
  if (a[3].map != a[5].map || a[4].map != a[5].map || a[8].map != a[5].map)
    die("get_bedge_q: every cell argument must use the same boundary-edge-to-cell map");
  const double t = gbl_double(a[7], "get_bedge_q", 7);
  op_set cells = a[5].map->to;
  if (s.b2c != a[5].map || s.bnum != a[0].dat || s.btype) {
    s.b2c = a[5].map;
    s.bnum = a[0].dat;
    s.btype = nullptr;
    s.mesh = false;
  }
  ensure_mesh(cells);
  set_vortex_exact_state();
  const int np = Np(), nf = a[8].dat->dim / 4;
  check(dg_get_bedge_q_exact(s.ctx, dptr(a[5], cells, np, name, 5), dptr(a[6], cells, np, name, 6), t,
                             dptr(a[8], cells, 4 * nf, name, 8)),
        name);
  sync_and_time(name, t0);
}

// l2_error(x, y, q, t, four OP_INC globals, error) (vortex/src/vortex.cpp:
// 314-324): the squared nodal errors against the exact vortex into error,
// their sums added to the globals
void vortex_l2_error(char const *name, op_set cells, const op_arg *a) {
  auto t0 = std::chrono::steady_clock::now();
  ensure_mesh(cells);
  set_vortex_exact_state();
  const int np = Np();
  for (int k : {0, 1, 2, 8}) need_direct(a[k], "l2_error", k);
  const double t = gbl_double(a[3], "l2_error", 3);
  double err[4];
  for (int k = 0; k < 4; k++) {
    err[k] = gbl_double(a[4 + k], "l2_error", 4 + k);
    if (a[4 + k].acc != OP_INC) die("l2_error: arguments 4-7 must be OP_INC globals");
  }
  check(dg_l2_error(st().ctx, dptr(a[0], cells, np, name, 0), dptr(a[1], cells, np, name, 1),
                    dptr(a[2], cells, 4 * np, name, 2), t, dptr(a[8], cells, 4 * np, name, 8), err),
        name);
  for (int k = 0; k < 4; k++) *(double *)a[4 + k].gbl = err[k];
  sync_and_time(name, t0);
}

}  // namespace

// src/airfoil.cpp's op_decl_const(dim, type, &var): the name is the variable.
void op_decl_const_expr(int dim, char const *type, int type_size, char *data, char const *expr) {
  std::string e;
  for (const char *p = expr ? expr : ""; *p; p++)
    if (!std::isspace((unsigned char)*p)) e += *p;
  while (!e.empty() && (e.front() == '&' || e.front() == '(')) e.erase(0, 1);
  while (!e.empty() && e.back() == ')') e.pop_back();
  bool ok = !e.empty() && (std::isalpha((unsigned char)e[0]) || e[0] == '_');
  for (char ch : e) ok = ok && (std::isalnum((unsigned char)ch) || ch == '_');
  if (!ok)
    die(std::string("op_decl_const: cannot take a constant name from '") + (expr ? expr : "") +
        "'; use op_decl_const2(name, dim, type, ptr)");
  op_decl_const_char(dim, type, type_size, data, e.c_str());
}

// src/airfoil.cpp's op_par_loop(kernel, "name", set, args...): the loop by name
// (and vortex/src/vortex.cpp's, which differ in their argument lists).
void op_par_loop_by_name(char const *name, op_set set, std::initializer_list<op_arg> args) {
  const std::string n = name ? name : "";
  const op_arg *a = args.begin();
  const size_t na = args.size();
  auto need = [&](size_t k, size_t vortex_form = 0) {
    if (na != k)
      die("op_par_loop(" + n + "): " + std::to_string(na) + " arguments, the loop takes " + std::to_string(k) +
          (vortex_form ? " (" + std::to_string(vortex_form) + " in the vortex case's form)" : std::string()));
  };
  // the vortex case's forms, told apart by their argument counts
  if (n == "set_ic" && na == 5) return vortex_set_ic(name, set, a);
  if (n == "get_neighbour_q" && na == 11) return vortex_get_neighbour_q(name, set, a);
  if (n == "get_bedge_q" && na == 9) return vortex_get_bedge_q(name, set, a);
  if (n == "l2_error") {
    need(9);
    return vortex_l2_error(name, set, a);
  }
  if (n == "init_grid") {
    need(18);
    op_par_loop_init_grid(name, set, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11],
                          a[12], a[13], a[14], a[15], a[16], a[17]);
  } else if (n == "set_ic") {
    need(2, 5);
    op_par_loop_set_ic(name, set, a[0], a[1]);
  } else if (n == "neighbour_zero") {
    need(1);
    op_par_loop_neighbour_zero(name, set, a[0]);
  } else if (n == "calc_dt") {
    need(3);
    op_par_loop_calc_dt(name, set, a[0], a[1], a[2]);
  } else if (n == "get_neighbour_q") {
    need(9, 11);
    op_par_loop_get_neighbour_q(name, set, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]);
  } else if (n == "get_bedge_q") {
    need(6, 9);
    op_par_loop_get_bedge_q(name, set, a[0], a[1], a[2], a[3], a[4], a[5]);
  } else if (n == "euler_rhs") {
    need(10);
    op_par_loop_euler_rhs(name, set, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9]);
  } else if (n == "set_workingQ") {
    need(6);
    op_par_loop_set_workingQ(name, set, a[0], a[1], a[2], a[3], a[4], a[5]);
  } else if (n == "update_Q") {
    need(6);
    op_par_loop_update_Q(name, set, a[0], a[1], a[2], a[3], a[4], a[5]);
  } else {
    die("op_par_loop: no loop named '" + n + "' in this interface");
  }
}
