// The SSPRK3 driver of src/airfoil.cpp:204-334 on the device, for one mesh
// or for a batch of independent cases (meshes of any sizes, each with its
// own free stream) concatenated into one set of op_dats.
//
// Fused mode runs three stage kernels per step over all cases at once; dt1
// is kept per case in two ping-pong slots on the device, so a run never
// synchronises with the host, and two steps (one per slot parity) can be
// captured once into a hipGraph and replayed.  Reference mode (one case)
// runs the nine per-loop kernels of the reference sequence with dt on the
// host.  Every case's arithmetic depends only on its own cells and dt, so a
// case run inside a batch is bit-identical to the same case run alone.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/dg_euler.h"
#include "dg_ctx.h"
#include "dg_internal.h"
#include "dg_mesh.h"
#include "dg_ops.h"

using dg::fail;
using dg::hipfail;

struct dg_solver {
  dg_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int ncases = 0;
  int K = 0;  // total cells
  std::vector<int> cell_off, tile_off, wall_off;
  std::vector<dg::Phys> phys;
  std::vector<double> alpha;
  int *d_cell_off = nullptr, *d_tile_off = nullptr, *d_tile_case = nullptr;
  dg::Phys *d_phys = nullptr;
  // concatenated connectivity (global cell indices)
  int *d_fmap = nullptr, *d_b2c = nullptr, *d_bnum = nullptr, *d_btype = nullptr;
  int nbedges = 0;
  // boundary edges of type 4 (exact vortex state, single-case solvers): each
  // fused stage first stores their states at the stage time in exteriorQ
  bool has_exact = false;
  std::map<std::string, double *> dats;
  double *q = nullptr, *w = nullptr, *rk[3] = {nullptr, nullptr, nullptr}, *ext = nullptr;
  unsigned long long *dt1 = nullptr;  // [2][ncases]
  double *t_dev = nullptr;            // [ncases]
  unsigned int *flag = nullptr;
  long long step = 0;
  bool dt_valid = false;
  double cfl_last = 1.0;
  // profiling
  bool profiling = false;
  struct Rec { std::string name; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, int>> ktimes;
  // wall faces (bc 2) of all cases, case ranges in wall_off
  int nwall = 0;
  std::vector<int> h_wcell, h_wface;
  int *d_wcell = nullptr, *d_wface = nullptr;
  double *d_facew = nullptr, *d_wforce = nullptr, *d_wcp = nullptr;
  // two-step graph
  hipGraphExec_t gexec = nullptr;
  double gcfl = -1.0;
  // launch shape of the matrix-core stage kernel (autotuned or set)
  int wpb = 1, walk = 0, max_blocks = 0;  // as dg::Launch's defaults
  // stage kernels read the per-cell metric record (DG_GEOM_CELL) or the
  // per-node factors, normals and fscale (DG_GEOM_NODE)
  bool cell_geometry = false;
  // vertices and cells of every case (host), for the vertex export
  std::vector<int> node_off;         // [ncases + 1]
  std::vector<double> h_coords;      // [nodes][2], cases concatenated
  std::vector<int> h_cells;          // [K][3], node indices local to the case
  std::vector<std::vector<long long>> vrec;  // per case: q record of each vertex
};

namespace {

dg::Launch launch_s(const dg_solver *s) {
  dg::Launch L = dg::launch_of(s->ctx);
  L.stream = s->stream;
  L.wpb = s->wpb;
  L.walk = s->walk;
  if (s->max_blocks > 0) L.max_blocks = s->max_blocks;
  return L;
}

hipEvent_t take_event(dg_solver *s) {
  if (!s->pool.empty()) {
    hipEvent_t e = s->pool.back();
    s->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

template <class F>
hipError_t timed(dg_solver *s, const char *name, F &&f) {
  if (!s->profiling) return f();
  dg_solver::Rec r{name, take_event(s), take_event(s)};
  (void)hipEventRecord(r.a, s->stream);
  hipError_t e = f();
  (void)hipEventRecord(r.b, s->stream);
  s->pending.push_back(r);
  return e;
}

void drain_profile(dg_solver *s) {
  if (s->pending.empty()) return;
  (void)hipStreamSynchronize(s->stream);
  for (auto &r : s->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto &t = s->ktimes[r.name];
      t.first += ms;
      t.second += 1;
    }
    s->pool.push_back(r.a);
    s->pool.push_back(r.b);
  }
  s->pending.clear();
}

double *dat(dg_solver *s, const char *n) { return s->dats[n]; }

void drop_graph(dg_solver *s) {
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  s->gexec = nullptr;
  s->gcfl = -1.0;
}

template <class T>
hipError_t upload(T **dst, const std::vector<T> &v) {
  size_t n = v.empty() ? 1 : v.size();
  hipError_t e = hipMalloc(dst, sizeof(T) * n);
  if (e == hipSuccess && !v.empty()) e = hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
  return e;
}

// dt1 of every case from the current q into slot p.
hipError_t compute_dt1(dg_solver *s, int p) {
  const dg::Launch L = launch_s(s);
  dg_ctx *c = s->ctx;
  unsigned long long *slot = s->dt1 + (size_t)p * s->ncases;
  hipError_t e = hipMemsetAsync(slot, 0, sizeof(unsigned long long) * s->ncases, s->stream);
  for (int k = 0; k < s->ncases && e == hipSuccess; k++) {
    dg::Launch Lk = L;
    Lk.ph = s->phys[k];
    const size_t off = s->cell_off[k];
    e = timed(s, "calc_dt", [&] {
      return dg::launch_calc_dt(Lk, s->cell_off[k + 1] - s->cell_off[k], s->q + off * c->Np * 4,
                                dat(s, "fscale") + off * c->NF, slot + k);
    });
  }
  return e;
}

dg::StageArgs stage_args(dg_solver *s, double cfl) {
  dg::StageArgs a{};
  a.rx = dat(s, "rx"); a.ry = dat(s, "ry"); a.sx = dat(s, "sx"); a.sy = dat(s, "sy");
  a.fscale = dat(s, "fscale"); a.nx = dat(s, "nx"); a.ny = dat(s, "ny");
  a.cmet = s->cell_geometry ? dat(s, "cmet") : nullptr;
  a.fmap = s->d_fmap; a.q = s->q; a.cfl = cfl; a.t_dev = s->t_dev; a.flag = s->flag;
  a.ncells = s->K;
  a.ncases = s->ncases;
  a.cell_off = s->d_cell_off;
  a.tile_off = s->d_tile_off;
  a.tile_case = s->ncases > 1 ? s->d_tile_case : nullptr;  // one case: no lookup at all
  a.phys = s->d_phys;
  a.ntiles_hint = s->tile_off.back();
  return a;
}

// Exact vortex states of the type-4 boundary slots at the fused stage's time
// t + frac dt (t and dt read on the device), stored in exteriorQ.
hipError_t exact_states(dg_solver *s, const dg::Launch &L, const dg::StageArgs &a, double frac) {
  if (!s->has_exact) return hipSuccess;
  return dg::launch_exact_bc(L, s->nbedges, s->d_b2c, s->d_bnum, s->d_btype, dat(s, "x"), dat(s, "y"),
                             s->ctx->vortex, 0.0, s->t_dev, a.dt1_cur, a.cfl, nullptr, frac, false, s->ext);
}

// One fused SSPRK3 step (three launches, plus one small one per stage on
// meshes with exact-state boundaries) for the step parity p.  The stages
// evaluate the RHS at t, t + dt and t + dt/2 (vortex/src/vortex.cpp:228).
hipError_t fused_step(dg_solver *s, int p, double cfl) {
  const dg::Launch L = launch_s(s);
  dg::StageArgs a = stage_args(s, cfl);
  a.dt1_cur = s->dt1 + (size_t)p * s->ncases;
  a.dt1_next = s->dt1 + (size_t)(p ^ 1) * s->ncases;
  if (s->has_exact) a.ext = s->ext;
  double *wA = s->w, *wB = s->rk[0];
  a.w = s->q; a.out = wA;
  hipError_t e = exact_states(s, L, a, 0.0);
  if (e == hipSuccess) e = timed(s, "rhs_stage0", [&] { return dg::launch_rhs(L, a, true, dg::EPI_STAGE0); });
  if (e != hipSuccess) return e;
  a.w = wA; a.out = wB;
  e = exact_states(s, L, a, 1.0);
  if (e == hipSuccess) e = timed(s, "rhs_stage1", [&] { return dg::launch_rhs(L, a, true, dg::EPI_STAGE1); });
  if (e != hipSuccess) return e;
  a.w = wB; a.out = s->q;
  e = exact_states(s, L, a, 0.5);
  if (e != hipSuccess) return e;
  return timed(s, "rhs_stage2", [&] { return dg::launch_rhs(L, a, true, dg::EPI_STAGE2); });
}

}  // namespace

extern "C" {

void dg_solver_destroy(dg_solver *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_graph(s);
  for (auto &kv : s->dats) (void)hipFree(kv.second);
  void *ptrs[] = {s->dt1, s->t_dev, s->flag, s->d_cell_off, s->d_tile_off, s->d_tile_case, s->d_phys, s->d_fmap,
                  s->d_b2c, s->d_bnum, s->d_btype, s->d_wcell, s->d_wface, s->d_facew,
                  s->d_wforce, s->d_wcp};
  for (void *p : ptrs) (void)hipFree(p);
  for (auto &r : s->pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : s->pool) (void)hipEventDestroy(e);
  if (s->own_stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int dg_solver_create_batch(dg_ctx *c, int ncases, const dg_mesh *const *meshes, const double *mach,
                           const double *alpha_deg, dg_solver **out) {
  if (!c || !meshes || !out || ncases < 1)
    return fail(c, DG_ERR_ARG, "dg_solver_create_batch: null argument or ncases < 1");
  *out = nullptr;
  dg_solver *s = new dg_solver();
  s->ctx = c;
  s->ncases = ncases;
  const int Np = c->Np, NF = c->NF, TE = dg::kMfmaTile;
  // case table, concatenated connectivity
  std::vector<int> fmap_all, b2c, bnum, btype;
  std::vector<double> cell_xy;
  s->cell_off.push_back(0);
  s->tile_off.push_back(0);
  s->wall_off.push_back(0);
  s->node_off.push_back(0);
  long long ktot = 0;
  for (int k = 0; k < ncases; k++) {
    if (!meshes[k]) { delete s; return fail(c, DG_ERR_ARG, "dg_solver_create_batch: null mesh"); }
    const dg::Mesh &M = meshes[k]->m;
    std::vector<int> fm;
    std::string err;
    if (!dg::build_face_map(M.ncells, M.nedges, M.edge2cells.data(), M.edgeNum.data(), M.nbedges,
                            M.bedge2cells.data(), M.bedgeNum.data(), M.bedge_type.data(),
                            M.cells.data(), fm, err)) {
      delete s;
      return fail(c, DG_ERR_MESH, "dg_solver_create_batch: case " + std::to_string(k) + ": " + err);
    }
    const int off = (int)ktot;
    ktot += M.ncells;
    if (ktot * 4 >= (1ll << 31)) { delete s; return fail(c, DG_ERR_ARG, "dg_solver_create_batch: too many cells"); }
    for (int v : fm) fmap_all.push_back(v >= 0 ? v + 4 * off : v);
    for (int b = 0; b < M.nbedges; b++) {
      b2c.push_back(M.bedge2cells[b] + off);
      bnum.push_back(M.bedgeNum[b]);
      btype.push_back(M.bedge_type[b]);
      s->has_exact |= M.bedge_type[b] == dg::kBcExact;
      if (M.bedge_type[b] == DG_BC_WALL) {
        s->h_wcell.push_back(M.bedge2cells[b] + off);
        s->h_wface.push_back(M.bedgeNum[b]);
      }
    }
    cell_xy.insert(cell_xy.end(), M.cell_xy.begin(), M.cell_xy.end());
    s->h_coords.insert(s->h_coords.end(), M.coords.begin(), M.coords.end());
    s->h_cells.insert(s->h_cells.end(), M.cells.begin(), M.cells.end());
    s->node_off.push_back(s->node_off.back() + M.nnodes);
    s->cell_off.push_back((int)ktot);
    s->tile_off.push_back(s->tile_off.back() + (M.ncells + TE - 1) / TE);
    s->wall_off.push_back((int)s->h_wcell.size());
    if (mach && alpha_deg) {
      s->phys.push_back(dg::freestream(c->gam, mach[k], alpha_deg[k], c->bc_p, c->bc_r));
      s->alpha.push_back(alpha_deg[k]);
    } else {
      s->phys.push_back(dg::launch_of(c).ph);
      s->alpha.push_back(c->bc_alpha);
    }
  }
  if (s->has_exact && ncases > 1) {
    delete s;
    return fail(c, DG_ERR_ARG, "dg_solver_create_batch: exact-state boundaries (type 4) run in single-case "
                               "solvers");
  }
  s->K = (int)ktot;
  s->nbedges = (int)b2c.size();
  s->nwall = (int)s->h_wcell.size();
  hipError_t e = hipSuccess;
  if (c->stream) {
    s->stream = c->stream;
  } else {
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    s->own_stream = e == hipSuccess;
  }
  const size_t K = s->K;
  struct { const char *n; size_t sz; } layout[] = {
      {"q", K * Np * 4},   {"workingQ", K * Np * 4}, {"rk1", K * Np * 4}, {"rk2", K * Np * 4},
      {"rk3", K * Np * 4}, {"exteriorQ", K * NF * 4}, {"x", K * Np},      {"y", K * Np},
      {"xr", K * Np},      {"yr", K * Np},           {"xs", K * Np},      {"ys", K * Np},
      {"rx", K * Np},      {"ry", K * Np},           {"sx", K * Np},      {"sy", K * Np},
      {"nx", K * NF},      {"ny", K * NF},           {"fscale", K * NF},  {"nodeX", K * 3},
      {"nodeY", K * 3},    {"J", K},                 {"cell_xy", K * 6},  {"cmet", K * 4},
  };
  for (auto &l : layout) {
    if (e != hipSuccess) break;
    double *p = nullptr;
    e = hipMalloc(&p, sizeof(double) * (l.sz ? l.sz : 1));
    if (e == hipSuccess) s->dats[l.n] = p;
  }
  if (e == hipSuccess) e = hipMalloc(&s->dt1, 2 * sizeof(unsigned long long) * ncases);
  if (e == hipSuccess) e = hipMalloc(&s->t_dev, sizeof(double) * ncases);
  if (e == hipSuccess) e = hipMalloc(&s->flag, sizeof(unsigned int));
  if (e == hipSuccess) e = hipMemset(s->dt1, 0, 2 * sizeof(unsigned long long) * ncases);
  if (e == hipSuccess) e = hipMemset(s->t_dev, 0, sizeof(double) * ncases);
  if (e == hipSuccess) e = hipMemset(s->flag, 0, sizeof(unsigned int));
  if (e == hipSuccess) e = upload(&s->d_cell_off, s->cell_off);
  if (e == hipSuccess) e = upload(&s->d_tile_off, s->tile_off);
  if (e == hipSuccess) {
    // each tile's case, so that a stage kernel's tile finds its case with one load
    std::vector<int> tile_case(s->tile_off.back());
    for (int k = 0; k < ncases; k++)
      for (int t = s->tile_off[k]; t < s->tile_off[k + 1]; t++) tile_case[t] = k;
    e = upload(&s->d_tile_case, tile_case);
  }
  if (e == hipSuccess) e = upload(&s->d_phys, s->phys);
  if (e == hipSuccess) e = upload(&s->d_fmap, fmap_all);
  if (e == hipSuccess) e = upload(&s->d_b2c, b2c);
  if (e == hipSuccess) e = upload(&s->d_bnum, bnum);
  if (e == hipSuccess) e = upload(&s->d_btype, btype);
  if (e == hipSuccess) e = upload(&s->d_wcell, s->h_wcell);
  if (e == hipSuccess) e = upload(&s->d_wface, s->h_wface);
  if (e == hipSuccess && K > 0)
    e = hipMemcpy(dat(s, "cell_xy"), cell_xy.data(), sizeof(double) * K * 6, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    dg::Operators o;
    const char *oe = nullptr;
    dg::build_operators(c->N, o, &oe);
    std::vector<double> fw(c->Nfp, 0.0);
    for (int i = 0; i < c->Nfp; i++)
      for (int j = 0; j < c->Nfp; j++) fw[i] += o.face_mass[i * c->Nfp + j];
    e = upload(&s->d_facew, fw);
    const size_t nw = s->nwall > 0 ? s->nwall : 1;
    if (e == hipSuccess) e = hipMalloc(&s->d_wforce, sizeof(double) * 2 * nw);
    if (e == hipSuccess) e = hipMalloc(&s->d_wcp, sizeof(double) * nw * c->Nfp);
  }
  if (e != hipSuccess) {
    dg_solver_destroy(s);
    return hipfail(c, e, "dg_solver_create: allocation");
  }
  s->q = dat(s, "q"); s->w = dat(s, "workingQ"); s->ext = dat(s, "exteriorQ");
  s->rk[0] = dat(s, "rk1"); s->rk[1] = dat(s, "rk2"); s->rk[2] = dat(s, "rk3");
  dg::Launch L = launch_s(s);
  e = dg::launch_init_grid(L, s->K, dat(s, "cell_xy"), dat(s, "nodeX"), dat(s, "nodeY"),
                           dat(s, "x"), dat(s, "y"), dat(s, "xr"), dat(s, "yr"), dat(s, "xs"),
                           dat(s, "ys"), dat(s, "rx"), dat(s, "ry"), dat(s, "sx"), dat(s, "sy"),
                           dat(s, "nx"), dat(s, "ny"), dat(s, "fscale"), dat(s, "J"));
  if (e == hipSuccess)
    e = dg::launch_cell_metrics(L, s->K, dat(s, "rx"), dat(s, "ry"), dat(s, "sx"), dat(s, "sy"), dat(s, "cmet"));
  for (int k = 0; k < ncases && e == hipSuccess; k++) {
    dg::Launch Lk = L;
    Lk.ph = s->phys[k];
    const size_t off = (size_t)s->cell_off[k] * Np * 4;
    e = dg::launch_set_ic(Lk, s->cell_off[k + 1] - s->cell_off[k], s->q + off, s->w + off);
  }
  if (e == hipSuccess) e = dg::launch_zero(L, K * NF * 4, s->ext);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    dg_solver_destroy(s);
    return hipfail(c, e, "dg_solver_create: set-up kernels");
  }
  *out = s;
  return DG_OK;
}

int dg_solver_create(dg_ctx *c, const dg_mesh *mesh, dg_solver **out) {
  if (!c || !mesh || !out) return fail(c, DG_ERR_ARG, "dg_solver_create: null argument");
  const dg::Mesh &M = mesh->m;
  // the context's declared mesh is this one, as op_decl_map would make it
  int rc = dg_decl_mesh(c, M.ncells, M.cells.data(), M.nedges, M.edge2cells.data(),
                        M.edgeNum.data(), M.nbedges, M.bedge2cells.data(), M.bedgeNum.data(),
                        M.bedge_type.data());
  if (rc != DG_OK) return rc;
  const dg_mesh *ms[1] = {mesh};
  return dg_solver_create_batch(c, 1, ms, nullptr, nullptr, out);
}

int dg_solver_ncases(const dg_solver *s, int *ncases) {
  if (!s || !ncases) return fail(nullptr, DG_ERR_ARG, "dg_solver_ncases: null argument");
  *ncases = s->ncases;
  return DG_OK;
}

int dg_solver_set_vortex(dg_solver *s, double uinf, double beta, double x0, double y0,
                         double period_x) {
  if (!s) return fail(nullptr, DG_ERR_ARG, "dg_solver_set_vortex: null solver");
  if (!(period_x >= 0.0)) return fail(s->ctx, DG_ERR_ARG, "dg_solver_set_vortex: period_x must be >= 0");
  HIPCHECK(s->ctx, dg::launch_set_vortex(launch_s(s), s->K, dat(s, "x"), dat(s, "y"), uinf, beta,
                                         x0, y0, period_x, 0.0, s->q, s->w));
  // the same vortex is the exact state of type-4 boundaries
  s->ctx->vortex.uinf = uinf; s->ctx->vortex.beta = beta; s->ctx->vortex.x0 = x0;
  s->ctx->vortex.y0 = y0; s->ctx->vortex.period = period_x;
  s->dt_valid = false;
  return DG_OK;
}

static int case_range(dg_solver *s, int k, size_t *off, size_t *n, const char *fn) {
  if (!s || k < 0 || k >= s->ncases)
    return fail(s ? s->ctx : nullptr, DG_ERR_ARG, std::string(fn) + ": bad case index");
  *off = s->cell_off[k];
  *n = (size_t)(s->cell_off[k + 1] - s->cell_off[k]);
  return DG_OK;
}

int dg_solver_case_set_q(dg_solver *s, int k, const double *host_q) {
  size_t off, n;
  int rc = case_range(s, k, &off, &n, "dg_solver_case_set_q");
  if (rc != DG_OK) return rc;
  if (!host_q) return fail(s->ctx, DG_ERR_ARG, "dg_solver_case_set_q: null array");
  const size_t st = (size_t)s->ctx->Np * 4;
  HIPCHECK(s->ctx, hipStreamSynchronize(s->stream));
  HIPCHECK(s->ctx, hipMemcpy(s->q + off * st, host_q, sizeof(double) * n * st, hipMemcpyHostToDevice));
  HIPCHECK(s->ctx, hipMemcpy(s->w + off * st, host_q, sizeof(double) * n * st, hipMemcpyHostToDevice));
  s->dt_valid = false;
  return DG_OK;
}

int dg_solver_case_get_q(dg_solver *s, int k, double *host_q) {
  size_t off, n;
  int rc = case_range(s, k, &off, &n, "dg_solver_case_get_q");
  if (rc != DG_OK) return rc;
  if (!host_q) return fail(s->ctx, DG_ERR_ARG, "dg_solver_case_get_q: null array");
  const size_t st = (size_t)s->ctx->Np * 4;
  HIPCHECK(s->ctx, hipStreamSynchronize(s->stream));
  HIPCHECK(s->ctx, hipMemcpy(host_q, s->q + off * st, sizeof(double) * n * st, hipMemcpyDeviceToHost));
  return DG_OK;
}

int dg_solver_set_q(dg_solver *s, const double *host_q) {
  if (!s) return fail(nullptr, DG_ERR_ARG, "dg_solver_set_q: null solver");
  for (int k = 0; k < s->ncases; k++) {
    int rc = dg_solver_case_set_q(s, k, host_q ? host_q + (size_t)s->cell_off[k] * s->ctx->Np * 4 : nullptr);
    if (rc != DG_OK) return rc;
  }
  return DG_OK;
}

int dg_solver_get_q(dg_solver *s, double *host_q) {
  if (!s) return fail(nullptr, DG_ERR_ARG, "dg_solver_get_q: null solver");
  for (int k = 0; k < s->ncases; k++) {
    int rc = dg_solver_case_get_q(s, k, host_q ? host_q + (size_t)s->cell_off[k] * s->ctx->Np * 4 : nullptr);
    if (rc != DG_OK) return rc;
  }
  return DG_OK;
}

int dg_solver_dat(dg_solver *s, const char *name, void **p) {
  if (!s || !name || !p) return fail(s ? s->ctx : nullptr, DG_ERR_ARG, "dg_solver_dat: null argument");
  auto it = s->dats.find(name);
  if (it == s->dats.end())
    return fail(s->ctx, DG_ERR_NAME, std::string("dg_solver_dat: unknown op_dat '") + name + "'");
  *p = it->second;
  return DG_OK;
}

int dg_solver_stream(dg_solver *s, void **stream) {
  if (!s || !stream) return fail(nullptr, DG_ERR_ARG, "dg_solver_stream: null argument");
  *stream = (void *)s->stream;
  return DG_OK;
}

int dg_solver_reset_time(dg_solver *s) {
  if (!s) return fail(nullptr, DG_ERR_ARG, "dg_solver_reset_time: null solver");
  HIPCHECK(s->ctx, hipStreamSynchronize(s->stream));
  HIPCHECK(s->ctx, hipMemset(s->t_dev, 0, sizeof(double) * s->ncases));
  HIPCHECK(s->ctx, hipMemset(s->flag, 0, sizeof(unsigned int)));
  s->dt_valid = false;
  s->step = 0;
  drop_graph(s);
  return DG_OK;
}

static int run_impl(dg_solver *s, int nsteps, double cfl, int mode, bool graph) {
  if (!s) return fail(nullptr, DG_ERR_ARG, "dg_solver_run: null solver");
  dg_ctx *c = s->ctx;
  if (nsteps < 0 || !(cfl > 0.0)) return fail(c, DG_ERR_ARG, "dg_solver_run: need nsteps >= 0, cfl > 0");
  if (mode != DG_MODE_FUSED && mode != DG_MODE_REFERENCE)
    return fail(c, DG_ERR_ARG, "dg_solver_run: unknown mode");
  if (mode == DG_MODE_REFERENCE && s->ncases != 1)
    return fail(c, DG_ERR_ARG, "dg_solver_run: the reference loop sequence runs one case");
  const dg::Launch L = launch_s(s);
  const int K = s->K;
  const size_t nstate = (size_t)K * c->Np * 4;
  s->cfl_last = cfl;
  if (!s->dt_valid) {
    HIPCHECK(c, compute_dt1(s, (int)(s->step & 1)));
    s->dt_valid = true;
  }
  if (mode == DG_MODE_FUSED) {
    int n = nsteps;
    if (graph && !s->profiling && n >= 2) {
      if ((s->step & 1) && n > 0) {  // align to an even step
        HIPCHECK(c, fused_step(s, 1, cfl));
        s->step++;
        n--;
      }
      if (!s->gexec || s->gcfl != cfl) {
        drop_graph(s);
        hipGraph_t g = nullptr;
        HIPCHECK(c, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e0 = fused_step(s, 0, cfl);
        hipError_t e1 = e0 == hipSuccess ? fused_step(s, 1, cfl) : e0;
        hipError_t ec = hipStreamEndCapture(s->stream, &g);
        if (e1 != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hipfail(c, e1, "dg_solver_run: capture"); }
        if (ec != hipSuccess) return hipfail(c, ec, "dg_solver_run: hipStreamEndCapture");
        hipError_t ei = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) { s->gexec = nullptr; return hipfail(c, ei, "dg_solver_run: hipGraphInstantiate"); }
        s->gcfl = cfl;
      }
      for (; n >= 2; n -= 2) {
        HIPCHECK(c, hipGraphLaunch(s->gexec, s->stream));
        s->step += 2;
      }
    }
    for (; n > 0; n--) {
      HIPCHECK(c, fused_step(s, (int)(s->step & 1), cfl));
      s->step++;
    }
    // keep the reference invariant workingQ == q after a step
    if (nsteps > 0) HIPCHECK(c, dg::launch_copy(L, nstate, s->q, s->w));
    return DG_OK;
  }
  // DG_MODE_REFERENCE: the nine loops of src/airfoil.cpp:252-334, dt on the host.
  // exteriorQ is zero between loops (euler_rhs zeroes it on exit); fused
  // steps leave the exact states of type-4 slots in it, so clear it first.
  if (s->has_exact) HIPCHECK(c, dg::launch_zero(L, (size_t)K * c->NF * 4, s->ext));
  unsigned long long *slot[2] = {s->dt1, s->dt1 + 1};
  double t = 0.0, dt1 = 0.0;
  unsigned long long bits = 0;
  HIPCHECK(c, hipMemcpyAsync(&t, s->t_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHECK(c, hipMemcpyAsync(&bits, slot[s->step & 1], sizeof(bits), hipMemcpyDeviceToHost, s->stream));
  HIPCHECK(c, hipStreamSynchronize(s->stream));
  std::memcpy(&dt1, &bits, sizeof(dt1));
  double dt = cfl / dt1;
  double *fscale = dat(s, "fscale");
  const double frac[3] = {0.0, 1.0, 0.5};  // stage times (vortex/src/vortex.cpp:228,248-258)
  for (int n = 0; n < nsteps; n++) {
    for (int j = 0; j < 3; j++) {
      HIPCHECK(c, timed(s, "get_neighbour_q", [&] { return dg::launch_get_neighbour_q(L, K, s->d_fmap, s->w, s->ext); }));
      HIPCHECK(c, timed(s, "get_bedge_q", [&] {
        return dg::launch_get_bedge_q(L, s->nbedges, s->d_b2c, s->d_bnum, s->d_btype, dat(s, "nx"),
                                      dat(s, "ny"), s->w, s->ext);
      }));
      if (s->has_exact)
        HIPCHECK(c, timed(s, "get_bedge_q_exact", [&] {
          return dg::launch_exact_bc(L, s->nbedges, s->d_b2c, s->d_bnum, s->d_btype, dat(s, "x"), dat(s, "y"),
                                     c->vortex, t + frac[j] * dt, nullptr, nullptr, 0.0, nullptr, 0.0, true,
                                     s->ext);
        }));
      dg::StageArgs a{};
      a.w = s->w; a.ext = s->ext; a.rx = dat(s, "rx"); a.ry = dat(s, "ry"); a.sx = dat(s, "sx");
      a.sy = dat(s, "sy"); a.fscale = fscale; a.nx = dat(s, "nx"); a.ny = dat(s, "ny");
      a.out = s->rk[j]; a.ncells = K;
      HIPCHECK(c, timed(s, "euler_rhs", [&] { return dg::launch_rhs(L, a, false, dg::EPI_RHS); }));
      if (j != 2)
        HIPCHECK(c, timed(s, "set_workingQ", [&] {
          return dg::launch_set_workingQ(L, K, dt, j, s->q, s->rk[0], s->rk[1], s->w);
        }));
    }
    HIPCHECK(c, timed(s, "update_Q", [&] {
      return dg::launch_update_Q(L, K, dt, s->q, s->rk[0], s->rk[1], s->rk[2], s->w);
    }));
    t += dt;
    s->step++;
    unsigned long long *nxt = slot[s->step & 1];
    HIPCHECK(c, hipMemsetAsync(nxt, 0, sizeof(unsigned long long), s->stream));
    HIPCHECK(c, timed(s, "calc_dt", [&] { return dg::launch_calc_dt(L, K, s->q, fscale, nxt); }));
    HIPCHECK(c, hipMemcpyAsync(&bits, nxt, sizeof(bits), hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(c, hipStreamSynchronize(s->stream));
    std::memcpy(&dt1, &bits, sizeof(dt1));
    dt = cfl / dt1;
  }
  HIPCHECK(c, hipMemcpyAsync(s->t_dev, &t, sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIPCHECK(c, hipStreamSynchronize(s->stream));
  return DG_OK;
}

int dg_solver_run(dg_solver *s, int nsteps, double cfl, int mode) {
  return run_impl(s, nsteps, cfl, mode, false);
}

int dg_solver_run_graph(dg_solver *s, int nsteps, double cfl) {
  return run_impl(s, nsteps, cfl, DG_MODE_FUSED, true);
}

int dg_solver_case_time(dg_solver *s, int k, double *t, double *dt) {
  size_t off, n;
  int rc = case_range(s, k, &off, &n, "dg_solver_case_time");
  if (rc != DG_OK) return rc;
  dg_ctx *c = s->ctx;
  double tv = 0.0;
  unsigned long long bits = 0;
  HIPCHECK(c, hipMemcpyAsync(&tv, s->t_dev + k, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHECK(c, hipMemcpyAsync(&bits, s->dt1 + (size_t)(s->step & 1) * s->ncases + k, sizeof(bits),
                             hipMemcpyDeviceToHost, s->stream));
  HIPCHECK(c, hipStreamSynchronize(s->stream));
  double dt1;
  std::memcpy(&dt1, &bits, sizeof(dt1));
  if (t) *t = tv;
  if (dt) *dt = s->dt_valid ? s->cfl_last / dt1 : 0.0;
  return DG_OK;
}

int dg_solver_time(dg_solver *s, double *t, double *dt) { return dg_solver_case_time(s, 0, t, dt); }

int dg_solver_check(dg_solver *s) {
  if (!s) return fail(nullptr, DG_ERR_ARG, "dg_solver_check: null solver");
  unsigned int f = 0;
  HIPCHECK(s->ctx, hipMemcpyAsync(&f, s->flag, sizeof(f), hipMemcpyDeviceToHost, s->stream));
  HIPCHECK(s->ctx, hipStreamSynchronize(s->stream));
  if (f) return fail(s->ctx, DG_ERR_NONFINITE, "non-finite or non-positive-density state detected");
  return DG_OK;
}

int dg_solver_case_forces(dg_solver *s, int k, double chord, double *cl, double *cd, double *fxy) {
  size_t off, n;
  int rc = case_range(s, k, &off, &n, "dg_solver_case_forces");
  if (rc != DG_OK) return rc;
  dg_ctx *c = s->ctx;
  if (!(chord > 0.0)) return fail(c, DG_ERR_ARG, "dg_solver_forces: chord must be positive");
  const dg::Phys &ph = s->phys[k];
  const double vinf2 = ph.bc_u * ph.bc_u + ph.bc_v * ph.bc_v;
  const double qdyn = 0.5 * ph.bc_r * (vinf2 > 0.0 ? vinf2 : 1.0);
  const int w0 = s->wall_off[k], nw = s->wall_off[k + 1] - s->wall_off[k];
  dg::Launch L = launch_s(s);
  L.ph = ph;
  HIPCHECK(c, dg::launch_wall_forces(L, nw, s->d_wcell + w0, s->d_wface + w0, s->q, dat(s, "nx"),
                                     dat(s, "ny"), dat(s, "fscale"), dat(s, "J"), s->d_facew,
                                     ph.bc_p, qdyn, s->d_wforce + 2 * (size_t)w0,
                                     s->d_wcp + (size_t)w0 * c->Nfp));
  std::vector<double> f(2 * (size_t)nw);
  if (nw > 0)
    HIPCHECK(c, hipMemcpyAsync(f.data(), s->d_wforce + 2 * (size_t)w0, sizeof(double) * f.size(),
                               hipMemcpyDeviceToHost, s->stream));
  HIPCHECK(c, hipStreamSynchronize(s->stream));
  double fx = 0.0, fy = 0.0;  // summed in wall-face order: deterministic
  for (int b = 0; b < nw; b++) { fx += f[2 * b]; fy += f[2 * b + 1]; }
  const double al = s->alpha[k] * M_PI / 180.0;
  const double lift = fy * std::cos(al) - fx * std::sin(al);
  const double drag = fx * std::cos(al) + fy * std::sin(al);
  if (cl) *cl = lift / (qdyn * chord);
  if (cd) *cd = drag / (qdyn * chord);
  if (fxy) { fxy[0] = fx; fxy[1] = fy; }
  return DG_OK;
}

int dg_solver_forces(dg_solver *s, double chord, double *cl, double *cd, double *fxy) {
  return dg_solver_case_forces(s, 0, chord, cl, cd, fxy);
}

int dg_solver_wall_cp(dg_solver *s, int *nwall, double *x, double *cp) {
  if (!s || !nwall) return fail(s ? s->ctx : nullptr, DG_ERR_ARG, "dg_solver_wall_cp: null argument");
  dg_ctx *c = s->ctx;
  const int nw = s->wall_off[1] - s->wall_off[0];
  *nwall = nw;
  if (!x && !cp) return DG_OK;
  int rc = dg_solver_case_forces(s, 0, 1.0, nullptr, nullptr, nullptr);  // refreshes the Cp buffer
  if (rc != DG_OK) return rc;
  const size_t n = (size_t)nw * c->Nfp;
  if (cp && n) HIPCHECK(c, hipMemcpy(cp, s->d_wcp, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (x && n) {
    std::vector<double> xs((size_t)s->cell_off[1] * c->Np);
    HIPCHECK(c, hipMemcpy(xs.data(), dat(s, "x"), sizeof(double) * xs.size(), hipMemcpyDeviceToHost));
    for (int b = 0; b < nw; b++)
      for (int i = 0; i < c->Nfp; i++)
        x[(size_t)b * c->Nfp + i] = xs[(size_t)s->h_wcell[b] * c->Np + c->FMASK[s->h_wface[b] * c->Nfp + i]];
  }
  return DG_OK;
}

int dg_solver_profile(dg_solver *s, int enable) {
  if (!s) return fail(nullptr, DG_ERR_ARG, "dg_solver_profile: null solver");
  drain_profile(s);
  s->profiling = enable != 0;
  s->ktimes.clear();
  return DG_OK;
}

int dg_solver_kernel_time(dg_solver *s, const char *kernel, double *avg_ms, int *count) {
  if (!s || !kernel) return fail(nullptr, DG_ERR_ARG, "dg_solver_kernel_time: null argument");
  drain_profile(s);
  auto it = s->ktimes.find(kernel);
  if (it == s->ktimes.end() || it->second.second == 0) {
    if (avg_ms) *avg_ms = 0.0;
    if (count) *count = 0;
    return DG_OK;
  }
  if (avg_ms) *avg_ms = it->second.first / it->second.second;
  if (count) *count = it->second.second;
  return DG_OK;
}

// Vertex values of case k as src/save_solution.cpp:13-47 forms them: each
// vertex takes the state at the matching corner node (0, N, Np-1 for cell
// vertices 0, 1, 2) of the last cell, in cell order, that contains it;
// density, velocity and pressure from that state; vertices of no cell stay 0.
static int vertex_fields(dg_solver *s, int k, std::vector<double> &rho, std::vector<double> &vx,
                         std::vector<double> &vy, std::vector<double> &p) {
  dg_ctx *c = s->ctx;
  const int nn = s->node_off[k + 1] - s->node_off[k];
  const int c0 = s->cell_off[k], kc = s->cell_off[k + 1] - s->cell_off[k];
  if ((int)s->vrec.size() < s->ncases) s->vrec.resize(s->ncases);
  std::vector<long long> &rec = s->vrec[k];
  if ((int)rec.size() != nn) {
    rec.assign(nn, -1);
    const int corner[3] = {0, c->N, c->Np - 1};
    for (int e = 0; e < kc; e++)
      for (int t = 0; t < 3; t++)
        rec[s->h_cells[3 * ((size_t)c0 + e) + t]] = ((long long)c0 + e) * c->Np + corner[t];
  }
  rho.assign(nn, 0.0); vx.assign(nn, 0.0); vy.assign(nn, 0.0); p.assign(nn, 0.0);
  if (nn == 0) return DG_OK;
  long long *d_rec = nullptr;
  double *d_out = nullptr;
  std::vector<double> out(4 * (size_t)nn);
  hipError_t e = hipMalloc(&d_rec, sizeof(long long) * nn);
  if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * out.size());
  if (e == hipSuccess) e = hipMemcpyAsync(d_rec, rec.data(), sizeof(long long) * nn, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = dg::launch_gather_records(launch_s(s), nn, d_rec, s->q, d_out);
  if (e == hipSuccess) e = hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(d_rec);
  (void)hipFree(d_out);
  if (e != hipSuccess) return hipfail(c, e, "dg_solver_vertex_fields");
  const double gam = s->phys[k].gam;
  for (int i = 0; i < nn; i++) {
    if (rec[i] < 0) continue;
    const double *q = &out[4 * (size_t)i];
    rho[i] = q[0];
    vx[i] = q[1] / rho[i];
    vy[i] = q[2] / rho[i];
    p[i] = (gam - 1) * (q[3] - 0.5 * (q[1] * vx[i] + q[2] * vy[i]));
  }
  return DG_OK;
}

int dg_solver_vertex_fields(dg_solver *s, int k, int *nnodes, double *rho, double *vx, double *vy, double *p) {
  size_t off, n;
  int rc = case_range(s, k, &off, &n, "dg_solver_vertex_fields");
  if (rc != DG_OK) return rc;
  const int nn = s->node_off[k + 1] - s->node_off[k];
  if (nnodes) *nnodes = nn;
  if (!rho && !vx && !vy && !p) return DG_OK;
  std::vector<double> a, b, cc, d;
  rc = vertex_fields(s, k, a, b, cc, d);
  if (rc != DG_OK) return rc;
  if (rho) std::memcpy(rho, a.data(), sizeof(double) * nn);
  if (vx) std::memcpy(vx, b.data(), sizeof(double) * nn);
  if (vy) std::memcpy(vy, cc.data(), sizeof(double) * nn);
  if (p) std::memcpy(p, d.data(), sizeof(double) * nn);
  return DG_OK;
}

int dg_solver_write_vtk(dg_solver *s, int k, const char *path) {
  size_t off, n;
  int rc = case_range(s, k, &off, &n, "dg_solver_write_vtk");
  if (rc != DG_OK) return rc;
  if (!path) return fail(s->ctx, DG_ERR_ARG, "dg_solver_write_vtk: null path");
  std::vector<double> rho, vx, vy, p;
  rc = vertex_fields(s, k, rho, vx, vy, p);
  if (rc != DG_OK) return rc;
  FILE *f = std::fopen(path, "w");
  if (!f) return fail(s->ctx, DG_ERR_ARG, std::string("dg_solver_write_vtk: cannot open ") + path);
  const int nn = s->node_off[k + 1] - s->node_off[k];
  const int c0 = s->cell_off[k], kc = s->cell_off[k + 1] - s->cell_off[k];
  const double *xy = &s->h_coords[2 * (size_t)s->node_off[k]];
  std::fprintf(f, "# vtk DataFile Version 3.0\nDG Euler solution (vertex values)\nASCII\nDATASET UNSTRUCTURED_GRID\n");
  std::fprintf(f, "POINTS %d double\n", nn);
  for (int i = 0; i < nn; i++) std::fprintf(f, "%.17g %.17g 0\n", xy[2 * i], xy[2 * i + 1]);
  std::fprintf(f, "CELLS %d %d\n", kc, 4 * kc);
  for (int e = 0; e < kc; e++) {
    const int *cl = &s->h_cells[3 * ((size_t)c0 + e)];
    std::fprintf(f, "3 %d %d %d\n", cl[0], cl[1], cl[2]);
  }
  std::fprintf(f, "CELL_TYPES %d\n", kc);
  for (int e = 0; e < kc; e++) std::fprintf(f, "5\n");
  std::fprintf(f, "POINT_DATA %d\n", nn);
  const char *names[4] = {"Density", "Pressure", "VelocityX", "VelocityY"};
  const std::vector<double> *vals[4] = {&rho, &p, &vx, &vy};
  for (int a = 0; a < 4; a++) {
    std::fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", names[a]);
    for (int i = 0; i < nn; i++) std::fprintf(f, "%.17g\n", (*vals[a])[i]);
  }
  const bool ok = std::fclose(f) == 0;
  if (!ok) return fail(s->ctx, DG_ERR_ARG, std::string("dg_solver_write_vtk: write failed: ") + path);
  return DG_OK;
}

int dg_solver_set_launch(dg_solver *s, int wpb, int walk, int max_blocks) {
  if (!s || (wpb != 1 && wpb != 2 && wpb != 4) || (walk != 0 && walk != 1) || max_blocks < 0)
    return fail(s ? s->ctx : nullptr, DG_ERR_ARG, "dg_solver_set_launch: wpb in {1,2,4}, walk in {0,1}, max_blocks >= 0");
  drop_graph(s);
  s->wpb = wpb; s->walk = walk; s->max_blocks = max_blocks;
  return DG_OK;
}

int dg_solver_set_geometry(dg_solver *s, int mode) {
  if (!s || (mode != DG_GEOM_CELL && mode != DG_GEOM_NODE))
    return fail(s ? s->ctx : nullptr, DG_ERR_ARG, "dg_solver_set_geometry: mode is DG_GEOM_CELL or DG_GEOM_NODE");
  drop_graph(s);
  s->cell_geometry = mode == DG_GEOM_CELL;
  return DG_OK;
}

int dg_solver_launch_config(const dg_solver *s, int *wpb, int *walk, int *max_blocks) {
  if (!s) return fail(nullptr, DG_ERR_ARG, "dg_solver_launch_config: null solver");
  if (wpb) *wpb = s->wpb;
  if (walk) *walk = s->walk;
  if (max_blocks) *max_blocks = s->max_blocks;
  return DG_OK;
}

// Times trial_steps fused steps for every launch shape (waves per workgroup
// 1/2/4 x XCD-chunked or round-robin walk x 4/8/12 resident waves per CU)
// on the solver's own cases, restores the state, time and dt it started
// from, and keeps the fastest shape.  Every shape runs the same per-cell
// arithmetic, so the results do not depend on the choice.
int dg_solver_autotune(dg_solver *s, int trial_steps, double cfl, double *table, int *ntable) {
  if (!s || trial_steps < 1 || !(cfl > 0.0)) return fail(s ? s->ctx : nullptr, DG_ERR_ARG, "dg_solver_autotune: bad argument");
  dg_ctx *c = s->ctx;
  const size_t nstate = (size_t)s->K * c->Np * 4;
  if (!s->dt_valid) {
    HIPCHECK(c, compute_dt1(s, (int)(s->step & 1)));
    s->dt_valid = true;
  }
  drop_graph(s);
  double *q0 = nullptr, *t0 = nullptr;
  unsigned long long *d0 = nullptr;
  HIPCHECK(c, hipMalloc(&q0, sizeof(double) * (nstate ? nstate : 1)));
  hipError_t e = hipMalloc(&t0, sizeof(double) * s->ncases);
  if (e == hipSuccess) e = hipMalloc(&d0, 2 * sizeof(unsigned long long) * s->ncases);
  if (e == hipSuccess) e = hipMemcpyAsync(q0, s->q, sizeof(double) * nstate, hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(t0, s->t_dev, sizeof(double) * s->ncases, hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d0, s->dt1, 2 * sizeof(unsigned long long) * s->ncases, hipMemcpyDeviceToDevice, s->stream);
  const long long step0 = s->step;
  const int w0 = s->wpb, k0 = s->walk, b0 = s->max_blocks;
  const int wpbs[3] = {1, 2, 4}, waves[3] = {4, 8, 12};
  double best = 1e300;
  int bw = w0, bk = k0, bb = b0, n = 0;
  hipEvent_t ea = nullptr, eb = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&ea);
  if (e == hipSuccess) e = hipEventCreate(&eb);
  const int cus = c->cus > 0 ? c->cus : 256;
  for (int iw = 0; iw < 3 && e == hipSuccess; iw++)
    for (int walk = 0; walk < 2 && e == hipSuccess; walk++)
      for (int iv = 0; iv < 3 && e == hipSuccess; iv++) {
        s->wpb = wpbs[iw];
        s->walk = walk;
        s->max_blocks = cus * waves[iv] / wpbs[iw];
        e = fused_step(s, (int)(s->step & 1), cfl);  // warm
        s->step++;
        if (e == hipSuccess) e = hipEventRecord(ea, s->stream);
        for (int k = 0; k < trial_steps && e == hipSuccess; k++) {
          e = fused_step(s, (int)(s->step & 1), cfl);
          s->step++;
        }
        if (e == hipSuccess) e = hipEventRecord(eb, s->stream);
        if (e == hipSuccess) e = hipEventSynchronize(eb);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ea, eb);
        const double per = ms / trial_steps;
        if (table) {
          table[4 * n] = s->wpb; table[4 * n + 1] = walk; table[4 * n + 2] = waves[iv]; table[4 * n + 3] = per;
        }
        n++;
        if (e == hipSuccess && per < best) { best = per; bw = s->wpb; bk = walk; bb = s->max_blocks; }
      }
  // restore the starting state, time, dt slots and step parity
  if (e == hipSuccess) e = hipMemcpyAsync(s->q, q0, sizeof(double) * nstate, hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(s->w, q0, sizeof(double) * nstate, hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(s->t_dev, t0, sizeof(double) * s->ncases, hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->dt1, d0, 2 * sizeof(unsigned long long) * s->ncases, hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  s->step = step0;
  if (ea) (void)hipEventDestroy(ea);
  if (eb) (void)hipEventDestroy(eb);
  (void)hipFree(q0); (void)hipFree(t0); (void)hipFree(d0);
  if (e != hipSuccess) {
    s->wpb = w0; s->walk = k0; s->max_blocks = b0;
    return hipfail(c, e, "dg_solver_autotune");
  }
  s->wpb = bw; s->walk = bk; s->max_blocks = bb;
  if (ntable) *ntable = n;
  return DG_OK;
}

}  // extern "C"
