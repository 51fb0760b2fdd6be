// C ABI of the DG Euler solver (declared in include/dg_euler.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dg_euler.h"
#include "dg_ctx.h"
#include "dg_internal.h"
#include "dg_mesh.h"
#include "dg_ops.h"

namespace {
thread_local std::string g_err;
}

// Device scratch of the no-allocation entry point dg_rhs_adjoint_noalloc: a
// file-scope array every GPU holds its own copy of, for the life of the
// process.  Host code never passes the array's name to a kernel or copy (on
// the host it is only a stand-in for the symbol): dg_ctx_create looks up each
// device's address of it once, with that device current, and everything
// uses that address.
static __device__ __align__(256) double scratch[DG_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Address of this device's copy of `scratch` (the device must be current).
hipError_t scratch_lookup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  double *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_ptr[device] = ptr;
  return e;
}
}  // namespace

namespace dg {

int fail(dg_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  g_err = msg;
  return code;
}

int hipfail(dg_ctx *c, hipError_t e, const char *what) {
  return fail(c, DG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

double *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

Launch launch_of(const dg_ctx *c) {
  Launch L;
  L.N = c->N; L.Np = c->Np; L.Nfp = c->Nfp; L.NF = c->NF;
  L.stream = c->stream;
  L.ops = c->d_ops;
  L.fmask = c->d_fmask;
  L.ph = {c->gam, c->bc_r, c->bc_u, c->bc_v, c->bc_e, c->bc_p};
  uint64_t mask = 0, packed = 0;
  for (int f : c->FMASK) mask |= (1ull << f);
  for (size_t m = 0; m < c->FMASK.size() && m < 16; m++) packed |= (uint64_t)(c->FMASK[m] & 15) << (4 * m);
  L.facenodes = mask;
  L.fmask4 = packed;
  L.cus = c->cus;
  L.max_blocks = c->max_blocks;
  return L;
}

Phys freestream(double gam, double mach, double alpha_deg, double p, double rho) {
  Phys ph;
  ph.gam = gam; ph.bc_p = p; ph.bc_r = rho;
  const double a = std::sqrt(gam * p / rho) * mach;
  if (alpha_deg != 0.0) {
    ph.bc_u = std::sin((M_PI / 2.0) - (alpha_deg * M_PI / 180.0)) * a;
    ph.bc_v = std::cos((M_PI / 2.0) - (alpha_deg * M_PI / 180.0)) * a;
    ph.bc_e = p / (gam - 1.0) + 0.5 * rho * (ph.bc_u * ph.bc_u + ph.bc_v * ph.bc_v);
  } else {
    ph.bc_u = a;
    ph.bc_v = 0.0;
    ph.bc_e = p / (gam - 1.0) + 0.5 * rho * (ph.bc_u * ph.bc_u);
  }
  return ph;
}

}  // namespace dg

namespace {
using dg::fail;
using dg::hipfail;
using dg::launch_of;

int upload_ops(dg_ctx *c) {
  const int Np = c->Np, NF = c->NF;
  std::vector<double> buf(dg::ops_size(Np, NF));
  double *p = buf.data();
  std::memcpy(p, c->Drw.data(), sizeof(double) * Np * Np); p += Np * Np;
  std::memcpy(p, c->Dsw.data(), sizeof(double) * Np * Np); p += Np * Np;
  std::memcpy(p, c->LIFT.data(), sizeof(double) * Np * NF); p += Np * NF;
  std::memcpy(p, c->Dr.data(), sizeof(double) * Np * Np); p += Np * Np;
  std::memcpy(p, c->Ds.data(), sizeof(double) * Np * Np); p += Np * Np;
  std::memcpy(p, c->r.data(), sizeof(double) * Np); p += Np;
  std::memcpy(p, c->s.data(), sizeof(double) * Np);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(c->d_ops, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(c->d_fmask, c->FMASK.data(), sizeof(int) * NF, hipMemcpyHostToDevice));
  // [Np][2]: the face slots m with FMASK[m] == node (-1 when fewer), read by
  // the adjoint's cell pass
  std::vector<int> slots(2 * Np, -1);
  for (int m = 0; m < NF; m++) {
    const int nd = c->FMASK[m];
    if (slots[2 * nd] < 0) slots[2 * nd] = m; else slots[2 * nd + 1] = m;
  }
  HIPCHECK(c, hipMemcpy(c->d_node_slots, slots.data(), sizeof(int) * slots.size(), hipMemcpyHostToDevice));
  return DG_OK;
}

bool mesh_declared(const dg_ctx *c) { return c->ncells >= 0; }

// State records are moved as 16-byte vectors.
bool al16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

extern "C" {

const char *dg_last_error(const dg_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int dg_generate_operators(int order, double *r, double *s, double *Dr, double *Ds, double *Drw,
                          double *Dsw, double *LIFT, int *FMASK, double *mass, double *face_mass) {
  dg::Operators o;
  const char *e = nullptr;
  if (order < 1 || order > 8 || !dg::build_operators(order, o, &e))
    return fail(nullptr, DG_ERR_ARG, std::string("dg_generate_operators: ") + (e ? e : "bad order"));
  auto put = [](double *dst, const std::vector<double> &v) {
    if (dst) std::memcpy(dst, v.data(), sizeof(double) * v.size());
  };
  put(r, o.r); put(s, o.s); put(Dr, o.Dr); put(Ds, o.Ds); put(Drw, o.Drw); put(Dsw, o.Dsw);
  put(LIFT, o.LIFT); put(mass, o.mass); put(face_mass, o.face_mass);
  if (FMASK) std::memcpy(FMASK, o.FMASK.data(), sizeof(int) * o.FMASK.size());
  return DG_OK;
}

int dg_ctx_create(int order, dg_ctx **out) {
  if (!out) return fail(nullptr, DG_ERR_ARG, "dg_ctx_create: null output pointer");
  *out = nullptr;
  if (order < 1 || order > 4)
    return fail(nullptr, DG_ERR_ARG, "dg_ctx_create: order must be in [1, 4] (the calc_dt "
                                     "face-scale indexing of the reference needs Np <= 3*Nfp)");
  dg::Operators ops;
  const char *e = nullptr;
  if (!dg::build_operators(order, ops, &e)) return fail(nullptr, DG_ERR_ARG, e ? e : "operators");
  dg_ctx *c = new dg_ctx();
  c->N = order; c->Np = ops.Np; c->Nfp = ops.Nfp; c->NF = 3 * ops.Nfp;
  c->r = ops.r; c->s = ops.s; c->Dr = ops.Dr; c->Ds = ops.Ds; c->Drw = ops.Drw;
  c->Dsw = ops.Dsw; c->LIFT = ops.LIFT; c->FMASK = ops.FMASK; c->mass = ops.mass;
  c->ones.assign(c->Np, 1.0);
  hipError_t he = hipGetDevice(&c->device);
  if (he == hipSuccess) {
    hipDeviceProp_t prop;
    he = hipGetDeviceProperties(&prop, c->device);
    if (he == hipSuccess) c->cus = prop.multiProcessorCount;
  }
  if (he == hipSuccess) he = hipMalloc(&c->d_ops, sizeof(double) * dg::ops_size(c->Np, c->NF));
  if (he == hipSuccess) he = hipMalloc(&c->d_fmask, sizeof(int) * c->NF);
  if (he == hipSuccess) he = hipMalloc(&c->d_red, sizeof(unsigned long long) * 4);
  if (he == hipSuccess) he = hipMalloc(&c->d_node_slots, sizeof(int) * 2 * c->Np);
  if (he == hipSuccess) he = hipMalloc(&c->d_l2, sizeof(double) * (c->Np * c->Np + 8 * dg::kL2Blocks));
  if (he == hipSuccess)
    he = hipMemcpy(c->d_l2, c->mass.data(), sizeof(double) * c->Np * c->Np, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = scratch_lookup(c->device);
  if (he != hipSuccess) {
    std::string msg = std::string("dg_ctx_create: ") + hipGetErrorString(he);
    dg_ctx_destroy(c);
    return fail(nullptr, DG_ERR_HIP, msg);
  }
  // free stream of src/airfoil.cpp:92-99 (alpha = 0)
  c->bc_u = std::sqrt(c->gam * c->bc_p / c->bc_r) * c->bc_mach;
  c->bc_v = 0.0;
  c->bc_e = c->bc_p / (c->gam - 1.0) + 0.5 * c->bc_r * (c->bc_u * c->bc_u);
  int rc = upload_ops(c);
  if (rc != DG_OK) {
    std::string msg = c->err;
    dg_ctx_destroy(c);
    return fail(nullptr, rc, msg);
  }
  *out = c;
  return DG_OK;
}

void dg_ctx_destroy(dg_ctx *c) {
  if (!c) return;
  (void)hipFree(c->d_ops); (void)hipFree(c->d_fmask); (void)hipFree(c->d_red);
  (void)hipFree(c->d_fmap); (void)hipFree(c->d_b2c); (void)hipFree(c->d_bnum);
  (void)hipFree(c->d_btype); (void)hipFree(c->d_node_slots); (void)hipFree(c->d_l2);
  delete c;
}

int dg_order(const dg_ctx *c, int *N, int *Np, int *Nfp) {
  if (!c) return fail(nullptr, DG_ERR_ARG, "dg_order: null context");
  if (N) *N = c->N;
  if (Np) *Np = c->Np;
  if (Nfp) *Nfp = c->Nfp;
  return DG_OK;
}

int dg_set_max_blocks(dg_ctx *c, int max_blocks) {
  if (!c) return fail(nullptr, DG_ERR_ARG, "dg_set_max_blocks: null context");
  if (max_blocks < 0) return fail(c, DG_ERR_ARG, "dg_set_max_blocks: max_blocks must be >= 0");
  c->max_blocks = max_blocks;
  return DG_OK;
}

int dg_set_stream(dg_ctx *c, void *s) {
  if (!c) return fail(nullptr, DG_ERR_ARG, "dg_set_stream: null context");
  c->stream = (hipStream_t)s;
  return DG_OK;
}

namespace {
struct ConstSlot {
  std::vector<double> *vec;
  double *scalar;
  int dim;
};

bool find_const(dg_ctx *c, const char *name, ConstSlot &s) {
  const int Np = c->Np, NF = c->NF;
  struct { const char *n; std::vector<double> *v; double *sc; int dim; } tab[] = {
      {"gam", nullptr, &c->gam, 1},         {"bc_mach", nullptr, &c->bc_mach, 1},
      {"bc_alpha", nullptr, &c->bc_alpha, 1}, {"bc_p", nullptr, &c->bc_p, 1},
      {"bc_r", nullptr, &c->bc_r, 1},       {"bc_u", nullptr, &c->bc_u, 1},
      {"bc_v", nullptr, &c->bc_v, 1},       {"bc_e", nullptr, &c->bc_e, 1},
      {"ones", &c->ones, nullptr, Np},      {"r", &c->r, nullptr, Np},
      {"s", &c->s, nullptr, Np},            {"Dr", &c->Dr, nullptr, Np * Np},
      {"Ds", &c->Ds, nullptr, Np * Np},     {"Drw", &c->Drw, nullptr, Np * Np},
      {"Dsw", &c->Dsw, nullptr, Np * Np},   {"LIFT", &c->LIFT, nullptr, Np * NF},
      // the vortex app's constants (vortex/src/vortex.cpp:171-174)
      {"vortex_x0", nullptr, &c->vortex.x0, 1}, {"vortex_y0", nullptr, &c->vortex.y0, 1},
      {"vortex_beta", nullptr, &c->vortex.beta, 1},
  };
  for (auto &t : tab)
    if (std::strcmp(t.n, name) == 0) {
      s = {t.v, t.sc, t.dim};
      return true;
    }
  return false;
}
}  // namespace

int dg_set_const(dg_ctx *c, const char *name, int dim, const void *data) {
  if (!c || !name || !data) return fail(c, DG_ERR_ARG, "dg_set_const: null argument");
  if (std::strcmp(name, "FMASK") == 0) {
    if (dim != c->NF) return fail(c, DG_ERR_ARG, "dg_set_const: FMASK needs dim 3*Nfp");
    const int *f = (const int *)data;
    for (int i = 0; i < dim; i++)
      if (f[i] < 0 || f[i] >= c->Np) return fail(c, DG_ERR_ARG, "dg_set_const: FMASK entry out of range");
    c->FMASK.assign(f, f + dim);
    return upload_ops(c);
  }
  ConstSlot s;
  if (!find_const(c, name, s))
    return fail(c, DG_ERR_NAME, std::string("dg_set_const: unknown constant name '") + name + "'");
  if (dim != s.dim)
    return fail(c, DG_ERR_ARG, std::string("dg_set_const: wrong dim for '") + name + "'");
  const double *d = (const double *)data;
  if (s.scalar) {
    *s.scalar = d[0];
    return DG_OK;
  }
  s.vec->assign(d, d + dim);
  return upload_ops(c);
}

int dg_get_const(const dg_ctx *cc, const char *name, int dim, void *out) {
  dg_ctx *c = const_cast<dg_ctx *>(cc);
  if (!c || !name || !out) return fail(c, DG_ERR_ARG, "dg_get_const: null argument");
  if (std::strcmp(name, "FMASK") == 0) {
    if (dim != c->NF) return fail(c, DG_ERR_ARG, "dg_get_const: FMASK needs dim 3*Nfp");
    std::memcpy(out, c->FMASK.data(), sizeof(int) * dim);
    return DG_OK;
  }
  ConstSlot s;
  if (!find_const(c, name, s))
    return fail(c, DG_ERR_NAME, std::string("dg_get_const: unknown constant name '") + name + "'");
  if (dim != s.dim) return fail(c, DG_ERR_ARG, std::string("dg_get_const: wrong dim for '") + name + "'");
  if (s.scalar)
    *(double *)out = *s.scalar;
  else
    std::memcpy(out, s.vec->data(), sizeof(double) * dim);
  return DG_OK;
}

int dg_set_freestream(dg_ctx *c, double gam, double mach, double alpha_deg, double p, double rho) {
  if (!c) return fail(nullptr, DG_ERR_ARG, "dg_set_freestream: null context");
  if (!(gam > 1.0) || !(p > 0.0) || !(rho > 0.0) || !(mach >= 0.0))
    return fail(c, DG_ERR_ARG, "dg_set_freestream: need gam > 1, p > 0, rho > 0, mach >= 0");
  const dg::Phys ph = dg::freestream(gam, mach, alpha_deg, p, rho);
  c->gam = gam; c->bc_mach = mach; c->bc_alpha = alpha_deg; c->bc_p = p; c->bc_r = rho;
  c->bc_u = ph.bc_u; c->bc_v = ph.bc_v; c->bc_e = ph.bc_e;
  return DG_OK;
}

int dg_decl_mesh(dg_ctx *c, int ncells, const int *cell2nodes, int nedges, const int *edge2cells,
                 const int *edgeNum, int nbedges, const int *bedge2cells, const int *bedgeNum,
                 const int *bedge_type) {
  if (!c) return fail(nullptr, DG_ERR_ARG, "dg_decl_mesh: null context");
  if (ncells < 0 || nedges < 0 || nbedges < 0 || (nedges > 0 && (!edge2cells || !edgeNum)) ||
      (nbedges > 0 && (!bedge2cells || !bedgeNum || !bedge_type)))
    return fail(c, DG_ERR_ARG, "dg_decl_mesh: bad sizes or null arrays");
  if ((long long)ncells * 4 >= (1ll << 31))
    return fail(c, DG_ERR_ARG, "dg_decl_mesh: too many cells for the 32-bit face map");
  std::vector<int> fmap;
  std::string err;
  if (!dg::build_face_map(ncells, nedges, edge2cells, edgeNum, nbedges, bedge2cells, bedgeNum,
                          bedge_type, cell2nodes, fmap, err))
    return fail(c, DG_ERR_MESH, "dg_decl_mesh: " + err);
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(c->d_fmap); (void)hipFree(c->d_b2c); (void)hipFree(c->d_bnum);
  (void)hipFree(c->d_btype);
  c->d_fmap = c->d_b2c = c->d_bnum = c->d_btype = nullptr;
  c->ncells = -1;
  if (ncells > 0) {
    HIPCHECK(c, hipMalloc(&c->d_fmap, sizeof(int) * fmap.size()));
    HIPCHECK(c, hipMemcpy(c->d_fmap, fmap.data(), sizeof(int) * fmap.size(), hipMemcpyHostToDevice));
  }
  if (nbedges > 0) {
    HIPCHECK(c, hipMalloc(&c->d_b2c, sizeof(int) * nbedges));
    HIPCHECK(c, hipMalloc(&c->d_bnum, sizeof(int) * nbedges));
    HIPCHECK(c, hipMalloc(&c->d_btype, sizeof(int) * nbedges));
    HIPCHECK(c, hipMemcpy(c->d_b2c, bedge2cells, sizeof(int) * nbedges, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_bnum, bedgeNum, sizeof(int) * nbedges, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_btype, bedge_type, sizeof(int) * nbedges, hipMemcpyHostToDevice));
  }
  c->ncells = ncells; c->nedges = nedges; c->nbedges = nbedges;
  c->has_exact = false;
  for (int b = 0; b < nbedges; b++) c->has_exact |= bedge_type[b] == dg::kBcExact;
  return DG_OK;
}

#define NEED_MESH(c, fn)                                                              \
  do {                                                                                \
    if (!(c)) return fail(nullptr, DG_ERR_ARG, fn ": null context");                  \
    if (!mesh_declared(c)) return fail((c), DG_ERR_STATE, fn ": no mesh declared");   \
  } while (0)

int dg_init_grid(dg_ctx *c, const double *cell_xy, double *nodeX, double *nodeY, double *x,
                 double *y, double *xr, double *yr, double *xs, double *ys, double *rx, double *ry,
                 double *sx, double *sy, double *nx, double *ny, double *fscale) {
  NEED_MESH(c, "dg_init_grid");
  if (c->ncells > 0 && (!cell_xy || !nodeX || !nodeY || !x || !y || !xr || !yr || !xs || !ys ||
                        !rx || !ry || !sx || !sy || !nx || !ny || !fscale))
    return fail(c, DG_ERR_ARG, "dg_init_grid: null array");
  HIPCHECK(c, dg::launch_init_grid(launch_of(c), c->ncells, cell_xy, nodeX, nodeY, x, y, xr, yr,
                                   xs, ys, rx, ry, sx, sy, nx, ny, fscale, nullptr));
  return DG_OK;
}

int dg_set_ic(dg_ctx *c, double *q, double *w) {
  NEED_MESH(c, "dg_set_ic");
  if (c->ncells > 0 && (!q || !w)) return fail(c, DG_ERR_ARG, "dg_set_ic: null array");
  HIPCHECK(c, dg::launch_set_ic(launch_of(c), c->ncells, q, w));
  return DG_OK;
}

int dg_neighbour_zero(dg_ctx *c, double *ext) {
  NEED_MESH(c, "dg_neighbour_zero");
  if (c->ncells > 0 && !ext) return fail(c, DG_ERR_ARG, "dg_neighbour_zero: null array");
  HIPCHECK(c, dg::launch_zero(launch_of(c), (size_t)c->ncells * c->NF * 4, ext));
  return DG_OK;
}

int dg_get_neighbour_q(dg_ctx *c, const double *nodeX, const double *nodeY, const double *w,
                       double *ext) {
  (void)nodeX; (void)nodeY;
  NEED_MESH(c, "dg_get_neighbour_q");
  if (c->ncells > 0 && (!w || !ext)) return fail(c, DG_ERR_ARG, "dg_get_neighbour_q: null array");
  if (!al16(w) || !al16(ext)) return fail(c, DG_ERR_ARG, "dg_get_neighbour_q: arrays must be 16-byte aligned");
  HIPCHECK(c, dg::launch_get_neighbour_q(launch_of(c), c->ncells, c->d_fmap, w, ext));
  return DG_OK;
}

int dg_get_bedge_q(dg_ctx *c, const double *nx, const double *ny, const double *w, double *ext) {
  NEED_MESH(c, "dg_get_bedge_q");
  if (c->nbedges > 0 && (!nx || !ny || !w || !ext))
    return fail(c, DG_ERR_ARG, "dg_get_bedge_q: null array");
  HIPCHECK(c, dg::launch_get_bedge_q(launch_of(c), c->nbedges, c->d_b2c, c->d_bnum, c->d_btype, nx,
                                     ny, w, ext));
  return DG_OK;
}

int dg_get_bedge_q_exact(dg_ctx *c, const double *x, const double *y, double t, double *ext) {
  NEED_MESH(c, "dg_get_bedge_q_exact");
  if (c->nbedges > 0 && (!x || !y || !ext)) return fail(c, DG_ERR_ARG, "dg_get_bedge_q_exact: null array");
  HIPCHECK(c, dg::launch_exact_bc(launch_of(c), c->nbedges, c->d_b2c, c->d_bnum, nullptr, x, y, c->vortex, t,
                                  nullptr, nullptr, 0.0, nullptr, 0.0, true, ext));
  return DG_OK;
}

int dg_set_ic_vortex(dg_ctx *c, double t, const double *x, const double *y, double *q, double *w) {
  NEED_MESH(c, "dg_set_ic_vortex");
  if (c->ncells > 0 && (!x || !y || !q || !w)) return fail(c, DG_ERR_ARG, "dg_set_ic_vortex: null array");
  const dg::Vortex &v = c->vortex;
  HIPCHECK(c, dg::launch_set_vortex(launch_of(c), c->ncells, x, y, v.uinf, v.beta, v.x0, v.y0, v.period, t, q, w));
  return DG_OK;
}

int dg_set_exact_state(dg_ctx *c, double uinf, double beta, double x0, double y0, double period_x) {
  if (!c) return fail(nullptr, DG_ERR_ARG, "dg_set_exact_state: null context");
  if (!(period_x >= 0.0)) return fail(c, DG_ERR_ARG, "dg_set_exact_state: period_x must be >= 0");
  c->vortex.uinf = uinf; c->vortex.beta = beta; c->vortex.x0 = x0; c->vortex.y0 = y0;
  c->vortex.period = period_x;
  return DG_OK;
}

// Errors against the exact vortex: the squared nodal errors (optional) and
// the workgroup partials, summed here in workgroup order.
static int l2_errors(dg_ctx *c, const char *who, const double *x, const double *y, const double *xr,
                     const double *yr, const double *xs, const double *ys, const double *q, double t,
                     double *error, double sq[4], double wsum[4]) {
  if (!c) return fail(nullptr, DG_ERR_ARG, std::string(who) + ": null context");
  if (!mesh_declared(c)) return fail(c, DG_ERR_STATE, std::string(who) + ": no mesh declared");
  if (c->ncells > 0 && (!x || !y || !q)) return fail(c, DG_ERR_ARG, std::string(who) + ": null array");
  if (!al16(q) || (error && !al16(error)))
    return fail(c, DG_ERR_ARG, std::string(who) + ": q and error must be 16-byte aligned");
  const int g = dg::l2_error_blocks(c->Np, c->ncells);
  for (int k = 0; k < 4; k++) sq[k] = wsum[k] = 0.0;
  if (g == 0) return DG_OK;
  double *partial = c->d_l2 + c->Np * c->Np;
  HIPCHECK(c, dg::launch_l2_error(launch_of(c), c->ncells, c->vortex, t, x, y, q, xr, yr, xs, ys, c->d_l2,
                                  error, partial));
  std::vector<double> h(8 * (size_t)g);
  HIPCHECK(c, hipMemcpyAsync(h.data(), partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int b = 0; b < g; b++)
    for (int k = 0; k < 4; k++) {
      sq[k] += h[8 * b + k];
      wsum[k] += h[8 * b + 4 + k];
    }
  return DG_OK;
}

int dg_l2_error(dg_ctx *c, const double *x, const double *y, const double *q, double t, double *error,
                double err[4]) {
  if (!err) return fail(c, DG_ERR_ARG, "dg_l2_error: null err");
  double sq[4], ws[4];
  int rc = l2_errors(c, "dg_l2_error", x, y, nullptr, nullptr, nullptr, nullptr, q, t, error, sq, ws);
  if (rc != DG_OK) return rc;
  for (int k = 0; k < 4; k++) err[k] += sq[k];
  return DG_OK;
}

int dg_l2_error_mass(dg_ctx *c, const double *x, const double *y, const double *xr, const double *yr,
                     const double *xs, const double *ys, const double *q, double t, double err[4]) {
  if (!err) return fail(c, DG_ERR_ARG, "dg_l2_error_mass: null err");
  if (c && c->ncells > 0 && (!xr || !yr || !xs || !ys))
    return fail(c, DG_ERR_ARG, "dg_l2_error_mass: null metric array");
  double sq[4], ws[4];
  int rc = l2_errors(c, "dg_l2_error_mass", x, y, xr, yr, xs, ys, q, t, nullptr, sq, ws);
  if (rc != DG_OK) return rc;
  for (int k = 0; k < 4; k++) err[k] = std::sqrt(ws[k]);
  return DG_OK;
}

int dg_euler_rhs(dg_ctx *c, const double *w, double *ext, const double *rx, const double *ry,
                 const double *sx, const double *sy, const double *fscale, const double *nx,
                 const double *ny, double *qRHS) {
  NEED_MESH(c, "dg_euler_rhs");
  if (c->ncells > 0 && (!w || !ext || !rx || !ry || !sx || !sy || !fscale || !nx || !ny || !qRHS))
    return fail(c, DG_ERR_ARG, "dg_euler_rhs: null array");
  if (!al16(w) || !al16(ext) || !al16(qRHS))
    return fail(c, DG_ERR_ARG, "dg_euler_rhs: state arrays must be 16-byte aligned");
  dg::StageArgs a{};
  a.w = w; a.ext = ext; a.rx = rx; a.ry = ry; a.sx = sx; a.sy = sy;
  a.fscale = fscale; a.nx = nx; a.ny = ny; a.out = qRHS; a.ncells = c->ncells;
  HIPCHECK(c, dg::launch_rhs(launch_of(c), a, false, dg::EPI_RHS));
  return DG_OK;
}

int dg_rhs_fused(dg_ctx *c, const double *w, const double *rx, const double *ry, const double *sx,
                 const double *sy, const double *fscale, const double *nx, const double *ny,
                 double *qRHS) {
  NEED_MESH(c, "dg_rhs_fused");
  if (c->ncells > 0 && (!w || !rx || !ry || !sx || !sy || !fscale || !nx || !ny || !qRHS))
    return fail(c, DG_ERR_ARG, "dg_rhs_fused: null array");
  if (!al16(w) || !al16(qRHS)) return fail(c, DG_ERR_ARG, "dg_rhs_fused: state arrays must be 16-byte aligned");
  if (c->has_exact)
    return fail(c, DG_ERR_ARG, "dg_rhs_fused: the mesh has exact-state boundaries (type 4), whose states come "
                               "from dg_get_bedge_q_exact; use dg_get_neighbour_q, dg_get_bedge_q, "
                               "dg_get_bedge_q_exact and dg_euler_rhs");
  dg::StageArgs a{};
  a.w = w; a.rx = rx; a.ry = ry; a.sx = sx; a.sy = sy; a.fscale = fscale; a.nx = nx; a.ny = ny;
  a.fmap = c->d_fmap; a.out = qRHS; a.ncells = c->ncells;
  HIPCHECK(c, dg::launch_rhs(launch_of(c), a, true, dg::EPI_RHS));
  return DG_OK;
}

int dg_set_workingQ(dg_ctx *c, double dt, int stage, const double *q, const double *k1,
                    const double *k2, double *w) {
  NEED_MESH(c, "dg_set_workingQ");
  if (c->ncells > 0 && (!q || !k1 || !w || (stage != 0 && !k2)))
    return fail(c, DG_ERR_ARG, "dg_set_workingQ: null array");
  HIPCHECK(c, dg::launch_set_workingQ(launch_of(c), c->ncells, dt, stage, q, k1, k2 ? k2 : k1, w));
  return DG_OK;
}

int dg_update_Q(dg_ctx *c, double dt, double *q, const double *k1, const double *k2,
                const double *k3, double *w) {
  NEED_MESH(c, "dg_update_Q");
  if (c->ncells > 0 && (!q || !k1 || !k2 || !k3 || !w))
    return fail(c, DG_ERR_ARG, "dg_update_Q: null array");
  HIPCHECK(c, dg::launch_update_Q(launch_of(c), c->ncells, dt, q, k1, k2, k3, w));
  return DG_OK;
}

int dg_calc_dt(dg_ctx *c, const double *q, const double *fscale, double *dt1) {
  NEED_MESH(c, "dg_calc_dt");
  if (!dt1 || (c->ncells > 0 && (!q || !fscale))) return fail(c, DG_ERR_ARG, "dg_calc_dt: null argument");
  HIPCHECK(c, hipMemsetAsync(c->d_red, 0, sizeof(unsigned long long), c->stream));
  HIPCHECK(c, dg::launch_calc_dt(launch_of(c), c->ncells, q, fscale, c->d_red));
  unsigned long long bits = 0;
  HIPCHECK(c, hipMemcpyAsync(&bits, c->d_red, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  double v;
  std::memcpy(&v, &bits, sizeof(v));
  if (v > *dt1 || std::isnan(v)) *dt1 = v;
  return DG_OK;
}

/* ---------------------------- meshes ---------------------------- */

int dg_mesh_rect(int nx, int ny, double x0, double y0, double Lx, double Ly, uint64_t seed, double jitter,
                 int bc_type, dg_mesh **out) {
  if (!out) return fail(nullptr, DG_ERR_ARG, "dg_mesh_rect: null output");
  *out = nullptr;
  if (nx < 1 || ny < 1 || !(Lx > 0) || !(Ly > 0) || !(jitter >= 0.0) || jitter >= 0.5 || bc_type < 0 ||
      bc_type > 4)
    return fail(nullptr, DG_ERR_ARG, "dg_mesh_rect: need nx, ny >= 1, L > 0, 0 <= jitter < 0.5, type in [0, 4]");
  dg_mesh *m = new dg_mesh();
  std::string err;
  if (!dg::gen_rect(nx, ny, x0, y0, Lx, Ly, seed, jitter, bc_type, m->m, err)) {
    delete m;
    return fail(nullptr, DG_ERR_MESH, "dg_mesh_rect: " + err);
  }
  *out = m;
  return DG_OK;
}

int dg_mesh_periodic(int nx, int ny, double x0, double y0, double Lx, double Ly, uint64_t seed,
                     double jitter, dg_mesh **out) {
  if (!out) return fail(nullptr, DG_ERR_ARG, "dg_mesh_periodic: null output");
  *out = nullptr;
  if (nx < 3 || ny < 3 || !(Lx > 0) || !(Ly > 0) || !(jitter >= 0.0) || jitter >= 0.5)
    return fail(nullptr, DG_ERR_ARG, "dg_mesh_periodic: need nx, ny >= 3, L > 0, 0 <= jitter < 0.5");
  dg_mesh *m = new dg_mesh();
  std::string err;
  if (!dg::gen_periodic(nx, ny, x0, y0, Lx, Ly, seed, jitter, m->m, err)) {
    delete m;
    return fail(nullptr, DG_ERR_MESH, "dg_mesh_periodic: " + err);
  }
  *out = m;
  return DG_OK;
}

int dg_mesh_ogrid(int ncirc, int nrad, double stretch, double rfar, dg_mesh **out) {
  if (!out) return fail(nullptr, DG_ERR_ARG, "dg_mesh_ogrid: null output");
  *out = nullptr;
  dg_mesh *m = new dg_mesh();
  std::string err;
  if (!dg::gen_ogrid(ncirc, nrad, stretch, rfar, m->m, err)) {
    delete m;
    return fail(nullptr, DG_ERR_MESH, "dg_mesh_ogrid: " + err);
  }
  *out = m;
  return DG_OK;
}

int dg_mesh_check(int ncells, const int *cell2nodes, int nedges, const int *edge2cells,
                  const int *edgeNum, int nbedges, const int *bedge2cells, const int *bedgeNum,
                  const int *bedge_type) {
  if (ncells < 0 || nedges < 0 || nbedges < 0 || (nedges > 0 && (!edge2cells || !edgeNum)) ||
      (nbedges > 0 && (!bedge2cells || !bedgeNum || !bedge_type)))
    return fail(nullptr, DG_ERR_ARG, "dg_mesh_check: bad sizes or null arrays");
  std::vector<int> fmap;
  std::string err;
  if (!dg::build_face_map(ncells, nedges, edge2cells, edgeNum, nbedges, bedge2cells, bedgeNum,
                          bedge_type, cell2nodes, fmap, err))
    return fail(nullptr, DG_ERR_MESH, "dg_mesh_check: " + err);
  return DG_OK;
}

int dg_mesh_read_vtk(const char *path, dg_mesh **out) {
  if (!out || !path) return fail(nullptr, DG_ERR_ARG, "dg_mesh_read_vtk: null argument");
  *out = nullptr;
  dg_mesh *m = new dg_mesh();
  std::string err;
  if (!dg::read_vtk(path, m->m, err)) {
    delete m;
    return fail(nullptr, DG_ERR_MESH, "dg_mesh_read_vtk: " + err);
  }
  *out = m;
  return DG_OK;
}

void dg_mesh_destroy(dg_mesh *m) { delete m; }

int dg_mesh_sizes(const dg_mesh *m, int *nnodes, int *ncells, int *nedges, int *nbedges) {
  if (!m) return fail(nullptr, DG_ERR_ARG, "dg_mesh_sizes: null mesh");
  if (nnodes) *nnodes = m->m.nnodes;
  if (ncells) *ncells = m->m.ncells;
  if (nedges) *nedges = m->m.nedges;
  if (nbedges) *nbedges = m->m.nbedges;
  return DG_OK;
}

int dg_mesh_set_bedge_types(dg_mesh *m, int nbedges, const int *types) {
  if (!m || (nbedges > 0 && !types)) return fail(nullptr, DG_ERR_ARG, "dg_mesh_set_bedge_types: null argument");
  if (nbedges != m->m.nbedges)
    return fail(nullptr, DG_ERR_ARG, "dg_mesh_set_bedge_types: nbedges must equal the mesh's " +
                                         std::to_string(m->m.nbedges) + " boundary edges");
  for (int b = 0; b < nbedges; b++)
    if (types[b] < DG_BC_INFLOW || types[b] > DG_BC_EXACT)
      return fail(nullptr, DG_ERR_ARG, "dg_mesh_set_bedge_types: type " + std::to_string(types[b]) +
                                           " of boundary edge " + std::to_string(b) + " is not in [0, 4]");
  std::copy(types, types + nbedges, m->m.bedge_type.begin());
  return DG_OK;
}

int dg_mesh_get(const dg_mesh *m, const char *field, void *out) {
  if (!m || !field || !out) return fail(nullptr, DG_ERR_ARG, "dg_mesh_get: null argument");
  const dg::Mesh &M = m->m;
  auto cpy_i = [&](const std::vector<int> &v) { std::memcpy(out, v.data(), sizeof(int) * v.size()); };
  auto cpy_d = [&](const std::vector<double> &v) {
    std::memcpy(out, v.data(), sizeof(double) * v.size());
  };
  std::string f(field);
  if (f == "coords") cpy_d(M.coords);
  else if (f == "cell_xy") cpy_d(M.cell_xy);
  else if (f == "cells") cpy_i(M.cells);
  else if (f == "edge2cells") cpy_i(M.edge2cells);
  else if (f == "edgeNum") cpy_i(M.edgeNum);
  else if (f == "edge2nodes") cpy_i(M.edge2nodes);
  else if (f == "bedge2cells") cpy_i(M.bedge2cells);
  else if (f == "bedgeNum") cpy_i(M.bedgeNum);
  else if (f == "bedge_type") cpy_i(M.bedge_type);
  else if (f == "bedge2nodes") cpy_i(M.bedge2nodes);
  else return fail(nullptr, DG_ERR_NAME, "dg_mesh_get: unknown field '" + f + "'");
  return DG_OK;
}

}  // extern "C"

