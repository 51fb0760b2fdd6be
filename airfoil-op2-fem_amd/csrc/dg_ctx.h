// Context and mesh objects of the C ABI, shared by its implementation files.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/dg_euler.h"
#include "dg_internal.h"
#include "dg_mesh.h"

struct dg_ctx {
  int N = 0, Np = 0, Nfp = 0, NF = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // host copies of the constants (op_decl_const names)
  std::vector<double> r, s, Dr, Ds, Drw, Dsw, LIFT, ones, mass;
  std::vector<int> FMASK;
  double gam = 1.4, bc_mach = 0.4, bc_alpha = 0.0, bc_p = 1.0, bc_r = 1.0, bc_u = 0.0,
         bc_v = 0.0, bc_e = 0.0;
  double *d_ops = nullptr;
  int *d_fmask = nullptr;
  int cus = 0;
  int max_blocks = 0;
  // declared mesh
  int ncells = -1, nedges = 0, nbedges = 0;
  int *d_fmap = nullptr, *d_b2c = nullptr, *d_bnum = nullptr, *d_btype = nullptr;
  unsigned long long *d_red = nullptr;
  double *d_l2 = nullptr;  // [Np*Np] mass matrix, then [kL2Blocks][8] error partials
  int *d_node_slots = nullptr;  // [Np][2] face slots reading each node (adjoint)
  // exact isentropic-vortex state of boundary type 4 (vortex/src/get_bedge_q.h)
  dg::Vortex vortex;
  bool has_exact = false;  // the declared mesh has boundary edges of type 4
};

struct dg_mesh {
  dg::Mesh m;
};

namespace dg {
// Records the message on ctx (if any) and as the thread's last error.
int fail(dg_ctx *c, int code, const std::string &msg);
int hipfail(dg_ctx *c, hipError_t e, const char *what);
Launch launch_of(const dg_ctx *c);
// Free-stream physics constants for (gam, mach, alpha, p, rho) as
// src/airfoil.cpp:79-99.
Phys freestream(double gam, double mach, double alpha_deg, double p, double rho);
// This device's address of the static scratch of dg_rhs_adjoint_noalloc
// (DG_NOALLOC_SCRATCH_BYTES), looked up by dg_ctx_create; null before any
// context was created on the device.
double *noalloc_scratch(int device);
}  // namespace dg

#define HIPCHECK(ctx, call)                                          \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) return dg::hipfail((ctx), e_, #call);      \
  } while (0)
