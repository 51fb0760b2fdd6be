// Host check of boundary_vjp (the hand-written reverse of the boundary
// states, csrc/dg_bc.h) against boundary_t in dual arithmetic: for random
// states, normals, free-stream velocities, exterior cotangents bp and input
// directions dx, bp . (J dx) from one dual evaluation must equal
// (J^T bp) . dx from the reverse, in every branch (types 0-3; type 3
// subsonic and supersonic, inflow and outflow).
#include <cmath>
#include <cstdio>
#include <random>

#include "../../airfoil-op2-fem_amd/csrc/dg_bc.h"

struct Ph {
  double gam, bc_r, bc_u, bc_v, bc_e, bc_p;
};

int main() {
  std::mt19937_64 rng(20261204);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  const double g = 1.4;
  int fails = 0, count[4] = {0, 0, 0, 0}, regime[4] = {0, 0, 0, 0};  // regime: type 3 branch
  double worst = 0.0;
  for (int it = 0; it < 4000; it++) {
    const int type = it % 4;
    const double mach = 0.2 + 0.6 * (0.5 + 0.5 * U(rng));
    const double al = 0.3 * U(rng);
    Ph ph{g, 1.0 + 0.1 * U(rng), 0.0, 0.0, 0.0, 1.0 / g};
    ph.bc_u = mach * std::sqrt(g * ph.bc_p / ph.bc_r) * std::cos(al);
    ph.bc_v = mach * std::sqrt(g * ph.bc_p / ph.bc_r) * std::sin(al);
    ph.bc_e = ph.bc_p / (g - 1.0) + 0.5 * ph.bc_r * (ph.bc_u * ph.bc_u + ph.bc_v * ph.bc_v);
    const double th = 3.14159265358979 * U(rng);
    const double nx = std::cos(th), ny = std::sin(th);
    // interior state: speeds from well below to above the sound speed, so
    // that type 3 runs through all its branches
    const double rho = 1.0 + 0.3 * U(rng), p = (1.0 + 0.3 * U(rng)) / g, c = std::sqrt(g * p / rho);
    const double vn = c * 1.6 * U(rng), vt = c * U(rng);
    const double u = vn * nx - vt * ny, v = vn * ny + vt * nx;
    const double q[4] = {rho, rho * u, rho * v, p / (g - 1.0) + 0.5 * rho * (u * u + v * v)};
    double bp[4], dx[8];
    for (double &x : bp) x = U(rng);
    for (double &x : dx) x = U(rng);
    // forward derivative along dx
    dg::Dual dq[4], qb[4];
    for (int k = 0; k < 4; k++) dq[k] = dg::Dual(q[k], dx[k]);
    dg::boundary_t<dg::Dual>(ph, type, dg::Dual(ph.bc_u, dx[6]), dg::Dual(ph.bc_v, dx[7]), dq,
                                 dg::Dual(nx, dx[4]), dg::Dual(ny, dx[5]), qb);
    double lhs = 0.0, scale = 0.0;
    for (int k = 0; k < 4; k++) {
      lhs += bp[k] * qb[k].d;
      scale += std::fabs(bp[k] * qb[k].d);
    }
    // reverse
    double bq[4] = {0, 0, 0, 0}, bnx = 0, bny = 0, bu = 0, bv = 0;
    dg::boundary_vjp(ph, type, ph.bc_u, ph.bc_v, q, nx, ny, bp, bq, bnx, bny, bu, bv);
    const double rhs = bq[0] * dx[0] + bq[1] * dx[1] + bq[2] * dx[2] + bq[3] * dx[3] + bnx * dx[4] + bny * dx[5] +
                       bu * dx[6] + bv * dx[7];
    const double err = std::fabs(lhs - rhs) / (scale + 1e-300);
    if (err > worst) worst = err;
    if (!(err <= 1e-12)) {
      if (fails < 10) std::fprintf(stderr, "type %d: %.17g vs %.17g (rel %.3g)\n", type, lhs, rhs, err);
      fails++;
    }
    count[type]++;
    if (type == 3) {
      const double ci = c, vni = vn;
      const double cinf = std::sqrt(g * ph.bc_p / ph.bc_r), vninf = ph.bc_u * nx + ph.bc_v * ny;
      const double vnb = 0.5 * (vni + 2.0 * ci / (g - 1.0) + vninf - 2.0 * cinf / (g - 1.0));
      const int r = std::fabs(vni) >= ci ? (vni < 0.0 ? 0 : 1) : (vnb > 0.0 ? 2 : 3);
      regime[r]++;
    }
  }
  std::printf("cases per type %d %d %d %d; far field supersonic in/out %d %d, subsonic out/in %d %d; worst %.3g\n",
              count[0], count[1], count[2], count[3], regime[0], regime[1], regime[2], regime[3], worst);
  if (fails || regime[0] < 50 || regime[1] < 50 || regime[2] < 50 || regime[3] < 50) return 1;
  std::printf("bc reverse ok\n");
  return 0;
}
