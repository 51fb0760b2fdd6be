// Nodal DG operator generator (host C++), Hesthaven & Warburton (2008)
// chapters 3 and 6: Gauss-Lobatto nodes, warp & blend triangle nodes with the
// alpha-optimised blend table, the orthonormal simplex basis, and the
// derived differentiation, weak-differentiation and lift matrices.
#include "dg_ops.h"

#include <cmath>
#include <cstring>

namespace dg {
namespace {

using Mat = std::vector<double>;  // row-major

const double kAlphaOpt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751,
                              0.9800, 1.0999, 1.2832, 1.3648, 1.4773,
                              1.4959, 1.5743, 1.5770, 1.6223, 1.6258};

// Orthonormal Jacobi polynomial P_n^{(a,b)}(x) by the three-term recurrence.
double jacobi_p(double x, double a, double b, int n) {
  double g0 = std::pow(2.0, a + b + 1) / (a + b + 1) * std::tgamma(a + 1) *
              std::tgamma(b + 1) / std::tgamma(a + b + 1);
  double p0 = 1.0 / std::sqrt(g0);
  if (n == 0) return p0;
  double g1 = (a + 1) * (b + 1) / (a + b + 3) * g0;
  double p1 = ((a + b + 2) * x / 2 + (a - b) / 2) / std::sqrt(g1);
  if (n == 1) return p1;
  double aold = 2 / (2 + a + b) * std::sqrt((a + 1) * (b + 1) / (a + b + 3));
  double pm1 = p0, pc = p1;
  for (int i = 1; i < n; i++) {
    double h1 = 2 * i + a + b;
    double anew = 2 / (h1 + 2) *
                  std::sqrt((i + 1) * (i + 1 + a + b) * (i + 1 + a) * (i + 1 + b) /
                            (h1 + 1) / (h1 + 3));
    double bnew = -(a * a - b * b) / h1 / (h1 + 2);
    double pn = 1 / anew * (-aold * pm1 + (x - bnew) * pc);
    pm1 = pc;
    pc = pn;
    aold = anew;
  }
  return pc;
}

double grad_jacobi_p(double x, double a, double b, int n) {
  if (n == 0) return 0.0;
  return std::sqrt(n * (n + a + b + 1)) * jacobi_p(x, a + 1, b + 1, n - 1);
}

// Roots of P_n^{(a,b)} by Newton iteration with deflation of the roots
// already found, started from Chebyshev-Gauss points.
std::vector<double> jacobi_roots(double a, double b, int n) {
  std::vector<double> x(n);
  for (int k = 0; k < n; k++) {
    double r = -std::cos((2.0 * k + 1) * M_PI / (2.0 * n));
    if (k > 0 && r < x[k - 1]) r = x[k - 1] + 1e-3;
    for (int it = 0; it < 100; it++) {
      double p = jacobi_p(r, a, b, n);
      double dp = grad_jacobi_p(r, a, b, n);
      double s = 0.0;
      for (int j = 0; j < k; j++) s += 1.0 / (r - x[j]);
      double delta = p / (dp - s * p);
      r -= delta;
      if (std::fabs(delta) < 1e-16) break;
    }
    x[k] = r;
  }
  // insertion sort (roots come out ascending, this guards against a swap)
  for (int i = 1; i < n; i++)
    for (int j = i; j > 0 && x[j] < x[j - 1]; j--) std::swap(x[j], x[j - 1]);
  return x;
}

std::vector<double> jacobi_gl(int n) {
  std::vector<double> x;
  x.push_back(-1.0);
  if (n > 1) {
    std::vector<double> xi = jacobi_roots(1.0, 1.0, n - 1);
    x.insert(x.end(), xi.begin(), xi.end());
  }
  x.push_back(1.0);
  return x;
}

// Dense LU with partial pivoting; solves A X = B in place (B is n x m).
bool lu_solve(int n, Mat A, Mat &B, int m) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (std::fabs(A[i * n + k]) > std::fabs(A[p * n + k])) p = i;
    if (A[p * n + k] == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; j++) std::swap(A[k * n + j], A[p * n + j]);
      for (int j = 0; j < m; j++) std::swap(B[k * m + j], B[p * m + j]);
    }
    for (int i = k + 1; i < n; i++) {
      double f = A[i * n + k] / A[k * n + k];
      A[i * n + k] = f;
      for (int j = k + 1; j < n; j++) A[i * n + j] -= f * A[k * n + j];
      for (int j = 0; j < m; j++) B[i * m + j] -= f * B[k * m + j];
    }
  }
  for (int i = n - 1; i >= 0; i--)
    for (int j = 0; j < m; j++) {
      double s = B[i * m + j];
      for (int k = i + 1; k < n; k++) s -= A[i * n + k] * B[k * m + j];
      B[i * m + j] = s / A[i * n + i];
    }
  return true;
}

Mat inverse(int n, const Mat &A) {
  Mat I(n * n, 0.0);
  for (int i = 0; i < n; i++) I[i * n + i] = 1.0;
  lu_solve(n, A, I, n);
  return I;
}

Mat matmul(int n, int k, int m, const Mat &A, const Mat &B) {
  Mat C(n * m, 0.0);
  for (int i = 0; i < n; i++)
    for (int l = 0; l < k; l++) {
      double a = A[i * k + l];
      for (int j = 0; j < m; j++) C[i * m + j] += a * B[l * m + j];
    }
  return C;
}

Mat transpose(int n, int m, const Mat &A) {
  Mat T(n * m);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < m; j++) T[j * n + i] = A[i * m + j];
  return T;
}

Mat vandermonde1d(int N, const std::vector<double> &r) {
  int n = (int)r.size();
  Mat V(n * (N + 1));
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= N; j++) V[i * (N + 1) + j] = jacobi_p(r[i], 0, 0, j);
  return V;
}

std::vector<double> warp_factor(int N, const std::vector<double> &rout) {
  std::vector<double> lglr = jacobi_gl(N);
  std::vector<double> req(N + 1);
  for (int i = 0; i <= N; i++) req[i] = -1.0 + 2.0 * i / N;
  Mat veq = vandermonde1d(N, req);  // (N+1) x (N+1)
  int nr = (int)rout.size();
  Mat pmat((N + 1) * nr);
  for (int i = 0; i <= N; i++)
    for (int j = 0; j < nr; j++) pmat[i * nr + j] = jacobi_p(rout[j], 0, 0, i);
  Mat lmat = pmat;
  lu_solve(N + 1, transpose(N + 1, N + 1, veq), lmat, nr);  // veq^T \ pmat
  std::vector<double> warp(nr, 0.0);
  for (int j = 0; j < nr; j++) {
    double w = 0.0;
    for (int i = 0; i <= N; i++) w += lmat[i * nr + j] * (lglr[i] - req[i]);
    double zerof = std::fabs(rout[j]) < 1.0 - 1.0e-10 ? 1.0 : 0.0;
    double sf = 1.0 - (zerof * rout[j]) * (zerof * rout[j]);
    warp[j] = w / sf + w * (zerof - 1);
  }
  return warp;
}

void nodes2d(int N, std::vector<double> &r, std::vector<double> &s) {
  double alpha = N < 16 ? kAlphaOpt[N - 1] : 5.0 / 3.0;
  std::vector<double> L1, L2, L3;
  for (int i = 1; i <= N + 1; i++)
    for (int j = 1; j <= N + 2 - i; j++) {
      L1.push_back((i - 1) / (double)N);
      L3.push_back((j - 1) / (double)N);
    }
  int np = (int)L1.size();
  L2.resize(np);
  std::vector<double> d1(np), d2(np), d3(np);
  for (int k = 0; k < np; k++) {
    L2[k] = 1.0 - L1[k] - L3[k];
    d1[k] = L3[k] - L2[k];
    d2[k] = L1[k] - L3[k];
    d3[k] = L2[k] - L1[k];
  }
  std::vector<double> wf1 = warp_factor(N, d1), wf2 = warp_factor(N, d2),
                      wf3 = warp_factor(N, d3);
  r.resize(np);
  s.resize(np);
  for (int k = 0; k < np; k++) {
    double x = -L2[k] + L3[k];
    double y = (-L2[k] - L3[k] + 2 * L1[k]) / std::sqrt(3.0);
    double w1 = 4 * L2[k] * L3[k] * wf1[k] * (1 + (alpha * L1[k]) * (alpha * L1[k]));
    double w2 = 4 * L1[k] * L3[k] * wf2[k] * (1 + (alpha * L2[k]) * (alpha * L2[k]));
    double w3 = 4 * L1[k] * L2[k] * wf3[k] * (1 + (alpha * L3[k]) * (alpha * L3[k]));
    x = x + w1 + std::cos(2 * M_PI / 3) * w2 + std::cos(4 * M_PI / 3) * w3;
    y = y + std::sin(2 * M_PI / 3) * w2 + std::sin(4 * M_PI / 3) * w3;
    double l1 = (std::sqrt(3.0) * y + 1.0) / 3.0;
    double l2 = (-3.0 * x - std::sqrt(3.0) * y + 2.0) / 6.0;
    double l3 = (3.0 * x - std::sqrt(3.0) * y + 2.0) / 6.0;
    r[k] = -l2 + l3 - l1;
    s[k] = -l2 - l3 + l1;
  }
}

void rs_to_ab(double r, double s, double &a, double &b) {
  a = std::fabs(s - 1.0) > 1e-14 ? 2 * (1 + r) / (1 - s) - 1 : -1.0;
  b = s;
}

void vandermonde2d(int N, const std::vector<double> &r, const std::vector<double> &s,
                   Mat &V, Mat &Vr, Mat &Vs) {
  int np = (int)r.size();
  int nm = (N + 1) * (N + 2) / 2;
  V.assign(np * nm, 0.0);
  Vr.assign(np * nm, 0.0);
  Vs.assign(np * nm, 0.0);
  for (int k = 0; k < np; k++) {
    double a, b;
    rs_to_ab(r[k], s[k], a, b);
    int sk = 0;
    for (int i = 0; i <= N; i++)
      for (int j = 0; j <= N - i; j++, sk++) {
        double h1 = jacobi_p(a, 0, 0, i);
        double h2 = jacobi_p(b, 2 * i + 1, 0, j);
        V[k * nm + sk] = std::sqrt(2.0) * h1 * h2 * std::pow(1 - b, i);
        double fa = h1, dfa = grad_jacobi_p(a, 0, 0, i);
        double gb = h2, dgb = grad_jacobi_p(b, 2 * i + 1, 0, j);
        double dr = dfa * gb;
        if (i > 0) dr *= std::pow(0.5 * (1 - b), i - 1);
        double ds = dfa * (gb * (0.5 * (1 + a)));
        if (i > 0) ds *= std::pow(0.5 * (1 - b), i - 1);
        double tmp = dgb * std::pow(0.5 * (1 - b), i);
        if (i > 0) tmp -= 0.5 * i * gb * std::pow(0.5 * (1 - b), i - 1);
        ds += fa * tmp;
        Vr[k * nm + sk] = dr * std::pow(2.0, i + 0.5);
        Vs[k * nm + sk] = ds * std::pow(2.0, i + 0.5);
      }
  }
}

}  // namespace

bool build_operators(int N, Operators &o, const char **err) {
  if (N < 1 || N > 8) {
    if (err) *err = "order N must be in [1, 8]";
    return false;
  }
  o.N = N;
  o.Np = (N + 1) * (N + 2) / 2;
  o.Nfp = N + 1;
  const int np = o.Np, nfp = o.Nfp;
  nodes2d(N, o.r, o.s);
  Mat V, Vr, Vs;
  vandermonde2d(N, o.r, o.s, V, Vr, Vs);
  Mat Vinv = inverse(np, V);
  o.Dr = matmul(np, np, np, Vr, Vinv);
  o.Ds = matmul(np, np, np, Vs, Vinv);
  Mat Vt = transpose(np, np, V);
  Mat VVt = matmul(np, np, np, V, Vt);
  Mat VVt_inv = inverse(np, VVt);
  o.mass = VVt_inv;
  o.Drw = matmul(np, np, np, matmul(np, np, np, V, transpose(np, np, Vr)), VVt_inv);
  o.Dsw = matmul(np, np, np, matmul(np, np, np, V, transpose(np, np, Vs)), VVt_inv);

  const double tol = 1e-12;
  o.FMASK.clear();
  for (int k = 0; k < np; k++)
    if (std::fabs(o.s[k] + 1) < tol) o.FMASK.push_back(k);
  for (int k = 0; k < np; k++)
    if (std::fabs(o.r[k] + o.s[k]) < tol) o.FMASK.push_back(k);
  for (int k = np - 1; k >= 0; k--)
    if (std::fabs(o.r[k] + 1) < tol) o.FMASK.push_back(k);
  if ((int)o.FMASK.size() != 3 * nfp) {
    if (err) *err = "face node detection failed";
    return false;
  }
  Mat E(np * 3 * nfp, 0.0);
  for (int f = 0; f < 3; f++) {
    std::vector<double> coord(nfp);
    for (int i = 0; i < nfp; i++) {
      int k = o.FMASK[f * nfp + i];
      coord[i] = f == 2 ? o.s[k] : o.r[k];
    }
    Mat v1 = vandermonde1d(N, coord);
    Mat m1 = inverse(nfp, matmul(nfp, nfp, nfp, v1, transpose(nfp, nfp, v1)));
    if (f == 0) o.face_mass = m1;
    for (int i = 0; i < nfp; i++)
      for (int j = 0; j < nfp; j++)
        E[o.FMASK[f * nfp + i] * 3 * nfp + f * nfp + j] = m1[i * nfp + j];
  }
  o.LIFT = matmul(np, np, 3 * nfp, V, matmul(np, np, 3 * nfp, Vt, E));
  return true;
}

}  // namespace dg
