// Forward-mode dual numbers (value + one directional derivative), usable in
// host and device code.  The tangent-linear model runs the solver's point
// functions in this arithmetic; small Jacobians (boundary states, geometry)
// are assembled from one dual evaluation per input direction.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define DG_HD __host__ __device__
#else
#define DG_HD
#endif

namespace dg {

struct Dual {
  double v, d;
  DG_HD Dual() : v(0.0), d(0.0) {}
  DG_HD Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants promote implicitly
  DG_HD Dual(double x, double dx) : v(x), d(dx) {}
};

DG_HD inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
DG_HD inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
DG_HD inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
DG_HD inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
DG_HD inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
DG_HD inline Dual &operator+=(Dual &a, Dual b) { a = a + b; return a; }
DG_HD inline Dual &operator-=(Dual &a, Dual b) { a = a - b; return a; }
DG_HD inline Dual &operator*=(Dual &a, Dual b) { a = a * b; return a; }
DG_HD inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
DG_HD inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
DG_HD inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
DG_HD inline Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return Dual(s, a.d / (2.0 * s));
}
// derivative sign convention at 0: +1 (matches the reverse-mode code)
DG_HD inline Dual fabs(Dual a) { return a.v >= 0.0 ? a : -a; }
DG_HD inline Dual pow(Dual a, double e) {
  const double p = ::pow(a.v, e);
  return Dual(p, e * ::pow(a.v, e - 1.0) * a.d);
}
DG_HD inline Dual sin(Dual a) { return Dual(::sin(a.v), ::cos(a.v) * a.d); }
DG_HD inline Dual cos(Dual a) { return Dual(::cos(a.v), -::sin(a.v) * a.d); }
// double overloads, so that unqualified calls inside namespace dg keep
// resolving to the plain functions for double arguments
DG_HD inline double sqrt(double a) { return ::sqrt(a); }
DG_HD inline double fabs(double a) { return ::fabs(a); }
DG_HD inline double pow(double a, double e) { return ::pow(a, e); }
DG_HD inline double sin(double a) { return ::sin(a); }
DG_HD inline double cos(double a) { return ::cos(a); }
DG_HD inline double val(Dual a) { return a.v; }
DG_HD inline double val(double a) { return a; }

}  // namespace dg
