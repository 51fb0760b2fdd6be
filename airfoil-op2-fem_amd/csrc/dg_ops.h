// Nodal DG operators on the reference triangle, generated for any order N.
//
// The reference declares these as order-4 literal tables
// (src/constants/constant_{r,s,Dr,Ds,Drw,Dsw,LIFT}.h, FMASK in
// src/constants/constant_util.h:15) and uploads them with op_decl_const
// (src/airfoil.cpp:192-199).  This generator rebuilds them from the
// warp & blend construction of Hesthaven & Warburton (2008) so that the
// solver runs at any order; at N=4 it reproduces the reference tables to
// round-off.
#pragma once
#include <vector>

namespace dg {

struct Operators {
  int N = 0, Np = 0, Nfp = 0;
  std::vector<double> r, s;          // [Np]
  std::vector<double> Dr, Ds;        // [Np][Np] row-major
  std::vector<double> Drw, Dsw;      // [Np][Np] weak derivatives M^-1 S^T
  std::vector<double> LIFT;          // [Np][3*Nfp]
  std::vector<double> mass;          // [Np][Np] reference-element mass matrix
  std::vector<double> face_mass;     // [Nfp][Nfp] 1D face mass (length-2 face)
  std::vector<int> FMASK;            // [3*Nfp], faces traversed counter-clockwise
};

// Returns false (with a message) when N is outside the supported range.
bool build_operators(int N, Operators &out, const char **err);

}  // namespace dg
