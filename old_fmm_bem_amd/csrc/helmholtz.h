// helmholtz.h -- host side of the oscillatory Helmholtz kernel
// exp(i k r) / (4 pi r) with complex charges: radial functions and the dense
// translation operators on the split expansions described in helmholtz.cpp.
#pragma once

#include "fmmx_internal.h"

namespace fmmx {

// Operators need the local expansion's radial functions to stay away from
// their zeros on the projection sphere (radius half a box): k times the box
// size of the coarsest level (2) must stay below this.
constexpr double HH_KRB_MAX = 5.0;

// j^_n(x) = (2n+1)!! j_n(x) / x^n, n = 0..N (-> 1 as x -> 0)
void jhat_all(double x, int N, double* out);
// h~_n(x) = h_n(x) x^(n+1), h_n = j_n + i y_n (spherical Hankel, first kind)
void htil_all(double x, int N, cplx* out);

// Real (2 p^2) x (2 p^2) operators, row-major, on split expansions (the
// real-field part's slot_re / slot_im slots, then the imaginary-field
// part's), at kappa = k * (box size of the level):
//   M2L for target-minus-source offset (dx, dy, dz) in box units;
//   M2M from child octant `oct` to its parent and L2L from the parent to
//   that child, kappa_child = k * (child box size).
void hh_op_m2l(int p, double kappa, int dx, int dy, int dz, double* A);
void hh_op_m2m_l2l(int p, double kappa_child, int oct, double* Am2m, double* Al2l);

}  // namespace fmmx
