// yukawa.h -- host side of the Yukawa (screened Coulomb) kernel, neq 10 of
// the reference (kernel/Helmholtz_*.cxx): radial functions and the dense
// translation operators on the scaled bases described in yukawa.cpp.
#pragma once

namespace fmmx {

// Boxes with lambda * (box size) above this have no far field worth
// translating (exp(-50) ~ 2e-22 across one box): their operators are zero.
constexpr double YUK_LRB_MAX = 50.0;

// i~_n(x) = i_n(x) / x^n, n = 0..N (modified spherical Bessel, first kind)
void ibar_all(double x, int N, double* out);
// k~_n(x) = k_n(x) x^(n+1), n = 0..N, k_0(x) = (pi/2) exp(-x) / x
void kbar_all(double x, int N, double* out);

// Real p^2 x p^2 operators (row-major, slots of slot_re / slot_im) at
// lrb = lambda * (box size of the level):
//   M2L for target-minus-source offset (dx, dy, dz) in box units;
//   M2M from child octant `oct` to its parent and L2L from the parent to
//   that child, lrb_child = lambda * (child box size).
void yuk_op_m2l(int p, double lrb, int dx, int dy, int dz, double* A);
void yuk_op_m2m_l2l(int p, double lrb_child, int oct, double* Am2m, double* Al2l);

}  // namespace fmmx
