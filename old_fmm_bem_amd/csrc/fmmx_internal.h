// fmmx_internal.h -- shared definitions of the MI355X FMM engine.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fmmx.h"

namespace fmmx {

using cplx = std::complex<double>;

constexpr int NEBM = 189;  // include/constants.h:21, max far-list length
constexpr int NNEAR = 27;  // near list length bound (3x3x3)
constexpr int NRBM = 512;  // include/constants.h:22, lattice directions
constexpr int NOFF = 316;  // M2L offsets: [-3,3]^3 minus [-1,1]^3
constexpr int MAXP = FMMX_MAX_P;
// orders built in every precision (2..P_ALL) and the one order above them,
// FP64 and the EXACT convention only (the reference's mp = 18)
constexpr int P_ALL = 12, P_HI = 18;
constexpr float EPSF = 1e-6f;                      // include/constants.h:34
constexpr float PIF = 3.14159265358979323846f;     // include/constants.h:35

// Real coefficient basis of an expansion truncated at p degrees: the
// p(p+1)/2 packed complex coefficients (n, 0<=m<=n) of the reference
// (include/arrays.h:38) become p*p reals: slot nms = n(n+1)/2+m holds Re,
// slot p(p+1)/2 + n(n-1)/2 + m-1 holds Im for m >= 1 (Im of m = 0 is zero
// for expansions of real sources).
inline int nre(int p) { return p * (p + 1) / 2; }
inline int slot_re(int n, int m) { return n * (n + 1) / 2 + m; }
inline int slot_im(int p, int n, int m) { return nre(p) + n * (n - 1) / 2 + m - 1; }

// FP64 tables of the reference precompute (kernel/precalc.cxx:10-213) for p
// degrees in a convention.  Product copy: built at handle creation.
struct Tables {
  int p = 0, conv = 0;
  std::vector<double> fac, anm, sr;  // float-rounded in FMMX_CONV_REF
  std::vector<cplx> ynm;             // 4p^2
  std::vector<cplx> dnm;             // [2*NRBM][p][p*p]
  std::vector<uint8_t> flip;         // EXACT: direction along -z (rotation by pi)
  std::vector<double> corr;          // per packed (n,m): table fac / exact fac
};
void build_tables(Tables& T, int p, int conv);

// Dense real operators (row = output slot, column = input slot), p*p square,
// acting on scaled expansions: multipoles M~_n = M_n / rb^n and locals
// L~_n = L_n * rb^(n+1), rb the box size of the level the expansion lives on.
// M2M and L2L are level independent; M2L depends on the level only through
// the reference's eps added to the translation distance (kernel/m2l.cxx:54).
void op_m2m(const Tables& T, int octant, double* A);                 // child octant -> parent
void op_l2l(const Tables& T, int octant, double* A);                 // parent -> child octant
void op_m2l(const Tables& T, int dx, int dy, int dz, double rb, double* A);

// Offset index of (dx,dy,dz) in [-3,3]^3 with max|d| >= 2, 0..NOFF-1.
int offset_index(int dx, int dy, int dz);
void offset_vector(int idx, int* d);
// The NCANON offsets with 0 <= dx, dy, dz <= 3 and max >= 2 (every M2L offset
// is one of them reflected through coordinate planes).
constexpr int NCANON = 56;
int canon_index(int ax, int ay, int az);
void canon_vector(int idx, int* d);

std::string& last_error();
int fail(int code, const std::string& msg);

}  // namespace fmmx
