// Host numerics of the engine built with AddressSanitizer and
// UndefinedBehaviorSanitizer (g++, host code only): the FP64 tables and dense
// translation operators of every order and convention (precompute.cpp), the
// Yukawa and Helmholtz projection operators (yukawa.cpp, helmholtz.cpp), the
// offset tables, and the FP16 hi/lo operator split of the M2L kernels
// (split_host.cpp).  Any memory error or undefined behaviour aborts the run;
// the checks below also require finite results and exact table round trips.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../old_fmm_bem_amd/csrc/fmmx_internal.h"
#include "../../old_fmm_bem_amd/csrc/helmholtz.h"
#include "../../old_fmm_bem_amd/csrc/m2l.h"
#include "../../old_fmm_bem_amd/csrc/yukawa.h"

using namespace fmmx;

static int bad = 0;

static void finite(const std::vector<double>& a, const char* what, int p) {
  for (double v : a)
    if (!std::isfinite(v)) {
      std::printf("non-finite entry in %s at p = %d\n", what, p);
      bad++;
      return;
    }
}

int main() {
  // offset tables: index <-> vector round trips
  for (int i = 0; i < NOFF; i++) {
    int d[3];
    offset_vector(i, d);
    if (offset_index(d[0], d[1], d[2]) != i) bad++;
  }
  for (int i = 0; i < NCANON; i++) {
    int d[3];
    canon_vector(i, d);
    if (canon_index(d[0], d[1], d[2]) != i) bad++;
  }
  for (int conv : {FMMX_CONV_REF, FMMX_CONV_EXACT}) {
    for (int p = 2; p <= MAXP; p++) {
      Tables T;
      build_tables(T, p, conv);
      std::vector<double> A((size_t)p * p * p * p);
      for (int oct = 0; oct < 8; oct++) {
        op_m2m(T, oct, A.data());
        finite(A, "M2M", p);
        op_l2l(T, oct, A.data());
        finite(A, "L2L", p);
      }
      for (int i = 0; i < NOFF; i += p <= 6 ? 1 : 29) {
        int d[3];
        offset_vector(i, d);
        op_m2l(T, d[0], d[1], d[2], conv == FMMX_CONV_REF ? 0.0625 : 1.0, A.data());
        finite(A, "M2L", p);
      }
    }
  }
  for (int p : {4, 8, 12, 18}) {
    std::vector<double> A((size_t)p * p * p * p), B(A.size());
    for (double lrb : {0.3, 4.0, 30.0}) {
      yuk_op_m2l(p, lrb, 2, 0, -3, A.data());
      finite(A, "Yukawa M2L", p);
      yuk_op_m2m_l2l(p, lrb, 5, A.data(), B.data());
      finite(A, "Yukawa M2M", p);
      finite(B, "Yukawa L2L", p);
    }
    if (p > 12) continue;  // the Helmholtz kernels stop at p = 12
    std::vector<double> H(4 * A.size()), H2(H.size());
    for (double kappa : {0.5, 2.5}) {
      hh_op_m2l(p, kappa, 3, 1, -2, H.data());
      finite(H, "Helmholtz M2L", p);
      hh_op_m2m_l2l(p, kappa / 2, 6, H.data(), H2.data());
      finite(H, "Helmholtz M2M", p);
      finite(H2, "Helmholtz L2L", p);
    }
  }
  // FP16 split of three M2L operators in both fragment orders
  for (int p = 6; p <= MAXP; p++) {
    Tables T;
    build_tables(T, p, FMMX_CONV_EXACT);
    const size_t sq = (size_t)p * p * p * p;
    std::vector<double> A(sq), kmaj(3 * sq);
    for (int o = 0; o < 3; o++) {
      int d[3];
      offset_vector(o * 100, d);
      op_m2l(T, d[0], d[1], d[2], 1.0, A.data());
      for (int r = 0; r < p * p; r++)
        for (int k = 0; k < p * p; k++) kmaj[o * sq + (size_t)k * p * p + r] = A[(size_t)r * p * p + k];
    }
    const int e = split_scale_exp(kmaj.data(), kmaj.size());
    std::vector<uint4> h, w;
    split_fragments(kmaj.data(), 3, p, e, h);
    if (h.size() != 3 * split_frag_count(p)) bad++;
    if (split_wide_fits(p)) {
      split_fragments_wide(kmaj.data(), 3, p, e, w);
      if (w.empty()) bad++;
    }
  }
  std::printf("host checks: %d failures\n", bad);
  return bad ? 1 : 0;
}
