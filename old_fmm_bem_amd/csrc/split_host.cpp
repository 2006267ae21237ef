// split_host.cpp -- host side of the error-compensated FP16 M2L
// (m2l_split.hip): the operators' FP16 hi/lo split in the kernels' fragment
// orders and the power-of-two scale that keeps them in range.  Plain host
// C++ (no device code), so that it also builds under the sanitizers.
#include <cmath>
#include <vector>

#include "m2l.h"

namespace fmmx {

namespace {

// IEEE binary16 of x, rounded to nearest even (host)
uint16_t to_half(double x) {
  const uint16_t sgn = std::signbit(x) ? 0x8000 : 0;
  double a = std::fabs(x);
  if (a == 0.0) return sgn;
  if (!(a < 65520.0)) return sgn | 0x7c00;
  int e;
  (void)std::frexp(a, &e);  // a = m 2^e, m in [0.5, 1)
  const int ex = e - 1;      // a in [2^ex, 2^(ex+1))
  if (ex < -14) return sgn | (uint16_t)std::nearbyint(std::ldexp(a, 24));  // subnormal (1024: smallest normal)
  double q = std::nearbyint(std::ldexp(a, 10 - ex));
  int be = ex + 15;
  if (q >= 2048.0) {
    q = 1024.0;
    be++;
  }
  if (be >= 31) return sgn | 0x7c00;
  return sgn | (uint16_t)(be << 10) | (uint16_t)(q - 1024.0);
}

double from_half(uint16_t h) {
  const int be = (h >> 10) & 31, mant = h & 1023;
  const double v = be == 0 ? std::ldexp((double)mant, -24) : std::ldexp(1024.0 + mant, be - 25);
  return (h & 0x8000) ? -v : v;
}

}  // namespace

// two operator buffers of the 32x32x16 form fit one CU's LDS
bool split_wide_fits(int p) {
  if (p < 6 || p > 12) return false;
  const int P2 = p * p, MT = (P2 + 31) / 32, KS = (P2 + 1) / 2, SEG = KS + (KS & 1), KH = (3 * SEG + 7) / 8;
  return (size_t)2 * KH * MT * 64 * 16 <= 160 * 1024 - 64;
}

// two operator buffers of the 16x16x32 form fit one CU's LDS
bool split_db_fits(int p) {
  const int P2 = p * p, MT = (P2 + 15) / 16, KS = (P2 + 3) / 4, SEG = KS + (KS & 1), KH = (3 * SEG + 7) / 8;
  return (size_t)2 * KH * MT * 64 * 16 <= 160 * 1024 - 64;
}

int split_scale_exp(const double* ops, size_t n) {
  double mx = 0;
  for (size_t i = 0; i < n; i++) mx = std::max(mx, std::fabs(ops[i]));
  if (!(mx > 0)) return 0;
  int e;
  (void)std::frexp(mx, &e);
  return SPLIT_TOP - e;
}

void split_fragments(const double* kmaj, int count, int p, int sa_e, std::vector<uint4>& out) {
  const int P2 = p * p, MT = (P2 + 15) / 16, KS = (P2 + 3) / 4, SEG = KS + (KS & 1), KH = (3 * SEG + 7) / 8;
  const size_t FRH = (size_t)KH * MT * 64;
  out.assign((size_t)count * FRH, make_uint4(0u, 0u, 0u, 0u));
  for (int o = 0; o < count; o++) {
    const double* A = kmaj + (size_t)o * P2 * P2;  // element (i, k) at A[k * P2 + i]
    for (int ks = 0; ks < KH; ks++)
      for (int mt = 0; mt < MT; mt++)
        for (int l = 0; l < 64; l++) {
          const int i = 16 * mt + (l & 15), g = l >> 4;
          uint16_t h[8] = {};
          for (int t = 0; t < 8; t++) {
            const int s = 8 * ks + t, seg = s / SEG, u = s % SEG, k = g * KS + u;
            if (i >= P2 || seg >= 3 || u >= KS || k >= P2) continue;
            const double a = std::ldexp(A[(size_t)k * P2 + i], sa_e);
            const uint16_t hi = to_half(a);
            h[t] = seg == 1 ? to_half(a - from_half(hi)) : hi;
          }
          out[(size_t)o * FRH + ((size_t)ks * MT + mt) * 64 + l] =
              make_uint4(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16, h[4] | (uint32_t)h[5] << 16,
                         h[6] | (uint32_t)h[7] << 16);
        }
  }
}

void split_fragments_wide(const double* kmaj, int count, int p, int sa_e, std::vector<uint4>& out) {
  const int P2 = p * p, MT = (P2 + 31) / 32, KS = (P2 + 1) / 2, SEG = KS + (KS & 1), KH = (3 * SEG + 7) / 8;
  const size_t FRW = (size_t)KH * MT * 64;
  out.assign((size_t)count * FRW, make_uint4(0u, 0u, 0u, 0u));
  for (int o = 0; o < count; o++) {
    const double* A = kmaj + (size_t)o * P2 * P2;  // element (i, k) at A[k * P2 + i]
    for (int ks = 0; ks < KH; ks++)
      for (int mt = 0; mt < MT; mt++)
        for (int l = 0; l < 64; l++) {
          const int i = 32 * mt + (l & 31), g = l >> 5;
          uint16_t h[8] = {};
          for (int t = 0; t < 8; t++) {
            const int s = 8 * ks + t, seg = s / SEG, u = s % SEG, k = g * KS + u;
            if (i >= P2 || seg >= 3 || u >= KS || k >= P2) continue;
            const double a = std::ldexp(A[(size_t)k * P2 + i], sa_e);
            const uint16_t hi = to_half(a);
            h[t] = seg == 1 ? to_half(a - from_half(hi)) : hi;
          }
          out[(size_t)o * FRW + ((size_t)ks * MT + mt) * 64 + l] =
              make_uint4(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16, h[4] | (uint32_t)h[5] << 16,
                         h[6] | (uint32_t)h[7] << 16);
        }
  }
}

}  // namespace fmmx
