// kernels.hip -- FMM phase kernels for CDNA4 (gfx950), wave64.
//
//   P2P  kernel/G_p2p.cxx, Gni_p2p.cxx, Gnj_p2p.cxx: one workgroup per target
//        leaf box (two lanes per target for small leaves), source particles
//        of the near boxes staged through LDS, FP32 (or FP64) pair math, FP64
//        per-target sums.
//   P2M  kernel/G_p2m.cxx, Gn_p2m.cxx; L2P kernel/G_l2p.cxx, Gn_l2p.cxx: one
//        wave per leaf box (two for small leaves); solid harmonics from a
//        Cartesian recurrence (no acos/atan per term); L2P writes the outputs
//        in the caller's order.
//   M2M/L2L kernel/m2m.cxx, l2l.cxx: dense p^2 x p^2 operators per child
//        octant (level independent in the scaled basis), on MFMA in FP32.
//   M2L  kernel/m2l.cxx: targets of one parity class share the same 189
//        lattice offsets, so a tile of targets runs as 189 GEMMs
//        (operator [p^2 x p^2] x gathered source multipoles [p^2 x tile]);
//        FP32 on MFMA, every level in one launch.
//   direct sums kernel/G_dir.cxx, Gni_dir.cxx, Gnj_dir.cxx.
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "engine.h"
#include "m2l.h"
#include "tree_util.h"

namespace fmmx {

namespace {

// ---------------------------------------------------------------------
// Coefficients of the solid-harmonic recurrences, S_n^m = rho^n fac P_n^m e^{im beta}
// with fac = sqrt((n-m)!/(n+m)!) and the Condon-Shortley phase
// (kernel/multipole.cxx:7-35):
//   S_m^m     = -a_m (x+iy) S_{m-1}^{m-1},      a_m = sqrt((2m-1)/(2m))
//   S_{m+1}^m = b_m z S_m^m,                    b_m = sqrt(2m+1)
//   S_{n+1}^m = c_nm z S_n^m - d_nm r^2 S_{n-1}^m
// and of their derivatives (d/dx + i d/dy) S_n^m = gp_nm S_{n-1}^{m+1},
// (d/dx - i d/dy) S_n^m = -gm_nm S_{n-1}^{m-1}, d/dz S_n^m = gz_nm S_{n-1}^m.
constexpr double csqrt(double x) {
  if (x <= 0) return 0;
  double r = x > 1 ? x : 1;
  for (int i = 0; i < 200; i++) r = 0.5 * (r + x / r);
  return r;
}
struct HarmCoef {
  double a[MAXP + 2], b[MAXP + 2], c[MAXP + 2][MAXP + 2], d[MAXP + 2][MAXP + 2];
  double gp[MAXP + 2][MAXP + 2], gm[MAXP + 2][MAXP + 2], gz[MAXP + 2][MAXP + 2];
};
constexpr HarmCoef make_coef() {
  HarmCoef h{};
  for (int m = 0; m <= MAXP + 1; m++) {
    h.a[m] = m > 0 ? csqrt((2.0 * m - 1) / (2.0 * m)) : 0;
    h.b[m] = csqrt(2.0 * m + 1);
  }
  for (int n = 0; n <= MAXP + 1; n++)
    for (int m = 0; m <= MAXP + 1; m++) {
      double den = (double)(n + 1 - m) * (n + 1 + m);
      h.c[n][m] = den > 0 ? (2.0 * n + 1) / csqrt(den) : 0;
      h.d[n][m] = den > 0 ? csqrt((double)(n - m) * (n + m) / den) : 0;
      h.gp[n][m] = (n - m) * (n - m - 1) > 0 ? csqrt((double)(n - m) * (n - m - 1)) : 0;
      h.gm[n][m] = (n + m) * (n + m - 1) > 0 ? csqrt((double)(n + m) * (n + m - 1)) : 0;
      h.gz[n][m] = (n - m) * (n + m) > 0 ? csqrt((double)(n - m) * (n + m)) : 0;
    }
  return h;
}
__constant__ HarmCoef c_hc = make_coef();

__host__ __device__ constexpr int sl(int n, int m) { return n * (n + 1) / 2 + m; }
// the same coefficients rounded to float once, read by uniform (scalar) loads
struct HarmCoefF {
  float a[MAXP + 2], b[MAXP + 2], c[MAXP + 2][MAXP + 2], d[MAXP + 2][MAXP + 2];
};
constexpr HarmCoefF make_coef_f() {
  const HarmCoef h = make_coef();
  HarmCoefF f{};
  for (int m = 0; m <= MAXP + 1; m++) {
    f.a[m] = (float)h.a[m];
    f.b[m] = (float)h.b[m];
    for (int n = 0; n <= MAXP + 1; n++) {
      f.c[n][m] = (float)h.c[n][m];
      f.d[n][m] = (float)h.d[n][m];
    }
  }
  return f;
}
__constant__ HarmCoefF c_hcf = make_coef_f();
template <typename T>
struct HC {
  static __device__ __forceinline__ T a(int m) { return T(c_hc.a[m]); }
  static __device__ __forceinline__ T b(int m) { return T(c_hc.b[m]); }
  static __device__ __forceinline__ T c(int n, int m) { return T(c_hc.c[n][m]); }
  static __device__ __forceinline__ T d(int n, int m) { return T(c_hc.d[n][m]); }
};
template <>
struct HC<float> {
  static __device__ __forceinline__ float a(int m) { return c_hcf.a[m]; }
  static __device__ __forceinline__ float b(int m) { return c_hcf.b[m]; }
  static __device__ __forceinline__ float c(int n, int m) { return c_hcf.c[n][m]; }
  static __device__ __forceinline__ float d(int n, int m) { return c_hcf.d[n][m]; }
};



template <int P, typename T>
__device__ __forceinline__ void solid_harmonics(T x, T y, T z, T* Sr, T* Si) {
  const T r2 = x * x + y * y + z * z;
  Sr[0] = T(1);
  Si[0] = T(0);
#pragma unroll
  for (int m = 0; m < P; m++) {
    if (m > 0) {
      const T a = -T(c_hc.a[m]);
      const int pm = sl(m - 1, m - 1);
      Sr[sl(m, m)] = a * (x * Sr[pm] - y * Si[pm]);
      Si[sl(m, m)] = a * (x * Si[pm] + y * Sr[pm]);
    }
    if (m + 1 < P) {
      const T b = T(c_hc.b[m]) * z;
      Sr[sl(m + 1, m)] = b * Sr[sl(m, m)];
      Si[sl(m + 1, m)] = b * Si[sl(m, m)];
    }
#pragma unroll
    for (int n = m + 1; n + 1 < P; n++) {
      const T c = T(c_hc.c[n][m]) * z, d = T(c_hc.d[n][m]) * r2;
      Sr[sl(n + 1, m)] = c * Sr[sl(n, m)] - d * Sr[sl(n - 1, m)];
      Si[sl(n + 1, m)] = c * Si[sl(n, m)] - d * Si[sl(n - 1, m)];
    }
  }
}

// Gradient of S_n^m (n >= 1) in terms of degree n-1 values:
// dS/dx = (Ap + Am)/2, dS/dy = -i (Ap - Am)/2, dS/dz = gz S_{n-1}^m
template <int P, typename T>
__device__ __forceinline__ void harm_grad(const T* Sr, const T* Si, int n, int m, T& pr, T& pi, T& mr, T& mi,
                                          T& zr, T& zi) {
  pr = pi = mr = mi = zr = zi = T(0);
  if (m + 1 <= n - 1) {
    const T g = T(c_hc.gp[n][m]);
    pr = g * Sr[sl(n - 1, m + 1)];
    pi = g * Si[sl(n - 1, m + 1)];
  }
  if (m >= 1) {
    const T g = -T(c_hc.gm[n][m]);
    mr = g * Sr[sl(n - 1, m - 1)];
    mi = g * Si[sl(n - 1, m - 1)];
  } else if (n >= 2) {
    const T g = T(c_hc.gp[n][0]);  // m = 0: conj(S_{n-1}^1) term
    mr = g * Sr[sl(n - 1, 1)];
    mi = -g * Si[sl(n - 1, 1)];
  }
  if (m <= n - 1) {
    const T g = T(c_hc.gz[n][m]);
    zr = g * Sr[sl(n - 1, m)];
    zi = g * Si[sl(n - 1, m)];
  }
}

// Evaluation point of an expansion in units of the box size.  REF: the
// reference's angles (kernel/G_p2m.cxx:31-41) with eps added to the radius
// and float pi in the azimuth branches; the harmonics at these angles equal
// the solid harmonics at the point below.
__device__ __forceinline__ void eff_point(int conv, double dx, double dy, double dz, double rbinv, double& ux,
                                          double& uy, double& uz) {
  if (conv == FMMX_CONV_EXACT) {
    ux = dx * rbinv;
    uy = dy * rbinv;
    uz = dz * rbinv;
    return;
  }
  const double eps = (double)EPSF, pif = (double)PIF;
  double rh = sqrt(dx * dx + dy * dy + dz * dz) + eps;
  double ca = cos(acos(dz / rh));
  double sa = sqrt((1 - ca) * (1 + ca));
  double be;
  if (fabs(dx) + fabs(dy) < eps) be = 0;
  else if (fabs(dx) < eps) be = dy / fabs(dy) * pif * 0.5;
  else if (dx > 0) be = atan(dy / dx);
  else be = atan(dy / dx) + pif;
  double sb, cb;
  sincos(be, &sb, &cb);
  ux = rh * sa * cb * rbinv;
  uy = rh * sa * sb * rbinv;
  uz = rh * ca * rbinv;
}

// i~_n(x) = i_n(x) / x^n for n = 0..N (modified spherical Bessel function
// of the first kind over x^n): power series for N+1 and N, then the stable
// backward recurrence i~_{n-1} = (2n+1) i~_n + x^2 i~_{n+1}
template <int N>
__device__ inline void ibar_dev(double x, double* out) {
  const double h = 0.5 * x * x;
  double top[2];
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int n = N + 1 - u;
    double df = 1;
    for (int k = 3; k <= 2 * n + 1; k += 2) df *= k;
    double t = 1.0 / df, sum = t;
    for (int k = 1; k < 400; k++) {
      t *= h / (k * (2.0 * n + 2.0 * k + 1.0));
      sum += t;
      if (t < 1e-18 * sum) break;
    }
    top[u] = sum;
  }
  double a = top[0], b = top[1];
  out[N] = b;
#pragma unroll
  for (int n = N; n >= 1; n--) {
    const double c = (2 * n + 1) * b + x * x * a;
    a = b;
    b = c;
    out[n - 1] = c;
  }
}

// j^_n(x) = (2n+1)!! j_n(x) / x^n for n = 0..N (spherical Bessel function of
// the first kind, scaled to 1 at x = 0): power series for N+1 and N, then
// the stable backward recurrence j^_{n-1} = j^_n - x^2 j^_{n+1}/((2n+1)(2n+3))
template <int N>
__device__ inline void jhat_dev(double x, double* out) {
  const double h = -0.5 * x * x;
  double top[2];
#pragma unroll
  for (int u = 0; u < 2; u++) {
    const int n = N + 1 - u;
    double t = 1, sum = 1, tmax = 1;
    for (int k = 1; k < 400; k++) {
      t *= h / (k * (2.0 * n + 2.0 * k + 1.0));
      sum += t;
      tmax = fmax(tmax, fabs(t));
      if (fabs(t) < 1e-18 * tmax) break;
    }
    top[u] = sum;
  }
  double a = top[0], b = top[1];
  out[N] = b;
#pragma unroll
  for (int n = N; n >= 1; n--) {
    const double c = b - x * x * a / ((2.0 * n + 1.0) * (2.0 * n + 3.0));
    a = b;
    b = c;
    out[n - 1] = c;
  }
}

__device__ __forceinline__ void box_center(uint32_t key, int lev, double xmin, double ymin, double zmin, double rb,
                                           double& cx, double& cy, double& cz) {
  int ix, iy, iz;
  morton_decode(key, lev, ix, iy, iz);
  cx = xmin + (ix + 0.5) * rb;
  cy = ymin + (iy + 0.5) * rb;
  cz = zmin + (iz + 0.5) * rb;
}

// ---------------------------------------------------------------------
// P2M: one wave per source leaf box
//   YK 1: Yukawa multipoles, M~_n^m = sum q i~_n(lambda rho)/i~_n(lambda rb)
//       (rho/rb)^n conj(Y_n^m): the Laplace harmonics times a radial weight
//       per degree (lrb = lambda rb)
//   YK 2: Helmholtz multipoles of one part of the charges (real or
//       imaginary), M_n^m = sum q j^_n(k rho) (rho/rb)^n conj(Y_n^m)
//       (lrb = k rb)
//   Box b's multipole starts at M + b * mstride + moff.
template <int P, typename T, int YK = 0, int HALF = 0, int CONV = -1, int GN = -1>
__global__ __launch_bounds__(64) void k_p2m(int nb, const uint32_t* __restrict__ bkey, const int* __restrict__ bfirst,
                                            const int* __restrict__ blast, const float* __restrict__ x,
                                            const float* __restrict__ y, const float* __restrict__ z,
                                            const float* __restrict__ nx, const float* __restrict__ ny,
                                            const float* __restrict__ nz, const float* __restrict__ q, int gn,
                                            int conv, int lev, double xmin, double ymin, double zmin, double rb,
                                            const double* __restrict__ corr, T* __restrict__ M, double lrb,
                                            int mstride, int moff) {
  // HALF: two leaves per wave, rows 0-31 for the first and 32-63 for the
  // second, for leaves of at most about 32 particles
  constexpr int NR = P * (P + 1) / 2, P2 = P * P, NL = HALF ? 2 : 1, STEP = 64 / NL;
  constexpr int NC = (NL * P2 + 63) / 64;
  __shared__ T red[64][P2 + 1];
  const int lane = threadIdx.x, h = HALF ? lane >> 5 : 0, sub = HALF ? lane & 31 : lane;
  const int b = HALF ? 2 * (int)blockIdx.x + h : (int)blockIdx.x;
  const bool valid = b < nb;
  if (!HALF && !valid) return;
  double cx = 0, cy = 0, cz = 0;
  if (valid) box_center(bkey[b], lev, xmin, ymin, zmin, rb, cx, cy, cz);
  const double rbinv = 1.0 / rb;
  T acc[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) acc[c] = T(0);
  const int i0 = valid ? bfirst[b] : 0, i1 = valid ? blast[b] : -1;
  for (int base = 0;; base += STEP) {
    if (__ballot(i0 + base <= i1) == 0) break;  // one wave per block: uniform
    const int i = i0 + base + sub;
    T* row = red[lane];
    if (YK == 0 && CONV == FMMX_CONV_EXACT && GN == 0 && i <= i1) {
      // charges in the EXACT convention (unit table ratios): the harmonics
      // column by column straight into the LDS row; the unrolled recurrence
      // uses literal coefficients and keeps only its state live
      double ux, uy, uz;
      eff_point(FMMX_CONV_EXACT, (double)x[i] - cx, (double)y[i] - cy, (double)z[i] - cz, rbinv, ux, uy, uz);
      const T tx = T(ux), ty = T(uy), tz = T(uz), r2 = tx * tx + ty * ty + tz * tz, qi = T(q[i]);
      constexpr HarmCoef hc = make_coef();
      auto put = [&](int n, int m, T sr, T si) {
        row[sl(n, m)] = qi * sr;
        if (m > 0) row[NR + n * (n - 1) / 2 + m - 1] = -qi * si;
      };
      T dr = T(1), di = T(0);
#pragma unroll
      for (int m = 0; m < P; m++) {
        if (m > 0) {
          const T a = -T(hc.a[m]), nr = a * (tx * dr - ty * di), ni = a * (tx * di + ty * dr);
          dr = nr;
          di = ni;
        }
        put(m, m, dr, di);
        if (m + 1 < P) {
          const T bz = T(hc.b[m]) * tz;
          T pr = dr, pi = di, sr = bz * dr, si = bz * di;
          put(m + 1, m, sr, si);
#pragma unroll
          for (int n = m + 2; n < P; n++) {
            const T cz = T(hc.c[n - 1][m]) * tz, d = T(hc.d[n - 1][m]) * r2;
            const T nr = cz * sr - d * pr, ni = cz * si - d * pi;
            pr = sr;
            pi = si;
            sr = nr;
            si = ni;
            put(n, m, sr, si);
          }
        }
      }
    } else if (i <= i1) {
      const int cv = CONV >= 0 ? CONV : conv;
      const int gnv = GN >= 0 ? GN : gn;
      double ux, uy, uz;
      eff_point(cv, (double)x[i] - cx, (double)y[i] - cy, (double)z[i] - cz, rbinv, ux, uy, uz);
      T Sr[NR], Si[NR];
      solid_harmonics<P, T>(T(ux), T(uy), T(uz), Sr, Si);
      if (YK) {
        double ib0[P], ibx[P];
        const double lr = lrb * sqrt(ux * ux + uy * uy + uz * uz);
        if (YK == 2) {
          jhat_dev<P - 1>(lr, ibx);
        } else {
          ibar_dev<P - 1>(lrb, ib0);
          ibar_dev<P - 1>(lr, ibx);
        }
        const double qd = (double)q[i];
#pragma unroll
        for (int n = 0; n < P; n++) {
          const T w = T(YK == 2 ? qd * ibx[n] : qd * ibx[n] / ib0[n]);
#pragma unroll
          for (int m = 0; m <= n; m++) {
            row[sl(n, m)] = w * Sr[sl(n, m)];
            if (m > 0) row[NR + n * (n - 1) / 2 + m - 1] = -w * Si[sl(n, m)];
          }
        }
      } else if (!gnv) {
        const T qi = T(q[i]);
#pragma unroll
        for (int n = 0; n < P; n++)
#pragma unroll
          for (int m = 0; m <= n; m++) {
            const T w = qi * T(corr[sl(n, m)]);
            row[sl(n, m)] = w * Sr[sl(n, m)];
            if (m > 0) row[NR + n * (n - 1) / 2 + m - 1] = -w * Si[sl(n, m)];
          }
      } else {
        // dipole q*n (the float product vj*gxj of kernel/Gn_p2m.cxx:113)
        const T qx = T(q[i] * nx[i]), qy = T(q[i] * ny[i]), qz = T(q[i] * nz[i]);
        row[0] = T(0);
#pragma unroll
        for (int n = 1; n < P; n++)
#pragma unroll
          for (int m = 0; m <= n; m++) {
            T pr, pi, mr, mi, zr, zi;
            harm_grad<P, T>(Sr, Si, n, m, pr, pi, mr, mi, zr, zi);
            // n . grad S = qx (Ap+Am)/2 + qy (-i)(Ap-Am)/2 + qz dS/dz
            const T dr = T(0.5) * qx * (pr + mr) + T(0.5) * qy * (pi - mi) + qz * zr;
            const T di = T(0.5) * qx * (pi + mi) - T(0.5) * qy * (pr - mr) + qz * zi;
            const T w = T(corr[sl(n, m)] * rbinv);
            row[sl(n, m)] = w * dr;
            if (m > 0) row[NR + n * (n - 1) / 2 + m - 1] = -w * di;
          }
      }
    } else {
#pragma unroll
      for (int c = 0; c < P2; c++) row[c] = T(0);
    }
    __syncthreads();
    // column sums: entry idx = (leaf of the wave, column) over that leaf's rows
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const int idx = lane + 64 * c, hh = idx / P2, col = idx % P2;
      if (hh < NL) {
        T s = T(0);
        for (int r = 0; r < STEP; r++) s += red[hh * STEP + r][col];
        acc[c] += s;
      }
    }
    __syncthreads();
  }
  using ML = MLayout<P>;
  const int b0 = HALF ? 2 * (int)blockIdx.x : (int)blockIdx.x;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int idx = lane + 64 * c, hh = idx / P2, col = idx % P2;
    if (hh < NL && b0 + hh < nb) M[(size_t)(b0 + hh) * mstride + moff + ML::at(col)] = acc[c];
  }
  // zero the padding slots (M2M applies its operator to whole stored vectors)
  for (int c2 = lane; c2 < NL * ML::VS; c2 += 64) {
    const int hh = c2 / ML::VS, c = c2 % ML::VS;
    if (b0 + hh < nb && (c % ML::KSP >= ML::KS || (c / ML::KSP) * ML::KS + c % ML::KSP >= P2))
      M[(size_t)(b0 + hh) * mstride + moff + c] = T(0);
  }
}

// P2M above order 12 (FP64, EXACT convention): a 64 x P^2 table of particle
// terms (166 KB at P = 18) no longer fits LDS, so each lane keeps its
// particle's harmonics in registers (spilled as the compiler needs) and the
// table goes through LDS 64 columns at a time, each column summed by one lane
// in fixed row order (deterministic).  YK 1: Yukawa weights as k_p2m;
// GN 1: the dipole form of k_p2m (Gnj).
template <int P, int YK, int GN>
__global__ __launch_bounds__(64) void k_p2m_hi(int nb, const uint32_t* __restrict__ bkey,
                                               const int* __restrict__ bfirst, const int* __restrict__ blast,
                                               const float* __restrict__ x, const float* __restrict__ y,
                                               const float* __restrict__ z, const float* __restrict__ nx,
                                               const float* __restrict__ ny, const float* __restrict__ nz,
                                               const float* __restrict__ q, int lev, double xmin, double ymin,
                                               double zmin, double rb, double* __restrict__ M, double lrb) {
  constexpr int NR = P * (P + 1) / 2, P2 = P * P, CH = 64, NC = (P2 + CH - 1) / CH;
  __shared__ double red[64][CH + 1];
  const int lane = threadIdx.x, b = blockIdx.x;
  if (b >= nb) return;  // one wave per block: uniform
  double cx, cy, cz;
  box_center(bkey[b], lev, xmin, ymin, zmin, rb, cx, cy, cz);
  const double rbinv = 1.0 / rb;
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) acc[c] = 0;
  double ib0[P];
  if (YK) ibar_dev<P - 1>(lrb, ib0);
  const int i0 = bfirst[b], i1 = blast[b];
  for (int base = i0; base <= i1; base += 64) {
    const int i = base + lane;
    const bool on = i <= i1;
    double Sr[NR], Si[NR], w[P];
    const double ux = on ? ((double)x[i] - cx) * rbinv : 0, uy = on ? ((double)y[i] - cy) * rbinv : 0,
                 uz = on ? ((double)z[i] - cz) * rbinv : 0;
    solid_harmonics<P, double>(ux, uy, uz, Sr, Si);
    const double qi = on ? (double)q[i] : 0;
    if (YK) {
      double ibx[P];
      ibar_dev<P - 1>(lrb * sqrt(ux * ux + uy * uy + uz * uz), ibx);
#pragma unroll
      for (int n = 0; n < P; n++) w[n] = qi * ibx[n] / ib0[n];
    } else {
#pragma unroll
      for (int n = 0; n < P; n++) w[n] = GN ? rbinv : qi;
    }
    // dipole q*n: the float product vj*gxj of kernel/Gn_p2m.cxx:113, as k_p2m
    const double qx = GN && on ? (double)(q[i] * nx[i]) : 0, qy = GN && on ? (double)(q[i] * ny[i]) : 0,
                 qz = GN && on ? (double)(q[i] * nz[i]) : 0;
    // column col of this particle's row: Re slots sl(n, m), then Im slots
    auto term = [&](int n, int m, bool im) -> double {
      if (GN) {
        if (n == 0) return 0;
        double pr, pi, mr, mi, zr, zi;
        harm_grad<P, double>(Sr, Si, n, m, pr, pi, mr, mi, zr, zi);
        if (!im) return w[n] * (0.5 * qx * (pr + mr) + 0.5 * qy * (pi - mi) + qz * zr);
        return -w[n] * (0.5 * qx * (pi + mi) - 0.5 * qy * (pr - mr) + qz * zi);
      }
      return im ? -w[n] * Si[sl(n, m)] : w[n] * Sr[sl(n, m)];
    };
#pragma unroll
    for (int c = 0; c < NC; c++) {
#pragma unroll
      for (int n = 0; n < P; n++)
#pragma unroll
        for (int m = 0; m <= n; m++) {
          const int cr = sl(n, m), cim = NR + n * (n - 1) / 2 + m - 1;
          if (cr >= c * CH && cr < (c + 1) * CH) red[lane][cr - c * CH] = term(n, m, false);
          if (m > 0 && cim >= c * CH && cim < (c + 1) * CH) red[lane][cim - c * CH] = term(n, m, true);
        }
      __syncthreads();
      if (c * CH + lane < P2) {
        double s = 0;
        for (int r = 0; r < 64; r++) s += red[r][lane];
        acc[c] += s;
      }
      __syncthreads();
    }
  }
  using ML = MLayout<P>;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int col = c * CH + lane;
    if (col < P2) M[(size_t)b * ML::VS + ML::at(col)] = acc[c];
  }
  for (int c = lane; c < ML::VS; c += 64)
    if (c % ML::KSP >= ML::KS || (c / ML::KSP) * ML::KS + c % ML::KSP >= P2) M[(size_t)b * ML::VS + c] = 0;
}

// P2M of charges in the EXACT convention, FP32 (the mixed engine's G, Gni):
// each lane's particle terms stay in registers and the per-leaf column sums
// go through LDS in chunks of CH columns, so a workgroup holds 8-17 KB of
// LDS instead of a whole 64 x P^2 table (four to eight times the
// workgroups per CU of k_p2m).  Same leaves-per-wave forms as k_p2m.
template <int P, int HALF>
__global__ __launch_bounds__(64) void k_p2m_exact(int nb, const uint32_t* __restrict__ bkey,
                                                  const int* __restrict__ bfirst, const int* __restrict__ blast,
                                                  const float* __restrict__ x, const float* __restrict__ y,
                                                  const float* __restrict__ z, const float* __restrict__ q, int lev,
                                                  double xmin, double ymin, double zmin, double rb,
                                                  float* __restrict__ M, int mstride, int moff) {
  constexpr int NR = P * (P + 1) / 2, P2 = P * P, NL = HALF ? 2 : 1, STEP = 64 / NL;
  constexpr int CH = 64 / NL, NQ = (P2 + CH - 1) / CH;
  __shared__ float red[64][CH + 1];
  const int lane = threadIdx.x, h = HALF ? lane >> 5 : 0, sub = HALF ? lane & 31 : lane;
  const int b = HALF ? 2 * (int)blockIdx.x + h : (int)blockIdx.x;
  const bool valid = b < nb;
  if (!HALF && !valid) return;
  double cx = 0, cy = 0, cz = 0;
  if (valid) box_center(bkey[b], lev, xmin, ymin, zmin, rb, cx, cy, cz);
  const double rbinv = 1.0 / rb;
  float acc[NQ];
#pragma unroll
  for (int k = 0; k < NQ; k++) acc[k] = 0.f;
  // column sums: lane -> (leaf cl, column cc of the chunk)
  const int cl = lane / CH, cc = lane % CH;
  const int i0 = valid ? bfirst[b] : 0, i1 = valid ? blast[b] : -1;
  constexpr HarmCoef hc = make_coef();
  for (int base = 0;; base += STEP) {
    if (__ballot(i0 + base <= i1) == 0) break;  // one wave per block: uniform
    const int i = i0 + base + sub;
    float v[P2];
#pragma unroll
    for (int c = 0; c < P2; c++) v[c] = 0.f;
    if (i <= i1) {
      double ux, uy, uz;
      eff_point(FMMX_CONV_EXACT, (double)x[i] - cx, (double)y[i] - cy, (double)z[i] - cz, rbinv, ux, uy, uz);
      const float tx = (float)ux, ty = (float)uy, tz = (float)uz, r2 = tx * tx + ty * ty + tz * tz, qi = q[i];
      auto put = [&](int n, int m, float sr, float si) {
        v[sl(n, m)] = qi * sr;
        if (m > 0) v[NR + n * (n - 1) / 2 + m - 1] = -qi * si;
      };
      float dr = 1.f, di = 0.f;
#pragma unroll
      for (int m = 0; m < P; m++) {
        if (m > 0) {
          const float a = -(float)hc.a[m], nr = a * (tx * dr - ty * di), ni = a * (tx * di + ty * dr);
          dr = nr;
          di = ni;
        }
        put(m, m, dr, di);
        if (m + 1 < P) {
          const float bz = (float)hc.b[m] * tz;
          float pr = dr, pi = di, sr = bz * dr, si = bz * di;
          put(m + 1, m, sr, si);
#pragma unroll
          for (int n = m + 2; n < P; n++) {
            const float cz = (float)hc.c[n - 1][m] * tz, d = (float)hc.d[n - 1][m] * r2;
            const float nr = cz * sr - d * pr, ni = cz * si - d * pi;
            pr = sr;
            pi = si;
            sr = nr;
            si = ni;
            put(n, m, sr, si);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NQ; k++) {
#pragma unroll
      for (int c = 0; c < CH; c++) red[lane][c] = k * CH + c < P2 ? v[k * CH + c] : 0.f;
      __syncthreads();
      float s0 = 0.f, s1 = 0.f;
#pragma unroll 8
      for (int r = 0; r < STEP; r += 2) {
        s0 += red[cl * STEP + r][cc];
        s1 += red[cl * STEP + r + 1][cc];
      }
      acc[k] += s0 + s1;
      __syncthreads();
    }
  }
  using ML = MLayout<P>;
  const int b0 = HALF ? 2 * (int)blockIdx.x : (int)blockIdx.x;
  if (b0 + cl < nb) {
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const int col = k * CH + cc;
      if (col < P2) M[(size_t)(b0 + cl) * mstride + moff + ML::at(col)] = acc[k];
    }
  }
  // zero the padding slots (M2M applies its operator to whole stored vectors)
  for (int c2 = lane; c2 < NL * ML::VS; c2 += 64) {
    const int hh = c2 / ML::VS, c = c2 % ML::VS;
    if (b0 + hh < nb && (c % ML::KSP >= ML::KS || (c / ML::KSP) * ML::KS + c % ML::KSP >= P2))
      M[(size_t)(b0 + hh) * mstride + moff + c] = 0.f;
  }
}

// The far-field result of target i (sorted position) added to its near-field
// sums: stored back in place, or, with out.perm set, written straight to the
// caller's order (out.v[perm[i]], ...; the gradient outputs may be NULL),
// which saves the separate scatter pass over the outputs.
struct L2POut {
  const uint32_t* perm;
  double *v, *gx, *gy, *gz;
};
__device__ __forceinline__ void l2p_store(const L2POut& out, int i, int gn, int grad, double s1, double s2,
                                          double phi, double gx, double gy, double gz,
                                          const float* __restrict__ nx, const float* __restrict__ ny,
                                          const float* __restrict__ nz, double* __restrict__ vs,
                                          double* __restrict__ gxs, double* __restrict__ gys,
                                          double* __restrict__ gzs) {
  if (gn) {
    const double v = vs[i] - s2 * ((double)nx[i] * gx + (double)ny[i] * gy + (double)nz[i] * gz);
    if (!out.perm) vs[i] = v;
    else if (out.v) out.v[out.perm[i]] = v;
    return;
  }
  const double v = vs[i] + s1 * phi;
  if (!out.perm) {
    vs[i] = v;
    if (grad) {
      gxs[i] += s2 * gx;
      gys[i] += s2 * gy;
      gzs[i] += s2 * gz;
    }
    return;
  }
  const uint32_t j = out.perm[i];
  if (out.v) out.v[j] = v;
  if (grad) {
    if (out.gx) out.gx[j] = gxs[i] + s2 * gx;
    if (out.gy) out.gy[j] = gys[i] + s2 * gy;
    if (out.gz) out.gz[j] = gzs[i] + s2 * gz;
  }
}

// ---------------------------------------------------------------------
// L2P: one wave per target leaf box; adds to the P2P results (sorted order)
//   YK 1: Yukawa locals, phi = sum_n h_n(r) sum_m L~ S_n^m(u) with
//       h_n = i~_n(lambda r)/i~_n(lambda rb); grad adds
//       lrb^2 i~_{n+1}(lambda r)/i~_n(lambda rb) (sum_m L~ S_n^m) u
//   YK 2: one part (Re or Im of the field) of Helmholtz locals,
//       phi = c4/rb sum_n j^_n(k r) sum_m L S_n^m(u), potential only
//   Box b's local starts at L + b * lstride + loff.
template <int P, typename T, int YK = 0, int HALF = 0, int CONV = -1>
__global__ __launch_bounds__(64) void k_l2p(int nb, const uint32_t* __restrict__ bkey, const int* __restrict__ bfirst,
                                            const int* __restrict__ blast, const float* __restrict__ x,
                                            const float* __restrict__ y, const float* __restrict__ z,
                                            const float* __restrict__ nx, const float* __restrict__ ny,
                                            const float* __restrict__ nz, const T* __restrict__ L, int gn, int grad,
                                            int conv, int lev, double xmin, double ymin, double zmin, double rb,
                                            const double* __restrict__ corr, double c4, double* __restrict__ vs,
                                            double* __restrict__ gxs, double* __restrict__ gys,
                                            double* __restrict__ gzs, double lrb, int lstride, int loff,
                                            L2POut out) {
  // HALF: two leaves per wave (32 lanes each), for leaves of at most about
  // 32 targets where a whole wave per leaf would idle half its lanes
  constexpr int NR = P * (P + 1) / 2, P2 = P * P, NL = HALF ? 2 : 1, STEP = HALF ? 32 : 64;
  __shared__ T Lrs[NL][NR], Lis[NL][NR];
  const int h = HALF ? (int)(threadIdx.x >> 5) : 0, lane = HALF ? (int)(threadIdx.x & 31) : (int)threadIdx.x;
  const int b = HALF ? 2 * (int)blockIdx.x + h : (int)blockIdx.x;
  const bool valid = b < nb;
  T* Lr = Lrs[h];
  T* Li = Lis[h];
  // weights: 1 for m = 0, 2 for m > 0 (kernel/G_l2p.cxx:246-260); table ratio in REF
  for (int s = lane; valid && s < NR; s += STEP) {
    int n = 0;
    while (sl(n + 1, 0) <= s) n++;
    const int m = s - sl(n, 0);
    const double w = (m == 0 ? 1.0 : 2.0) * (YK ? 1.0 : corr[s]);
    Lr[s] = T(w * (double)L[(size_t)b * lstride + loff + s]);
    Li[s] = m == 0 ? T(0) : T(w * (double)L[(size_t)b * lstride + loff + NR + n * (n - 1) / 2 + m - 1]);
  }
  __syncthreads();
  if constexpr (YK == 0) {
    // The gradient of sum L_n^m S_n^m is a sum over S_{n-1}^{m'} with
    // coefficients linear in L (the derivative rules of the harmonics above),
    // so per leaf the table C[s] = (L, X/2, Y/2, Z) is built once and each
    // target streams its harmonics column by column against it, keeping only
    // the recurrence state live:
    //   gx = Re sum S X/2, gy = Im sum S Y/2, gz = Re sum S Z,
    //   X = gp(n,m-1) L_n^{m-1} - gm(n,m+1) L_n^{m+1} + [m=1] gp(n,0) conj(L_n^0),
    //   Y = gp(n,m-1) L_n^{m-1} + gm(n,m+1) L_n^{m+1} + [m=1] gp(n,0) conj(L_n^0),
    //   Z = gz(n,m) L_n^m, with n one above the degree of S.
    __shared__ __align__(16) T Cs[NL][NR][8];
    for (int s = lane; valid && s < NR; s += STEP) {
      int n = 0;
      while (sl(n + 1, 0) <= s) n++;
      const int m = s - sl(n, 0);
      T* c = Cs[h][s];
      c[0] = Lr[s];
      c[1] = Li[s];
      T xr = 0, xi = 0, yr = 0, yi = 0, zr = 0, zi = 0;
      if (n <= P - 2) {
        const int nu = n + 1;
        if (m >= 1) {
          const T g = T(c_hc.gp[nu][m - 1]), ar = g * Lr[sl(nu, m - 1)], ai = g * Li[sl(nu, m - 1)];
          xr += ar;
          xi += ai;
          yr += ar;
          yi += ai;
        }
        {
          const T g = T(c_hc.gm[nu][m + 1]), ar = g * Lr[sl(nu, m + 1)], ai = g * Li[sl(nu, m + 1)];
          xr -= ar;
          xi -= ai;
          yr += ar;
          yi += ai;
        }
        if (m == 1) {
          const T g = T(c_hc.gp[nu][0]), ar = g * Lr[sl(nu, 0)], ai = -g * Li[sl(nu, 0)];
          xr += ar;
          xi += ai;
          yr += ar;
          yi += ai;
        }
        const T g = T(c_hc.gz[nu][m]);
        zr = g * Lr[sl(nu, m)];
        zi = g * Li[sl(nu, m)];
      }
      c[2] = T(0.5) * xr;
      c[3] = T(0.5) * xi;
      c[4] = T(0.5) * yi;
      c[5] = T(0.5) * yr;
      c[6] = zr;
      c[7] = zi;
    }
    __syncthreads();
    const double rbinv = 1.0 / rb;
    // a compile-time convention keeps the REF angle code (double acos,
    // atan, sincos) out of the EXACT kernel's register budget
    const int cv = CONV >= 0 ? CONV : conv;
    // target i of the leaf whose table is Cs[hs], centre (cx, cy, cz)
    auto target = [&](int i, int hs, double cx, double cy, double cz) {
      double ux, uy, uz;
      eff_point(cv, (double)x[i] - cx, (double)y[i] - cy, (double)z[i] - cz, rbinv, ux, uy, uz);
      const T tx = T(ux), ty = T(uy), tz = T(uz), r2 = tx * tx + ty * ty + tz * tz;
      T phi = T(0), gx = T(0), gy = T(0), gz = T(0);
      // the table holds zero gradient entries at degree p-1
      auto term = [&](T sr, T si, int s) {
        const T* c = Cs[hs][s];
        phi += sr * c[0] - si * c[1];
        gx += sr * c[2] - si * c[3];
        gy += sr * c[4] + si * c[5];
        gz += sr * c[6] - si * c[7];
      };
      // harmonics column by column (runtime loops: only the recurrence state
      // is live, and the table reads cannot be hoisted out of the target loop)
      T dr = T(1), di = T(0);  // S_m^m
#pragma unroll 1
      for (int m = 0; m < P; m++) {
        if (m > 0) {
          const T a = -HC<T>::a(m), nr = a * (tx * dr - ty * di), ni = a * (tx * di + ty * dr);
          dr = nr;
          di = ni;
        }
        term(dr, di, sl(m, m));
        if (m + 1 < P) {
          const T bz = HC<T>::b(m) * tz;
          T pr = dr, pi = di, sr = bz * dr, si = bz * di;  // S_{n-1}^m, S_n^m
          term(sr, si, sl(m + 1, m));
#pragma unroll 1
          for (int n = m + 2; n < P; n++) {
            const T cz = HC<T>::c(n - 1, m) * tz, d = HC<T>::d(n - 1, m) * r2;
            const T nr = cz * sr - d * pr, ni = cz * si - d * pi;
            pr = sr;
            pi = si;
            sr = nr;
            si = ni;
            term(sr, si, sl(n, m));
          }
        }
      }
      l2p_store(out, i, gn, grad, c4 * rbinv, c4 * rbinv * rbinv, (double)phi, (double)gx, (double)gy, (double)gz, nx, ny,
                nz, vs, gxs, gys, gzs);
    };
    if constexpr (HALF != 0) {
      // the targets of the wave's two leaves, one list over all 64 lanes: a
      // leaf of 33-40 targets beside a smaller one still takes one pass
      const int bA = 2 * (int)blockIdx.x, bB = bA + 1;  // bA < nb: the grid is (nb + 1) / 2
      const int a0 = bfirst[bA], nA = max(0, blast[bA] - a0 + 1);
      const bool vB = bB < nb;
      const int b0 = vB ? bfirst[bB] : 0, nB = vB ? max(0, blast[bB] - b0 + 1) : 0;
      double axc, ayc, azc, bxc = 0, byc = 0, bzc = 0;
      box_center(bkey[bA], lev, xmin, ymin, zmin, rb, axc, ayc, azc);
      if (vB) box_center(bkey[bB], lev, xmin, ymin, zmin, rb, bxc, byc, bzc);
      for (int t = (int)threadIdx.x; t < nA + nB; t += 64) {
        const bool hb = t >= nA;
        target(hb ? b0 + (t - nA) : a0 + t, hb ? 1 : 0, hb ? bxc : axc, hb ? byc : ayc, hb ? bzc : azc);
      }
    } else {
      if (!valid) return;
      double cx, cy, cz;
      box_center(bkey[b], lev, xmin, ymin, zmin, rb, cx, cy, cz);
      const int i0 = bfirst[b], i1 = blast[b];
      for (int i = i0 + lane; i <= i1; i += STEP) target(i, h, cx, cy, cz);
    }
    return;
  }
  if (!valid) return;
  double cx, cy, cz;
  box_center(bkey[b], lev, xmin, ymin, zmin, rb, cx, cy, cz);
  const double rbinv = 1.0 / rb;
  const int i0 = bfirst[b], i1 = blast[b];
  for (int i = i0 + lane; i <= i1; i += STEP) {
    double ux, uy, uz;
    eff_point(conv, (double)x[i] - cx, (double)y[i] - cy, (double)z[i] - cz, rbinv, ux, uy, uz);
    T Sr[NR], Si[NR];
    solid_harmonics<P, T>(T(ux), T(uy), T(uz), Sr, Si);
    T phi = T(0), gx = T(0), gy = T(0), gz = T(0);
    if constexpr (YK == 2) {
      double jn[P];
      jhat_dev<P - 1>(lrb * sqrt(ux * ux + uy * uy + uz * uz), jn);
#pragma unroll
      for (int n = 0; n < P; n++) {
        T pn = T(0);
#pragma unroll
        for (int m = 0; m <= n; m++) pn += Sr[sl(n, m)] * Lr[sl(n, m)] - Si[sl(n, m)] * Li[sl(n, m)];
        phi += T(jn[n]) * pn;
      }
    } else if constexpr (YK) {
      double hn[P + 1], h0[P];
      ibar_dev<P - 1>(lrb, h0);
      ibar_dev<P>(lrb * sqrt(ux * ux + uy * uy + uz * uz), hn);
#pragma unroll
      for (int n = 0; n < P; n++) {
        T pn = T(0), gxn = T(0), gyn = T(0), gzn = T(0);
#pragma unroll
        for (int m = 0; m <= n; m++) {
          const int s = sl(n, m);
          const T lr = Lr[s], li = Li[s];
          pn += Sr[s] * lr - Si[s] * li;
          if (n >= 1) {
            T pr, pi, mr, mi, zr, zi;
            harm_grad<P, T>(Sr, Si, n, m, pr, pi, mr, mi, zr, zi);
            const T apr = pr * lr - pi * li, api = pr * li + pi * lr;
            const T amr = mr * lr - mi * li, ami = mr * li + mi * lr;
            gxn += T(0.5) * (apr + amr);
            gyn += T(0.5) * (api - ami);
            gzn += zr * lr - zi * li;
          }
        }
        const T h = T(hn[n] / h0[n]), e = T(lrb * lrb * hn[n + 1] / h0[n]) * pn;
        phi += h * pn;
        gx += h * gxn + e * T(ux);
        gy += h * gyn + e * T(uy);
        gz += h * gzn + e * T(uz);
      }
    } else {
#pragma unroll
      for (int n = 0; n < P; n++)
#pragma unroll
        for (int m = 0; m <= n; m++) {
          const int s = sl(n, m);
          const T lr = Lr[s], li = Li[s];
          phi += Sr[s] * lr - Si[s] * li;
          if (n >= 1) {
            T pr, pi, mr, mi, zr, zi;
            harm_grad<P, T>(Sr, Si, n, m, pr, pi, mr, mi, zr, zi);
            // Re(Ap L), Im(Ap L), Re(Am L), Im(Am L)
            const T apr = pr * lr - pi * li, api = pr * li + pi * lr;
            const T amr = mr * lr - mi * li, ami = mr * li + mi * lr;
            gx += T(0.5) * (apr + amr);
            gy += T(0.5) * (api - ami);
            gz += zr * lr - zi * li;
          }
        }
    }
    l2p_store(out, i, gn, grad, c4 * rbinv, c4 * rbinv * rbinv, (double)phi, (double)gx, (double)gy, (double)gz, nx, ny,
              nz, vs, gxs, gys, gzs);
  }
}

// ---------------------------------------------------------------------
// M2M and L2L as small batched GEMMs: a workgroup (128 threads) owns 16
// output boxes; per child octant the operator (VS x VS, k-major) is staged
// in LDS in K-chunks and applied to the 16 gathered input vectors, each
// thread holding a (rows x 4 boxes) register block.
//   M2M: output parent b, input = its child with octant o, summed over the
//        octants present (kernel/m2m.cxx:115-155).
//   L2L: the 16 children of a tile share one parity class, hence one octant
//        and one operator; output child = operator x its parent
//        (kernel/l2l.cxx:203-237).
constexpr int TR_TB = 16, TR_KC = 32, TR_BST = TR_TB + 4, TR_NT = 128;

template <int VS, typename T, int L2L>
__global__ __launch_bounds__(TR_NT) void k_translate(int nout, const int* __restrict__ cfirst,
                                                     const uint32_t* __restrict__ key, const int* __restrict__ parent,
                                                     const int* __restrict__ cls_cnt, const uint32_t* __restrict__ ord,
                                                     const T* __restrict__ in, T* __restrict__ out,
                                                     const T* __restrict__ opT) {
  // M2M: blockIdx.y = child octant; this block writes that octant's partial
  // product into out + octant * nout * VS (summed afterwards in fixed order)
  constexpr int RPT = (VS + 31) / 32, KC = VS > 160 ? TR_KC / 2 : TR_KC;  // LDS within 64 KB
  __shared__ T As[KC][VS];
  __shared__ T Bs[KC][TR_BST];
  __shared__ int src[TR_TB], dst[TR_TB];
  __shared__ int s_oct;
  const int tid = threadIdx.x;
  const int tx = tid & 3, ty = tid >> 2;  // 4 boxes x 32 row groups
  if (tid == 0) {
    int oct = -1;
    if (L2L) {
      // tile -> (class, index) over the class-sorted child list
      int t = blockIdx.x, start = 0;
      for (int c = 0; c < 8; c++) {
        const int cnt = cls_cnt[c], nt = (cnt + TR_TB - 1) / TR_TB;
        if (oct < 0 && t < nt) {
          // class bits (x,y,z) -> octant bits (y | x<<1 | z<<2)
          oct = ((c >> 1) & 1) | ((c & 1) << 1) | (c & 4);
          start += t * TR_TB;
          s_oct = oct;
          for (int k = 0; k < TR_TB; k++) {
            const int i = t * TR_TB + k;
            dst[k] = i < cnt ? (int)ord[start - t * TR_TB + i] : -1;
          }
        } else if (oct < 0) {
          t -= nt;
          start += cnt;
        }
      }
      if (oct < 0) s_oct = -1;
    } else {
      s_oct = 0;
      for (int k = 0; k < TR_TB; k++) {
        const int b = blockIdx.x * TR_TB + k;
        dst[k] = b < nout ? b : -1;
      }
    }
  }
  __syncthreads();
  if (s_oct < 0) return;
  T acc[RPT][4];
#pragma unroll
  for (int r = 0; r < RPT; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[r][c] = T(0);
  const int o_begin = L2L ? s_oct : (int)blockIdx.y, o_end = o_begin + 1;
  if (!L2L) out += (size_t)blockIdx.y * nout * VS;
  for (int o = o_begin; o < o_end; o++) {
    __syncthreads();
    if (tid < TR_TB) {
      const int b = dst[tid];
      int s = -1;
      if (b >= 0) {
        if (L2L) {
          s = parent[b];
        } else {
          for (int c = cfirst[b]; c < cfirst[b + 1]; c++)
            if ((int)(key[c] & 7) == o) s = c;
        }
      }
      src[tid] = s;
    }
    __syncthreads();
    const T* A = opT + (size_t)o * VS * VS;
    for (int k0 = 0; k0 < VS; k0 += KC) {
      const int kc = min(KC, VS - k0);
      __syncthreads();
      for (int e = tid; e < kc * VS; e += TR_NT) As[e / VS][e % VS] = A[(size_t)k0 * VS + e];
      for (int e = tid; e < kc * TR_TB; e += TR_NT) {
        const int t = e / kc, k = e % kc, s = src[t];
        Bs[k][t] = s >= 0 ? in[(size_t)s * VS + k0 + k] : T(0);
      }
      __syncthreads();
      for (int k = 0; k < kc; k++) {
        T bv[4];
#pragma unroll
        for (int c = 0; c < 4; c++) bv[c] = Bs[k][tx * 4 + c];
#pragma unroll
        for (int r = 0; r < RPT; r++) {
          const int row = ty + 32 * r;
          const T av = row < VS ? As[k][row] : T(0);
#pragma unroll
          for (int c = 0; c < 4; c++) acc[r][c] += av * bv[c];
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int b = dst[tx * 4 + c];
    if (b < 0) continue;
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int row = ty + 32 * r;
      if (row < VS) out[(size_t)b * VS + row] = acc[r][c];
    }
  }
}

// v_mfma 16x16x4 in FP32 and FP64
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  using v4 = f32x4;
  static __device__ __forceinline__ v4 fma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<double> {
  using v4 = f64x4;
  static __device__ __forceinline__ v4 fma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// ---------------------------------------------------------------------
// M2M and L2L of the FP32 engine on MFMA (v_mfma_f32_16x16x4f32), operators
// in the fragment order of the M2L kernel ([ks][mt][lane] holds row
// 16 mt + lane % 16, column (lane / 16) KS + ks of the p^2 x p^2 operator).
//   M2M: a workgroup of 8 waves owns 16 parent boxes; wave w multiplies the
//        octant-w operator with the 16 children of octant w (zero where a
//        parent has none) and the 8 products are summed in LDS in octant
//        order, so the result does not depend on scheduling
//        (kernel/m2m.cxx:115-155).
//   L2L: a wave owns 16 children of one parity class, which share one octant
//        and so one operator; each child's locals = operator x its parent's
//        (kernel/l2l.cxx:203-237).  The children's locals are written, or
//        added to when the level's M2L ran first (add).
template <int P, int L2L>
__global__ __launch_bounds__(L2L ? 256 : 512) void k_translate_mfma(int nout, const int* __restrict__ cfirst,
                                                                    const uint32_t* __restrict__ key,
                                                                    const int* __restrict__ parent,
                                                                    const int* __restrict__ cls_cnt,
                                                                    const uint32_t* __restrict__ ord,
                                                                    const float* __restrict__ in,
                                                                    float* __restrict__ out,
                                                                    const float* __restrict__ opF, int add) {
  using ML = MLayout<P>;
  constexpr int P2 = P * P, MT = (P2 + 15) / 16, KS = (P2 + 3) / 4, FR = KS * MT * 64;
  using V4 = f32x4;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, j = l & 15, kq = l >> 4;
  int dst = -1, src = -1, oct = 0;
  if (L2L) {
    // wave -> (class, tile) over the class-sorted child list
    int t = blockIdx.x * 4 + w, start = 0, c = 0;
    for (; c < 8; c++) {
      const int cnt = cls_cnt[c], nt = (cnt + 15) / 16;
      if (t < nt) break;
      t -= nt;
      start += cnt;
    }
    if (c == 8) return;  // no barrier below in this mode
    // class bits (x, y, z) -> octant bits (y | x << 1 | z << 2)
    oct = ((c >> 1) & 1) | ((c & 1) << 1) | (c & 4);
    const int i = t * 16 + j;
    if (i < cls_cnt[c]) {
      dst = (int)ord[start + i];
      src = parent[dst];
    }
  } else {
    oct = w;
    const int b = blockIdx.x * 16 + j;
    if (b < nout) {
      dst = b;
      for (int c = cfirst[b]; c < cfirst[b + 1]; c++)
        if ((int)(key[c] & 7) == oct) src = c;
    }
  }
  V4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; mt++) acc[mt] = V4{0.f, 0.f, 0.f, 0.f};
  if (__ballot(src >= 0) != 0) {
    float bv[KS];
    if (L2L) {
      const float* ps = in + (size_t)(src >= 0 ? src : 0) * P2 + kq * KS;
#pragma unroll
      for (int ks = 0; ks < KS; ks++) bv[ks] = (src >= 0 && kq * KS + ks < P2) ? ps[ks] : 0.f;
    } else {
      const float4* ms = reinterpret_cast<const float4*>(in + (size_t)(src >= 0 ? src : 0) * ML::VS + kq * ML::KSP);
#pragma unroll
      for (int v4 = 0; v4 < ML::KSP / 4; v4++) {
        const float4 m4 = src >= 0 ? ms[v4] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (4 * v4 + 0 < KS) bv[4 * v4 + 0] = (kq * KS + 4 * v4 + 0 < P2) ? m4.x : 0.f;
        if (4 * v4 + 1 < KS) bv[4 * v4 + 1] = (kq * KS + 4 * v4 + 1 < P2) ? m4.y : 0.f;
        if (4 * v4 + 2 < KS) bv[4 * v4 + 2] = (kq * KS + 4 * v4 + 2 < P2) ? m4.z : 0.f;
        if (4 * v4 + 3 < KS) bv[4 * v4 + 3] = (kq * KS + 4 * v4 + 3 < P2) ? m4.w : 0.f;
      }
    }
    const float* A = opF + (size_t)oct * FR + l;
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      float af[MT];
#pragma unroll
      for (int mt = 0; mt < MT; mt++) af[mt] = A[(ks * MT + mt) * 64];
#pragma unroll
      for (int mt = 0; mt < MT; mt++) acc[mt] = Mfma<float>::fma(af[mt], bv[ks], acc[mt]);
    }
  }
  if (L2L) {
    if (dst < 0) return;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * mt + 4 * kq + r;
        if (i < P2) {
          float* o = out + (size_t)dst * P2 + i;
          *o = add ? *o + acc[mt][r] : acc[mt][r];
        }
      }
    return;
  }
  // M2M: the 8 octant products through LDS, summed in octant order
  __shared__ float red[8][MT * 4 * 64];
#pragma unroll
  for (int mt = 0; mt < MT; mt++)
#pragma unroll
    for (int r = 0; r < 4; r++) red[w][(mt * 4 + r) * 64 + l] = acc[mt][r];
  __syncthreads();
  for (int e = tid; e < 16 * ML::VS; e += 512) {
    const int jb = e / ML::VS, pos = e % ML::VS, b = blockIdx.x * 16 + jb;
    if (b >= nout) continue;
    // padded position -> slot (pads are written as zeros)
    const int seg = pos / ML::KSP, ks = pos % ML::KSP, i = seg * KS + ks;
    float v = 0.f;
    if (ks < KS && i < P2) {
      const int q = ((i >> 4) * 4 + (i & 3)) * 64 + ((i >> 2) & 3) * 16 + jb;
#pragma unroll
      for (int o = 0; o < 8; o++) v += red[o][q];
    }
    out[(size_t)b * ML::VS + pos] = v;
  }
}

// ---------------------------------------------------------------------
// M2L on MFMA.  A workgroup (WPG waves) owns a tile of 16 WPG targets of one
// parity class at a level; all of them see the same 189 lattice offsets
// (fmm/ijbox.cxx:366-401).  Per offset the operator (p^2 x p^2, stored by the
// host in MFMA fragment order) is staged once in LDS and each wave runs
// ceil(p^2/16) x ceil(p^2/4) v_mfma 16x16x4 steps for its 16 targets, the
// source multipoles (gathered through the dense key->box table) feeding the
// B operand straight from registers.  With nsplit > 1 the offsets are split
// over workgroups and the partial sums reduced by k_m2l_reduce in fixed
// order (deterministic).
constexpr int M2L_TN = 64;  // targets per workgroup of the FP64 vector kernel

// FP32 MFMA kernel: NT N-tiles (16 targets each) per wave, 4 waves, so a
// workgroup owns 64*NT targets.  With PA the next offset's operator
// fragments are prefetched into registers while the current offset runs on
// the matrix cores; the next offset's source multipoles always are
// (software pipelining).  The operator goes through one LDS buffer shared by
// the 4 waves.
template <int P, int NT, int PA, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_m2l(M2LLaunch A) {
  // this workgroup's level (uniform selects: no dynamic indexing of the
  // kernel arguments)
  M2LLev Q = A.lv[0];
#pragma unroll
  for (int k = 1; k < MAXL + 1; k++)
    if (k < A.nlv && (int)blockIdx.x >= A.lv[k].wg0) Q = A.lv[k];
  const int lev = Q.lev, nb = Q.nb, nsplit = Q.nsplit, init = Q.init, region = Q.region, gsh = Q.gsh;
  const int* __restrict__ cls_cnt = Q.cls_cnt;
  const uint32_t* __restrict__ ord = Q.ord;
  const uint32_t* __restrict__ ikey = Q.ikey;
  const int* __restrict__ dense_j = Q.dense_j;
  const float* __restrict__ M = Q.M;
  float* __restrict__ L = Q.L;
  float* __restrict__ part = Q.part;
  const float* __restrict__ opF = Q.opF;
  const int4* __restrict__ offtab = Q.offtab;
  unsigned long long* __restrict__ npairs = Q.npairs;
  const int bid = (int)blockIdx.x - Q.wg0, nwg = Q.nwg;
  using T = float;
  using ML = MLayout<P>;
  constexpr int P2 = P * P, MT = (P2 + 15) / 16, KS = (P2 + 3) / 4, FR = KS * MT * 64;
  constexpr int NTH = 64 * WPG;          // threads per workgroup
  constexpr int TN = 16 * NT * WPG;      // targets per workgroup
  constexpr int NV = FR / 4;             // float4 per operator
  constexpr int NVT = (NV + NTH - 1) / NTH;  // float4 per thread
  using V4 = f32x4;
  // PA 2: two operator buffers (the next offset's operator is written to one
  // while the MFMAs read the other; one barrier per offset)
  constexpr int NBUF = PA == 2 ? 2 : 1;
  __shared__ __align__(16) T As[NBUF * FR];
  __shared__ int s_class, s_start, s_n;
  const int tid = threadIdx.x;
  // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs,
  // so workgroup b runs on XCD b%8; class b%8 keeps each XCD's L2 on one
  // parity class's 189 operators.  Work item k = b/8 of that class, split
  // major: the workgroups resident on an XCD at a time share one split's
  // 189/nsplit operators, which then stay in its L2 across tiles.
  // With region set, workgroup b runs on XCD b%8 for a block of tiles
  // instead: XCD x owns tiles [x*tpr, (x+1)*tpr) of every class, so the 8
  // siblings of a box (one per class) run on one XCD and share their source
  // multipoles in its L2; items on an XCD go tile-major, then split, then
  // class.
  int cls0, split, tile;
  // (level ranges start at multiples of 8, so bid % 8 is still the XCD)
  if (region) {
    const int xcd = bid & 7, item = bid >> 3, tpr = (nwg >> 6) / nsplit;
    cls0 = item & 7;
    split = (item >> 3) % nsplit;
    tile = xcd * tpr + (item >> 3) / nsplit;
  } else {
    const int item = bid >> 3, tpc = (nwg >> 3) / nsplit;
    cls0 = bid & 7;
    split = item / tpc;
    tile = item % tpc;
  }
  if (tid == 0) {
    int start = 0;
    for (int c = 0; c < cls0; c++) start += cls_cnt[c];
    const int cnt = cls_cnt[cls0], nt = (cnt + TN - 1) / TN;
    s_class = tile < nt ? cls0 : -1;
    s_start = start + tile * TN;
    s_n = tile < nt ? min(TN, cnt - tile * TN) : 0;
  }
  __syncthreads();
  const int cls = s_class, ntile = s_n;
  if (cls < 0) return;
  const int w = tid >> 6, l = tid & 63, j = l & 15, kq = l >> 4;
  const int g = 1 << lev;
  int tord[NT], tx[NT], ty[NT], tz[NT];
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const int t = (w * NT + n) * 16 + j;
    tord[n] = -1;
    tx[n] = ty[n] = tz[n] = 0;
    if (t < ntile) {
      tord[n] = (int)ord[s_start + t];
      morton_decode(ikey[tord[n]], lev, tx[n], ty[n], tz[n]);
    }
  }
  const bool wave_on = w * NT * 16 < ntile;
  const int per = (NEBM + nsplit - 1) / nsplit, o0 = split * per, o1 = min(NEBM, o0 + per);
  V4 acc[NT][MT];
#pragma unroll
  for (int n = 0; n < NT; n++)
#pragma unroll
    for (int mt = 0; mt < MT; mt++) acc[n][mt] = V4{0.f, 0.f, 0.f, 0.f};
  unsigned long long mypairs = 0;
  float4 areg[PA ? NVT : 1];  // PA 1: unsigned operator prefetch; PA 2: signed
  // Reflected offsets use the operator of the canonical offset with element
  // (row i, col k) times T_r(i) T_r(k).  The float4s a thread stages sit at
  // fixed places, so their sign patterns are worked out once: nibble r of
  // sgn[v] flags the 4 floats of the thread's v-th float4 under reflection r.
  // Fragment element q = (ks * MT + mt) * 64 + lane holds row 16 mt + lane % 16
  // and col (lane / 16) * KS + ks.
  uint32_t sgn[PA == 1 ? 1 : NVT];
  if (PA != 1) {
    constexpr ReflMask<P> RM{};
#pragma unroll
    for (int v = 0; v < NVT; v++) {
      sgn[v] = 0;
      const int e = tid + NTH * v;
      if (e < NV) {
        const int q = 4 * e, lane = q & 63, mt = (q >> 6) % MT, ks = (q >> 6) / MT;
        const int col = (lane >> 4) * KS + ks, row0 = 16 * mt + (lane & 15);
        for (int r = 1; r < 8; r++) {
          const bool cn = col < P2 && RM.neg(r, col);
          for (int jj = 0; jj < 4; jj++)
            if (row0 + jj < P2 && RM.neg(r, row0 + jj) != cn) sgn[v] |= 1u << (4 * r + jj);
        }
      }
    }
  }
  T bcur[NT][KS];
#define FMMX_LOAD_B(o_, dst_)                                                                 \
  {                                                                                           \
    const int4 off_ = offtab[cls * NEBM + (o_)];                                              \
    _Pragma("unroll") for (int n = 0; n < NT; n++) {                                          \
      int s_ = -1;                                                                            \
      if (tord[n] >= 0) {                                                                     \
        const int jx = tx[n] - off_.x, jy = ty[n] - off_.y, jz = tz[n] - off_.z;              \
        if (jx >= 0 && jx < g && jy >= 0 && jy < g && jz >= 0 && jz < g &&                    \
            (gsh < 0 || !((jx ^ tx[n]) >> gsh | (jy ^ ty[n]) >> gsh | (jz ^ tz[n]) >> gsh)))   \
          s_ = dense_j[morton_encode(jx, jy, jz, lev)];                                       \
      }                                                                                       \
      if (s_ >= 0 && kq == 0) mypairs++;                                                      \
      any_ = any_ || s_ >= 0;                                                                 \
      const float4* ms_ = reinterpret_cast<const float4*>(M + (size_t)(s_ >= 0 ? s_ : 0) * ML::VS + kq * ML::KSP); \
      _Pragma("unroll") for (int v4 = 0; v4 < ML::KSP / 4; v4++) {                            \
        const float4 m4 = s_ >= 0 ? ms_[v4] : make_float4(0.f, 0.f, 0.f, 0.f);                \
        if (4 * v4 + 0 < KS) dst_[n][4 * v4 + 0] = (kq * KS + 4 * v4 + 0 < P2) ? m4.x : 0.f;  \
        if (4 * v4 + 1 < KS) dst_[n][4 * v4 + 1] = (kq * KS + 4 * v4 + 1 < P2) ? m4.y : 0.f;  \
        if (4 * v4 + 2 < KS) dst_[n][4 * v4 + 2] = (kq * KS + 4 * v4 + 2 < P2) ? m4.z : 0.f;  \
        if (4 * v4 + 3 < KS) dst_[n][4 * v4 + 3] = (kq * KS + 4 * v4 + 3 < P2) ? m4.w : 0.f;  \
      }                                                                                       \
    }                                                                                         \
  }
#define FMMX_LOAD_A(o_)                                                                       \
  {                                                                                           \
    const float4* src_ = reinterpret_cast<const float4*>(opF + (size_t)(offtab[cls * NEBM + (o_)].w & 0xffff) * FR); \
    _Pragma("unroll") for (int v = 0; v < NVT; v++) {                                         \
      const int e = tid + NTH * v;                                                            \
      if (e < NV) areg[v] = src_[e];                                                          \
    }                                                                                         \
  }
#define FMMX_LOAD_AS(o_)                                                                      \
  {                                                                                           \
    const int w_ = offtab[cls * NEBM + (o_)].w, r_ = w_ >> 16;                                \
    const float4* src_ = reinterpret_cast<const float4*>(opF + (size_t)(w_ & 0xffff) * FR);   \
    _Pragma("unroll") for (int v = 0; v < NVT; v++) {                                         \
      const int e = tid + NTH * v;                                                            \
      if (e < NV) {                                                                           \
        float4 x_ = src_[e];                                                                  \
        const uint32_t nib_ = sgn[v] >> (4 * r_);                                             \
        x_.x = __uint_as_float(__float_as_uint(x_.x) ^ ((nib_ & 1u) << 31));                  \
        x_.y = __uint_as_float(__float_as_uint(x_.y) ^ (((nib_ >> 1) & 1u) << 31));           \
        x_.z = __uint_as_float(__float_as_uint(x_.z) ^ (((nib_ >> 2) & 1u) << 31));           \
        x_.w = __uint_as_float(__float_as_uint(x_.w) ^ (((nib_ >> 3) & 1u) << 31));           \
        areg[v] = x_;                                                                         \
      }                                                                                       \
    }                                                                                         \
  }
#define FMMX_STORE_A(buf_)                                                                    \
  {                                                                                           \
    _Pragma("unroll") for (int v = 0; v < NVT; v++) {                                         \
      const int e = tid + NTH * v;                                                            \
      if (e < NV) reinterpret_cast<float4*>(As + (buf_) * FR)[e] = areg[v];                   \
    }                                                                                         \
  }
  if constexpr (PA == 2) {
    // double-buffered operator: load offset o+1's operator (signed) and
    // source multipoles into registers, run offset o's MFMAs from one LDS
    // buffer, write o+1's operator into the other, one barrier
    bool cur_any = false;
    if (o0 < o1) {
      FMMX_LOAD_AS(o0);
      FMMX_STORE_A(0);
      bool any_ = false;
      FMMX_LOAD_B(o0, bcur);
      cur_any = __ballot(any_) != 0;
    }
    __syncthreads();
    int buf = 0;
    for (int o = o0; o < o1; o++) {
      T bnext[NT][KS];
      bool next_any = false;
      if (o + 1 < o1) {
        FMMX_LOAD_AS(o + 1);
        bool any_ = false;
        FMMX_LOAD_B(o + 1, bnext);
        next_any = __ballot(any_) != 0;
      }
      if (wave_on && cur_any) {
        const T* Ab = As + buf * FR;
      T af[2][MT];
#pragma unroll
      for (int mt = 0; mt < MT; mt++) af[0][mt] = Ab[mt * 64 + l];
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        if (ks + 1 < KS) {
#pragma unroll
          for (int mt = 0; mt < MT; mt++) af[(ks + 1) & 1][mt] = Ab[((ks + 1) * MT + mt) * 64 + l];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
          for (int n = 0; n < NT; n++) acc[n][mt] = Mfma<T>::fma(af[ks & 1][mt], bcur[n][ks], acc[n][mt]);
        __builtin_amdgcn_sched_barrier(0);
      }
      }
      if (o + 1 < o1) FMMX_STORE_A(buf ^ 1);
      __syncthreads();
      if (o + 1 < o1) {
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
          for (int ks = 0; ks < KS; ks++) bcur[n][ks] = bnext[n][ks];
        cur_any = next_any;
      }
      buf ^= 1;
    }
  } else {
  // a wave skips the MFMAs of an offset for which none of its 16 targets has
  // a source box (boundary tiles); the vote keeps the branch wave-uniform
  bool cur_any = false;
  if (o0 < o1) {
    if (PA) FMMX_LOAD_A(o0);
    bool any_ = false;
    FMMX_LOAD_B(o0, bcur);
    cur_any = __ballot(any_) != 0;
  }
  for (int o = o0; o < o1; o++) {
    __syncthreads();  // every wave is done with the previous operator
    if (PA) {
#pragma unroll
      for (int v = 0; v < NVT; v++) {
        const int e = tid + NTH * v;
        if (e < NV) reinterpret_cast<float4*>(As)[e] = areg[v];
      }
    } else {
      const int w = offtab[cls * NEBM + o].w, refl = w >> 16;
      const float4* src = reinterpret_cast<const float4*>(opF + (size_t)(w & 0xffff) * FR);
#pragma unroll
      for (int v = 0; v < NVT; v++) {
        const int e = tid + NTH * v;
        if (e < NV) {
          float4 x = src[e];
          const uint32_t nib = sgn[v] >> (4 * refl);
          x.x = __uint_as_float(__float_as_uint(x.x) ^ ((nib & 1u) << 31));
          x.y = __uint_as_float(__float_as_uint(x.y) ^ (((nib >> 1) & 1u) << 31));
          x.z = __uint_as_float(__float_as_uint(x.z) ^ (((nib >> 2) & 1u) << 31));
          x.w = __uint_as_float(__float_as_uint(x.w) ^ (((nib >> 3) & 1u) << 31));
          reinterpret_cast<float4*>(As)[e] = x;
        }
      }
    }
    __syncthreads();
    T bnext[NT][KS];
    bool next_any = false;
    if (o + 1 < o1) {
      if (PA) FMMX_LOAD_A(o + 1);
      bool any_ = false;
      FMMX_LOAD_B(o + 1, bnext);
      next_any = __ballot(any_) != 0;
    }
    if (wave_on && cur_any) {
      // operator fragments of step ks+1 are read from LDS while the MFMAs of
      // step ks run; the scheduling barriers keep that order, so the LDS
      // latency sits behind MT matrix instructions instead of in front of one
      const T* Ab = As;
      T af[2][MT];
#pragma unroll
      for (int mt = 0; mt < MT; mt++) af[0][mt] = Ab[mt * 64 + l];
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        if (ks + 1 < KS) {
#pragma unroll
          for (int mt = 0; mt < MT; mt++) af[(ks + 1) & 1][mt] = Ab[((ks + 1) * MT + mt) * 64 + l];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
          for (int n = 0; n < NT; n++) acc[n][mt] = Mfma<T>::fma(af[ks & 1][mt], bcur[n][ks], acc[n][mt]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (o + 1 < o1) {
#pragma unroll
      for (int n = 0; n < NT; n++)
#pragma unroll
        for (int ks = 0; ks < KS; ks++) bcur[n][ks] = bnext[n][ks];
      cur_any = next_any;
    }
  }
  }
#undef FMMX_LOAD_AS
#undef FMMX_STORE_A
#undef FMMX_LOAD_A
#undef FMMX_LOAD_B
#pragma unroll
  for (int n = 0; n < NT; n++) {
    if (tord[n] < 0) continue;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * mt + 4 * kq + r;
        if (i < P2) {
          const T v = acc[n][mt][r];
          if (nsplit == 1) {
            T* dst = L + (size_t)tord[n] * P2 + i;
            *dst = init ? v : *dst + v;
          } else {
            part[((size_t)split * nb + tord[n]) * P2 + i] = v;
          }
        }
      }
  }
  if (mypairs) atomicAdd(npairs, mypairs);
}

// Helmholtz M2L of the FP32 engine on MFMA (v_mfma_f32_16x16x4f32): split
// complex expansions, D = 2 p^2 rows ([Re; Im] locals) and columns ([R; I]
// multipoles), operators per level in fragment order.  The whole operator
// (D^2 floats, 332 KB at p = 12) exceeds LDS, so it is staged in chunks of
// HH_KC k-steps shared by the workgroup's waves; each wave keeps its 16
// targets' source multipoles (D / 4 per lane) in registers and skips the
// products of an offset for which none of its targets has a source box.  Tile,
// class, split and XCD mapping as k_m2l.
constexpr int HH_KC = 8;
template <int P, int WPG>
__global__ __launch_bounds__(64 * WPG) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_m2l_hh(int lev, int nb, int nsplit, const int* __restrict__ cls_cnt,
                                                     const uint32_t* __restrict__ ord,
                                                     const uint32_t* __restrict__ ikey,
                                                     const int* __restrict__ dense_j, const float* __restrict__ M,
                                                     float* __restrict__ L, float* __restrict__ part,
                                                     const float* __restrict__ opF, const int4* __restrict__ offtab,
                                                     int init, int region, int gsh,
                                                     unsigned long long* __restrict__ npairs) {
  using ML = MLayout<P>;
  constexpr int P2 = P * P, D = 2 * P2, MT = (D + 15) / 16, KS = D / 4, FR = KS * MT * 64;
  constexpr int NCH = (KS + HH_KC - 1) / HH_KC, CHF = HH_KC * MT * 64;  // chunks, floats per chunk
  constexpr int NTH = 64 * WPG, TN = 16 * WPG;
  static_assert(P % 2 == 0 && KS * 2 == P2, "each lane's k-range lies in one part");
  using V4 = f32x4;
  // two chunk buffers: the next chunk is loaded into registers while the
  // current one feeds the MFMAs, then stored into the other buffer (one
  // barrier per chunk)
  __shared__ __align__(16) float As[2 * CHF];
  __shared__ int s_class, s_start, s_n;
  const int tid = threadIdx.x;
  int cls0, split, tile;
  if (region) {
    const int xcd = blockIdx.x & 7, item = blockIdx.x >> 3, tpr = (int)(gridDim.x >> 6) / nsplit;
    cls0 = item & 7;
    split = (item >> 3) % nsplit;
    tile = xcd * tpr + (item >> 3) / nsplit;
  } else {
    const int item = blockIdx.x >> 3, tpc = (int)(gridDim.x >> 3) / nsplit;
    cls0 = blockIdx.x & 7;
    split = item / tpc;
    tile = item % tpc;
  }
  if (tid == 0) {
    int start = 0;
    for (int c = 0; c < cls0; c++) start += cls_cnt[c];
    const int cnt = cls_cnt[cls0], nt = (cnt + TN - 1) / TN;
    s_class = tile < nt ? cls0 : -1;
    s_start = start + tile * TN;
    s_n = tile < nt ? min(TN, cnt - tile * TN) : 0;
  }
  __syncthreads();
  const int cls = s_class, ntile = s_n;
  if (cls < 0) return;
  const int w = tid >> 6, l = tid & 63, j = l & 15, kq = l >> 4;
  const int g = 1 << lev;
  const int t = w * 16 + j;
  int tord = -1, tx = 0, ty = 0, tz = 0;
  if (t < ntile) {
    tord = (int)ord[s_start + t];
    morton_decode(ikey[tord], lev, tx, ty, tz);
  }
  const bool wave_on = w * 16 < ntile;
  const int per = (NEBM + nsplit - 1) / nsplit, o0 = split * per, o1 = min(NEBM, o0 + per);
  V4 acc[MT];
#pragma unroll
  for (int mt = 0; mt < MT; mt++) acc[mt] = V4{0.f, 0.f, 0.f, 0.f};
  unsigned long long mypairs = 0;
  // lane kq holds k = kq KS + ks: part kq / 2, slots (kq % 2) KS + ks
  const int bpart = kq >> 1, bslot = (kq & 1) * KS;
  constexpr int NV4 = (CHF / 4 + NTH - 1) / NTH;  // float4 per thread per chunk
  V4 areg[NV4];
#define FMMX_HH_OP(o_) (opF + (size_t)(offtab[cls * NEBM + (o_)].w & 0xffff) * FR)
#define FMMX_HH_LOAD(a_, c_)                                                                     \
  {                                                                                              \
    const int nf4_ = min(HH_KC, KS - (c_) * HH_KC) * MT * 16;                                     \
    const V4* src_ = reinterpret_cast<const V4*>((a_) + (size_t)(c_) * CHF);             \
    _Pragma("unroll") for (int v = 0; v < NV4; v++) {                                            \
      const int e_ = tid + NTH * v;                                                              \
      if (e_ < nf4_) areg[v] = src_[e_];                                                         \
    }                                                                                            \
  }
#define FMMX_HH_STORE(buf_, c_)                                                                  \
  {                                                                                              \
    const int nf4_ = min(HH_KC, KS - (c_) * HH_KC) * MT * 16;                                     \
    V4* dst_ = reinterpret_cast<V4*>(As + (buf_) * CHF);                                 \
    _Pragma("unroll") for (int v = 0; v < NV4; v++) {                                            \
      const int e_ = tid + NTH * v;                                                              \
      if (e_ < nf4_) dst_[e_] = areg[v];                                                         \
    }                                                                                            \
  }
  if (o0 < o1) {
    FMMX_HH_LOAD(FMMX_HH_OP(o0), 0);
    FMMX_HH_STORE(0, 0);
  }
  __syncthreads();
  int buf = 0;
  for (int o = o0; o < o1; o++) {
    const int4 off = offtab[cls * NEBM + o];
    int s_ = -1;
    if (tord >= 0) {
      const int jx = tx - off.x, jy = ty - off.y, jz = tz - off.z;
      if (jx >= 0 && jx < g && jy >= 0 && jy < g && jz >= 0 && jz < g &&
          (gsh < 0 || !((jx ^ tx) >> gsh | (jy ^ ty) >> gsh | (jz ^ tz) >> gsh)))
        s_ = dense_j[morton_encode(jx, jy, jz, lev)];
    }
    if (s_ >= 0 && kq == 0) mypairs++;
    const bool any = wave_on && __ballot(s_ >= 0) != 0;
    float b[KS];
    if (any) {
      const float* ms = M + (size_t)(s_ >= 0 ? s_ : 0) * 2 * ML::VS + bpart * ML::VS;
#pragma unroll
      for (int ks = 0; ks < KS; ks++) b[ks] = s_ >= 0 ? ms[ML::at(bslot + ks)] : 0.f;
    }
    const float* Acur = FMMX_HH_OP(o);
    const float* Anext = o + 1 < o1 ? FMMX_HH_OP(o + 1) : nullptr;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      // the next chunk: of this offset, or the first of the next offset
      const bool has_next = c + 1 < NCH || Anext != nullptr;
      if (c + 1 < NCH) FMMX_HH_LOAD(Acur, c + 1)
      else if (Anext) FMMX_HH_LOAD(Anext, 0)
      if (any) {
        const float* Ab = As + buf * CHF;
#pragma unroll
        for (int kk = 0; kk < HH_KC; kk++) {
          const int ks = c * HH_KC + kk;
          if (ks < KS) {
            float af[MT];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) af[mt] = Ab[(kk * MT + mt) * 64 + l];
#pragma unroll
            for (int mt = 0; mt < MT; mt++) acc[mt] = Mfma<float>::fma(af[mt], b[ks], acc[mt]);
          }
        }
      }
      if (has_next) FMMX_HH_STORE(buf ^ 1, c + 1 < NCH ? c + 1 : 0)
      __syncthreads();
      buf ^= 1;
    }
  }
#undef FMMX_HH_OP
#undef FMMX_HH_LOAD
#undef FMMX_HH_STORE
  if (tord >= 0) {
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * mt + 4 * kq + r;
        if (i < D) {
          const float v = acc[mt][r];
          if (nsplit == 1) {
            float* dst = L + (size_t)tord * D + i;
            *dst = init ? v : *dst + v;
          } else {
            part[((size_t)split * nb + tord) * D + i] = v;
          }
        }
      }
  }
  if (mypairs) atomicAdd(npairs, mypairs);
}

// FP64 M2L on the vector FMA units (MI355X: FP64 vector peak = FP64 matrix
// peak).  Same tiles and offset split as k_m2l; per offset the operator
// (k-major) and the gathered source multipoles go through LDS in K-chunks
// and each thread accumulates a (rows x 4 targets) register block.
constexpr int M2LV_KC = 32, M2LV_BST = M2L_TN + 4;

// PARTS 2 (Helmholtz): split expansions, D = 2 p^2 rows and k-indices; the
// multipole part t of box s sits at s * 2 VS + t * VS in the padded layout.
template <int P, typename T, int PARTS = 1>
__global__ __launch_bounds__(256) void k_m2l_vec(int lev, int nb, int nsplit, const int* __restrict__ cls_cnt,
                                                 const uint32_t* __restrict__ ord, const uint32_t* __restrict__ ikey,
                                                 const int* __restrict__ dense_j, const T* __restrict__ M,
                                                 T* __restrict__ L, T* __restrict__ part, const T* __restrict__ opT,
                                                 const int4* __restrict__ offtab, int init, int gsh,
                                                 unsigned long long* __restrict__ npairs) {
  using ML = MLayout<P>;
  constexpr int P2 = P * P, D = PARTS * P2, RPT = (D + 15) / 16;
  constexpr int KC = D * (int)sizeof(T) > 1200 ? M2LV_KC / 2 : M2LV_KC;  // LDS within 64 KB
  __shared__ T As[KC][D];
  __shared__ T Bs[KC][M2LV_BST];
  __shared__ int tpos[M2L_TN][3];
  __shared__ int tord[M2L_TN];
  __shared__ int src[M2L_TN];
  __shared__ int s_class, s_start, s_n;
  const int tid = threadIdx.x;
  const int tile = blockIdx.x / nsplit, split = blockIdx.x % nsplit;
  if (tid == 0) {
    int w = tile, start = 0, cls = -1, n = 0;
    for (int c = 0; c < 8; c++) {
      const int cnt = cls_cnt[c], nt = (cnt + M2L_TN - 1) / M2L_TN;
      if (cls < 0 && w < nt) {
        cls = c;
        n = min(M2L_TN, cnt - w * M2L_TN);
        start += w * M2L_TN;
      } else if (cls < 0) {
        w -= nt;
        start += cnt;
      }
    }
    s_class = cls;
    s_start = start;
    s_n = n;
  }
  __syncthreads();
  const int cls = s_class, ntile = s_n;
  if (cls < 0) return;
  const int g = 1 << lev;
  if (tid < M2L_TN) {
    int ix = -1000, iy = -1000, iz = -1000, o = -1;
    if (tid < ntile) {
      o = (int)ord[s_start + tid];
      morton_decode(ikey[o], lev, ix, iy, iz);
    }
    tpos[tid][0] = ix;
    tpos[tid][1] = iy;
    tpos[tid][2] = iz;
    tord[tid] = o;
  }
  const int tx = tid & 15, ty = tid >> 4;
  T acc[RPT][4];
#pragma unroll
  for (int r = 0; r < RPT; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[r][c] = T(0);
  unsigned long long mypairs = 0;
  const int per = (NEBM + nsplit - 1) / nsplit, o0 = split * per, o1 = min(NEBM, o0 + per);
  for (int oi = o0; oi < o1; oi++) {
    const int4 off = offtab[cls * NEBM + oi];
    __syncthreads();
    int has = 0;
    if (tid < M2L_TN) {
      int s = -1;
      if (tid < ntile) {
        const int ix = tpos[tid][0], iy = tpos[tid][1], iz = tpos[tid][2];
        const int jx = ix - off.x, jy = iy - off.y, jz = iz - off.z;
        if (jx >= 0 && jx < g && jy >= 0 && jy < g && jz >= 0 && jz < g &&
            (gsh < 0 || !((jx ^ ix) >> gsh | (jy ^ iy) >> gsh | (jz ^ iz) >> gsh)))
          s = dense_j[morton_encode(jx, jy, jz, lev)];
      }
      src[tid] = s;
      has = s >= 0;
      if (s >= 0) mypairs++;
    }
    // no target of the tile has a source box at this offset (domain edge,
    // sparse surfaces): neither stage the operator nor multiply
    if (!__syncthreads_or(has)) continue;
    const T* A = opT + (size_t)off.w * D * D;
    for (int k0 = 0; k0 < D; k0 += KC) {
      const int kc = min(KC, D - k0);
      __syncthreads();
      for (int e = tid; e < kc * D; e += 256) As[e / D][e % D] = A[(size_t)k0 * D + e];
      for (int e = tid; e < kc * M2L_TN; e += 256) {
        const int t = e / kc, k = k0 + e % kc, s = src[t];
        const int part = PARTS == 1 ? 0 : k / P2, kk = PARTS == 1 ? k : k % P2;
        Bs[e % kc][t] = s >= 0 ? M[(size_t)s * PARTS * ML::VS + part * ML::VS + ML::at(kk)] : T(0);
      }
      __syncthreads();
      for (int k = 0; k < kc; k++) {
        T bv[4];
#pragma unroll
        for (int c = 0; c < 4; c++) bv[c] = Bs[k][tx * 4 + c];
#pragma unroll
        for (int r = 0; r < RPT; r++) {
          const int row = ty + 16 * r;
          const T av = row < D ? As[k][row] : T(0);
#pragma unroll
          for (int c = 0; c < 4; c++) acc[r][c] += av * bv[c];
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int t = tx * 4 + c;
    if (t >= ntile) continue;
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      const int row = ty + 16 * r;
      if (row >= D) continue;
      if (nsplit == 1) {
        T* dst = L + (size_t)tord[t] * D + row;
        *dst = init ? acc[r][c] : *dst + acc[r][c];
      } else {
        part[((size_t)split * nb + tord[t]) * D + row] = acc[r][c];
      }
    }
  }
  if (tid < M2L_TN && mypairs) atomicAdd(npairs, mypairs);
}

template <typename T>
__global__ void k_m2l_reduce(long n, int nsplit, const T* __restrict__ part, T* __restrict__ L, int init) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  T s = init ? T(0) : L[e];
  for (int k = 0; k < nsplit; k++) s += part[(size_t)k * n + e];
  L[e] = s;
}

// the reductions of several levels in one launch: elements [off[k], off[k+1])
// belong to level k (same order of the split sums as k_m2l_reduce, init 1)
struct ReduceLevels {
  const float* part[MAXL + 1];
  float* L[MAXL + 1];
  long off[MAXL + 2];
  int nsplit[MAXL + 1];
  int nl;
};
__global__ void k_m2l_reduce_levels(ReduceLevels R) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.off[R.nl]) return;
  int k = 0;
  while (k + 1 < R.nl && t >= R.off[k + 1]) k++;
  const long n = R.off[k + 1] - R.off[k], e = t - R.off[k];
  const float* part = R.part[k];
  float s = 0.f;
  for (int j = 0; j < R.nsplit[k]; j++) s += part[(size_t)j * n + e];
  R.L[k][e] = s;
}

// ---------------------------------------------------------------------
// reciprocal square root: v_rsq_f32 (1 ulp) in FP32, full precision in FP64
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq(double x) { return rsqrt(x); }
__device__ __forceinline__ float yexp(float x) { return __expf(x); }
__device__ __forceinline__ double yexp(double x) { return exp(x); }

// ---------------------------------------------------------------------
// P2P: one workgroup per target leaf box (chunks of BS targets), sources of
// each near box staged through LDS in tiles of BS.
template <typename T, int NEQ, int CONV, bool SELF = true>
__device__ __forceinline__ void pair(float xi, float yi, float zi, float nxi, float nyi, float nzi, float4 s,
                                     float4 sn, T& p, T& fx, T& fy, T& fz, T lam = T(0)) {
  T dx, dy, dz;
  if (CONV == FMMX_CONV_REF || sizeof(T) == 4) {
    dx = T(xi - s.x);  // float subtraction as kernel/G_p2p.cxx:21-23
    dy = T(yi - s.y);
    dz = T(zi - s.z);
  } else {
    dx = T(xi) - T(s.x);
    dy = T(yi) - T(s.y);
    dz = T(zi) - T(s.z);
  }
  T r2 = dx * dx + dy * dy + dz * dz;
  T inv;
  if (CONV == FMMX_CONV_REF) {
    r2 += T(EPSF);
    inv = rsq(r2);
  } else {
    inv = rsq(r2);
    if (SELF) inv = r2 > T(0) ? inv : T(0);  // coincident pair excluded (select, no branch)
  }
  const T qi = T(s.w) * inv;
  if (NEQ == FMMX_HELMHOLTZ) {
    // (qr + i qi) exp(i k r) / r: real part into p, imaginary part into fx
    // (sn.x carries the imaginary part of the charge, lam the wavenumber)
    const T r = r2 * inv;
    T sk, ck;
    if constexpr (sizeof(T) == 4) {
      // hardware sine/cosine (argument in revolutions): near-field k r is
      // O(1), where their absolute error is far below the FP32 path's
      const float t = float(lam * r) * 0.15915494309189535f;
      sk = __builtin_amdgcn_sinf(t);
      ck = __builtin_amdgcn_cosf(t);
    } else {
      sincos(lam * r, &sk, &ck);
    }
    const T qim = T(sn.x) * inv;
    p += qi * ck - qim * sk;
    fx += qi * sk + qim * ck;
  } else if (NEQ == FMMX_YUKAWA) {
    // exp(-lambda r)/r (kernel/Helmholtz_G_p2p.cxx:22 without its constant
    // (pi/2)/lambda); gradient -(1 + lambda r) e d / r^3
    const T r = r2 * inv;
    const T e = qi * yexp(-lam * r);
    p += e;
    const T rs = e * (T(1) + lam * r) * inv * inv;
    fx -= dx * rs;
    fy -= dy * rs;
    fz -= dz * rs;
  } else if (NEQ == FMMX_G) {
    p += qi;
    const T rs = qi * inv * inv;
    fx -= dx * rs;
    fy -= dy * rs;
    fz -= dz * rs;
  } else if (NEQ == FMMX_GNI) {
    p += (dx * T(nxi) + dy * T(nyi) + dz * T(nzi)) * (qi * inv * inv);
  } else {
    p += (dx * T(sn.x) + dy * T(sn.y) + dz * T(sn.z)) * (qi * inv * inv);
  }
}

// TPT lanes per target (2: leaves of at most about 32 targets, one wave per
// workgroup): the lanes of a target take alternating staged sources and add
// their sums at the end, so a half-empty leaf does not idle half the wave.
// With TPT 2 each pass over the leaf's targets picks its own lanes per target
// (2, 4, 8 or 16) from the targets it has left, so a pass of a few targets
// (a leaf of 33-40, or of 10) still keeps the wave's lanes busy.
template <typename T, int NEQ, int CONV, int BS, int TPT = 1>
__global__ __launch_bounds__(BS) void k_p2p(int nbi, const uint32_t* __restrict__ ikey, const uint32_t* __restrict__ jkey,
                                            const int* __restrict__ ifirst, const int* __restrict__ ilast,
                                            const int* __restrict__ nij, const int* __restrict__ neij,
                                            const int* __restrict__ jfirst, const int* __restrict__ jlast,
                                            const float* __restrict__ xi, const float* __restrict__ yi,
                                            const float* __restrict__ zi, const float* __restrict__ nxi,
                                            const float* __restrict__ nyi, const float* __restrict__ nzi,
                                            const float* __restrict__ xj, const float* __restrict__ yj,
                                            const float* __restrict__ zj, const float* __restrict__ q,
                                            const float* __restrict__ nxj, const float* __restrict__ nyj,
                                            const float* __restrict__ nzj, double c4, double* __restrict__ vs,
                                            double* __restrict__ gxs, double* __restrict__ gys,
                                            double* __restrict__ gzs, double lamd) {
  constexpr bool GRAD = NEQ == FMMX_G || NEQ == FMMX_YUKAWA;
  // Helmholtz: nxj holds the imaginary parts of the charges, gxs receives the
  // imaginary part of the potential
  constexpr bool HH = NEQ == FMMX_HELMHOLTZ, SN = NEQ == FMMX_GNJ || HH;
  const T lam = T(lamd);
  // staging slots (SB = SPT x BS, each thread stages SPT of them) and near
  // boxes packed per round: the one-wave form stages up to 256 sources of up
  // to 8 boxes, so a leaf's 27 near boxes take a few rounds instead of ~14
  constexpr int SPT = TPT == 2 ? 4 : 1, SB = SPT * BS, KB = TPT == 2 ? 8 : 4;
  __shared__ float4 sp[SB];
  __shared__ float4 sn[SN ? SB : 1];
  static_assert(TPT == 1 || (TPT == 2 && BS == 64), "two lanes per target within one wave");
  const int b = blockIdx.x, tid = threadIdx.x;
  if (b >= nbi) return;
  const int i0 = ifirst[b], i1 = ilast[b], nn = nij[b];
  for (int base = i0, tb = BS; base <= i1; base += tb) {
    // lanes per target of this pass (uniform): TPT, or up to 16 (TPT 2) for
    // the targets left
    int tpt = TPT;
    if (TPT == 2) {
      const int m = i1 - base + 1;
      tpt = m > 16 ? 2 : m > 8 ? 4 : m > 4 ? 8 : 16;
    }
    tb = BS / tpt;  // targets of this pass
    const int tt = tid % tb, part = tid / tb;
    const int i = base + tt;
    const bool act = i <= i1;
    float px = 0, py = 0, pz = 0, qx = 0, qy = 0, qz = 0;
    if (act) {
      px = xi[i];
      py = yi[i];
      pz = zi[i];
      if (NEQ == FMMX_GNI) {
        qx = nxi[i];
        qy = nyi[i];
        qz = nzi[i];
      }
    }
    double ap = 0, ax = 0, ay = 0, az = 0;
    const uint32_t kb = ikey[b];
    for (int e = 0; e < nn;) {
      // one staging round: near boxes e .. e + nbox - 1 packed into the SB
      // slots when they fit together (fewer barriers and shorter tails for
      // small leaves), else box e's sources SB at a time.  The target's own
      // box keeps a round to itself: only its pairs need the coincident test
      int jj = neij[b * NNEAR + e];
      int j0 = jfirst[jj], j1 = jlast[jj];
      int nbox = 1, cnt0 = j1 - j0 + 1, tot = cnt0;
      int s0[KB], c0[KB];
      s0[0] = j0;
      c0[0] = cnt0;
      const bool self = jkey[jj] == kb;
      bool packing = cnt0 <= SB && !self;
#pragma unroll
      for (int k = 1; k < KB; k++) {
        s0[k] = 0;
        c0[k] = 0;
        if (packing && e + k < nn) {
          const int jn = neij[b * NNEAR + e + k];
          const int a0 = jfirst[jn], c = jlast[jn] - a0 + 1;
          if (tot + c <= SB && jkey[jn] != kb) {
            s0[k] = a0;
            c0[k] = c;
            tot += c;
            nbox = k + 1;
          } else {
            packing = false;
          }
        }
      }
      const int rounds = cnt0 <= SB ? 1 : (cnt0 + SB - 1) / SB;
      for (int rr = 0; rr < rounds; rr++) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < SPT; u++) {
          const int slot = tid + u * BS;
          int js = -1;
          if (cnt0 <= SB) {
            // slot -> (packed box, offset)
            int t = slot;
#pragma unroll
            for (int k = 0; k < KB; k++) {
              if (js < 0 && k < nbox) {
                if (t < c0[k]) js = s0[k] + t;
                else t -= c0[k];
              }
            }
          } else if (rr * SB + slot < cnt0) {
            js = j0 + rr * SB + slot;
          }
          if (js >= 0) {
            sp[slot] = make_float4(xj[js], yj[js], zj[js], q[js]);
            if (NEQ == FMMX_GNJ) sn[slot] = make_float4(nxj[js], nyj[js], nzj[js], 0.f);
            if (HH) sn[slot] = make_float4(nxj[js], 0.f, 0.f, 0.f);
          }
        }
        __syncthreads();
        const int cnt = cnt0 <= SB ? tot : min(SB, cnt0 - rr * SB);
        T p = T(0), fx = T(0), fy = T(0), fz = T(0);
        // staged sources part, part + L, part + 2L, ... of the L lanes of a
        // target (L a compile-time stride: LDS offsets stay immediates)
        auto sweep = [&](auto lc) {
          constexpr int L = decltype(lc)::value;
          int s = part;
          // coincident points can only share a box key; other box pairs skip the test
          if (CONV == FMMX_CONV_EXACT && !self) {
            for (; s + 3 * L < cnt; s += 4 * L) {
#pragma unroll
              for (int u = 0; u < 4; u++)
                pair<T, NEQ, CONV, false>(px, py, pz, qx, qy, qz, sp[s + u * L], SN ? sn[s + u * L] : sp[0], p, fx,
                                          fy, fz, lam);
            }
            for (; s < cnt; s += L)
              pair<T, NEQ, CONV, false>(px, py, pz, qx, qy, qz, sp[s], SN ? sn[s] : sp[0], p, fx, fy, fz, lam);
          } else {
            for (; s + 3 * L < cnt; s += 4 * L) {
#pragma unroll
              for (int u = 0; u < 4; u++)
                pair<T, NEQ, CONV>(px, py, pz, qx, qy, qz, sp[s + u * L], SN ? sn[s + u * L] : sp[0], p, fx, fy,
                                   fz, lam);
            }
            for (; s < cnt; s += L)
              pair<T, NEQ, CONV>(px, py, pz, qx, qy, qz, sp[s], SN ? sn[s] : sp[0], p, fx, fy, fz, lam);
          }
        };
        if constexpr (TPT == 1) {
          sweep(std::integral_constant<int, 1>{});
        } else {
          if (tpt == 2) sweep(std::integral_constant<int, 2>{});
          else if (tpt == 4) sweep(std::integral_constant<int, 4>{});
          else if (tpt == 8) sweep(std::integral_constant<int, 8>{});
          else sweep(std::integral_constant<int, 16>{});
        }
        ap += (double)p;
        if (GRAD || HH) {
          ax += (double)fx;
          ay += (double)fy;
          az += (double)fz;
        }
      }
      e += nbox;
    }
    if (TPT == 2) {
      // the lanes of a target are tt, tt + tb, tt + 2 tb, ... of the one
      // wave: butterfly over the lane bits at and above tb (fixed order)
      for (int d = tb; d < 64; d <<= 1) {
        ap += __shfl_xor(ap, d, 64);
        if (GRAD || HH) {
          ax += __shfl_xor(ax, d, 64);
          ay += __shfl_xor(ay, d, 64);
          az += __shfl_xor(az, d, 64);
        }
      }
    }
    if (act && part == 0) {
      vs[i] = c4 * ap;
      if (HH) gxs[i] = c4 * ax;
      if (GRAD) {
        gxs[i] = c4 * ax;
        gys[i] = c4 * ay;
        gzs[i] = c4 * az;
      }
    }
  }
}

// direct sums: block of 64 targets x a slice of the sources, FP64 pair math,
// partial sums combined with atomics (outputs zeroed first)
template <int NEQ, int CONV>
__global__ __launch_bounds__(64) void k_direct(int ni, const float* __restrict__ xi, const float* __restrict__ yi,
                                               const float* __restrict__ zi, const float* __restrict__ nxi,
                                               const float* __restrict__ nyi, const float* __restrict__ nzi, int nj,
                                               const float* __restrict__ xj, const float* __restrict__ yj,
                                               const float* __restrict__ zj, const float* __restrict__ q,
                                               const float* __restrict__ nxj, const float* __restrict__ nyj,
                                               const float* __restrict__ nzj, int slice, double c4,
                                               double* __restrict__ v, double* __restrict__ gx,
                                               double* __restrict__ gy, double* __restrict__ gz, double lam) {
  constexpr bool GRAD = NEQ == FMMX_G || NEQ == FMMX_YUKAWA, HH = NEQ == FMMX_HELMHOLTZ;
  __shared__ float4 sp[64];
  __shared__ float4 sn[64];
  const int i = blockIdx.x * 64 + threadIdx.x;
  const bool act = i < ni;
  float px = 0, py = 0, pz = 0, qx = 0, qy = 0, qz = 0;
  if (act) {
    px = xi[i];
    py = yi[i];
    pz = zi[i];
    if (NEQ == FMMX_GNI) {
      qx = nxi[i];
      qy = nyi[i];
      qz = nzi[i];
    }
  }
  const int j0 = blockIdx.y * slice, j1 = min(nj, j0 + slice);
  double p = 0, fx = 0, fy = 0, fz = 0;
  for (int jb = j0; jb < j1; jb += 64) {
    const int js = jb + threadIdx.x;
    __syncthreads();
    if (js < j1) {
      sp[threadIdx.x] = make_float4(xj[js], yj[js], zj[js], q[js]);
      if (NEQ == FMMX_GNJ) sn[threadIdx.x] = make_float4(nxj[js], nyj[js], nzj[js], 0.f);
      if (HH) sn[threadIdx.x] = make_float4(nxj[js], 0.f, 0.f, 0.f);  // imaginary charge
    }
    __syncthreads();
    const int cnt = min(64, j1 - jb);
    for (int s = 0; s < cnt; s++)
      pair<double, NEQ, CONV>(px, py, pz, qx, qy, qz, sp[s], sn[s], p, fx, fy, fz, lam);
  }
  if (act) {
    atomicAdd(&v[i], c4 * p);
    if (HH) atomicAdd(&gx[i], c4 * fx);
    if (GRAD) {
      atomicAdd(&gx[i], c4 * fx);
      atomicAdd(&gy[i], c4 * fy);
      atomicAdd(&gz[i], c4 * fz);
    }
  }
}

template <typename T>
__global__ void k_zero_t(long n, T* __restrict__ a) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = T(0);
}
__global__ void k_zero_d(long n, double* __restrict__ a) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

template <class T>
struct tag {
  using type = T;
};

template <typename F>
int dispatch_p(int p, F&& f) {
  switch (p) {
#define FMMX_CASE(P) \
  case P:            \
    return f(std::integral_constant<int, P>{});
    FMMX_CASE(2) FMMX_CASE(3) FMMX_CASE(4) FMMX_CASE(5) FMMX_CASE(6) FMMX_CASE(7) FMMX_CASE(8) FMMX_CASE(9)
    FMMX_CASE(10) FMMX_CASE(11) FMMX_CASE(12)
#undef FMMX_CASE
  }
  return fail(FMMX_EUNSUPPORTED, "unsupported p");
}

// the orders the Helmholtz kernels are built for
template <typename F>
int dispatch_hp(int p, F&& f) {
  switch (p) {
#define FMMX_CASE(P) \
  case P:            \
    return f(std::integral_constant<int, P>{});
    FMMX_CASE(4) FMMX_CASE(6) FMMX_CASE(8) FMMX_CASE(10) FMMX_CASE(12)
#undef FMMX_CASE
  }
  return fail(FMMX_EUNSUPPORTED, "Helmholtz needs p in {4, 6, 8, 10, 12}");
}

// the order above P_ALL: FP64 and the EXACT convention (checked when the
// handle is created), on the vector-unit kernels
template <typename F>
int dispatch_hi(int p, F&& f) {
  if (p == P_HI) return f(std::integral_constant<int, P_HI>{});
  return fail(FMMX_EUNSUPPORTED, "unsupported p");
}

}  // namespace

// ---------------------------------------------------------------------
// launchers

int Engine::run_p2p(int neq) {
  const int nbi = si.nb[lmax];
  if (nbi == 0) return FMMX_OK;
  const double cc = neq == FMMX_YUKAWA ? c_yukawa() : c4();
  const bool ref = cfg.convention == FMMX_CONV_REF;
  const double mean = (double)si.n / nbi;
  // Helmholtz: the imaginary charges ride in the source-normal slot, k in lambda's
  const float* nxj = neq == FMMX_HELMHOLTZ ? qs2.as<float>() : sj.nx.as<float>();
  const double kl = neq == FMMX_HELMHOLTZ ? hk : lam;
  auto go = [&](auto tag_t, auto neq_c, auto conv_c, auto bs_c, auto tpt_c) -> int {
    using T = typename decltype(tag_t)::type;
    constexpr int NEQ = decltype(neq_c)::value, CONV = decltype(conv_c)::value, BS = decltype(bs_c)::value;
    constexpr int TPT = decltype(tpt_c)::value;
    k_p2p<T, NEQ, CONV, BS, TPT><<<nbi, BS, 0, stream>>>(
        nbi, si.bkey[lmax].as<uint32_t>(), sj.bkey[lmax].as<uint32_t>(), si.bfirst[lmax].as<int>(), si.blast.as<int>(), nij.as<int>(), neij.as<int>(), sj.bfirst[lmax].as<int>(),
        sj.blast.as<int>(), si.x.as<float>(), si.y.as<float>(), si.z.as<float>(), si.nx.as<float>(), si.ny.as<float>(),
        si.nz.as<float>(), sj.x.as<float>(), sj.y.as<float>(), sj.z.as<float>(), qs.as<float>(), nxj,
        sj.ny.as<float>(), sj.nz.as<float>(), cc, vs.as<double>(), gxs.as<double>(), gys.as<double>(),
        gzs.as<double>(), kl);
    return FMMX_OK;
  };
  // leaves of at most `small` targets on average: one wave per leaf with
  // lanes per target chosen per pass (TPT 2; measured 36-47% less P2P time
  // than one lane per target in 64/128-lane blocks at 46-61 targets per
  // leaf); FMMX_P2P_SMALL sets the bound
  double small = 96.0;
  if (const char* e = getenv("FMMX_P2P_SMALL")) small = atof(e);
  auto by_bs = [&](auto t, auto nq, auto cv) -> int {
    using one = std::integral_constant<int, 1>;
    if (mean <= small) return go(t, nq, cv, std::integral_constant<int, 64>{}, std::integral_constant<int, 2>{});
    if (mean > 96) return go(t, nq, cv, std::integral_constant<int, 256>{}, one{});
    if (mean > 40) return go(t, nq, cv, std::integral_constant<int, 128>{}, one{});
    return go(t, nq, cv, std::integral_constant<int, 64>{}, one{});
  };
  auto by_conv = [&](auto t, auto nq) -> int {
    return ref ? by_bs(t, nq, std::integral_constant<int, FMMX_CONV_REF>{})
               : by_bs(t, nq, std::integral_constant<int, FMMX_CONV_EXACT>{});
  };
  auto by_neq = [&](auto t) -> int {
    if (neq == FMMX_HELMHOLTZ) return by_conv(t, std::integral_constant<int, FMMX_HELMHOLTZ>{});
    if (neq == FMMX_YUKAWA) return by_conv(t, std::integral_constant<int, FMMX_YUKAWA>{});
    if (neq == FMMX_G) return by_conv(t, std::integral_constant<int, FMMX_G>{});
    if (neq == FMMX_GNI) return by_conv(t, std::integral_constant<int, FMMX_GNI>{});
    return by_conv(t, std::integral_constant<int, FMMX_GNJ>{});
  };
  int rc = fp64 ? by_neq(tag<double>{}) : by_neq(tag<float>{});
  if (rc) return rc;
  FMMX_HIP(hipGetLastError());
  return FMMX_OK;
}

int Engine::run_p2m(int neq) {
  const int nb = sj.nb[lmax];
  if (nb == 0) return FMMX_OK;
  const double* corr = corr_d.as<double>();
  if (p > P_ALL) {
    return dispatch_hi(p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value;
#define FMMX_P2M_HI(YK_, GN_)                                                                                      \
  k_p2m_hi<P, YK_, GN_><<<nb, 64, 0, stream>>>(nb, sj.bkey[lmax].as<uint32_t>(), sj.bfirst[lmax].as<int>(),        \
                                               sj.blast.as<int>(), sj.x.as<float>(), sj.y.as<float>(),             \
                                               sj.z.as<float>(), sj.nx.as<float>(), sj.ny.as<float>(),             \
                                               sj.nz.as<float>(), qs.as<float>(), lmax, xmin, ymin, zmin, rb(lmax), \
                                               mexp[lmax].as<double>(), lam * rb(lmax))
      if (yk) FMMX_P2M_HI(1, 0);
      else if (neq == FMMX_GNJ) FMMX_P2M_HI(0, 1);
      else FMMX_P2M_HI(0, 0);
#undef FMMX_P2M_HI
      FMMX_HIP(hipGetLastError());
      return FMMX_OK;
    });
  }
  return dispatch_p(p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    const double lrb = (hh ? hk : lam) * rb(lmax);
    // two leaves per wave when leaves hold at most about 32 particles
    const bool half = (double)sj.n / nb <= 32.0;
#define FMMX_P2M_HC(T_, YK_, H_, C_, G_, Q_, STRIDE_, OFF_)                                                       \
  k_p2m<P, T_, YK_, H_, C_, G_><<<H_ ? (nb + 1) / 2 : nb, 64, 0, stream>>>(                                       \
      nb, sj.bkey[lmax].as<uint32_t>(), sj.bfirst[lmax].as<int>(), sj.blast.as<int>(), sj.x.as<float>(),         \
      sj.y.as<float>(), sj.z.as<float>(), sj.nx.as<float>(), sj.ny.as<float>(), sj.nz.as<float>(), Q_,            \
      neq == FMMX_GNJ, cfg.convention, lmax, xmin, ymin, zmin, rb(lmax), corr, mexp[lmax].as<T_>(), lrb, STRIDE_, \
      OFF_)
    // Laplace charges in the EXACT convention: the streaming form
#define FMMX_P2M_H(T_, YK_, H_, Q_, STRIDE_, OFF_)                                                 \
  do {                                                                                            \
    if constexpr (YK_ == 0) {                                                                     \
      if (neq != FMMX_GNJ && cfg.convention == FMMX_CONV_EXACT && std::is_same_v<T_, float>)      \
        k_p2m_exact<P, H_><<<H_ ? (nb + 1) / 2 : nb, 64, 0, stream>>>(                            \
            nb, sj.bkey[lmax].as<uint32_t>(), sj.bfirst[lmax].as<int>(), sj.blast.as<int>(),      \
            sj.x.as<float>(), sj.y.as<float>(), sj.z.as<float>(), Q_, lmax, xmin, ymin, zmin,     \
            rb(lmax), (float*)mexp[lmax].as<T_>(), STRIDE_, OFF_);                                \
      else if (neq != FMMX_GNJ && cfg.convention == FMMX_CONV_EXACT)                              \
        FMMX_P2M_HC(T_, 0, H_, FMMX_CONV_EXACT, 0, Q_, STRIDE_, OFF_);                            \
      else                                                                                        \
        FMMX_P2M_HC(T_, 0, H_, -1, -1, Q_, STRIDE_, OFF_);                                        \
    } else {                                                                                      \
      FMMX_P2M_HC(T_, YK_, H_, -1, -1, Q_, STRIDE_, OFF_);                                        \
    }                                                                                             \
  } while (0)
#define FMMX_P2M(T_, YK_, Q_, STRIDE_, OFF_)                 \
  do {                                                       \
    if (half) FMMX_P2M_H(T_, YK_, 1, Q_, STRIDE_, OFF_);     \
    else FMMX_P2M_H(T_, YK_, 0, Q_, STRIDE_, OFF_);          \
  } while (0)
    if (hh) {
      // the multipoles of the real and of the imaginary charges, side by side
      if constexpr (P % 2 == 0 && P >= 4) {
        if (fp64) {
          FMMX_P2M(double, 2, qs.as<float>(), 2 * VS, 0);
          FMMX_P2M(double, 2, qs2.as<float>(), 2 * VS, VS);
        } else {
          FMMX_P2M(float, 2, qs.as<float>(), 2 * VS, 0);
          FMMX_P2M(float, 2, qs2.as<float>(), 2 * VS, VS);
        }
      } else {
        return fail(FMMX_EUNSUPPORTED, "Helmholtz needs p in {4, 6, 8, 10, 12}");
      }
    } else if (fp64) {
      if (yk) FMMX_P2M(double, 1, qs.as<float>(), VS, 0);
      else FMMX_P2M(double, 0, qs.as<float>(), VS, 0);
    } else {
      if (yk) FMMX_P2M(float, 1, qs.as<float>(), VS, 0);
      else FMMX_P2M(float, 0, qs.as<float>(), VS, 0);
    }
#undef FMMX_P2M
#undef FMMX_P2M_H
#undef FMMX_P2M_HC
    FMMX_HIP(hipGetLastError());
    return FMMX_OK;
  });
}

int Engine::run_l2p(int neq, double* const* dst) {
  const int nb = si.nb[lmax];
  if (nb == 0) return FMMX_OK;
  // with dst (v, gx, gy, gz in the caller's order) the results go there
  L2POut lout{};
  if (dst && !hh) lout = L2POut{si.perm.as<uint32_t>(), dst[0], dst[1], dst[2], dst[3]};
  const double* corr = corr_d.as<double>();
  if (p > P_ALL) {
    return dispatch_hi(p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value;
      const double lrb = lam * rb(lmax), cl = yk ? rb(lmax) : c4();
      const int grad = neq == FMMX_G || neq == FMMX_YUKAWA;
#define FMMX_L2P_HI(YK_, C_)                                                                                    \
  k_l2p<P, double, YK_, 0, C_><<<nb, 64, 0, stream>>>(                                                          \
      nb, si.bkey[lmax].as<uint32_t>(), si.bfirst[lmax].as<int>(), si.blast.as<int>(), si.x.as<float>(),      \
      si.y.as<float>(), si.z.as<float>(), si.nx.as<float>(), si.ny.as<float>(), si.nz.as<float>(),             \
      lexp[lmax].as<double>(), neq == FMMX_GNI, grad, cfg.convention, lmax, xmin, ymin, zmin, rb(lmax), corr, \
      cl, vs.as<double>(), gxs.as<double>(), gys.as<double>(), gzs.as<double>(), lrb, P * P, 0, lout)
      if (yk) FMMX_L2P_HI(1, -1);
      else FMMX_L2P_HI(0, FMMX_CONV_EXACT);
#undef FMMX_L2P_HI
      FMMX_HIP(hipGetLastError());
      return FMMX_OK;
    });
  }
  return dispatch_p(p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    // Yukawa locals carry the kernel's constant: phi = sum L~ h_n S_n(u)
    const double lrb = (hh ? hk : lam) * rb(lmax), cl = yk ? rb(lmax) : c4();
    const int grad = neq == FMMX_G || neq == FMMX_YUKAWA;
    // two leaves per wave when leaves hold at most about 32 targets
    const bool half = (double)si.n / nb <= 32.0;
#define FMMX_L2P_HC(T_, YK_, H_, C_, V_, STRIDE_, OFF_)                                                           \
  k_l2p<P, T_, YK_, H_, C_><<<H_ ? (nb + 1) / 2 : nb, 64, 0, stream>>>(                                          \
      nb, si.bkey[lmax].as<uint32_t>(), si.bfirst[lmax].as<int>(), si.blast.as<int>(), si.x.as<float>(),        \
      si.y.as<float>(), si.z.as<float>(), si.nx.as<float>(), si.ny.as<float>(), si.nz.as<float>(),               \
      lexp[lmax].as<T_>(), neq == FMMX_GNI, grad, cfg.convention, lmax, xmin, ymin, zmin, rb(lmax), corr, cl, V_, \
      gxs.as<double>(), gys.as<double>(), gzs.as<double>(), lrb, STRIDE_, OFF_, lout)
    // Laplace: the convention as a template argument
#define FMMX_L2P_H(T_, YK_, H_, V_, STRIDE_, OFF_)                                          \
  do {                                                                                     \
    if constexpr (YK_ != 0) FMMX_L2P_HC(T_, YK_, H_, -1, V_, STRIDE_, OFF_);               \
    else if (cfg.convention == FMMX_CONV_REF)                                              \
      FMMX_L2P_HC(T_, 0, H_, FMMX_CONV_REF, V_, STRIDE_, OFF_);                            \
    else FMMX_L2P_HC(T_, 0, H_, FMMX_CONV_EXACT, V_, STRIDE_, OFF_);                       \
  } while (0)
#define FMMX_L2P(T_, YK_, V_, STRIDE_, OFF_)               \
  do {                                                     \
    if (half) FMMX_L2P_H(T_, YK_, 1, V_, STRIDE_, OFF_);   \
    else FMMX_L2P_H(T_, YK_, 0, V_, STRIDE_, OFF_);        \
  } while (0)
    if (hh) {
      // Re of the potential into vs, Im into gxs
      if constexpr (P % 2 == 0 && P >= 4) {
        if (fp64) {
          FMMX_L2P(double, 2, vs.as<double>(), 2 * P2, 0);
          FMMX_L2P(double, 2, gxs.as<double>(), 2 * P2, P2);
        } else {
          FMMX_L2P(float, 2, vs.as<double>(), 2 * P2, 0);
          FMMX_L2P(float, 2, gxs.as<double>(), 2 * P2, P2);
        }
      } else {
        return fail(FMMX_EUNSUPPORTED, "Helmholtz needs p in {4, 6, 8, 10, 12}");
      }
    } else if (fp64) {
      if (yk) FMMX_L2P(double, 1, vs.as<double>(), P2, 0);
      else FMMX_L2P(double, 0, vs.as<double>(), P2, 0);
    } else {
      if (yk) FMMX_L2P(float, 1, vs.as<double>(), P2, 0);
      else FMMX_L2P(float, 0, vs.as<double>(), P2, 0);
    }
#undef FMMX_L2P
#undef FMMX_L2P_H
#undef FMMX_L2P_HC
    FMMX_HIP(hipGetLastError());
    return FMMX_OK;
  });
}

int Engine::run_m2m(int lev) {
  const int nbp = sj.nb[lev];
  if (nbp == 0) return FMMX_OK;
  if (!fp64 && !hh) {
    return dispatch_p(p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value;
      k_translate_mfma<P, 0><<<(nbp + 15) / 16, 512, 0, stream>>>(
          nbp, sj.cfirst[lev].as<int>(), sj.bkey[lev + 1].as<uint32_t>(), nullptr, nullptr, nullptr,
          mexp[lev + 1].as<float>(), mexp[lev].as<float>(), m2m_ops<float>(lev), 0);
      FMMX_HIP(hipGetLastError());
      return FMMX_OK;
    });
  }
  const dim3 grid((nbp + TR_TB - 1) / TR_TB, 8);
  const size_t es = fp64 ? 8 : 4;
  const long n = (long)nbp * VS * parts();
  int rc = m2m_part.ensure((size_t)8 * n * es);
  if (rc) return rc;
  if (hh) {
    return dispatch_hp(p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value, V = 2 * MLayout<P>::VS;
      if (fp64) {
        k_translate<V, double, 0><<<grid, TR_NT, 0, stream>>>(
            nbp, sj.cfirst[lev].as<int>(), sj.bkey[lev + 1].as<uint32_t>(), nullptr, nullptr, nullptr,
            mexp[lev + 1].as<double>(), m2m_part.as<double>(), m2m_ops<double>(lev));
        k_m2l_reduce<double><<<(int)((n + 255) / 256), 256, 0, stream>>>(n, 8, m2m_part.as<double>(),
                                                                         mexp[lev].as<double>(), 1);
      } else {
        k_translate<V, float, 0><<<grid, TR_NT, 0, stream>>>(
            nbp, sj.cfirst[lev].as<int>(), sj.bkey[lev + 1].as<uint32_t>(), nullptr, nullptr, nullptr,
            mexp[lev + 1].as<float>(), m2m_part.as<float>(), m2m_ops<float>(lev));
        k_m2l_reduce<float><<<(int)((n + 255) / 256), 256, 0, stream>>>(n, 8, m2m_part.as<float>(),
                                                                        mexp[lev].as<float>(), 1);
      }
      FMMX_HIP(hipGetLastError());
      return FMMX_OK;
    });
  }
  auto m2m = [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value, V = MLayout<P>::VS;
    if (fp64) {
      k_translate<V, double, 0><<<grid, TR_NT, 0, stream>>>(
          nbp, sj.cfirst[lev].as<int>(), sj.bkey[lev + 1].as<uint32_t>(), nullptr, nullptr, nullptr,
          mexp[lev + 1].as<double>(), m2m_part.as<double>(), m2m_ops<double>(lev));
      k_m2l_reduce<double><<<(int)((n + 255) / 256), 256, 0, stream>>>(n, 8, m2m_part.as<double>(),
                                                                       mexp[lev].as<double>(), 1);
    } else if constexpr (P <= P_ALL) {
      k_translate<V, float, 0><<<grid, TR_NT, 0, stream>>>(
          nbp, sj.cfirst[lev].as<int>(), sj.bkey[lev + 1].as<uint32_t>(), nullptr, nullptr, nullptr,
          mexp[lev + 1].as<float>(), m2m_part.as<float>(), m2m_ops<float>(lev));
      k_m2l_reduce<float><<<(int)((n + 255) / 256), 256, 0, stream>>>(n, 8, m2m_part.as<float>(),
                                                                      mexp[lev].as<float>(), 1);
    }
    FMMX_HIP(hipGetLastError());
    return FMMX_OK;
  };
  return p > P_ALL ? dispatch_hi(p, m2m) : dispatch_p(p, m2m);
}

int Engine::run_l2l(int lev, bool add) {
  const int nbc = si.nb[lev];
  if (nbc == 0) return FMMX_OK;
  if (!fp64 && !hh) {
    // one wave per 16 children of a class: at most nbc / 16 + 8 waves
    return dispatch_p(p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value;
      k_translate_mfma<P, 1><<<(nbc / 16 + 8 + 3) / 4, 256, 0, stream>>>(
          nbc, nullptr, nullptr, si.parent[lev].as<int>(), cls_cnt[lev].as<int>(), cls_ord[lev].as<uint32_t>(),
          lexp[lev - 1].as<float>(), lexp[lev].as<float>(), l2l_ops<float>(lev), add ? 1 : 0);
      FMMX_HIP(hipGetLastError());
      return FMMX_OK;
    });
  }
  if (add) return fail(FMMX_ESTATE, "L2L add form is FP32 Laplace/Yukawa only");
  const int grid = (nbc + TR_TB - 1) / TR_TB + 8;
  if (hh) {
    return dispatch_hp(p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value, Q = 2 * P * P;
      if (fp64)
        k_translate<Q, double, 1><<<grid, TR_NT, 0, stream>>>(
            nbc, nullptr, si.bkey[lev].as<uint32_t>(), si.parent[lev].as<int>(), cls_cnt[lev].as<int>(),
            cls_ord[lev].as<uint32_t>(), lexp[lev - 1].as<double>(), lexp[lev].as<double>(), l2l_ops<double>(lev));
      else
        k_translate<Q, float, 1><<<grid, TR_NT, 0, stream>>>(
            nbc, nullptr, si.bkey[lev].as<uint32_t>(), si.parent[lev].as<int>(), cls_cnt[lev].as<int>(),
            cls_ord[lev].as<uint32_t>(), lexp[lev - 1].as<float>(), lexp[lev].as<float>(), l2l_ops<float>(lev));
      FMMX_HIP(hipGetLastError());
      return FMMX_OK;
    });
  }
  auto l2l = [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value, Q = P * P;
    if (fp64)
      k_translate<Q, double, 1><<<grid, TR_NT, 0, stream>>>(
          nbc, nullptr, si.bkey[lev].as<uint32_t>(), si.parent[lev].as<int>(), cls_cnt[lev].as<int>(),
          cls_ord[lev].as<uint32_t>(), lexp[lev - 1].as<double>(), lexp[lev].as<double>(), l2l_ops<double>(lev));
    else if constexpr (P <= P_ALL)
      k_translate<Q, float, 1><<<grid, TR_NT, 0, stream>>>(
          nbc, nullptr, si.bkey[lev].as<uint32_t>(), si.parent[lev].as<int>(), cls_cnt[lev].as<int>(),
          cls_ord[lev].as<uint32_t>(), lexp[lev - 1].as<float>(), lexp[lev].as<float>(), l2l_ops<float>(lev));
    FMMX_HIP(hipGetLastError());
    return FMMX_OK;
  };
  return p > P_ALL ? dispatch_hi(p, l2l) : dispatch_p(p, l2l);
}

// Offset splits of the M2L launch at a level (partial sums when > 1).
Engine::M2LVariant Engine::m2l_variant(int lev) const {
  M2LVariant v = tune[lev];
  if (lev != lmax && m2l_all_levels()) {
    // one launch for all levels: the leaf level's shape, this level's splits
    const M2LVariant lf = m2l_variant(lmax);
    v.nt = lf.nt;
    v.pa = lf.pa;
    v.wpg = lf.wpg;
  }
  if (v.nt <= 0) v.nt = 1;
  // default: the FP16 split kernel where it applies (FP32 Laplace engine,
  // EXACT, p 6..12), else LDS staging
  if (v.pa < 0) v.pa = (split_ok && !yk && !hh && !fp64) ? (wide_ok ? 5 : 4) : 0;
  if (v.wpg <= 0) v.wpg = 8;
  if (const char* e = getenv("FMMX_M2L_NT")) v.nt = atoi(e) == 4 ? 4 : atoi(e) == 2 ? 2 : 1;
  if (const char* e = getenv("FMMX_M2L_PA")) v.pa = std::max(0, std::min(5, atoi(e)));
  if (const char* e = getenv("FMMX_M2L_WPG")) v.wpg = atoi(e) == 4 ? 4 : atoi(e) == 16 ? 16 : 8;
  if (v.pa >= 3) {
    // FP16 split kernel: Laplace FP32 engine, p 6..12; two operator buffers
    // with one N-tile and 8 waves where they fit LDS
    if (!split_ok || yk || hh || fp64) v.pa = 0;
    else if (v.pa == 5 && !wide_ok) v.pa = 4;
    if (v.pa == 4 && !split_db_fits(p)) v.pa = 3;
    // 32 targets per wave: 8 waves (two buffers) or 4 (one); 64 targets per
    // wave (nt 4): 4 waves, two buffers
    if (v.pa == 5) {
      if (v.nt == 4) v.wpg = 4;
      else v.nt = 2, v.wpg = v.wpg == 4 ? 4 : 8;
    }
    if (v.pa != 5 && v.nt == 4) v.nt = 2;
    if (v.pa == 4) v.wpg = v.nt == 2 ? 4 : 8;  // 8 waves of one N-tile or 4 of two
    if (v.pa == 3 && v.nt == 2) v.wpg = 4;
    if (v.wpg == 16) v.wpg = 8;
  }
  if (v.pa == 2 && (p > 11 || v.nt != 1)) v.pa = 0;  // two operator buffers fit LDS up to p = 11
  if (v.wpg == 16 && v.pa != 2) v.wpg = 8;
  // reflection-canonical operators are signed while staged: no register prefetch form
  if (v.pa == 1 && canon && !yk) v.pa = 0;
  if (v.pa < 3 && v.wpg == 16 && v.pa != 2) v.wpg = 8;
  if (const char* e = getenv("FMMX_M2L_SPLIT")) v.nsplit = atoi(e);
  if (const char* e = getenv("FMMX_M2L_SPLITS")) {
    // per level, comma-separated from level 2
    for (int l = 2; e && *e; l++) {
      if (l == lev) v.nsplit = atoi(e);
      e = strchr(e, ',');
      if (e) e++;
    }
  }
  if (v.nsplit <= 0) {
    // enough workgroups to cover the chip: split the 189 offsets when tiles are few
    if (fp64) {
      v.nsplit = 1024 / ((si.nb[lev] + M2L_TN - 1) / M2L_TN + 8);
    } else {
      int cmax = 0;
      for (int c = 0; c < 8; c++) cmax = std::max(cmax, cls_cnt_h[lev][c]);
      const int tpc = std::max(1, (cmax + 16 * v.wpg * v.nt - 1) / (16 * v.wpg * v.nt));
      v.nsplit = 1024 / (8 * tpc);
    }
  }
  v.nsplit = std::max(1, std::min(16, v.nsplit));
  return v;
}

// phase 0: the M2L launch and (with splits) the reduction into the locals;
// 1: the launch alone, partial sums only (needs m2l_splits(lev) > 1);
// 2: the reduction alone, adding the partials of phase 1 to the locals
// FP32 MFMA M2L: the launch parameters of one level, its workgroups from wg0
// (grid = 8 classes x tiles x splits; a block of tiles per XCD once every XCD
// gets at least one tile)
M2LLev Engine::m2l_desc(int lev, int nt, int wpg, int nsplit, float* part, int init, int wg0) {
  M2LLev d{};
  const int MT = (P2 + 15) / 16, KS = (P2 + 3) / 4;
  const size_t FR = (size_t)KS * MT * 64;
  const size_t opoff = (yk || m2l_per_level) ? (size_t)(lev - 2) * NOFF * FR : 0;
  // reflection-canonical operators (Laplace, EXACT) are signed while staged
  const bool use_canon = canon && !yk;
  int cmax = 0;
  for (int c = 0; c < 8; c++) cmax = std::max(cmax, cls_cnt_h[lev][c]);
  const int tpc = std::max(1, (cmax + 16 * wpg * nt - 1) / (16 * wpg * nt));
  int region = tpc >= 8 ? 1 : 0;
  if (const char* e = getenv("FMMX_M2L_REGION")) region = region && atoi(e) != 0;
  d.cls_cnt = cls_cnt[lev].as<int>();
  d.ord = cls_ord[lev].as<uint32_t>();
  d.ikey = si.bkey[lev].as<uint32_t>();
  d.dense_j = sj.dense[lev].as<int>();
  d.M = mexp[lev].as<float>();
  d.L = lexp[lev].as<float>();
  d.part = part;
  d.opF = (yk ? yop_m2l : op_m2l).as<float>() + opoff;
  // FP16 split form: per-level operator sets in REF, one set otherwise
  const size_t frh = split_ok ? split_frag_count(p) : 0;
  d.Mh = mexp_h[lev].as<uint4>();
  d.opH = op_m2l_h.as<uint4>() + (m2l_per_level ? (size_t)(lev - 2) * NOFF * frh : 0);
  d.mmax = mmax_buf.as<unsigned>() + lev;
  d.sa_e = sa_e;
  d.offtab = (use_canon ? offtab_c : offtab).as<int4>();
  d.npairs = pair_ctr.as<unsigned long long>() + 2 + lev;
  d.lev = lev;
  d.nb = si.nb[lev];
  d.nbj = sj.nb[lev];
  d.nsplit = nsplit;
  d.init = init;
  d.region = region;
  d.gsh = group_shift(lev);
  d.wg0 = wg0;
  d.nwg = region ? 64 * ((tpc + 7) / 8) * nsplit : 8 * tpc * nsplit;
  return d;
}

// the k_m2l instance of a launch shape (N-tiles per wave, operator staging,
// waves per workgroup)
template <int P>
void launch_m2l(const M2LLaunch& A, int nwg, int nt, int pa, int wpg, hipStream_t stream) {
#define FMMX_M2L_LAUNCH(NT_, PA_, W_) k_m2l<P, NT_, PA_, W_><<<nwg, 64 * W_, 0, stream>>>(A)
  if (pa == 2 && P > 11) pa = 0;  // two operator buffers fit the 160 KB of LDS up to p = 11
  if (pa == 2) {
    if constexpr (P <= 11) {
      if (wpg == 16) FMMX_M2L_LAUNCH(1, 2, 16);
      else FMMX_M2L_LAUNCH(1, 2, 8);
    }
  } else if (wpg == 8) {
    if (nt == 1 && !pa) FMMX_M2L_LAUNCH(1, 0, 8);
    else if (nt == 1) FMMX_M2L_LAUNCH(1, 1, 8);
    else FMMX_M2L_LAUNCH(2, 0, 8);
  } else {
    if (nt == 1 && !pa) FMMX_M2L_LAUNCH(1, 0, 4);
    else if (nt == 1) FMMX_M2L_LAUNCH(1, 1, 4);
    else FMMX_M2L_LAUNCH(2, 0, 4);
  }
#undef FMMX_M2L_LAUNCH
}

// FP32 MFMA M2L of all levels l0..lmax in one launch, in the leaf level's
// launch shape with each level's offset splits: levels with splits write
// partial sums (reduced afterwards in fixed order), the others write their
// locals; the L2L chain then adds the parents' locals level by level.
bool Engine::m2l_split(int lev) const { return !fp64 && !hh && !yk && m2l_variant(lev).pa >= 3; }

// The FP16 split kernel's inputs for levels [l0, l1]: the multipoles of every
// level split into hi/lo pairs scaled by 2^(14 - exponent of the level's
// largest magnitude), one pair of launches for all levels
int Engine::run_msplit(int l0, int l1, bool wide) {
  MSplitLevels S{};
  S.off[0] = 0;
  S.wide = wide ? 1 : 0;
  int rc = mmax_buf.ensure((MAXL + 1) * 4);
  if (rc) return rc;
  for (int lev = l0; lev <= l1; lev++) {
    const int nb = sj.nb[lev];
    if (nb == 0 || si.nb[lev] == 0) continue;
    S.M[S.nl] = mexp[lev].as<float>();
    S.Mh[S.nl] = mexp_h[lev].as<uint4>();
    S.mmax[S.nl] = mmax_buf.as<unsigned>() + lev;
    S.off[S.nl + 1] = S.off[S.nl] + nb + 1;  // and the zero box
    S.nl++;
  }
  if (S.nl == 0) return FMMX_OK;
  return launch_msplit(p, S, stream);
}

// FP32 M2L of all levels l0..lmax in one launch, in the leaf level's launch
// shape with each level's offset splits: levels with splits write partial
// sums (reduced afterwards in fixed order), the others write their locals;
// the L2L chain then adds the parents' locals level by level.  With the FP16
// split shape (staging 3, 4) the multipoles are split first.
int Engine::run_m2l_all(int l0) {
  const M2LVariant leaf = m2l_variant(lmax);
  const bool split = leaf.pa >= 3;
  M2LLaunch A{};
  size_t need = 0;
  for (int lev = l0; lev <= lmax; lev++)
    if (si.nb[lev] > 0 && sj.nb[lev] > 0 && m2l_splits(lev) > 1) need += (size_t)m2l_splits(lev) * si.nb[lev] * P2;
  int rc = m2l_part.ensure(std::max<size_t>(need, 1) * 4);
  if (rc) return rc;
  if (split && (rc = run_msplit(l0, lmax, leaf.pa == 5))) return rc;
  size_t off = 0;
  int wg = 0;
  for (int lev = l0; lev <= lmax; lev++) {
    const int nb = si.nb[lev];
    if (nb == 0) continue;
    const long nl = (long)nb * P2;
    if (sj.nb[lev] == 0) {
      k_zero_t<float><<<(int)((nl + 255) / 256), 256, 0, stream>>>(nl, lexp[lev].as<float>());
      continue;
    }
    const int ns = m2l_splits(lev);
    A.lv[A.nlv] = m2l_desc(lev, leaf.nt, leaf.wpg, ns, ns > 1 ? m2l_part.as<float>() + off : nullptr, 1, wg);
    if (leaf.pa == 4) A.lv[A.nlv].offtab = offtab_s.as<int4>();
    if (leaf.pa == 5) {
      A.lv[A.nlv].offtab = offtab.as<int4>();
      A.lv[A.nlv].opH = op_m2l_w.as<uint4>();
    }
    wg += A.lv[A.nlv].nwg;
    A.nlv++;
    if (ns > 1) off += (size_t)ns * nl;
  }
  if (A.nlv == 0) return FMMX_OK;
  return dispatch_p(p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    if (per_phase) FMMX_HIP(hipEventRecord(ev[8], stream));
    if (split) {
      int r2 = leaf.pa == 5 ? launch_m2l_wide(p, A, wg, leaf.wpg, leaf.nt, stream)
                            : launch_m2l_split(p, A, wg, leaf.nt, leaf.wpg, leaf.pa == 4, stream);
      if (r2) return r2;
    } else {
      launch_m2l<P>(A, wg, leaf.nt, leaf.pa, leaf.wpg, stream);
    }
    if (per_phase) {
      FMMX_HIP(hipEventRecord(ev[9], stream));
      m2l_timed = true;
    }
    // the partial sums of every level with splits, one launch
    ReduceLevels R{};
    R.off[0] = 0;
    for (int k = 0; k < A.nlv; k++) {
      const M2LLev& d = A.lv[k];
      if (d.nsplit < 2) continue;
      R.part[R.nl] = d.part;
      R.L[R.nl] = d.L;
      R.nsplit[R.nl] = d.nsplit;
      R.off[R.nl + 1] = R.off[R.nl] + (long)d.nb * P2;
      R.nl++;
    }
    if (R.nl > 0)
      k_m2l_reduce_levels<<<(int)((R.off[R.nl] + 255) / 256), 256, 0, stream>>>(R);
    FMMX_HIP(hipGetLastError());
    return FMMX_OK;
  });
}

int Engine::run_m2l(int lev, bool init, int phase) {
  const int nb = si.nb[lev];
  if (nb == 0) return FMMX_OK;
  const long nl = (long)nb * P2 * parts();
  if (sj.nb[lev] == 0) {
    if (init && phase != 1) {
      if (fp64) k_zero_t<double><<<(int)((nl + 255) / 256), 256, 0, stream>>>(nl, lexp[lev].as<double>());
      else k_zero_t<float><<<(int)((nl + 255) / 256), 256, 0, stream>>>(nl, lexp[lev].as<float>());
    }
    return FMMX_OK;
  }
  const int MT = (P2 + 15) / 16, KS = (P2 + 3) / 4;
  // FP32: MFMA fragment order; FP64: k-major (transposed) operators
  const size_t FR = fp64 ? (size_t)P2 * P2 : (size_t)KS * MT * 64;
  const size_t opoff = (yk || m2l_per_level) ? (size_t)(lev - 2) * NOFF * FR : 0;
  DevBuf& opm2l = yk ? yop_m2l : op_m2l;
  const int tiles = (nb + M2L_TN - 1) / M2L_TN + 8;
  // enough workgroups to cover the chip: split the 189 offsets when tiles are few
  int nsplit = m2l_splits(lev);
  const size_t es = fp64 ? 8 : 4;
  if (phase == 2) {
    if (nsplit < 2) return fail(FMMX_ESTATE, "M2L reduction without partial sums");
    if (fp64)
      k_m2l_reduce<double><<<(int)((nl + 255) / 256), 256, 0, stream>>>(nl, nsplit, part_buf->as<double>(),
                                                                        lexp[lev].as<double>(), init ? 1 : 0);
    else
      k_m2l_reduce<float><<<(int)((nl + 255) / 256), 256, 0, stream>>>(nl, nsplit, part_buf->as<float>(),
                                                                       lexp[lev].as<float>(), init ? 1 : 0);
    FMMX_HIP(hipGetLastError());
    return FMMX_OK;
  }
  if (phase == 1 && nsplit < 2) return fail(FMMX_ESTATE, "M2L partial sums need offset splits");
  if (nsplit > 1) {
    int rc = part_buf->ensure((size_t)nsplit * nl * es);
    if (rc) return rc;
  }
  int rc2 = FMMX_OK;
  const int gsh = group_shift(lev);
  if (hh) {
    // split complex expansions: D = 2 p^2, k-major operators per level, on the
    // vector units in both precisions
    const size_t hoff = (size_t)(lev - 2) * NOFF * 4 * P2 * P2;
    return dispatch_hp(p, [&](auto pc) -> int {
      constexpr int P = decltype(pc)::value;
      if (fp64)
        k_m2l_vec<P, double, 2><<<tiles * nsplit, 256, 0, stream>>>(
            lev, nb, nsplit, cls_cnt[lev].as<int>(), cls_ord[lev].as<uint32_t>(), si.bkey[lev].as<uint32_t>(),
            sj.dense[lev].as<int>(), mexp[lev].as<double>(), lexp[lev].as<double>(), part_buf->as<double>(),
            hop_m2l.as<double>() + hoff, offtab.as<int4>(), init ? 1 : 0, gsh, pair_ctr.as<unsigned long long>() + 2 + lev);
      else {
        // MFMA, 8 waves of 16 targets per workgroup, operator chunks in LDS
        constexpr int D = 2 * P * P, FRH = (D / 4) * ((D + 15) / 16) * 64;
        int cmax = 0;
        for (int c = 0; c < 8; c++) cmax = std::max(cmax, cls_cnt_h[lev][c]);
        const int tpc = std::max(1, (cmax + 127) / 128);
        const int region = tpc >= 8 ? 1 : 0;
        const int nwg = region ? 64 * ((tpc + 7) / 8) * nsplit : 8 * tpc * nsplit;
        k_m2l_hh<P, 8><<<nwg, 512, 0, stream>>>(
            lev, nb, nsplit, cls_cnt[lev].as<int>(), cls_ord[lev].as<uint32_t>(), si.bkey[lev].as<uint32_t>(),
            sj.dense[lev].as<int>(), mexp[lev].as<float>(), lexp[lev].as<float>(), part_buf->as<float>(),
            hop_m2l.as<float>() + (size_t)(lev - 2) * NOFF * FRH, offtab.as<int4>(), init ? 1 : 0, region, gsh,
            pair_ctr.as<unsigned long long>() + 2 + lev);
      }
      if (nsplit > 1 && phase == 0) {
        if (fp64)
          k_m2l_reduce<double><<<(int)((nl + 255) / 256), 256, 0, stream>>>(nl, nsplit, part_buf->as<double>(),
                                                                            lexp[lev].as<double>(), init ? 1 : 0);
        else
          k_m2l_reduce<float><<<(int)((nl + 255) / 256), 256, 0, stream>>>(nl, nsplit, part_buf->as<float>(),
                                                                           lexp[lev].as<float>(), init ? 1 : 0);
      }
      FMMX_HIP(hipGetLastError());
      return FMMX_OK;
    });
  }
  auto m2l = [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    if (fp64) {
      const bool tmark = per_phase && lev == lmax;
      if (tmark) FMMX_HIP(hipEventRecord(ev[8], stream));
      k_m2l_vec<P, double><<<tiles * nsplit, 256, 0, stream>>>(
          lev, nb, nsplit, cls_cnt[lev].as<int>(), cls_ord[lev].as<uint32_t>(), si.bkey[lev].as<uint32_t>(),
          sj.dense[lev].as<int>(), mexp[lev].as<double>(), lexp[lev].as<double>(), part_buf->as<double>(),
          opm2l.as<double>() + opoff, offtab.as<int4>(), init ? 1 : 0, gsh, pair_ctr.as<unsigned long long>() + 2 + lev);
      if (tmark) {
        FMMX_HIP(hipEventRecord(ev[9], stream));
        m2l_timed = true;
      }
      if (nsplit > 1 && phase == 0)
        k_m2l_reduce<double><<<(int)((nl + 255) / 256), 256, 0, stream>>>(nl, nsplit, part_buf->as<double>(),
                                                                          lexp[lev].as<double>(), init ? 1 : 0);
    } else if constexpr (P <= P_ALL) {
      const M2LVariant var = m2l_variant(lev);
      M2LLaunch A{};
      A.nlv = 1;
      A.lv[0] = m2l_desc(lev, var.nt, var.wpg, nsplit, part_buf->as<float>(), init ? 1 : 0, 0);
      if (var.pa == 4) A.lv[0].offtab = offtab_s.as<int4>();
      if (var.pa == 5) {
        A.lv[0].offtab = offtab.as<int4>();
        A.lv[0].opH = op_m2l_w.as<uint4>();
      }
      if (var.pa >= 3) {
        int r2 = run_msplit(lev, lev, var.pa == 5);
        if (r2) return r2;
      }
      const bool tmark = per_phase && lev == lmax;
      if (tmark) FMMX_HIP(hipEventRecord(ev[8], stream));
      if (var.pa >= 3) {
        int r2 = var.pa == 5 ? launch_m2l_wide(p, A, A.lv[0].nwg, var.wpg, var.nt, stream)
                             : launch_m2l_split(p, A, A.lv[0].nwg, var.nt, var.wpg, var.pa == 4, stream);
        if (r2) return r2;
      } else {
        launch_m2l<P>(A, A.lv[0].nwg, var.nt, var.pa, var.wpg, stream);
      }
      if (tmark) {
        FMMX_HIP(hipEventRecord(ev[9], stream));
        m2l_timed = true;
      }
      if (nsplit > 1 && phase == 0)
        k_m2l_reduce<float><<<(int)((nl + 255) / 256), 256, 0, stream>>>(nl, nsplit, part_buf->as<float>(),
                                                                         lexp[lev].as<float>(), init ? 1 : 0);
    }
    FMMX_HIP(hipGetLastError());
    return FMMX_OK;
  };
  return p > P_ALL ? dispatch_hi(p, m2l) : dispatch_p(p, m2l);
}

int Engine::run_direct(int neq, int ni, const float* xi, const float* yi, const float* zi, const float* nxi,
                       const float* nyi, const float* nzi, int nj, const float* xj, const float* yj, const float* zj,
                       const float* nxj, const float* nyj, const float* nzj, const float* qj, double* v, double* gx,
                       double* gy, double* gz) {
  if (ni == 0) return FMMX_OK;
  long nz4 = ni;
  k_zero_d<<<(int)((nz4 + 255) / 256), 256, 0, stream>>>(nz4, v);
  if (neq == FMMX_HELMHOLTZ) k_zero_d<<<(int)((nz4 + 255) / 256), 256, 0, stream>>>(nz4, gx);  // Im part
  if (neq == FMMX_G || neq == FMMX_YUKAWA) {
    k_zero_d<<<(int)((nz4 + 255) / 256), 256, 0, stream>>>(nz4, gx);
    k_zero_d<<<(int)((nz4 + 255) / 256), 256, 0, stream>>>(nz4, gy);
    k_zero_d<<<(int)((nz4 + 255) / 256), 256, 0, stream>>>(nz4, gz);
  }
  if (nj == 0) return FMMX_OK;
  const int tb = (ni + 63) / 64;
  int ns = std::max(1, std::min(1024, 2048 / tb));
  int slice = (nj + ns - 1) / ns;
  slice = ((slice + 63) / 64) * 64;
  ns = (nj + slice - 1) / slice;
  dim3 grid(tb, ns);
  const bool ref = cfg.convention == FMMX_CONV_REF;
  const double cc = neq == FMMX_YUKAWA ? c_yukawa() : c4();
  // Helmholtz: nxj carries the imaginary charges, gx receives Im of the potential
  const double kl = neq == FMMX_HELMHOLTZ ? hk : lam;
#define FMMX_DIR(NEQ, CONV)                                                                                      \
  k_direct<NEQ, CONV><<<grid, 64, 0, stream>>>(ni, xi, yi, zi, nxi, nyi, nzi, nj, xj, yj, zj, qj, nxj, nyj, nzj, \
                                               slice, cc, v, gx, gy, gz, kl)
  if (neq == FMMX_HELMHOLTZ) {
    if (ref) FMMX_DIR(FMMX_HELMHOLTZ, FMMX_CONV_REF);
    else FMMX_DIR(FMMX_HELMHOLTZ, FMMX_CONV_EXACT);
  } else if (neq == FMMX_YUKAWA) {
    if (ref) FMMX_DIR(FMMX_YUKAWA, FMMX_CONV_REF);
    else FMMX_DIR(FMMX_YUKAWA, FMMX_CONV_EXACT);
  } else if (ref) {
    if (neq == FMMX_G) FMMX_DIR(FMMX_G, FMMX_CONV_REF);
    else if (neq == FMMX_GNI) FMMX_DIR(FMMX_GNI, FMMX_CONV_REF);
    else FMMX_DIR(FMMX_GNJ, FMMX_CONV_REF);
  } else {
    if (neq == FMMX_G) FMMX_DIR(FMMX_G, FMMX_CONV_EXACT);
    else if (neq == FMMX_GNI) FMMX_DIR(FMMX_GNI, FMMX_CONV_EXACT);
    else FMMX_DIR(FMMX_GNJ, FMMX_CONV_EXACT);
  }
#undef FMMX_DIR
  FMMX_HIP(hipGetLastError());
  return FMMX_OK;
}

}  // namespace fmmx
