// projection.h -- host helpers for translation operators obtained by
// projecting fields onto spherical harmonics: associated Legendre functions
// with the engine's normalisation, Gauss-Legendre nodes, the harmonics of all
// orders at a point, and the projection itself (Gauss-Legendre in cos t,
// trapezoid in f).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "fmmx_internal.h"

namespace fmmx {
namespace proj {

inline int sl(int n, int m) { return n * (n + 1) / 2 + m; }
inline int ci(int n, int m) { return n * n + n + m; }  // all m = -n..n

// fac(n,m) P_n^m(c), 0 <= m <= n < p <= MAXP, Condon-Shortley phase
inline void legendre_fac(int p, double c, double* out) {
  const double s = std::sqrt(std::max(0.0, (1 - c) * (1 + c)));
  double P[MAXP * (MAXP + 1) / 2];
  double pmm = 1;
  for (int m = 0; m < p; m++) {
    if (m > 0) pmm *= -(2 * m - 1) * s;
    P[sl(m, m)] = pmm;
    if (m + 1 < p) P[sl(m + 1, m)] = c * (2 * m + 1) * pmm;
    for (int n = m + 2; n < p; n++)
      P[sl(n, m)] = ((2 * n - 1) * c * P[sl(n - 1, m)] - (n + m - 1) * P[sl(n - 2, m)]) / (n - m);
  }
  for (int n = 0; n < p; n++) {
    double f = 1;  // (n-m)!/(n+m)!
    for (int m = 0; m <= n; m++) {
      if (m > 0) f /= (double)(n + m) * (n - m + 1);
      out[sl(n, m)] = std::sqrt(f) * P[sl(n, m)];
    }
  }
}

inline void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1, p1 = z;
      for (int k = 2; k <= n; k++) {
        double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1, p1 = z;
    for (int k = 2; k <= n; k++) {
      double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1);
    x[i] = z;
    w[i] = 2 / ((1 - z * z) * dp * dp);
  }
}

// Project fields sampled on the sphere |s| = rad around the new centre onto
// fac P_j^k e^{-i k f}: out[(j, k)][col] = sum_q w_q f_col(q) conj(Y_j^k(s_q))
// for k = -kmax..kmax (all) or 0..kmax.  field(point, col_values) fills the
// ncol complex values at a point (relative to the new centre).
template <class Field>
void project(int p, int ncol, bool all_k, int nth, int nph, double rad, Field field, std::vector<cplx>& out) {
  std::vector<double> ct, wt;
  gauss_legendre(nth, ct, wt);
  const int nk = all_k ? 2 * p - 1 : p, k0 = all_k ? -(p - 1) : 0;
  std::vector<cplx> F((size_t)nk * ncol), f(ncol), eik(nph * nk);
  for (int l = 0; l < nph; l++) {
    const double ph = 2 * M_PI * l / nph;
    for (int k = 0; k < nk; k++) eik[l * nk + k] = std::polar(2 * M_PI / nph, -(k + k0) * ph);
  }
  out.assign((size_t)p * p * ncol, cplx(0));
  std::vector<double> Pt(p * (p + 1) / 2);
  for (int i = 0; i < nth; i++) {
    const double c = ct[i], s = std::sqrt(std::max(0.0, 1 - c * c));
    std::fill(F.begin(), F.end(), cplx(0));
    for (int l = 0; l < nph; l++) {
      const double ph = 2 * M_PI * l / nph;
      field(rad * s * std::cos(ph), rad * s * std::sin(ph), rad * c, f.data());
      for (int k = 0; k < nk; k++) {
        const cplx e = eik[l * nk + k];
        cplx* Fk = &F[(size_t)k * ncol];
        for (int col = 0; col < ncol; col++) Fk[col] += f[col] * e;
      }
    }
    legendre_fac(p, c, Pt.data());
    for (int j = 0; j < p; j++)
      for (int k = -j; k <= j; k++) {
        if (k - k0 < 0 || k - k0 >= nk) continue;
        const double y = wt[i] * Pt[sl(j, std::abs(k))];  // conj(Y_j^k) = fac P e^{-ikf}
        const cplx* Fk = &F[(size_t)(k - k0) * ncol];
        cplx* o = &out[(size_t)ci(j, k) * ncol];
        for (int col = 0; col < ncol; col++) o[col] += y * Fk[col];
      }
  }
}

// Y_n^m(z) for all m at a point: fac P_n^|m|(cos t) e^{i m f}
inline void harmonics_all(int p, double x, double y, double z, std::vector<double>& Pf, cplx* Y) {
  const double r = std::sqrt(x * x + y * y + z * z);
  legendre_fac(p, r > 0 ? z / r : 1.0, Pf.data());
  const double rxy = std::sqrt(x * x + y * y);
  const cplx e1 = rxy > 0 ? cplx(x / rxy, y / rxy) : cplx(1, 0);  // e^{i f}
  cplx em(1, 0);
  for (int m = 0; m < p; m++, em *= e1)
    for (int n = m; n < p; n++) {
      const cplx v = Pf[sl(n, m)] * em;
      Y[ci(n, m)] = v;
      Y[ci(n, -m)] = std::conj(v);
    }
}

}  // namespace proj
}  // namespace fmmx
