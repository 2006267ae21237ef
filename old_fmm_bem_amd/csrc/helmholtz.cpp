// helmholtz.cpp -- dense translation operators of the oscillatory Helmholtz
// FMM: kernel exp(i k r) / (4 pi r), complex charges.  The reference has no
// such kernel (its kernel/Helmholtz_*.cxx files are the screened kernel,
// neq 10); this one is new and pinned by direct sums.
//
// Addition theorem (Y_n^m = fac(n,m) P_n^|m|(cos t) e^{i m f} as everywhere
// in the engine, so sum_m Y_n^m(a) conj(Y_n^m(b)) = P_n(cos g)):
//   exp(i k |x-y|) / |x-y| = i k sum_n (2n+1) j_n(k r<) h_n(k r>) P_n(cos g).
// With box size rb, kappa = k rb, u = r / rb and the scaled radial functions
// j^_n(x) = (2n+1)!! j_n(x) / x^n and h~_n(x) = h_n(x) x^(n+1):
//   multipole  phi(x) = 1/(4 pi rb) sum M_n^m O_n^m(u),
//              O_n^m(u) = i h~_n(kappa |u|) / (2n-1)!! |u|^-(n+1) Y_n^m,
//              M_n^m = sum_j q_j j^_n(k rho_j) (rho_j/rb)^n conj(Y_n^m(y_j - c)),
//   local      phi(x) = 1/(4 pi rb) sum L_n^m j^_n(kappa |u|) |u|^n Y_n^m,
// which reduce to the Laplace engine's scaled expansions as k -> 0, so P2M
// and L2P are the Laplace harmonics times the per-degree weight j^_n.
//
// Complex charges q = qr + i qi give multipoles M = R + i I where R and I
// are the multipoles of the real charge sets qr and qi: each satisfies
// R_n^-m = conj(R_n^m) and is stored in the Laplace engine's real slots.
// Locals are split the same way into the expansions of Re(phi) and Im(phi).  Every operator below maps such a split pair (2 p^2 reals) to a
// split pair; it is the exact re-expansion of each basis function, obtained
// by projection onto the harmonics on a sphere around the new centre, as for
// the screened kernel (yukawa.cpp).
#include "helmholtz.h"

#include <cmath>
#include <complex>
#include <vector>

#include "projection.h"

namespace fmmx {

void jhat_all(double x, int N, double* out) {
  const double h = -0.5 * x * x;
  auto series = [&](int n) {
    double t = 1, s = 1, tmax = 1;
    for (int k = 1; k < 4000; k++) {
      t *= h / (k * (2.0 * n + 2.0 * k + 1.0));
      s += t;
      tmax = std::max(tmax, std::fabs(t));
      if (std::fabs(t) < 1e-18 * tmax) break;
    }
    return s;
  };
  double a = series(N + 1), b = series(N);
  out[N] = b;
  // j^_{n-1} = j^_n - x^2 j^_{n+1} / ((2n+1)(2n+3)), stable downwards
  for (int n = N; n >= 1; n--) {
    const double c = b - x * x * a / ((2.0 * n + 1.0) * (2.0 * n + 3.0));
    a = b;
    b = c;
    out[n - 1] = c;
  }
}

void htil_all(double x, int N, cplx* out) {
  const cplx e = std::polar(1.0, x);
  out[0] = cplx(0, -1) * e;  // h_0 = -i e^{ix} / x
  if (N >= 1) out[1] = -e * cplx(x, 1);
  for (int n = 1; n < N; n++) out[n + 1] = (2.0 * n + 1.0) * out[n] - x * x * out[n - 1];
}

namespace {

using proj::ci;
using proj::harmonics_all;
constexpr int NTH = 32, NPH = 48;  // quadrature on the projection spheres
constexpr double RAD = 0.5;        // local projections: radius in box units
constexpr double RBIG = 1.5;       // multipole projection (M2M): parent box units

double dfact(int n) {  // (2n-1)!!, (-1)!! = 1
  double f = 1;
  for (int k = 3; k <= 2 * n - 1; k += 2) f *= k;
  return f;
}

// complex operator C on all-m coefficients (row (j,k) at ci(j,k), column
// (n,m) at ci(n,m), row-major) -> real operator on split pairs
void split_real(int p, const std::vector<cplx>& C, double* A) {
  const int P2 = p * p, D = 2 * P2;
  std::fill(A, A + (size_t)D * D, 0.0);
  const cplx I(0, 1);
  for (int t = 0; t < 2; t++)
    for (int n = 0; n < p; n++)
      for (int m = 0; m <= n; m++)
        for (int im = 0; im < (m > 0 ? 2 : 1); im++) {
          const int col = t * P2 + (im ? slot_im(p, n, m) : slot_re(n, m));
          // unit real slot -> X_n^m = a, X_n^-m = conj(a) (Y_n^-m = conj(Y_n^m)
          // in this normalisation); part t = 1 enters the multipole times i
          cplx a = im ? I : cplx(1), am = std::conj(a);
          if (t == 1) {
            a *= I;
            am *= I;
          }
          for (int j = 0; j < p; j++)
            for (int k = 0; k <= j; k++) {
              const cplx* rp = &C[(size_t)ci(j, k) * P2];
              const cplx* rm = &C[(size_t)ci(j, -k) * P2];
              cplx lp = rp[ci(n, m)] * a, lm = rm[ci(n, m)] * a;
              if (m > 0) {
                lp += rp[ci(n, -m)] * am;
                lm += rm[ci(n, -m)] * am;
              }
              // Re(phi) and Im(phi) parts: (L_j^k +- conj(L_j^-k)) / (2, 2i)
              const cplx ca = 0.5 * (lp + std::conj(lm)), cb = (lp - std::conj(lm)) / (2.0 * I);
              A[(size_t)slot_re(j, k) * D + col] = ca.real();
              A[(size_t)(P2 + slot_re(j, k)) * D + col] = cb.real();
              if (k > 0) {
                A[(size_t)slot_im(p, j, k) * D + col] = ca.imag();
                A[(size_t)(P2 + slot_im(p, j, k)) * D + col] = cb.imag();
              }
            }
        }
}

}  // namespace

void hh_op_m2l(int p, double kappa, int dx, int dy, int dz, double* A) {
  const int P2 = p * p;
  std::vector<double> jr(p + 1), Pf(p * (p + 1) / 2), df(p);
  std::vector<cplx> ht(p + 1), Y(P2);
  for (int n = 0; n < p; n++) df[n] = dfact(n);
  jhat_all(kappa * RAD, p, jr.data());
  // columns: the field O_n^m(offset + s) of each multipole basis function
  auto field = [&](double sx, double sy, double sz, cplx* f) {
    const double zx = dx + sx, zy = dy + sy, zz = dz + sz, r = std::sqrt(zx * zx + zy * zy + zz * zz);
    htil_all(kappa * r, p, ht.data());
    harmonics_all(p, zx, zy, zz, Pf, Y.data());
    double rn = r;  // r^(n+1)
    for (int n = 0; n < p; n++, rn *= r) {
      const cplx rad = cplx(0, 1) * ht[n] / (df[n] * rn);
      for (int m = -n; m <= n; m++) f[ci(n, m)] = rad * Y[ci(n, m)];
    }
  };
  std::vector<cplx> C;
  proj::project(p, P2, true, NTH, NPH, RAD, field, C);
  double aj = 1;  // RAD^j
  for (int j = 0; j < p; j++, aj *= RAD) {
    const double s = (2 * j + 1) / (4 * M_PI) / (jr[j] * aj);
    for (int k = -j; k <= j; k++)
      for (int col = 0; col < P2; col++) C[(size_t)ci(j, k) * P2 + col] *= s;
  }
  split_real(p, C, A);
}

void hh_op_m2m_l2l(int p, double kc, int oct, double* Am2m, double* Al2l) {
  const int P2 = p * p;
  // child centre relative to the parent centre in child box units (octant
  // bits x = 2, y = 1, z = 4, the fmm/boxn.cxx bit order)
  const double cx = ((oct >> 1) & 1) ? 0.5 : -0.5, cy = (oct & 1) ? 0.5 : -0.5, cz = ((oct >> 2) & 1) ? 0.5 : -0.5;
  const double kp = 2 * kc;
  std::vector<double> jc(p + 1), jz(p + 1), Pf(p * (p + 1) / 2), df(p);
  std::vector<cplx> ht(p + 1), Y(P2);
  for (int n = 0; n < p; n++) df[n] = dfact(n);
  jhat_all(kc * RAD, p, jc.data());
  // L2L columns: the parent's local basis functions about the parent centre
  // on the child's sphere, j^_j(kappa_p |z|/2) (|z|/2)^j Y_j^k, z = c + s in
  // child units
  auto fl = [&](double sx, double sy, double sz, cplx* f) {
    const double zx = cx + sx, zy = cy + sy, zz = cz + sz, r = std::sqrt(zx * zx + zy * zy + zz * zz);
    jhat_all(kc * r, p, jz.data());
    harmonics_all(p, zx, zy, zz, Pf, Y.data());
    double rj = 1;
    for (int j = 0; j < p; j++, rj *= 0.5 * r)
      for (int k = -j; k <= j; k++) f[ci(j, k)] = jz[j] * rj * Y[ci(j, k)];
  };
  std::vector<cplx> U;
  proj::project(p, P2, true, NTH, NPH, RAD, fl, U);
  // child coefficients; the factor 1/2 = rb_child / rb_parent of the
  // 1/(4 pi rb) prefactors
  double an = 1;
  for (int n = 0; n < p; n++, an *= RAD) {
    const double s = 0.5 * (2 * n + 1) / (4 * M_PI) / (jc[n] * an);
    for (int m = -n; m <= n; m++)
      for (int col = 0; col < P2; col++) U[(size_t)ci(n, m) * P2 + col] *= s;
  }
  split_real(p, U, Al2l);
  // M2M columns: the child's multipole basis functions about the child
  // centre, on the sphere |s| = RBIG (parent units) around the parent centre:
  // z = 2 s - c in child units
  htil_all(kp * RBIG, p, ht.data());
  std::vector<cplx> hb(ht.begin(), ht.end()), hz(p + 1);
  auto fm = [&](double sx, double sy, double sz, cplx* f) {
    const double zx = 2 * sx - cx, zy = 2 * sy - cy, zz = 2 * sz - cz, r = std::sqrt(zx * zx + zy * zy + zz * zz);
    htil_all(kc * r, p, hz.data());
    harmonics_all(p, zx, zy, zz, Pf, Y.data());
    double rn = r;
    for (int n = 0; n < p; n++, rn *= r) {
      const cplx rad = cplx(0, 1) * hz[n] / (df[n] * rn);
      for (int m = -n; m <= n; m++) f[ci(n, m)] = rad * Y[ci(n, m)];
    }
  };
  std::vector<cplx> V;
  proj::project(p, P2, true, NTH, NPH, RBIG, fm, V);
  // parent coefficients: divide by the parent basis radial value at RBIG;
  // factor 2 = rb_parent / rb_child
  double rb = RBIG;  // RBIG^(n+1)
  for (int n = 0; n < p; n++, rb *= RBIG) {
    const cplx radial = cplx(0, 1) * hb[n] / (df[n] * rb);
    const cplx s = 2.0 * (2 * n + 1) / (4 * M_PI) / radial;
    for (int m = -n; m <= n; m++)
      for (int col = 0; col < P2; col++) V[(size_t)ci(n, m) * P2 + col] *= s;
  }
  split_real(p, V, Am2m);
}

}  // namespace fmmx
