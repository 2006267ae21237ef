// yukawa.cpp -- dense translation operators of the Yukawa (screened
// Coulomb) FMM: kernel k0(lambda r) = (pi/2) exp(-lambda r) / (lambda r)
// (kernel/Helmholtz_G_p2p.cxx:22, kernel/Helmholtz_G_dir.cxx:16).
//
// Expansions (m = -n..n, Y_n^m = fac(n,m) P_n^|m|(cos t) e^{i m f} with the
// Condon-Shortley phase, as the Laplace engine's harmonics):
//   multipole  phi(x) = sum M_n^m k_n(lambda|x-c|) Y_n^m,
//              M_n^m  = (2n+1) sum_j q_j i_n(lambda rho_j) conj(Y_n^m(y_j - c)),
//   local      phi(x) = sum L_n^m i_n(lambda|x-c|) Y_n^m(x - c),
// from the addition theorem k0(lambda|x-y|) = sum_n (2n+1) i_n k_n P_n(cos g).
// Stored scaled by the box size rb of their level: M~ = M / ((2n+1) i_n(lambda
// rb)), L~ = L i_n(lambda rb), so that P2M and L2P are the Laplace kernels'
// harmonics times per-degree radial weights.
//
// The reference translates with plane waves (kernel/Helmholtz_m2l.cxx) and
// rotations (kernel/Helmholtz_m2m.cxx).  Here every operator is the exact
// re-expansion of each basis function, obtained by projecting it onto the
// spherical harmonics on a sphere around the new centre (Gauss-Legendre in
// cos t, trapezoid in f: the integrands are analytic and the quadrature
// converges to rounding), which is what the truncated addition theorem gives.
#include <cmath>
#include <complex>
#include <thread>
#include <vector>

#include "fmmx_internal.h"
#include "projection.h"
#include "yukawa.h"

namespace fmmx {

void ibar_all(double x, int N, double* out) {
  auto series = [&](int n) {
    double df = 1;
    for (int k = 3; k <= 2 * n + 1; k += 2) df *= k;
    double t = 1.0 / df, s = t;
    const double h = 0.5 * x * x;
    for (int k = 1; k < 2000; k++) {
      t *= h / (k * (2.0 * n + 2.0 * k + 1.0));
      s += t;
      if (t < 1e-18 * s) break;
    }
    return s;
  };
  double a = series(N + 1), b = series(N);
  out[N] = b;
  for (int n = N; n >= 1; n--) {
    const double c = (2 * n + 1) * b + x * x * a;
    a = b;
    b = c;
    out[n - 1] = c;
  }
}

void kbar_all(double x, int N, double* out) {
  const double e = 0.5 * M_PI * std::exp(-x);
  out[0] = e;
  if (N >= 1) out[1] = e * (1 + x);
  for (int n = 1; n < N; n++) out[n + 1] = x * x * out[n - 1] + (2 * n + 1) * out[n];
}

namespace {

using proj::ci;
using proj::sl;
using proj::harmonics_all;
// quadrature on the projection sphere: 24 x 36 nodes up to p = 12, 2p x 3p
// above (Gauss-Legendre exact to degree 2 NTH - 1 >= 2(p-1) + the field's
// decay; the trapezoid rule aliases azimuthal orders NPH - p + 1 and up)
constexpr double RAD = 0.5;  // its radius in box units

template <class Field>
void project(int p, int ncol, bool all_k, Field field, std::vector<cplx>& out) {
  const int nth = p <= 12 ? 24 : 2 * p, nph = p <= 12 ? 36 : 3 * p;
  proj::project(p, ncol, all_k, nth, nph, RAD, field, out);
}

// complex operator C (rows (j, k >= 0) at ci(j,k), columns (n, m) all m at
// ci(n,m)) acting on coefficients with v_{n,-m} = conj(v_{n,m}) -> real
// p^2 x p^2 operator on the slots of slot_re / slot_im, row-major
void to_real(int p, const std::vector<cplx>& C, int ncol, double* A) {
  const int P2 = p * p;
  std::fill(A, A + (size_t)P2 * P2, 0.0);
  const cplx I(0, 1);
  for (int j = 0; j < p; j++)
    for (int k = 0; k <= j; k++) {
      const cplx* row = &C[(size_t)ci(j, k) * ncol];
      double* Ar = A + (size_t)slot_re(j, k) * P2;
      double* Ai = k > 0 ? A + (size_t)slot_im(p, j, k) * P2 : nullptr;
      for (int n = 0; n < p; n++)
        for (int m = 0; m <= n; m++) {
          const cplx da = m > 0 ? row[ci(n, m)] + row[ci(n, -m)] : row[ci(n, 0)];
          const cplx db = m > 0 ? I * (row[ci(n, m)] - row[ci(n, -m)]) : cplx(0);
          Ar[slot_re(n, m)] = da.real();
          if (m > 0) Ar[slot_im(p, n, m)] = db.real();
          if (Ai) {
            Ai[slot_re(n, m)] = da.imag();
            if (m > 0) Ai[slot_im(p, n, m)] = db.imag();
          }
        }
    }
}

}  // namespace

void yuk_op_m2l(int p, double lrb, int dx, int dy, int dz, double* A) {
  const int P2 = p * p;
  if (!(lrb <= YUK_LRB_MAX)) {
    std::fill(A, A + (size_t)P2 * P2, 0.0);
    return;
  }
  std::vector<double> ib(p + 2), ibr(p + 2), kb(p + 1), Pf(p * (p + 1) / 2);
  ibar_all(lrb, p, ib.data());
  ibar_all(lrb * RAD, p, ibr.data());
  std::vector<cplx> Y(P2);
  // columns: the field (2n+1) i_n(lrb) k_n(lrb |z|) Y_n^m(z) of each multipole
  // basis function about the source centre, z = target offset + s
  auto field = [&](double sx, double sy, double sz, cplx* f) {
    const double zx = dx + sx, zy = dy + sy, zz = dz + sz, r = std::sqrt(zx * zx + zy * zy + zz * zz);
    kbar_all(lrb * r, p - 1, kb.data());
    harmonics_all(p, zx, zy, zz, Pf, Y.data());
    double rn = r;  // r^(n+1)
    for (int n = 0; n < p; n++, rn *= r) {
      const double rad = (2 * n + 1) * ib[n] * kb[n] / (lrb * rn);
      for (int m = -n; m <= n; m++) f[ci(n, m)] = rad * Y[ci(n, m)];
    }
  };
  std::vector<cplx> C;
  project(p, P2, false, field, C);
  // L~_j = i_j(lrb) (2j+1) / (4 pi i_j(lrb RAD)) * projection
  double aj = 1;
  for (int j = 0; j < p; j++, aj *= RAD) {
    const double s = (2 * j + 1) / (4 * M_PI) * ib[j] / (ibr[j] * aj);
    for (int k = 0; k <= j; k++)
      for (int col = 0; col < P2; col++) C[(size_t)ci(j, k) * P2 + col] *= s;
  }
  to_real(p, C, P2, A);
}

void yuk_op_m2m_l2l(int p, double lrb_child, int oct, double* Am2m, double* Al2l) {
  const int P2 = p * p;
  if (!(lrb_child <= YUK_LRB_MAX)) {
    std::fill(Am2m, Am2m + (size_t)P2 * P2, 0.0);
    std::fill(Al2l, Al2l + (size_t)P2 * P2, 0.0);
    return;
  }
  // child centre relative to the parent centre, child box units (octant bits
  // x = 2, y = 1, z = 4, fmm/boxn.cxx bit order)
  const double cx = ((oct >> 1) & 1) ? 0.5 : -0.5, cy = (oct & 1) ? 0.5 : -0.5, cz = ((oct >> 2) & 1) ? 0.5 : -0.5;
  const double lc = lrb_child, lp = 2 * lrb_child;
  std::vector<double> ibc(p + 2), ibp(p + 2), ibr(p + 2), iz(p + 2), Pf(p * (p + 1) / 2);
  ibar_all(lc, p, ibc.data());
  ibar_all(lp, p, ibp.data());
  ibar_all(lc * RAD, p, ibr.data());
  std::vector<cplx> Y(P2);
  // columns: the parent's scaled regular basis functions (j, k) about the
  // parent centre, i_j(lambda |z|) / i_j(lambda rb_parent) Y_j^k, z = c + s
  auto field = [&](double sx, double sy, double sz, cplx* f) {
    const double zx = cx + sx, zy = cy + sy, zz = cz + sz, r = std::sqrt(zx * zx + zy * zy + zz * zz);
    ibar_all(lc * r, p, iz.data());
    harmonics_all(p, zx, zy, zz, Pf, Y.data());
    double rj = 1;  // (r / 2)^j: r in child units, the parent box is 2
    for (int j = 0; j < p; j++, rj *= 0.5 * r) {
      const double rad = iz[j] / ibp[j] * rj;
      for (int k = -j; k <= j; k++) f[ci(j, k)] = rad * Y[ci(j, k)];
    }
  };
  std::vector<cplx> U;
  project(p, P2, true, field, U);
  // L~_child(n, m) = i_n(lambda rb_c) (2n+1) / (4 pi i_n(lambda rb_c RAD)) * projection
  double an = 1;
  for (int n = 0; n < p; n++, an *= RAD) {
    const double s = (2 * n + 1) / (4 * M_PI) * ibc[n] / (ibr[n] * an);
    for (int m = -n; m <= n; m++)
      for (int col = 0; col < P2; col++) U[(size_t)ci(n, m) * P2 + col] *= s;
  }
  to_real(p, U, P2, Al2l);
  // M2M(j,k; n,m) = conj(L2L(n,m; j,k)) on the scaled bases
  std::vector<cplx> Mt((size_t)P2 * P2);
  for (int r = 0; r < P2; r++)
    for (int c = 0; c < P2; c++) Mt[(size_t)c * P2 + r] = std::conj(U[(size_t)r * P2 + c]);
  to_real(p, Mt, P2, Am2m);
}

}  // namespace fmmx
