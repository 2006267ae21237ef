// precompute.cpp -- FP64 host precompute: the reference's tables
// (kernel/precalc.cxx:10-213) and, from them, dense real translation
// operators.  The reference applies rotate / axial-translate / rotate-back
// per box pair (kernel/m2m.cxx:126-154, m2l.cxx:43-81, l2l.cxx:203-237);
// here the same composition is applied once per unit vector of the real
// coefficient basis, giving a p^2 x p^2 matrix per lattice offset that the
// GPU applies as a GEMM.  Every arithmetic step keeps the reference's
// precision boundaries so that FMMX_CONV_REF operators reproduce the
// reference's per-pair results to rounding.
#include <cmath>
#include <cstring>
#include <thread>

#include "fmmx_internal.h"

namespace fmmx {

namespace {

double factorial(int n) {  // kernel/factorial.cxx:1-10
  double f = 1;
  for (int i = 1; i <= n; i++) f = f * i;
  return f;
}

inline double tbl(int conv, double v) { return conv == FMMX_CONV_REF ? (double)(float)v : v; }

// Spherical angles of a lattice vector as kernel/precalc.cxx:79-93.
void angles(int conv, double x, double y, double z, double& rh, double& al, double& be) {
  double r = std::sqrt(x * x + y * y + z * z);
  if (conv == FMMX_CONV_REF) {
    rh = r + EPSF;
    al = std::acos(z / rh);
    if (std::abs(x) + std::abs(y) < EPSF) be = 0;
    else if (std::abs(x) < EPSF) be = y / std::abs(y) * PIF * 0.5;
    else if (x > 0) be = std::atan(y / x);
    else be = std::atan(y / x) + PIF;
  } else {
    rh = r;
    al = r > 0 ? std::acos(z / r) : 0.0;
    be = (x == 0 && y == 0) ? 0.0 : std::atan2(y, x);
  }
}

inline int nmk_index(int n, int m, int k) { return (4 * n * n * n + 6 * n * n + 5 * n) / 3 + m * (2 * n + 1) + k; }

// Lattice key <-> coordinates (fmm/boxc.cxx, fmm/boxn1.cxx with 3 levels).
void decode3(int key, int* c) {
  c[0] = c[1] = c[2] = 0;
  for (int i = 0; i < 3; i++) {
    c[1] |= ((key >> (3 * i)) & 1) << i;
    c[0] |= ((key >> (3 * i + 1)) & 1) << i;
    c[2] |= ((key >> (3 * i + 2)) & 1) << i;
  }
}
int encode3(int x, int y, int z) {
  int k = 0;
  for (int i = 0; i < 3; i++) {
    k |= ((y >> i) & 1) << (3 * i);
    k |= ((x >> i) & 1) << (3 * i + 1);
    k |= ((z >> i) & 1) << (3 * i + 2);
  }
  return k;
}

// Wigner rotation (kernel/precalc.cxx:95-146) into out[m*p*p + n(n+1)+k].
void wigner(int p, int conv, double al, double be, const std::vector<double>& a0,
            const std::vector<double>& a1, std::vector<double>& d, cplx* out) {
  std::vector<cplx> expbe(4 * p);
  double sc = conv == FMMX_CONV_REF ? std::sin(al) / (1 + std::cos(al)) : std::tan(al * 0.5);
  for (int n = 0; n < 4 * p - 3; n++) expbe[n] = std::exp(cplx(0.0, (n - 2 * p + 2) * be));
  for (int n = 0; n < p; n++) {
    d[nmk_index(n, n, n)] = std::pow(std::cos(al * 0.5), 2 * n);
    for (int k = n; k >= 1 - n; k--) {
      double ank = ((double)n + k) / (n - k + 1);
      d[nmk_index(n, n, k - 1)] = std::sqrt(ank) * std::tan(al * 0.5) * d[nmk_index(n, n, k)];
    }
    for (int m = n; m >= 1; m--)
      for (int k = m - 1; k >= 1 - m; k--)
        d[nmk_index(n, m - 1, k)] =
            a1[nmk_index(n, m, k)] * d[nmk_index(n, m, k + 1)] + a0[nmk_index(n, m, k)] * sc * d[nmk_index(n, m, k)];
  }
  for (int n = 1; n < p; n++)
    for (int m = 0; m <= n; m++) {
      int base = (4 * n * n * n + 6 * n * n + 5 * n) / 3;
      for (int k = -m; k <= -1; k++) {
        int i0 = nmk_index(n, m, k), i1 = base - k * (2 * n + 1) - m;
        d[i0] = std::pow(-1.0, k) * d[i0];
        d[i1] = std::pow(-1.0, m + k) * d[i0];
      }
      for (int k = 0; k <= m; k++) {
        int i0 = nmk_index(n, m, k), i1 = nmk_index(n, k, m), i2 = base - k * (2 * n + 1) - m;
        d[i1] = std::pow(-1.0, m + k) * d[i0];
        d[i2] = d[i1];
      }
    }
  int p2 = p * p;
  for (int n = 0; n < p; n++)
    for (int m = 0; m <= n; m++)
      for (int k = -n; k <= n; k++) out[m * p2 + n * (n + 1) + k] = d[nmk_index(n, m, k)] * expbe[k + m + 2 * p - 2];
}

// kernel/spharot.cxx:3-24
void spharot(const cplx* yx, cplx* yxr, int p, const cplx* d) {
  int p2 = p * p;
  for (int n = 0; n < p; n++)
    for (int m = 0; m <= n; m++) {
      cplx s = 0;
      for (int k = -n; k <= -1; k++) s += d[m * p2 + n * (n + 1) + k] * std::conj(yx[n * (n + 1) / 2 - k]);
      for (int k = 0; k <= n; k++) s += d[m * p2 + n * (n + 1) + k] * yx[n * (n + 1) / 2 + k];
      yxr[n * (n + 1) / 2 + m] = s;
    }
}

template <class F>
void parallel_for(int n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if ((int)nt > n) nt = n > 0 ? n : 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      for (int i = (int)t; i < n; i += (int)nt) f(i);
    });
  for (auto& x : th) x.join();
}

// Fill column `col` of a real operator from a complex map applied to a unit
// input: in_scale[n] multiplies the input of degree n, out_scale[j] the
// output of degree j.
template <class Map>
void build_dense(const Tables& T, const double* in_scale, const double* out_scale, Map map, double* A) {
  int p = T.p, P2 = p * p, nm = nre(p);
  std::vector<cplx> in(nm), out(nm);
  for (int col = 0; col < P2; col++) {
    std::fill(in.begin(), in.end(), cplx(0, 0));
    int n, m;
    bool im;
    if (col < nm) {
      n = 0;
      while ((n + 1) * (n + 2) / 2 <= col) n++;
      m = col - n * (n + 1) / 2;
      im = false;
    } else {
      int c = col - nm;
      n = 1;
      while (n * (n + 1) / 2 <= c) n++;
      m = c - n * (n - 1) / 2 + 1;
      im = true;
    }
    in[slot_re(n, m)] = im ? cplx(0, in_scale[n]) : cplx(in_scale[n], 0);
    map(in.data(), out.data());
    for (int j = 0; j < p; j++)
      for (int k = 0; k <= j; k++) {
        cplx v = out[slot_re(j, k)] * out_scale[j];
        A[(size_t)slot_re(j, k) * P2 + col] = v.real();
        if (k >= 1) A[(size_t)slot_im(p, j, k) * P2 + col] = v.imag();
      }
  }
}

}  // namespace

int canon_index(int ax, int ay, int az) {
  if (ax < 0 || ay < 0 || az < 0 || ax > 3 || ay > 3 || az > 3 || (ax <= 1 && ay <= 1 && az <= 1)) return -1;
  int idx = 0;
  for (int a = 0; a <= 3; a++)
    for (int b = 0; b <= 3; b++)
      for (int c = 0; c <= 3; c++) {
        if (a <= 1 && b <= 1 && c <= 1) continue;
        if (a == ax && b == ay && c == az) return idx;
        idx++;
      }
  return -1;
}

void canon_vector(int idx, int* d) {
  for (int a = 0; a <= 3; a++)
    for (int b = 0; b <= 3; b++)
      for (int c = 0; c <= 3; c++) {
        if (a <= 1 && b <= 1 && c <= 1) continue;
        if (idx-- == 0) {
          d[0] = a;
          d[1] = b;
          d[2] = c;
          return;
        }
      }
}

int offset_index(int dx, int dy, int dz) {
  int idx = 0;
  for (int x = -3; x <= 3; x++)
    for (int y = -3; y <= 3; y++)
      for (int z = -3; z <= 3; z++) {
        if (std::abs(x) <= 1 && std::abs(y) <= 1 && std::abs(z) <= 1) continue;
        if (x == dx && y == dy && z == dz) return idx;
        idx++;
      }
  return -1;
}

void offset_vector(int idx, int* d) {
  int i = 0;
  for (int x = -3; x <= 3; x++)
    for (int y = -3; y <= 3; y++)
      for (int z = -3; z <= 3; z++) {
        if (std::abs(x) <= 1 && std::abs(y) <= 1 && std::abs(z) <= 1) continue;
        if (i == idx) {
          d[0] = x; d[1] = y; d[2] = z;
          return;
        }
        i++;
      }
}

void build_tables(Tables& T, int p, int conv) {
  int p2 = p * p;
  T.p = p;
  T.conv = conv;
  T.fac.assign(4 * p2, 0);
  T.anm.assign(4 * p2, 0);
  T.sr.assign((size_t)p2 * p2, 0);
  T.ynm.assign(4 * p2, 0);
  T.dnm.assign((size_t)2 * NRBM * p * p2, 0);
  T.flip.assign(2 * NRBM, 0);
  T.corr.assign(nre(p), 1.0);
  for (int n = 0; n < 2 * p; n++)
    for (int m = -n; m <= n; m++) {
      int nm = n * n + n + m, am = std::abs(m);
      T.fac[nm] = tbl(conv, std::sqrt(factorial(n - am) / factorial(n + am)));
      T.anm[nm] = tbl(conv, std::pow(-1.0, n) / std::sqrt(factorial(n - m) * factorial(n + m)));
    }
  for (int n = 0; n < p; n++)
    for (int m = 0; m <= n; m++)
      T.corr[slot_re(n, m)] = T.fac[n * n + n + m] / std::sqrt(factorial(n - m) / factorial(n + m));
  for (int j = 0; j < p; j++)
    for (int k = -j; k <= j; k++) {
      int jk = j * j + j + k;
      for (int n = std::abs(k); n < p; n++) {
        int nk = n * n + n + k, jnk = (j + n) * (j + n) + j + n;
        T.sr[(size_t)jk * p2 + nk] = tbl(conv, std::pow(-1.0, j + k) * T.anm[nk] * T.anm[jk] / T.anm[jnk]);
      }
    }
  double pn = 1;
  for (int m = 0; m < 2 * p; m++) {
    double pv = pn, p1, p2v;
    int npn = m * m + 2 * m, nmn = m * m;
    T.ynm[npn] = T.fac[npn] * pv;
    T.ynm[nmn] = std::conj(T.ynm[npn]);
    p1 = pv;
    pv = (2 * m + 1) * pv;
    for (int n = m + 1; n < 2 * p; n++) {
      int npm = n * n + n + m, nmm = n * n + n - m;
      T.ynm[npm] = T.fac[npm] * pv;
      T.ynm[nmm] = std::conj(T.ynm[npm]);
      p2v = p1;
      p1 = pv;
      pv = ((2 * n + 1) * p1 - (n + m) * p2v) / (n - m + 1);
    }
    pn = 0;
  }
  size_t nk4 = (size_t)p2 * p2 + 64;
  std::vector<double> a0(nk4, 0), a1(nk4, 0);
  for (int n = 0; n < p; n++)
    for (int m = 1; m <= n; m++) {
      double anmd = n * (n + 1) - m * (m - 1);
      for (int k = 1 - m; k < m; k++) {
        int idx = nmk_index(n, m, k);
        a0[idx] = -(m + k) / std::sqrt(anmd);
        a1[idx] = std::sqrt(((double)(n * (n + 1) - k * (k + 1))) / (n * (n + 1) - m * (m - 1)));
      }
    }
  parallel_for(NRBM, [&](int i) {
    std::vector<double> d(nk4, 0);
    int c[3];
    decode3(i, c);
    double x = c[0] - 3, y = c[1] - 3, z = c[2] - 3, rh, al, be;
    angles(conv, x, y, z, rh, al, be);
    if (conv == FMMX_CONV_EXACT && x == 0 && y == 0 && z < 0) {
      // rotation by pi is singular in this recurrence: no rotation, and the
      // axial M2L translates along -z instead ((-1)^(j+n))
      al = 0;
      be = 0;
      T.flip[i] = 1;
    }
    wigner(p, conv, al, be, a0, a1, d, &T.dnm[(size_t)i * p * p2]);
    wigner(p, conv, -al, -be, a0, a1, d, &T.dnm[(size_t)(i + NRBM) * p * p2]);
  });
}

// kernel/m2m.cxx:126-154 as an operator on scaled multipoles
void op_m2m(const Tables& T, int oct, double* A) {
  int p = T.p, p2 = p * p;
  int cx = (oct >> 1) & 1, cy = oct & 1, cz = (oct >> 2) & 1;
  int je = encode3(4 - 2 * cx, 4 - 2 * cy, 4 - 2 * cz);
  // child box size 1/2, parent 1: rh = rb_parent*sqrt(3)/4 with rb_parent = 1
  double rh = 1.0 * std::sqrt(3.0) / 4;
  std::vector<double> sin_(p), sout(p);
  for (int n = 0; n < p; n++) {
    sin_[n] = std::pow(0.5, n);  // M_child = rb_c^n M~ (rb_c = 1/2)
    sout[n] = 1.0;               // M~_parent = M / 1^n
  }
  const cplx* d0 = &T.dnm[(size_t)je * p * p2];
  const cplx* d1 = &T.dnm[(size_t)(je + NRBM) * p * p2];
  build_dense(T, sin_.data(), sout.data(), [&](const cplx* in, cplx* out) {
    std::vector<cplx> b(nre(p)), bd(nre(p));
    spharot(in, b.data(), p, d0);
    for (int j = 0; j < p; j++)
      for (int k = 0; k <= j; k++) {
        int jk = j * j + j + k;
        cplx s = 0;
        for (int n = 0; n <= j - std::abs(k); n++) {
          int jnk = (j - n) * (j - n) + j - n + k, jnks = (j - n) * (j - n + 1) / 2 + k, nm = n * n + n;
          cplx cnm = std::pow(-1.0, n) * T.anm[nm] * T.anm[jnk] / T.anm[jk] * std::pow(rh, n) * T.ynm[nm];
          s += b[jnks] * cnm;
        }
        bd[slot_re(j, k)] = s;
      }
    spharot(bd.data(), out, p, d1);
  }, A);
}

// kernel/l2l.cxx:203-237 as an operator on scaled locals
void op_l2l(const Tables& T, int oct, double* A) {
  int p = T.p, p2 = p * p;
  int cx = (oct >> 1) & 1, cy = oct & 1, cz = (oct >> 2) & 1;
  int je = encode3(2 * cx + 2, 2 * cy + 2, 2 * cz + 2);
  double rh = 0.5 * std::sqrt(3.0) / 2;  // child box size 1/2
  std::vector<double> sin_(p), sout(p);
  for (int n = 0; n < p; n++) {
    sin_[n] = 1.0;                     // L_parent = L~ / 1^(n+1)
    sout[n] = std::pow(0.5, n + 1);    // L~_child = L * (1/2)^(n+1)
  }
  const cplx* d0 = &T.dnm[(size_t)je * p * p2];
  const cplx* d1 = &T.dnm[(size_t)(je + NRBM) * p * p2];
  build_dense(T, sin_.data(), sout.data(), [&](const cplx* in, cplx* out) {
    std::vector<cplx> a(nre(p)), ad(nre(p));
    spharot(in, a.data(), p, d0);
    for (int j = 0; j < p; j++)
      for (int k = 0; k <= j; k++) {
        int jk = j * j + j + k;
        cplx s = 0;
        for (int n = j; n < p; n++) {
          int jnk = (n - j) * (n - j) + n - j, nk = n * n + n + k, nks = n * (n + 1) / 2 + k;
          double a3;
          if (T.conv == FMMX_CONV_REF) {  // float operands in the reference
            float t = (float)T.anm[jnk] * (float)T.anm[jk];
            a3 = t / (float)T.anm[nk];
          } else {
            a3 = T.anm[jnk] * T.anm[jk] / T.anm[nk];
          }
          s += a[nks] * (a3 * std::pow(rh, n - j) * T.ynm[jnk]);
        }
        ad[slot_re(j, k)] = s;
      }
    spharot(ad.data(), out, p, d1);
  }, A);
}

// kernel/m2l.cxx:43-81 as an operator between scaled expansions at a level
// of box size rb (target minus source centre = (dx,dy,dz)*rb)
void op_m2l(const Tables& T, int dx, int dy, int dz, double rb, double* A) {
  int p = T.p, p2 = p * p;
  int je = encode3(dx + 3, dy + 3, dz + 3);
  double x = dx * rb, y = dy * rb, z = dz * rb;
  double rh = std::sqrt(x * x + y * y + z * z) + (T.conv == FMMX_CONV_REF ? (double)EPSF : 0.0);
  bool flip = T.flip[je] != 0;
  std::vector<double> sin_(p), sout(p);
  for (int n = 0; n < p; n++) {
    sin_[n] = std::pow(rb, n);       // M = rb^n M~
    sout[n] = std::pow(rb, n + 1);   // L~ = rb^(n+1) L
  }
  const cplx* d0 = &T.dnm[(size_t)je * p * p2];
  const cplx* d1 = &T.dnm[(size_t)(je + NRBM) * p * p2];
  build_dense(T, sin_.data(), sout.data(), [&](const cplx* in, cplx* out) {
    std::vector<cplx> b(nre(p)), ad(nre(p));
    spharot(in, b.data(), p, d0);
    double rhj = 1;
    for (int j = 0; j < p; j++) {
      double rhjk = rhj;
      rhj *= rh;
      for (int k = 0; k <= j; k++) {
        int jk = j * j + j + k;
        cplx s = 0;
        double rhjn = rhjk;
        rhjk *= rh;
        for (int n = std::abs(k); n < p; n++) {
          rhjn *= rh;
          int nk = n * n + n + k, nks = n * (n + 1) / 2 + k, jnk = (j + n) * (j + n) + j + n;
          cplx cnm = T.sr[(size_t)jk * p2 + nk] / rhjn * T.ynm[jnk];
          if (flip && ((j + n) & 1)) cnm = -cnm;
          s += b[nks] * cnm;
        }
        ad[slot_re(j, k)] = s;
      }
    }
    spharot(ad.data(), out, p, d1);
  }, A);
}

std::string& last_error() {
  static thread_local std::string e;
  return e;
}

int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}

}  // namespace fmmx
