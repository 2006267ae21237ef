/*
 * fmm_oracle.c -- CPU FP64 restatement of the reference Laplace FMM.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * the bench's cpu_baseline leg load this library; the product never does.
 *
 * Each function names the reference file:line it follows (paths relative to
 * the reference repository root).  The loop nests, the order of every
 * floating-point sum and the float/double boundaries of the reference are
 * kept so that, in ORC_CONV_REF, results match a build of the reference
 * sources to the last bit.  ORC_CONV_EXACT keeps the algorithm and removes
 * the eps/float-pi/float-table quirks (see fmm_oracle.h).
 *
 * Pinning (what checks this file against something other than itself):
 *   - tests/golden/ref_*.npz: vectors written by oracle/_ref (the reference
 *     sources compiled from /root/reference by oracle/Makefile.ref, mp=9)
 *     for the sort permutation, box lists, near and far interaction lists
 *     (values and order), precalc tables, multipoles after the upward pass,
 *     leaf local expansions and final potentials/gradients for neq 0/1/2.
 *   - closed forms (single charge, Wigner-d unitarity, Legendre identities)
 *     and brute-force direct sums for what the reference does not fix
 *     (ORC_CONV_EXACT, mp=10/12).
 */
#include "fmm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define NEBM 189  /* include/constants.h:21 */
#define NRBM 512  /* include/constants.h:22 */

/* include/constants.h:34-35: const float eps = 1e-6; const float pi = M_PI; */
static const float EPSF = 1e-6f;
static const float PIF = 3.14159265358979323846f;

static double now_s(void) {
  struct timeval tv;
  gettimeofday(&tv, 0);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static inline cplx mkc(double re, double im) { return CMPLX(re, im); }
/* exp(t*i) as std::exp(complex<double>(0,t)) evaluates it */
static inline cplx cis(double t) { return cexp(mkc(0.0, t)); }

/* ------------------------------------------------------------------ */
/* Morton helpers                                                      */
/* ------------------------------------------------------------------ */

/* fmm/boxn1.cxx:3-19: (x,y,z) -> key; x -> bit 3i+1, y -> 3i, z -> 3i+2 */
static int boxn1(int cx, int cy, int cz, int lev) {
  int n = 0;
  for (int i = 0; i < lev; i++) {
    n += (cx % 2) * (1 << (3 * i + 1)); cx /= 2;
    n += (cy % 2) * (1 << (3 * i));     cy /= 2;
    n += (cz % 2) * (1 << (3 * i + 2)); cz /= 2;
  }
  return n;
}

/* fmm/boxc.cxx:3-24 with ndi=3: key -> (x,y,z) */
static void boxc(int nf, int* nc) {
  int t[3] = {0, 0, 0};
  int nb = nf, k = 0, i = 0;
  while (nb != 0) {
    int j = 3 - k - 1;
    t[j] += (nb % 2) * (1 << i);
    nb /= 2;
    k = (k + 1) % 3;
    if (k == 0) i++;
  }
  nc[0] = t[1];
  nc[1] = t[2];
  nc[2] = t[0];
}

/* fmm/boxn.cxx:5-31: key per particle; (x-xmin) in float, times the scale
 * kl/rd, which fmm/boxn.cxx:9 computes as int/float (a float quotient)
 * before widening it to double */
static void boxn(int n, const float* x, const float* y, const float* z, int lmax,
                 float xmin, float ymin, float zmin, float rd, int* key) {
  int kl = 1 << lmax;
  double rb = (double)((float)kl / rd);
  for (int j = 0; j < n; j++) {
    int f[3];
    f[0] = (int)((double)(x[j] - xmin) * rb);
    f[1] = (int)((double)(y[j] - ymin) * rb);
    f[2] = (int)((double)(z[j] - zmin) * rb);
    for (int i = 0; i < 3; i++)
      if (f[i] >= kl) f[i]--;
    int nf = 0;
    for (int i = 0; i < lmax; i++) {
      nf += f[0] % 2 << (3 * i + 1); f[0] >>= 1;
      nf += f[1] % 2 << (3 * i);     f[1] >>= 1;
      nf += f[2] % 2 << (3 * i + 2); f[2] >>= 1;
    }
    key[j] = nf;
  }
}

/* sort/csort.cxx:5-17: stable counting sort; perm[sorted] = original */
static void csort(int np, int nbmax, int* na, int* perm) {
  int* nc = (int*)calloc((size_t)nbmax, sizeof(int));
  int* nd = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  for (int i = 0; i < np; i++) nc[na[i]]++;
  for (int i = 1; i < nbmax; i++) nc[i] += nc[i - 1];
  for (int i = np - 1; i >= 0; i--) {
    nc[na[i]]--;
    nd[nc[na[i]]] = na[i];
    perm[nc[na[i]]] = i;
  }
  for (int i = 0; i < np; i++) na[i] = nd[i];
  free(nc);
  free(nd);
}

/* ------------------------------------------------------------------ */
/* Tables: kernel/precalc.cxx:10-213, kernel/factorial.cxx:1-10        */
/* ------------------------------------------------------------------ */

static double factorial(int n) {
  double f = 1;
  for (int i = 1; i <= n; i++) f = f * i;
  return f;
}

typedef struct tables {
  int mp, conv;
  double *fac, *anm, *sr; /* float-rounded values in ORC_CONV_REF */
  cplx* ynm;              /* 4mp^2 */
  cplx* dnm;              /* [2*NRBM][mp][mp*mp] */
  unsigned char* flip;    /* EXACT only: lattice direction along -z */
} tables;

static inline double tbl(int conv, double v) { return conv == ORC_CONV_REF ? (double)(float)v : v; }

/* Spherical angles of a vector as kernel/G_p2m.cxx:31-41 (and precalc.cxx:83-93) */
static void angles(int conv, double x, double y, double z, double* rh, double* al, double* be) {
  double r = sqrt(x * x + y * y + z * z);
  if (conv == ORC_CONV_REF) {
    *rh = r + EPSF;
    *al = acos(z / *rh);
    if (fabs(x) + fabs(y) < EPSF) *be = 0;
    else if (fabs(x) < EPSF) *be = y / fabs(y) * PIF * 0.5;
    else if (x > 0) *be = atan(y / x);
    else *be = atan(y / x) + PIF;
  } else {
    *rh = r;
    *al = r > 0 ? acos(z / r) : 0.0;
    *be = (x == 0 && y == 0) ? 0.0 : atan2(y, x);
  }
}

#define NMK(n, m, k) ((4 * (n) * (n) * (n) + 6 * (n) * (n) + 5 * (n)) / 3 + (m) * (2 * (n) + 1) + (k))

/* kernel/precalc.cxx:95-146: Wigner rotation for angles (al,be) into dnm[mp][mp*mp] */
static void wigner(int mp, int conv, double al, double be, const double* anmk0, const double* anmk1,
                   double* dnmd, cplx* out) {
  cplx expbe[4 * 32];
  double sc = conv == ORC_CONV_REF ? sin(al) / (1 + cos(al)) : tan(al * 0.5);
  for (int n = 0; n < 4 * mp - 3; n++) expbe[n] = cis((n - 2 * mp + 2) * be);
  for (int n = 0; n < mp; n++) {
    dnmd[NMK(n, n, n)] = pow(cos(al * 0.5), 2 * n);
    for (int k = n; k >= 1 - n; k--) {
      double ank = ((double)n + k) / (n - k + 1);
      dnmd[NMK(n, n, k - 1)] = sqrt(ank) * tan(al * 0.5) * dnmd[NMK(n, n, k)];
    }
    for (int m = n; m >= 1; m--)
      for (int k = m - 1; k >= 1 - m; k--)
        dnmd[NMK(n, m - 1, k)] = anmk1[NMK(n, m, k)] * dnmd[NMK(n, m, k + 1)] +
                                 anmk0[NMK(n, m, k)] * sc * dnmd[NMK(n, m, k)];
  }
  for (int n = 1; n < mp; n++) {
    for (int m = 0; m <= n; m++) {
      for (int k = -m; k <= -1; k++) {
        double ek = pow(-1.0, k);
        int nmk = NMK(n, m, k);
        int nmk1 = (4 * n * n * n + 6 * n * n + 5 * n) / 3 - k * (2 * n + 1) - m;
        dnmd[nmk] = ek * dnmd[nmk];
        dnmd[nmk1] = pow(-1.0, m + k) * dnmd[nmk];
      }
      for (int k = 0; k <= m; k++) {
        int nmk = NMK(n, m, k);
        int nmk1 = NMK(n, k, m);
        int nmk2 = (4 * n * n * n + 6 * n * n + 5 * n) / 3 - k * (2 * n + 1) - m;
        dnmd[nmk1] = pow(-1.0, m + k) * dnmd[nmk];
        dnmd[nmk2] = dnmd[nmk1];
      }
    }
  }
  int mp2 = mp * mp;
  for (int n = 0; n < mp; n++)
    for (int m = 0; m <= n; m++)
      for (int k = -n; k <= n; k++)
        out[m * mp2 + n * (n + 1) + k] = dnmd[NMK(n, m, k)] * expbe[k + m + 2 * mp - 2];
}

static void tables_build(tables* T, int mp, int conv) {
  int mp2 = mp * mp, nfa = 4 * mp2;
  T->mp = mp;
  T->conv = conv;
  T->fac = (double*)calloc((size_t)nfa, sizeof(double));
  T->anm = (double*)calloc((size_t)nfa, sizeof(double));
  T->sr = (double*)calloc((size_t)mp2 * mp2, sizeof(double));
  T->ynm = (cplx*)calloc((size_t)nfa, sizeof(cplx));
  T->dnm = (cplx*)calloc((size_t)2 * NRBM * mp * mp2, sizeof(cplx));
  T->flip = (unsigned char*)calloc(2 * NRBM, 1);

  /* precalc.cxx:19-31 */
  for (int n = 0; n < 2 * mp; n++) {
    for (int m = -n; m <= n; m++) {
      int nm = n * n + n + m, am = abs(m);
      double fnmm = factorial(n - m), fnpm = factorial(n + m);
      double fnma = factorial(n - am), fnpa = factorial(n + am);
      T->fac[nm] = tbl(conv, sqrt(fnma / fnpa));
      double fad = sqrt(fnmm * fnpm);
      T->anm[nm] = tbl(conv, pow(-1.0, n) / fad);
    }
  }
  /* precalc.cxx:33-43 */
  for (int j = 0; j < mp; j++)
    for (int k = -j; k <= j; k++) {
      int jk = j * j + j + k;
      for (int n = abs(k); n < mp; n++) {
        int nk = n * n + n + k;
        int jnk = (j + n) * (j + n) + j + n;
        T->sr[jk * mp2 + nk] = tbl(conv, pow(-1.0, j + k) * T->anm[nk] * T->anm[jk] / T->anm[jnk]);
      }
    }
  /* precalc.cxx:45-64: Y_n^m at theta=0 */
  double pn = 1;
  for (int m = 0; m < 2 * mp; m++) {
    double p = pn, p1, p2;
    int npn = m * m + 2 * m, nmn = m * m;
    T->ynm[npn] = T->fac[npn] * p;
    T->ynm[nmn] = conj(T->ynm[npn]);
    p1 = p;
    p = (2 * m + 1) * p;
    for (int n = m + 1; n < 2 * mp; n++) {
      int npm = n * n + n + m, nmm = n * n + n - m;
      T->ynm[npm] = T->fac[npm] * p;
      T->ynm[nmm] = conj(T->ynm[npm]);
      p2 = p1;
      p1 = p;
      p = ((2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
    }
    pn = 0;
  }
  /* precalc.cxx:66-76 */
  int nk4 = mp2 * mp2 + 64;
  double* anmk0 = (double*)calloc((size_t)nk4, sizeof(double));
  double* anmk1 = (double*)calloc((size_t)nk4, sizeof(double));
  for (int n = 0; n < mp; n++)
    for (int m = 1; m <= n; m++) {
      double anmd = n * (n + 1) - m * (m - 1);
      for (int k = 1 - m; k < m; k++) {
        int nmk = NMK(n, m, k);
        double anmkd = ((double)(n * (n + 1) - k * (k + 1))) / (n * (n + 1) - m * (m - 1));
        anmk0[nmk] = -(m + k) / sqrt(anmd);
        anmk1[nmk] = sqrt(anmkd);
      }
    }
  /* precalc.cxx:78-203: rotations for the 512 lattice vectors (nc-3), and inverses */
#pragma omp parallel
  {
    double* dnmd = (double*)calloc((size_t)nk4, sizeof(double));
#pragma omp for schedule(static)
    for (int i = 0; i < NRBM; i++) {
      int nc[3];
      boxc(i, nc);
      double x = nc[0] - 3, y = nc[1] - 3, z = nc[2] - 3, rh, al, be;
      angles(conv, x, y, z, &rh, &al, &be);
      if (conv == ORC_CONV_EXACT && x == 0 && y == 0 && z < 0) {
        /* rotation by pi is singular in this recurrence: rotate by 0 and
           translate along -z instead (sign (-1)^(j+n) applied in m2l) */
        al = 0;
        be = 0;
        T->flip[i] = 1;
      }
      wigner(mp, conv, al, be, anmk0, anmk1, dnmd, T->dnm + (size_t)i * mp * mp2);
      wigner(mp, conv, -al, -be, anmk0, anmk1, dnmd, T->dnm + (size_t)(i + NRBM) * mp * mp2);
    }
    free(dnmd);
  }
  free(anmk0);
  free(anmk1);
}

static void tables_free(tables* T) {
  free(T->fac); free(T->anm); free(T->sr); free(T->ynm); free(T->dnm); free(T->flip);
}

/* kernel/spharot.cxx:3-24 */
static void spharot(const cplx* yx, cplx* yxr, int mp, const cplx* d /* [mp][mp*mp] */) {
  int mp2 = mp * mp;
  for (int n = 0; n < mp; n++)
    for (int m = 0; m <= n; m++) {
      cplx s = 0;
      for (int k = -n; k <= -1; k++) s += d[m * mp2 + n * (n + 1) + k] * conj(yx[n * (n + 1) / 2 - k]);
      for (int k = 0; k <= n; k++) s += d[m * mp2 + n * (n + 1) + k] * yx[n * (n + 1) / 2 + k];
      yxr[n * (n + 1) / 2 + m] = s;
    }
}

/* kernel/multipole.cxx:7-35 (bnm only; rh^n folded in) */
static void multipole(const double* fac, double rh, double al, int mp, cplx* bnm) {
  double xx = cos(al), s2 = sqrt((1 - xx) * (1 + xx)), fact = 1, pn = 1, rhm = 1;
  for (int m = 0; m < mp; m++) {
    double p = pn, p1, p2, rhn;
    int nm = m * m + 2 * m;
    bnm[nm] = rhm * fac[nm] * p;
    p1 = p;
    p = xx * (2 * m + 1) * p;
    rhm *= rh;
    rhn = rhm;
    for (int n = m + 1; n < mp; n++) {
      nm = n * n + n + m;
      bnm[nm] = rhn * fac[nm] * p;
      p2 = p1;
      p1 = p;
      p = (xx * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      rhn *= rh;
    }
    pn = -pn * fact * s2;
    fact += 2;
  }
}

/* kernel/multipoled.cxx:7-34: bnm and d/dtheta (bth), no radius */
static void multipoled(const double* fac, double th, int mp, cplx* bnm, cplx* bth) {
  double xx = cos(th), yy = sin(th), s2 = sqrt((1 - xx) * (1 + xx)), fact = 1, pn = 1;
  for (int m = 0; m < mp; m++) {
    double p = pn, p1, p2;
    int nm = m * m + 2 * m;
    bnm[nm] = fac[nm] * p;
    p1 = p;
    p = xx * (2 * m + 1) * p;
    bth[nm] = fac[nm] * (p - (m + 1) * xx * p1) / yy;
    for (int n = m + 1; n < mp; n++) {
      nm = n * n + n + m;
      bnm[nm] = fac[nm] * p;
      p2 = p1;
      p1 = p;
      p = (xx * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      bth[nm] = fac[nm] * ((n - m + 1) * p - (n + 1) * xx * p1) / yy;
    }
    pn = -pn * fact * s2;
    fact += 2;
  }
}

/* ------------------------------------------------------------------ */
/* Per-box operators                                                   */
/* ------------------------------------------------------------------ */

/* kernel/G_p2m.cxx:19-55 for one box */
static void p2m_G(const tables* T, int j0, int j1, const float* x, const float* y, const float* z,
                  const float* q, double cx, double cy, double cz, cplx* out) {
  int mp = T->mp, nmp = mp * (mp + 1) / 2;
  cplx bnm[4 * 32 * 32];
  for (int j = 0; j < nmp; j++) out[j] = 0;
  for (int j = j0; j <= j1; j++) {
    double dx = x[j] - cx, dy = y[j] - cy, dz = z[j] - cz, rh, al, be;
    angles(T->conv, dx, dy, dz, &rh, &al, &be);
    multipole(T->fac, rh, al, mp, bnm);
    for (int n = 0; n < mp; n++)
      for (int m = 0; m <= n; m++) {
        int nm = n * n + n + m, nms = n * (n + 1) / 2 + m;
        cplx eim = cis(-m * be);
        out[nms] += (mkc(q[j], 0.0) * bnm[nm]) * eim;
      }
  }
}

/* kernel/Gn_p2m.cxx:78-122 for one box (dipoles q*n) */
static void p2m_Gn(const tables* T, int j0, int j1, const float* x, const float* y, const float* z,
                   const float* nx, const float* ny, const float* nz, const float* q,
                   double cx, double cy, double cz, cplx* out) {
  int mp = T->mp, nmp = mp * (mp + 1) / 2;
  cplx bnm[4 * 32 * 32], bth[32 * 32];
  const cplx I1 = mkc(0.0, 1.0);
  for (int j = 0; j < nmp; j++) out[j] = 0;
  for (int j = j0; j <= j1; j++) {
    double dx = x[j] - cx, dy = y[j] - cy, dz = z[j] - cz, rh, al, be;
    angles(T->conv, dx, dy, dz, &rh, &al, &be);
    multipoled(T->fac, al, mp, bnm, bth);
    /* q*n in float as (vj[j]*gxj[j]) */
    float qnx = q[j] * nx[j], qny = q[j] * ny[j], qnz = q[j] * nz[j];
    for (int n = 0; n < mp; n++)
      for (int m = 0; m <= n; m++) {
        int nm = n * n + n + m, nms = n * (n + 1) / 2 + m;
        cplx eim = cis(-m * be);
        cplx brh = ((n * pow(rh, n - 1)) * bnm[nm]) * eim;
        cplx bal = (pow(rh, n) * bth[nm]) * eim;
        cplx bbe = (((-m * pow(rh, n)) * bnm[nm]) * eim) * I1;
        cplx bxd = sin(al) * cos(be) * brh + cos(al) * cos(be) / rh * bal - sin(be) / rh / sin(al) * bbe;
        cplx byd = sin(al) * sin(be) * brh + cos(al) * sin(be) / rh * bal + cos(be) / rh / sin(al) * bbe;
        cplx bzd = cos(al) * brh - sin(al) / rh * bal;
        out[nms] += mkc(qnx, 0.0) * bxd;
        out[nms] += mkc(qny, 0.0) * byd;
        out[nms] += mkc(qnz, 0.0) * bzd;
      }
  }
}

/* kernel/m2m.cxx:132-150: rotate, axial shift by rh, rotate back (je from child octant) */
static void m2m_one(const tables* T, int je, double rh, const cplx* in, cplx* out) {
  int mp = T->mp, mp2 = mp * mp;
  cplx bbx[32 * 33 / 2], bbxd[32 * 33 / 2];
  spharot(in, bbx, mp, T->dnm + (size_t)je * mp * mp2);
  for (int j = 0; j < mp; j++)
    for (int k = 0; k <= j; k++) {
      int jk = j * j + j + k, jks = j * (j + 1) / 2 + k;
      cplx s = 0;
      for (int n = 0; n <= j - abs(k); n++) {
        int jnk = (j - n) * (j - n) + j - n + k, jnks = (j - n) * (j - n + 1) / 2 + k, nm = n * n + n;
        cplx cnm = pow(-1.0, n) * T->anm[nm] * T->anm[jnk] / T->anm[jk] * pow(rh, n) * T->ynm[nm];
        s += bbx[jnks] * cnm;
      }
      bbxd[jks] = s;
    }
  spharot(bbxd, out, mp, T->dnm + (size_t)(je + NRBM) * mp * mp2);
}

/* kernel/m2l.cxx:43-78 for one (target, source) pair */
static void m2l_one(const tables* T, int je, double rh, const cplx* in, cplx* out) {
  int mp = T->mp, mp2 = mp * mp;
  cplx bbxd[32 * 33 / 2], aaxd[32 * 33 / 2];
  double sgn_flip = T->flip[je] ? -1.0 : 1.0;
  spharot(in, bbxd, mp, T->dnm + (size_t)je * mp * mp2);
  double rhj = 1;
  for (int j = 0; j < mp; j++) {
    double rhjk = rhj;
    rhj *= rh;
    for (int k = 0; k <= j; k++) {
      int jk = j * j + j + k, jks = j * (j + 1) / 2 + k;
      cplx s = 0;
      double rhjn = rhjk;
      rhjk *= rh;
      for (int n = abs(k); n < mp; n++) {
        rhjn *= rh;
        int nk = n * n + n + k, nks = n * (n + 1) / 2 + k;
        int jkn = jk * mp2 + nk, jnk = (j + n) * (j + n) + j + n;
        cplx cnm = T->sr[jkn] / rhjn * T->ynm[jnk];
        if (T->flip[je] && ((j + n) & 1)) cnm = sgn_flip * cnm;
        s += bbxd[nks] * cnm;
      }
      aaxd[jks] = s;
    }
  }
  spharot(aaxd, out, mp, T->dnm + (size_t)(je + NRBM) * mp * mp2);
}

/* kernel/l2l.cxx:214-233 */
static void l2l_one(const tables* T, int je, double rh, const cplx* in, cplx* out) {
  int mp = T->mp, mp2 = mp * mp;
  cplx aax[32 * 33 / 2], aaxd[32 * 33 / 2];
  spharot(in, aax, mp, T->dnm + (size_t)je * mp * mp2);
  for (int j = 0; j < mp; j++)
    for (int k = 0; k <= j; k++) {
      int jk = j * j + j + k, jks = j * (j + 1) / 2 + k;
      cplx s = 0;
      for (int n = j; n < mp; n++) {
        int jnk = (n - j) * (n - j) + n - j, nk = n * n + n + k, nks = n * (n + 1) / 2 + k;
        cplx cnm;
        /* anm is float in the reference: the first product and quotient
           are float operations there (kernel/l2l.cxx:227) */
        double a3;
        if (T->conv == ORC_CONV_REF) {
          float t = (float)T->anm[jnk] * (float)T->anm[jk];
          a3 = t / (float)T->anm[nk];
        } else {
          a3 = T->anm[jnk] * T->anm[jk] / T->anm[nk];
        }
        cnm = a3 * pow(rh, n - j) * T->ynm[jnk];
        s += aax[nks] * cnm;
      }
      aaxd[jks] = s;
    }
  spharot(aaxd, out, mp, T->dnm + (size_t)(je + NRBM) * mp * mp2);
}

/* kernel/G_l2p.cxx:220-269 for one box: returns unscaled (g, gx, gy, gz) sums
 * through the callback arrays with the 1/(4pi) factor and float rounding. */
/* outputs are float arrays in the reference (include/arrays.h:6): every += rounds */
static int g_fp64_out = 0;
static inline double oround(int conv, double v) { return (conv == ORC_CONV_REF && !g_fp64_out) ? (double)(float)v : v; }

static void l2p_G(const tables* T, int i0, int i1, const float* x, const float* y, const float* z,
                  double cx, double cy, double cz, const cplx* aax, int want_grad,
                  double* v, double* gxo, double* gyo, double* gzo) {
  int mp = T->mp, conv = T->conv;
  cplx bnm[4 * 32 * 32], bth[32 * 32];
  const cplx I1 = mkc(0.0, 1.0);
  for (int i = i0; i <= i1; i++) {
    double dx = x[i] - cx, dy = y[i] - cy, dz = z[i] - cz, r, th, ph;
    angles(conv, dx, dy, dz, &r, &th, &ph);
    double g = 0, gr = 0, gth = 0, gph = 0;
    multipoled(T->fac, th, mp, bnm, bth);
    for (int n = 0; n < mp; n++) {
      int nm = n * n + n, nms = n * (n + 1) / 2;
      cplx rs = pow(r, n) * bnm[nm];
      cplx rr = n * pow(r, n - 1) * bnm[nm];
      cplx rth = pow(r, n) * bth[nm];
      g += creal(rs * aax[nms]);
      gr += creal(rr * aax[nms]);
      gth += creal(rth * aax[nms]);
      for (int m = 1; m <= n; m++) {
        nm = n * n + n + m;
        nms = n * (n + 1) / 2 + m;
        cplx eim = cis(m * ph);
        rs = (pow(r, n) * bnm[nm]) * eim;
        rr = ((n * pow(r, n - 1)) * bnm[nm]) * eim;
        rth = (pow(r, n) * bth[nm]) * eim;
        cplx rph = (((m * pow(r, n)) * bnm[nm]) * eim) * I1;
        g += 2 * creal(rs * aax[nms]);
        gr += 2 * creal(rr * aax[nms]);
        gth += 2 * creal(rth * aax[nms]);
        gph += 2 * creal(rph * aax[nms]);
      }
    }
    double gx = sin(th) * cos(ph) * gr + cos(th) * cos(ph) / r * gth - sin(ph) / r / sin(th) * gph;
    double gy = sin(th) * sin(ph) * gr + cos(th) * sin(ph) / r * gth + cos(ph) / r / sin(th) * gph;
    double gz = cos(th) * gr - sin(th) / r * gth;
    double c4 = 0.25 / PIF;
    if (conv == ORC_CONV_EXACT) c4 = 0.25 / M_PI;
    if (want_grad) {
      gxo[i] = oround(conv, gxo[i] + c4 * gx);
      gyo[i] = oround(conv, gyo[i] + c4 * gy);
      gzo[i] = oround(conv, gzo[i] + c4 * gz);
    }
    v[i] = oround(conv, v[i] + c4 * g);
  }
}

/* kernel/Gn_l2p.cxx:300-342: vi -= n_i . grad(phi) / (4 pi) */
static void l2p_Gn(const tables* T, int i0, int i1, const float* x, const float* y, const float* z,
                   const float* nx, const float* ny, const float* nz, double cx, double cy, double cz,
                   const cplx* aax, double* v) {
  int mp = T->mp, conv = T->conv;
  cplx bnm[4 * 32 * 32], bth[32 * 32];
  const cplx I1 = mkc(0.0, 1.0);
  for (int i = i0; i <= i1; i++) {
    double dx = x[i] - cx, dy = y[i] - cy, dz = z[i] - cz, r, th, ph;
    angles(conv, dx, dy, dz, &r, &th, &ph);
    double gr = 0, gth = 0, gph = 0;
    multipoled(T->fac, th, mp, bnm, bth);
    for (int n = 0; n < mp; n++) {
      int nm = n * n + n, nms = n * (n + 1) / 2;
      cplx rr = n * pow(r, n - 1) * bnm[nm];
      cplx rth = pow(r, n) * bth[nm];
      gr += creal(rr * aax[nms]);
      gth += creal(rth * aax[nms]);
      for (int m = 1; m <= n; m++) {
        nm = n * n + n + m;
        nms = n * (n + 1) / 2 + m;
        cplx eim = cis(m * ph);
        rr = ((n * pow(r, n - 1)) * bnm[nm]) * eim;
        rth = (pow(r, n) * bth[nm]) * eim;
        cplx rph = (((m * pow(r, n)) * bnm[nm]) * eim) * I1;
        gr += 2 * creal(rr * aax[nms]);
        gth += 2 * creal(rth * aax[nms]);
        gph += 2 * creal(rph * aax[nms]);
      }
    }
    double gx = sin(th) * cos(ph) * gr + cos(th) * cos(ph) / r * gth - sin(ph) / r / sin(th) * gph;
    double gy = sin(th) * sin(ph) * gr + cos(th) * sin(ph) / r * gth + cos(ph) / r / sin(th) * gph;
    double gz = cos(th) * gr - sin(th) / r * gth;
    double c4 = conv == ORC_CONV_EXACT ? 0.25 / M_PI : 0.25 / PIF;
    v[i] = oround(conv, v[i] - c4 * (nx[i] * gx + ny[i] * gy + nz[i] * gz));
  }
}

/* ------------------------------------------------------------------ */
/* Pair kernels: kernel/G_p2p.cxx:20-30, Gni_p2p.cxx:58-65, Gnj_p2p.cxx:90-97 */
/* ------------------------------------------------------------------ */

typedef struct pacc { double p, fx, fy, fz; } pacc;

/* Accumulate targets i against sources [j0,j1]; the reference's per
 * (target, source box) double partial sum, added afterwards. */
static pacc pair_sum(int neq, int conv, int direct, float xi, float yi, float zi, float nxi, float nyi, float nzi,
                     int j0, int j1, const float* xj, const float* yj, const float* zj,
                     const float* nxj, const float* nyj, const float* nzj, const float* q) {
  pacc a = {0, 0, 0, 0};
  const double c4r = 0.25 / PIF, c4e = 0.25 / M_PI;
  for (int j = j0; j <= j1; j++) {
    double dx, dy, dz, rij;
    if (conv == ORC_CONV_REF) {
      dx = xi - xj[j]; dy = yi - yj[j]; dz = zi - zj[j]; /* float subtraction */
      rij = dx * dx + dy * dy + dz * dz + EPSF;
    } else {
      dx = (double)xi - xj[j]; dy = (double)yi - yj[j]; dz = (double)zi - zj[j];
      rij = dx * dx + dy * dy + dz * dz;
      if (rij == 0) continue;
    }
    if (neq == ORC_NEQ_G) {
      double c4 = conv == ORC_CONV_REF ? c4r : c4e;
      double rs = c4 * q[j] / rij / sqrt(rij);
      a.p += c4 * q[j] / sqrt(rij);
      a.fx -= dx * rs;
      a.fy -= dy * rs;
      a.fz -= dz * rs;
    } else {
      /* the P2P kernels divide by sqrt(rij) first, the direct sums by rij */
      double c4 = conv == ORC_CONV_REF ? c4r : c4e;
      double rs = direct ? c4 * q[j] / rij / sqrt(rij) : c4 * q[j] / sqrt(rij) / rij;
      if (neq == ORC_NEQ_GNI)
        a.p += (dx * nxi + dy * nyi + dz * nzi) * rs;
      else
        a.p += (dx * nxj[j] + dy * nyj[j] + dz * nzj[j]) * rs;
    }
  }
  return a;
}

/* ------------------------------------------------------------------ */
/* Tree, lists and driver                                              */
/* ------------------------------------------------------------------ */

struct orc_fmm {
  orc_opts o;
  int ni, nj, nbmax;
  /* sorted particle data (reference sorti/sortj) */
  float *xi, *yi, *zi, *nxi, *nyi, *nzi;
  float *xj, *yj, *zj, *nxj, *nyj, *nzj;
  int *nbi, *nbj;
  /* level-stacked box keys (leaf first) and level offsets, per side */
  int *nfi, *nfj, *nlbi, *nlbj; /* nlb*[lev-1] = offset of level lev */
  int *lbi_lev, *lbj_lev;       /* count of boxes at each level [lmax+1] */
  int nbneti, nbnetj, nbne;
  int *ndi0, *ndi1, *ndj0, *ndj1; /* leaf particle ranges */
  /* lists: near (leaf) and far per level */
  int *nij_near, *neij_near;
  int **nij_far, **neij_far; /* [lmax+1] */
  tables T;
  cplx *bx, *bx_dump, *ax, *ax_leaf;
  cplx** ax_lev; /* local expansions after each level's M2L */
  double t[8];
};

static int* dense_index(int nkeys, const int* keys, int nb) {
  int* nej = (int*)malloc(sizeof(int) * (size_t)nb);
  for (int i = 0; i < nb; i++) nej[i] = -1;
  for (int i = 0; i < nkeys; i++) nej[keys[i]] = i;
  return nej;
}

/* fmm/ijbox.cxx:9-117 for one level; neib 0 near list, 1 all non-adjacent,
 * 2 children of the parent's neighbours minus the adjacent ones */
static void ijbox(const int* fi, int lbi, const int* fj, int lbj, int lev, int neib, int* nij, int* neij) {
  int jxmin = 1000000, jxmax = -1000000, jymin = 1000000, jymax = -1000000, jzmin = 1000000, jzmax = -1000000;
  int nc[3];
  for (int jj = 0; jj < lbj; jj++) {
    boxc(fj[jj], nc);
    if (nc[0] < jxmin) jxmin = nc[0];
    if (nc[0] > jxmax) jxmax = nc[0];
    if (nc[1] < jymin) jymin = nc[1];
    if (nc[1] > jymax) jymax = nc[1];
    if (nc[2] < jzmin) jzmin = nc[2];
    if (nc[2] > jzmax) jzmax = nc[2];
  }
  int* nej = dense_index(lbj, fj, 1 << (3 * lev));
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#define MIN(a, b) ((a) < (b) ? (a) : (b))
  for (int ii = 0; ii < lbi; ii++) {
    int cnt = 0, *L = neij + (size_t)ii * NEBM;
    boxc(fi[ii], nc);
    int ix = nc[0], iy = nc[1], iz = nc[2];
    if (neib == 0) {
      for (int jx = MAX(ix - 1, jxmin); jx <= MIN(ix + 1, jxmax); jx++)
        for (int jy = MAX(iy - 1, jymin); jy <= MIN(iy + 1, jymax); jy++)
          for (int jz = MAX(iz - 1, jzmin); jz <= MIN(iz + 1, jzmax); jz++) {
            int jj = nej[boxn1(jx, jy, jz, lev)];
            if (jj != -1) L[cnt++] = jj;
          }
    } else if (neib == 1) {
      for (int jj = 0; jj < lbj; jj++) {
        boxc(fj[jj], nc);
        int jx = nc[0], jy = nc[1], jz = nc[2];
        if (jx < ix - 1 || ix + 1 < jx || jy < iy - 1 || iy + 1 < jy || jz < iz - 1 || iz + 1 < jz) L[cnt++] = jj;
      }
    } else {
      int ixp = (ix + 2) / 2, iyp = (iy + 2) / 2, izp = (iz + 2) / 2;
      for (int jxp = ixp - 1; jxp <= ixp + 1; jxp++)
        for (int jyp = iyp - 1; jyp <= iyp + 1; jyp++)
          for (int jzp = izp - 1; jzp <= izp + 1; jzp++)
            for (int jx = MAX(2 * jxp - 2, jxmin); jx <= MIN(2 * jxp - 1, jxmax); jx++)
              for (int jy = MAX(2 * jyp - 2, jymin); jy <= MIN(2 * jyp - 1, jymax); jy++)
                for (int jz = MAX(2 * jzp - 2, jzmin); jz <= MIN(2 * jzp - 1, jzmax); jz++)
                  if (jx < ix - 1 || ix + 1 < jx || jy < iy - 1 || iy + 1 < jy || jz < iz - 1 || iz + 1 < jz) {
                    int jj = nej[boxn1(jx, jy, jz, lev)];
                    if (jj != -1) L[cnt++] = jj;
                  }
    }
    nij[ii] = cnt;
  }
  free(nej);
}

static float* gather_f(int n, const float* src, const int* perm) {
  float* d = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) d[i] = src ? src[perm[i]] : 0.0f;
  return d;
}

/* fmm/boxdatai.cxx:10-34 / boxdataj.cxx:10-36 and boxm2mi/j.cxx:6-26 */
static void build_side(int n, const int* key, int lmax, int** nf_out, int** nlb_out, int* lbl,
                       int** nd0, int** nd1, int* nbnet) {
  int cap = n + 8;
  int* nf = (int*)malloc(sizeof(int) * (size_t)cap * (lmax + 1));
  int* nlb = (int*)calloc((size_t)lmax + 1, sizeof(int));
  int* d0 = (int*)malloc(sizeof(int) * (size_t)cap);
  int* d1 = (int*)malloc(sizeof(int) * (size_t)cap);
  int lb = 0, nbc = -1;
  for (int i = 0; i < n; i++)
    if (key[i] != nbc) {
      nf[lb] = key[i];
      d0[lb] = i;
      if (lb > 0) d1[lb - 1] = i - 1;
      nbc = key[i];
      lb++;
    }
  if (lb > 0) d1[lb - 1] = n - 1;
  nlb[lmax - 1] = 0;
  lbl[lmax] = lb;
  for (int lev = lmax - 1; lev >= 2; lev--) {
    nlb[lev - 1] = nlb[lev] + lb;
    int lbo = lb;
    lb = 0;
    nbc = -1;
    for (int ii = 0; ii < lbo; ii++) {
      int k = nf[ii + nlb[lev]] / 8;
      if (nbc != k) {
        nbc = k;
        nf[lb + nlb[lev - 1]] = k;
        lb++;
      }
    }
    lbl[lev] = lb;
  }
  *nbnet = (lmax > 2 ? nlb[1] : 0) + lbl[2];
  *nf_out = nf;
  *nlb_out = nlb;
  *nd0 = d0;
  *nd1 = d1;
}

orc_fmm* orc_new(const orc_opts* o, int ni, const float* xi, const float* yi, const float* zi,
                 const float* nxi, const float* nyi, const float* nzi, int nj, const float* xj,
                 const float* yj, const float* zj, const float* nxj, const float* nyj, const float* nzj) {
  if (!o || o->lmax < 2 || o->lmax > 10 || o->mp < 1 || o->mp > 30 || ni < 0 || nj < 0) return 0;
  orc_fmm* f = (orc_fmm*)calloc(1, sizeof(orc_fmm));
  f->o = *o;
#ifdef _OPENMP
  if (o->nthreads > 0) omp_set_num_threads(o->nthreads);
#endif
  double t0 = now_s();
  int lmax = o->lmax;
  f->ni = ni;
  f->nj = nj;
  f->nbmax = 1 << (3 * lmax);
  /* sort/sorti.cxx:29-52, sort/sortj.cxx:78-102 */
  int* ki = (int*)malloc(sizeof(int) * (size_t)(ni + 1));
  int* kj = (int*)malloc(sizeof(int) * (size_t)(nj + 1));
  f->nbi = (int*)malloc(sizeof(int) * (size_t)(ni + 1));
  f->nbj = (int*)malloc(sizeof(int) * (size_t)(nj + 1));
  boxn(ni, xi, yi, zi, lmax, o->xmin, o->ymin, o->zmin, o->rd, ki);
  csort(ni, f->nbmax, ki, f->nbi);
  boxn(nj, xj, yj, zj, lmax, o->xmin, o->ymin, o->zmin, o->rd, kj);
  csort(nj, f->nbmax, kj, f->nbj);
  f->xi = gather_f(ni, xi, f->nbi); f->yi = gather_f(ni, yi, f->nbi); f->zi = gather_f(ni, zi, f->nbi);
  f->nxi = gather_f(ni, nxi, f->nbi); f->nyi = gather_f(ni, nyi, f->nbi); f->nzi = gather_f(ni, nzi, f->nbi);
  f->xj = gather_f(nj, xj, f->nbj); f->yj = gather_f(nj, yj, f->nbj); f->zj = gather_f(nj, zj, f->nbj);
  f->nxj = gather_f(nj, nxj, f->nbj); f->nyj = gather_f(nj, nyj, f->nbj); f->nzj = gather_f(nj, nzj, f->nbj);
  /* box lists */
  f->lbi_lev = (int*)calloc((size_t)lmax + 1, sizeof(int));
  f->lbj_lev = (int*)calloc((size_t)lmax + 1, sizeof(int));
  build_side(ni, ki, lmax, &f->nfi, &f->nlbi, f->lbi_lev, &f->ndi0, &f->ndi1, &f->nbneti);
  build_side(nj, kj, lmax, &f->nfj, &f->nlbj, f->lbj_lev, &f->ndj0, &f->ndj1, &f->nbnetj);
  free(ki);
  free(kj);
  f->nbne = f->lbi_lev[lmax] > f->lbj_lev[lmax] ? f->lbi_lev[lmax] : f->lbj_lev[lmax];
  /* lists */
  int lbi = f->lbi_lev[lmax], lbj = f->lbj_lev[lmax];
  f->nij_near = (int*)calloc((size_t)lbi + 1, sizeof(int));
  f->neij_near = (int*)malloc(sizeof(int) * ((size_t)lbi + 1) * NEBM);
  ijbox(f->nfi, lbi, f->nfj, lbj, lmax, 0, f->nij_near, f->neij_near);
  f->nij_far = (int**)calloc((size_t)lmax + 1, sizeof(int*));
  f->neij_far = (int**)calloc((size_t)lmax + 1, sizeof(int*));
  for (int lev = 2; lev <= lmax; lev++) {
    int li = f->lbi_lev[lev], lj = f->lbj_lev[lev];
    f->nij_far[lev] = (int*)calloc((size_t)li + 1, sizeof(int));
    f->neij_far[lev] = (int*)malloc(sizeof(int) * ((size_t)li + 1) * NEBM);
    /* level 2: fmm/fmm.cxx:104 asks for neib 3, which fmm/ijbox.cxx never
       builds; the set it needs (all non-adjacent boxes) is neib 1. */
    ijbox(f->nfi + f->nlbi[lev - 1], li, f->nfj + f->nlbj[lev - 1], lj, lev, lev == 2 ? 1 : 2,
          f->nij_far[lev], f->neij_far[lev]);
  }
  f->t[7] = now_s() - t0;
  t0 = now_s();
  tables_build(&f->T, o->mp, o->conv);
  f->t[6] = now_s() - t0;
  int nmp = o->mp * (o->mp + 1) / 2;
  f->bx = (cplx*)calloc((size_t)(f->nbnetj + 1) * nmp, sizeof(cplx));
  f->bx_dump = (cplx*)calloc((size_t)(f->nbnetj + 1) * nmp, sizeof(cplx));
  f->ax = (cplx*)calloc((size_t)(f->nbne + 8) * nmp, sizeof(cplx));
  f->ax_leaf = (cplx*)calloc((size_t)(f->nbne + 8) * nmp, sizeof(cplx));
  f->ax_lev = (cplx**)calloc((size_t)lmax + 1, sizeof(cplx*));
  return f;
}

void orc_free(orc_fmm* f) {
  if (!f) return;
  free(f->xi); free(f->yi); free(f->zi); free(f->nxi); free(f->nyi); free(f->nzi);
  free(f->xj); free(f->yj); free(f->zj); free(f->nxj); free(f->nyj); free(f->nzj);
  free(f->nbi); free(f->nbj); free(f->nfi); free(f->nfj); free(f->nlbi); free(f->nlbj);
  free(f->lbi_lev); free(f->lbj_lev); free(f->ndi0); free(f->ndi1); free(f->ndj0); free(f->ndj1);
  free(f->nij_near); free(f->neij_near);
  for (int l = 0; l <= f->o.lmax; l++) {
    if (f->nij_far) free(f->nij_far[l]);
    if (f->neij_far) free(f->neij_far[l]);
  }
  free(f->nij_far); free(f->neij_far);
  tables_free(&f->T);
  free(f->bx); free(f->bx_dump); free(f->ax); free(f->ax_leaf);
  for (int l = 0; l <= f->o.lmax; l++) free(f->ax_lev[l]);
  free(f->ax_lev);
  free(f);
}

/* fmm/fmm.cxx:29-152 */
int orc_run(orc_fmm* f, const float* qj_in, double* vo, double* gxo, double* gyo, double* gzo) {
  const orc_opts* o = &f->o;
  const tables* T = &f->T;
  int lmax = o->lmax, mp = o->mp, nmp = mp * (mp + 1) / 2, neq = o->neq, conv = o->conv;
  int ni = f->ni, nj = f->nj;
  float* q = gather_f(nj, qj_in, f->nbj);
  g_fp64_out = o->fp64_out;
  double* v = (double*)calloc((size_t)ni + 1, sizeof(double));
  double* gx = (double*)calloc((size_t)ni + 1, sizeof(double));
  double* gy = (double*)calloc((size_t)ni + 1, sizeof(double));
  double* gz = (double*)calloc((size_t)ni + 1, sizeof(double));
  double t0;
  for (int k = 0; k < 6; k++) f->t[k] = 0;

  /* Step 1: P2P at lmax (fmm.cxx:44-65; kernel/G_p2p.cxx:12-36) */
  t0 = now_s();
  int lbi = f->lbi_lev[lmax];
#pragma omp parallel for schedule(dynamic, 4)
  for (int ii = 0; ii < lbi; ii++) {
    for (int ij = 0; ij < f->nij_near[ii]; ij++) {
      int jj = f->neij_near[(size_t)ii * NEBM + ij];
      for (int i = f->ndi0[ii]; i <= f->ndi1[ii]; i++) {
        pacc a = pair_sum(neq, conv, 0, f->xi[i], f->yi[i], f->zi[i], f->nxi[i], f->nyi[i], f->nzi[i],
                          f->ndj0[jj], f->ndj1[jj], f->xj, f->yj, f->zj, f->nxj, f->nyj, f->nzj, q);
        if (neq == ORC_NEQ_G) {
          gx[i] = oround(conv, gx[i] + a.fx);
          gy[i] = oround(conv, gy[i] + a.fy);
          gz[i] = oround(conv, gz[i] + a.fz);
        }
        v[i] = oround(conv, v[i] + a.p);
      }
    }
  }
  f->t[0] = now_s() - t0;

  /* Step 2: P2M at lmax (fmm.cxx:71-77) */
  t0 = now_s();
  int lbj = f->lbj_lev[lmax];
  double rbl = o->rd / (double)(1 << lmax);
  memset(f->bx, 0, sizeof(cplx) * (size_t)(f->nbnetj + 1) * nmp);
#pragma omp parallel for schedule(dynamic, 4)
  for (int jj = 0; jj < lbj; jj++) {
    int nc[3];
    boxc(f->nfj[jj], nc);
    double cx = o->xmin + (nc[0] + 0.5) * rbl, cy = o->ymin + (nc[1] + 0.5) * rbl, cz = o->zmin + (nc[2] + 0.5) * rbl;
    if (neq == ORC_NEQ_GNJ)
      p2m_Gn(T, f->ndj0[jj], f->ndj1[jj], f->xj, f->yj, f->zj, f->nxj, f->nyj, f->nzj, q, cx, cy, cz,
             f->bx + (size_t)jj * nmp);
    else
      p2m_G(T, f->ndj0[jj], f->ndj1[jj], f->xj, f->yj, f->zj, q, cx, cy, cz, f->bx + (size_t)jj * nmp);
  }
  f->t[1] = now_s() - t0;

  /* Step 3: M2M lmax-1 .. 2 (fmm.cxx:81-91; kernel/m2m.cxx:115-155) */
  t0 = now_s();
  for (int lev = lmax - 1; lev >= 2; lev--) {
    int lbp = f->lbj_lev[lev], lbc = f->lbj_lev[lev + 1];
    int offp = f->nlbj[lev - 1], offc = f->nlbj[lev];
    double rb = o->rd / (double)(1 << lev), rh = rb * sqrt(3.0) / 4;
    /* children of one parent are consecutive: first child of each parent */
    int* first = (int*)malloc(sizeof(int) * ((size_t)lbp + 1));
    int p = -1;
    for (int jj = 0; jj < lbc; jj++) {
      int k = f->nfj[offc + jj] / 8;
      if (p < 0 || f->nfj[offp + p] != k) first[++p] = jj;
    }
    first[lbp] = lbc;
#pragma omp parallel for schedule(dynamic, 4)
    for (int pp = 0; pp < lbp; pp++) {
      cplx* dst = f->bx + (size_t)(offp + pp) * nmp;
      cplx tmp[32 * 33 / 2];
      for (int j = 0; j < nmp; j++) dst[j] = 0;
      for (int jj = first[pp]; jj < first[pp + 1]; jj++) {
        int nc[3];
        boxc(f->nfj[offc + jj] % 8, nc);
        int je = boxn1(4 - nc[0] * 2, 4 - nc[1] * 2, 4 - nc[2] * 2, 3);
        m2m_one(T, je, rh, f->bx + (size_t)(offc + jj) * nmp, tmp);
        for (int j = 0; j < nmp; j++) dst[j] += tmp[j];
      }
    }
    free(first);
  }
  f->t[2] = now_s() - t0;
  memcpy(f->bx_dump, f->bx, sizeof(cplx) * (size_t)(f->nbnetj + 1) * nmp);

  /* Steps 4-6: M2L at 2, then L2L + M2L for 3..lmax (fmm.cxx:104-138) */
  for (int lev = 2; lev <= lmax; lev++) {
    int li = f->lbi_lev[lev];
    const int* fi = f->nfi + f->nlbi[lev - 1];
    const int* fj = f->nfj + f->nlbj[lev - 1];
    double rb = o->rd / (double)(1 << lev);
    if (lev == 2) {
      memset(f->ax, 0, sizeof(cplx) * (size_t)li * nmp);
    } else {
      /* kernel/l2l.cxx:172-243: parents at lev-1 hold ax; children overwrite */
      t0 = now_s();
      int lpar = f->lbi_lev[lev - 1];
      cplx* axo = (cplx*)malloc(sizeof(cplx) * ((size_t)lpar + 1) * nmp);
      memcpy(axo, f->ax, sizeof(cplx) * (size_t)lpar * nmp);
      const int* fp = f->nfi + f->nlbi[lev - 2];
      int* neo = dense_index(lpar, fp, 1 << (3 * (lev - 1)));
      double rh = rb * sqrt(3.0) / 2;
#pragma omp parallel for schedule(dynamic, 8)
      for (int ii = 0; ii < li; ii++) {
        int nc[3];
        boxc(fi[ii] % 8, nc);
        int je = boxn1(nc[0] * 2 + 2, nc[1] * 2 + 2, nc[2] * 2 + 2, 3);
        l2l_one(T, je, rh, axo + (size_t)neo[fi[ii] / 8] * nmp, f->ax + (size_t)ii * nmp);
      }
      free(neo);
      free(axo);
      f->t[4] += now_s() - t0;
    }
    t0 = now_s();
    /* kernel/m2l.cxx:30-83 */
#pragma omp parallel for schedule(dynamic, 4)
    for (int ii = 0; ii < li; ii++) {
      int nc[3];
      boxc(fi[ii], nc);
      int ix = nc[0], iy = nc[1], iz = nc[2];
      cplx tmp[32 * 33 / 2];
      cplx* dst = f->ax + (size_t)ii * nmp;
      for (int ij = 0; ij < f->nij_far[lev][ii]; ij++) {
        int jj = f->neij_far[lev][(size_t)ii * NEBM + ij];
        boxc(fj[jj], nc);
        int jx = nc[0], jy = nc[1], jz = nc[2];
        double xijc = (ix - jx) * rb, yijc = (iy - jy) * rb, zijc = (iz - jz) * rb;
        int je = boxn1(ix - jx + 3, iy - jy + 3, iz - jz + 3, 3);
        double rh = sqrt(xijc * xijc + yijc * yijc + zijc * zijc) + (conv == ORC_CONV_REF ? EPSF : 0.0);
        m2l_one(T, je, rh, f->bx + (size_t)(jj + f->nlbj[lev - 1]) * nmp, tmp);
        for (int j = 0; j < nmp; j++) dst[j] += tmp[j];
      }
    }
    f->t[3] += now_s() - t0;
    free(f->ax_lev[lev]);
    f->ax_lev[lev] = (cplx*)malloc(sizeof(cplx) * ((size_t)li + 1) * nmp);
    memcpy(f->ax_lev[lev], f->ax, sizeof(cplx) * (size_t)li * nmp);
  }
  memcpy(f->ax_leaf, f->ax, sizeof(cplx) * (size_t)lbi * nmp);

  /* Step 7: L2P at lmax (fmm.cxx:144-150) */
  t0 = now_s();
#pragma omp parallel for schedule(dynamic, 4)
  for (int ii = 0; ii < lbi; ii++) {
    int nc[3];
    boxc(f->nfi[ii], nc);
    double cx = o->xmin + (nc[0] + 0.5) * rbl, cy = o->ymin + (nc[1] + 0.5) * rbl, cz = o->zmin + (nc[2] + 0.5) * rbl;
    const cplx* a = f->ax + (size_t)ii * nmp;
    if (neq == ORC_NEQ_GNI)
      l2p_Gn(T, f->ndi0[ii], f->ndi1[ii], f->xi, f->yi, f->zi, f->nxi, f->nyi, f->nzi, cx, cy, cz, a, v);
    else
      l2p_G(T, f->ndi0[ii], f->ndi1[ii], f->xi, f->yi, f->zi, cx, cy, cz, a, neq == ORC_NEQ_G, v, gx, gy, gz);
  }
  f->t[5] = now_s() - t0;
  g_fp64_out = 0;

  /* sort/sorti.cxx:54-66: back to caller order */
  for (int i = 0; i < ni; i++) {
    int d = f->nbi[i];
    if (vo) vo[d] = v[i];
    if (gxo) gxo[d] = gx[i];
    if (gyo) gyo[d] = gy[i];
    if (gzo) gzo[d] = gz[i];
  }
  free(q); free(v); free(gx); free(gy); free(gz);
  return 0;
}

void orc_times(const orc_fmm* f, double* t8) {
  for (int k = 0; k < 8; k++) t8[k] = f->t[k];
}

int orc_info(const orc_fmm* f, int what, int lev) {
  switch (what) {
    case 0: return f->nbne;
    case 1: return f->nbnetj;
    case 2: return f->o.mp * (f->o.mp + 1) / 2;
    case 3: return (lev >= 2 && lev <= f->o.lmax) ? f->lbi_lev[lev] : -1;
    case 4: return (lev >= 2 && lev <= f->o.lmax) ? f->lbj_lev[lev] : -1;
    case 5: return f->nbneti;
  }
  return -1;
}

void orc_perm(const orc_fmm* f, int* nbi, int* nbj) {
  if (nbi) memcpy(nbi, f->nbi, sizeof(int) * (size_t)f->ni);
  if (nbj) memcpy(nbj, f->nbj, sizeof(int) * (size_t)f->nj);
}

static void boxes(const orc_fmm* f, int side, int lev, int* keys, int* first, int* last) {
  int lmax = f->o.lmax;
  const int* nf = side ? f->nfj : f->nfi;
  const int* nlb = side ? f->nlbj : f->nlbi;
  int lb = side ? f->lbj_lev[lev] : f->lbi_lev[lev];
  if (keys) memcpy(keys, nf + nlb[lev - 1], sizeof(int) * (size_t)lb);
  if (lev == lmax) {
    if (first) memcpy(first, side ? f->ndj0 : f->ndi0, sizeof(int) * (size_t)lb);
    if (last) memcpy(last, side ? f->ndj1 : f->ndi1, sizeof(int) * (size_t)lb);
  }
}
void orc_boxes_i(const orc_fmm* f, int lev, int* k, int* a, int* b) { boxes(f, 0, lev, k, a, b); }
void orc_boxes_j(const orc_fmm* f, int lev, int* k, int* a, int* b) { boxes(f, 1, lev, k, a, b); }

void orc_list(const orc_fmm* f, int lev, int kind, int* nij, int* neij) {
  int li = f->lbi_lev[lev];
  const int* n = kind == 0 ? f->nij_near : f->nij_far[lev];
  const int* e = kind == 0 ? f->neij_near : f->neij_far[lev];
  memcpy(nij, n, sizeof(int) * (size_t)li);
  memcpy(neij, e, sizeof(int) * (size_t)li * NEBM);
}

void orc_bx(const orc_fmm* f, double* out) {
  int nmp = f->o.mp * (f->o.mp + 1) / 2;
  memcpy(out, f->bx_dump, sizeof(cplx) * (size_t)f->nbnetj * nmp);
}
void orc_ax(const orc_fmm* f, double* out) {
  int nmp = f->o.mp * (f->o.mp + 1) / 2;
  memcpy(out, f->ax_leaf, sizeof(cplx) * (size_t)f->lbi_lev[f->o.lmax] * nmp);
}

void orc_ax_level(const orc_fmm* f, int lev, double* out) {
  int nmp = f->o.mp * (f->o.mp + 1) / 2;
  if (f->ax_lev[lev]) memcpy(out, f->ax_lev[lev], sizeof(cplx) * (size_t)f->lbi_lev[lev] * nmp);
}

void orc_precalc(int mp, int conv, double* fac, double* anm, double* sr, double* ynm, double* dnm) {
  tables T;
  tables_build(&T, mp, conv);
  int mp2 = mp * mp;
  if (fac) memcpy(fac, T.fac, sizeof(double) * 4 * mp2);
  if (anm) memcpy(anm, T.anm, sizeof(double) * 4 * mp2);
  if (sr) memcpy(sr, T.sr, sizeof(double) * mp2 * mp2);
  if (ynm) memcpy(ynm, T.ynm, sizeof(cplx) * 4 * mp2);
  if (dnm) memcpy(dnm, T.dnm, sizeof(cplx) * (size_t)2 * NRBM * mp * mp2);
  tables_free(&T);
}

/* kernel/G_dir.cxx:6-30, Gni_dir.cxx:6-21, Gnj_dir.cxx:6-21 */
void orc_direct(int neq, int conv, int nthreads, int ni, const float* xi, const float* yi, const float* zi,
                const float* nxi, const float* nyi, const float* nzi, int nj, const float* xj,
                const float* yj, const float* zj, const float* nxj, const float* nyj, const float* nzj,
                const float* qj, double* vi, double* gxi, double* gyi, double* gzi) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < ni; i++) {
    pacc a = pair_sum(neq, conv, 1, xi[i], yi[i], zi[i], nxi ? nxi[i] : 0.f, nyi ? nyi[i] : 0.f,
                      nzi ? nzi[i] : 0.f, 0, nj - 1, xj, yj, zj, nxj, nyj, nzj, qj);
    vi[i] = oround(conv, a.p);
    if (neq == ORC_NEQ_G) {
      if (gxi) gxi[i] = oround(conv, a.fx);
      if (gyi) gyi[i] = oround(conv, a.fy);
      if (gzi) gzi[i] = oround(conv, a.fz);
    }
  }
}

void orc_direct_helmholtz(int conv, int nthreads, double k, int ni, const float* xi, const float* yi,
                          const float* zi, int nj, const float* xj, const float* yj, const float* zj,
                          const float* qr, const float* qi, double* vr, double* vim) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const double c = conv == ORC_CONV_REF ? 0.25 / (double)PIF : 0.25 / M_PI;
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < ni; i++) {
    double sr = 0, si = 0;
    for (int j = 0; j < nj; j++) {
      double r2;
      if (conv == ORC_CONV_REF) {
        double dx = xi[i] - xj[j], dy = yi[i] - yj[j], dz = zi[i] - zj[j];
        r2 = dx * dx + dy * dy + dz * dz + EPSF;
      } else {
        double dx = (double)xi[i] - xj[j], dy = (double)yi[i] - yj[j], dz = (double)zi[i] - zj[j];
        r2 = dx * dx + dy * dy + dz * dz;
        if (r2 == 0) continue;
      }
      double r = sqrt(r2), ck = cos(k * r) / r, sk = sin(k * r) / r;
      sr += qr[j] * ck - qi[j] * sk;
      si += qr[j] * sk + qi[j] * ck;
    }
    vr[i] = c * sr;
    vim[i] = c * si;
  }
}

void orc_direct_yukawa(int conv, int nthreads, double lambda, int ni, const float* xi, const float* yi,
                       const float* zi, int nj, const float* xj, const float* yj, const float* zj,
                       const float* qj, double* vi, double* gxi, double* gyi, double* gzi) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < ni; i++) {
    double p = 0, fx = 0, fy = 0, fz = 0;
    for (int j = 0; j < nj; j++) {
      if (conv == ORC_CONV_REF) {
        /* kernel/Helmholtz_G_dir.cxx:13-17, float subtraction, the reference's order */
        double dx = xi[i] - xj[j], dy = yi[i] - yj[j], dz = zi[i] - zj[j];
        double r = sqrt(dx * dx + dy * dy + dz * dz + EPSF);
        p += 0.5 * PIF * qj[j] / r / lambda * exp(-lambda * r);
      } else {
        double dx = (double)xi[i] - xj[j], dy = (double)yi[i] - yj[j], dz = (double)zi[i] - zj[j];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 == 0) continue;
        double r = sqrt(r2), e = exp(-lambda * r) / r;
        p += qj[j] * e;
        double s = -qj[j] * (1 + lambda * r) * e / r2;
        fx += s * dx;
        fy += s * dy;
        fz += s * dz;
      }
    }
    if (conv == ORC_CONV_REF) {
      vi[i] = oround(conv, p);
    } else {
      double c = 0.5 * M_PI / lambda;
      vi[i] = c * p;
      if (gxi) gxi[i] = c * fx;
      if (gyi) gyi[i] = c * fy;
      if (gzi) gzi[i] = c * fz;
    }
  }
}

void orc_p2m_box(int mp, int conv, int n, const float* x, const float* y, const float* z, const float* q,
                 double cx, double cy, double cz, double* m) {
  tables T;
  tables_build(&T, mp, conv);
  p2m_G(&T, 0, n - 1, x, y, z, q, cx, cy, cz, (cplx*)m);
  tables_free(&T);
}

void orc_l2p_box(int mp, int conv, int n, const float* x, const float* y, const float* z, double cx,
                 double cy, double cz, const double* l, double* v, double* gx, double* gy, double* gz) {
  tables T;
  tables_build(&T, mp, conv);
  for (int i = 0; i < n; i++) v[i] = gx[i] = gy[i] = gz[i] = 0;
  l2p_G(&T, 0, n - 1, x, y, z, cx, cy, cz, (const cplx*)l, 1, v, gx, gy, gz);
  tables_free(&T);
}

void orc_m2l_pair(int mp, int conv, int dx, int dy, int dz, double rb, const double* m, double* l) {
  tables T;
  tables_build(&T, mp, conv);
  int je = boxn1(dx + 3, dy + 3, dz + 3, 3);
  double xijc = dx * rb, yijc = dy * rb, zijc = dz * rb;
  double rh = sqrt(xijc * xijc + yijc * yijc + zijc * zijc) + (conv == ORC_CONV_REF ? EPSF : 0.0);
  m2l_one(&T, je, rh, (const cplx*)m, (cplx*)l);
  tables_free(&T);
}
