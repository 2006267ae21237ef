// refshim.cpp -- the reference's C++ symbols (fmm, nlevel, boxpart, sorti,
// sortj, unsorti, unsortj, boxallocate, alloc, dealloc, G_dir, Gni_dir,
// Gnj_dir, Helmholtz_G_dir) over the caller's global arrays, served by one process-wide fmmx
// handle.  Built into libfmmx_ref.so; declarations and the list of
// deliberate differences are in include/fmmx_refshim.h.
//
// The particle arrays stay in the caller's host memory, as in the reference.
// The sort permutation and the FMM itself run on the GPU; the permutes of the
// caller's arrays (sort/sortvar.cxx) and the box counts (fmm/boxallocate.cxx)
// are host loops over those arrays, as in the reference.
#include "fmmx_refshim.h"

#include <sys/time.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Shim {
  fmmx_handle* h = nullptr;
  int p = 9;  // include/constants.h:12
  int precision = FMMX_FP64;
  int convention = FMMX_CONV_REF;
  int device = 0;
  bool configured = false;
  // host copy of what the engine's tree was last built from
  struct Key {
    int ni = -1, nj = -1, lmax = -1, nmode = -1;
    float xmin = 0, ymin = 0, zmin = 0, rd = 0;
    std::vector<float> a;
  } tree;
  std::vector<int> perm, keys;
  std::vector<float> tmp;
  std::vector<double> out[4];
  bool nek_owned = false;
};

Shim& S() {
  static Shim s;
  return s;
}

double now() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

void add_time(int slot, double t) {
  if (tfmm) tfmm[slot] += t;
}

[[noreturn]] void die(const char* fn, const std::string& msg) {
  std::fprintf(stderr, "fmmx_ref: %s: %s\n", fn, msg.c_str());
  std::exit(1);
}

void check(const char* fn, int rc) {
  if (rc != FMMX_OK) die(fn, fmmx_last_error());
}

void env_config(Shim& s) {
  if (s.configured) return;
  s.configured = true;
  if (const char* e = std::getenv("FMMX_REF_P")) s.p = std::atoi(e);
  if (const char* e = std::getenv("FMMX_REF_PRECISION"))
    s.precision = std::strcmp(e, "mixed") == 0 ? FMMX_MIXED : FMMX_FP64;
  if (const char* e = std::getenv("FMMX_REF_CONVENTION"))
    s.convention = std::strcmp(e, "exact") == 0 ? FMMX_CONV_EXACT : FMMX_CONV_REF;
  if (const char* e = std::getenv("FMMX_REF_DEVICE")) s.device = std::atoi(e);
}

// The handle, with the caller's current root cube and lmax (the reference
// reads xmin, ymin, zmin, rd and lmax from globals at every call).
fmmx_handle* handle(const char* fn) {
  Shim& s = S();
  env_config(s);
  if (!s.h) {
    fmmx_config c{};
    c.p = s.p;
    c.lmax = 0;
    c.precision = s.precision;
    c.convention = s.convention;
    c.rd = 0;
    check(fn, fmmx_create(&c, s.device, nullptr, &s.h));
  }
  if (lmax < 2 || lmax > 8) die(fn, "lmax " + std::to_string(lmax) + " outside [2, 8] (run nlevel first)");
  if (!(rd > 0)) die(fn, "root cube length rd must be > 0");
  check(fn, fmmx_set_domain(s.h, lmax, xmin, ymin, zmin, rd));
  return s.h;
}

// Sorted keys of n points (GPU: fmm/boxn.cxx + sort/csort.cxx); perm in s.perm.
void sorted_keys(const char* fn, int n, const float* x, const float* y, const float* z) {
  Shim& s = S();
  fmmx_handle* h = handle(fn);
  s.perm.resize(std::max(n, 1));
  s.keys.resize(std::max(n, 1));
  check(fn, fmmx_sort(h, n, x, y, z, s.keys.data(), s.perm.data(), 0));
}

// sort/sortvar.cxx:6-17 and 19-30
void sortvar(int n, float* var, const int* nvar) {
  if (!var) return;
  std::vector<float>& t = S().tmp;
  t.resize(std::max(n, 1));
  for (int i = 0; i < n; i++) t[i] = var[nvar[i]];
  std::memcpy(var, t.data(), sizeof(float) * n);
}
void unsortvar(int n, float* var, const int* nvar) {
  if (!var) return;
  std::vector<float>& t = S().tmp;
  t.resize(std::max(n, 1));
  for (int i = 0; i < n; i++) t[nvar[i]] = var[i];
  std::memcpy(var, t.data(), sizeof(float) * n);
}

// number of distinct keys at the leaf level and at levels lmax-1..2
// (fmm/boxallocate.cxx:29-44), keys sorted ascending
void count_boxes(const std::vector<int>& k, int n, int& leaf, int& total) {
  leaf = 0;
  for (int i = 0, c = -1; i < n; i++)
    if (k[i] != c) leaf++, c = k[i];
  total = leaf;
  for (int lev = lmax - 1; lev >= 2; lev--) {
    int sh = 3 * (lmax - lev);
    for (int i = 0, c = -1; i < n; i++)
      if ((k[i] >> sh) != c) total++, c = k[i] >> sh;
  }
}

bool tree_is_current(Shim::Key& t, int ni, int nj, int nmode, const float* const* a, const int* n) {
  size_t tot = 0;
  for (int k = 0; k < 12; k++)
    if (a[k]) tot += (size_t)n[k];
  bool same = t.ni == ni && t.nj == nj && t.lmax == lmax && t.nmode == nmode && t.xmin == xmin && t.ymin == ymin &&
              t.zmin == zmin && t.rd == rd && t.a.size() == tot;
  size_t o = 0;
  for (int k = 0; k < 12 && same; k++) {
    if (!a[k]) continue;
    same = std::memcmp(t.a.data() + o, a[k], sizeof(float) * n[k]) == 0;
    o += n[k];
  }
  if (same) return true;
  t.ni = ni; t.nj = nj; t.lmax = lmax; t.nmode = nmode;
  t.xmin = xmin; t.ymin = ymin; t.zmin = zmin; t.rd = rd;
  t.a.resize(tot);
  o = 0;
  for (int k = 0; k < 12; k++) {
    if (!a[k]) continue;
    std::memcpy(t.a.data() + o, a[k], sizeof(float) * n[k]);
    o += n[k];
  }
  return false;
}

// inclusive ranges n0..n1 / n2..n3 of kernel/G_dir.cxx:9,14
void direct(const char* fn, int neq, int n0, int n1, int n2, int n3) {
  Shim& s = S();
  const int ni = n1 - n0 + 1, nj = n3 - n2 + 1;
  if (ni <= 0) return;
  if (n0 < 0 || n2 < 0 || nj < 0) die(fn, "negative index range");
  env_config(s);
  if (!s.h) {
    fmmx_config c{};
    c.p = s.p;
    c.precision = s.precision;
    c.convention = s.convention;
    check(fn, fmmx_create(&c, s.device, nullptr, &s.h));
  }
  if (neq == FMMX_YUKAWA) check(fn, fmmx_set_yukawa(s.h, lambda));
  for (auto& v : s.out) v.resize(ni);
  const bool g = neq == FMMX_G;
  check(fn, fmmx_direct(s.h, neq, ni, xi + n0, yi + n0, zi + n0, neq == FMMX_GNI ? gxi + n0 : nullptr,
                        neq == FMMX_GNI ? gyi + n0 : nullptr, neq == FMMX_GNI ? gzi + n0 : nullptr, nj, xj + n2,
                        yj + n2, zj + n2, neq == FMMX_GNJ ? gxj + n2 : nullptr, neq == FMMX_GNJ ? gyj + n2 : nullptr,
                        neq == FMMX_GNJ ? gzj + n2 : nullptr, vj + n2, s.out[0].data(), g ? s.out[1].data() : nullptr,
                        g ? s.out[2].data() : nullptr, g ? s.out[3].data() : nullptr, 0));
  for (int i = 0; i < ni; i++) vi[n0 + i] = (float)s.out[0][i];
  if (g)
    for (int i = 0; i < ni; i++) {
      gxi[n0 + i] = (float)s.out[1][i];
      gyi[n0 + i] = (float)s.out[2][i];
      gzi[n0 + i] = (float)s.out[3][i];
    }
}

}  // namespace

extern "C" int fmmx_ref_configure(int p, int precision, int convention, int device) {
  Shim& s = S();
  if (p < 2 || p > FMMX_MAX_P || (precision != FMMX_MIXED && precision != FMMX_FP64) ||
      (convention != FMMX_CONV_REF && convention != FMMX_CONV_EXACT) || device < 0)
    return FMMX_EINVAL;
  if (s.h) {
    fmmx_destroy(s.h);
    s.h = nullptr;
  }
  s.p = p;
  s.precision = precision;
  s.convention = convention;
  s.device = device;
  s.configured = true;
  s.tree = Shim::Key{};
  return FMMX_OK;
}

extern "C" fmmx_handle* fmmx_ref_handle(void) { return S().h; }

// fmm/nlevel.cxx:5-30 ("gpu p=5" switch points, compared in float)
void nlevel(int nmax) {
  double t0 = now();
  const float sw[6] = {1e5f, 5e5f, 3e6f, 2e7f, 1.5e8f, 1.3e9f};
  int l = 8;
  for (int k = 0; k < 6; k++)
    if ((float)nmax < sw[k]) {
      l = k + 2;
      break;
    }
  lmax = l;
  std::printf("lmax   : %d\n", lmax);
  add_time(0, now() - t0);
}

// fmm/boxpart.cxx:11-67: nek numbers the occupied i boxes in key order, then
// the j-only boxes in key order; irank = 0 (one rank)
void boxpart(int mi, int mj) {
  double t0 = now();
  Shim& s = S();
  nbmax = 1 << 3 * lmax;
  if (s.nek_owned) {
    delete[] nek;
    delete[] irank;
  }
  nek = new int[nbmax];
  irank = new int[nbmax];
  s.nek_owned = true;
  std::fill(nek, nek + nbmax, -1);
  int lbk = 0;
  sorted_keys("boxpart", mi, xi, yi, zi);
  for (int i = 0, c = -1; i < mi; i++)
    if (s.keys[i] != c) c = s.keys[i], nek[c] = lbk++;
  sorted_keys("boxpart", mj, xj, yj, zj);
  for (int i = 0, c = -1; i < mj; i++)
    if (s.keys[i] != c) {
      c = s.keys[i];
      if (nek[c] == -1) nek[c] = lbk++;
    }
  for (int i = 0; i < lbk; i++) irank[i] = 0;
  add_time(4, now() - t0);
}

// sort/sorti.cxx:12-36
void sorti(int& mi) {
  double t0 = now();
  Shim& s = S();
  sorted_keys("sorti", mi, xi, yi, zi);
  std::memcpy(nbi, s.perm.data(), sizeof(int) * mi);
  add_time(6, now() - t0);
  t0 = now();
  for (float* v : {xi, yi, zi, gxi, gyi, gzi}) sortvar(mi, v, nbi);
  add_time(8, now() - t0);
}

// sort/sortj.cxx:12-37
void sortj(int& mj) {
  double t0 = now();
  Shim& s = S();
  sorted_keys("sortj", mj, xj, yj, zj);
  std::memcpy(nbj, s.perm.data(), sizeof(int) * mj);
  add_time(7, now() - t0);
  t0 = now();
  for (float* v : {xj, yj, zj, gxj, gyj, gzj, vj}) sortvar(mj, v, nbj);
  add_time(8, now() - t0);
}

// sort/sorti.cxx:38-50
void unsorti(int& mi) {
  double t0 = now();
  for (float* v : {xi, yi, zi, gxi, gyi, gzi, vi}) unsortvar(mi, v, nbi);
  add_time(25, now() - t0);
}

// sort/sortj.cxx:38-50
void unsortj(int& mj) {
  double t0 = now();
  for (float* v : {xj, yj, zj, gxj, gyj, gzj, vj}) unsortvar(mj, v, nbj);
  add_time(25, now() - t0);
}

// fmm/boxallocate.cxx:10-92: box counts of targets xi[0..n1) and sources
// xj[0..n3) (n0 and n2 offset only the reference's index array)
void boxallocate(int n0, int n1, int n2, int n3) {
  (void)n0;
  (void)n2;
  double t0 = now();
  Shim& s = S();
  nbmax = 1 << 3 * lmax;
  int lbi, ti, lbj, tj;
  sorted_keys("boxallocate", n1, xi, yi, zi);
  count_boxes(s.keys, n1, lbi, ti);
  sorted_keys("boxallocate", n3, xj, yj, zj);
  count_boxes(s.keys, n3, lbj, tj);
  nbne = std::max(lbi, lbj);
  nbnet = std::max(ti, tj);
  add_time(1, now() - t0);
}

// fmm/alloc.cxx:32-122: the lists, expansions and tables live in the
// engine; alloc(1) makes sure the handle exists
void alloc(int ialloc) {
  double t0 = now();
  if (ialloc == 1) (void)handle("alloc");
  add_time(2, now() - t0);
}

// fmm/alloc.cxx:124-214: frees what boxpart allocated; the engine keeps its
// device buffers for the next call (as the reference's GPU staging does)
void dealloc(int ialloc) {
  Shim& s = S();
  if (ialloc == 1 && s.nek_owned) {
    delete[] nek;
    delete[] irank;
    nek = nullptr;
    irank = nullptr;
    s.nek_owned = false;
  }
}

// fmm/fmm.cxx:29-152
void fmm(int mi, int mj, int neq) {
  const char* fn = "fmm";
  if ((neq < 0 || neq > 2) && neq != FMMX_YUKAWA)
    die(fn, "neq " + std::to_string(neq) + " is not provided (0 G, 1 Gni, 2 Gnj, 10 Yukawa)");
  Shim& s = S();
  fmmx_handle* h = handle(fn);
  if (neq == FMMX_YUKAWA) check(fn, fmmx_set_yukawa(h, lambda));
  const float* a[12] = {xi, yi, zi, nullptr, nullptr, nullptr, xj, yj, zj, nullptr, nullptr, nullptr};
  if (neq == FMMX_GNI) a[3] = gxi, a[4] = gyi, a[5] = gzi;
  if (neq == FMMX_GNJ) a[9] = gxj, a[10] = gyj, a[11] = gzj;
  int n[12];
  for (int k = 0; k < 12; k++) n[k] = k < 6 ? mi : mj;
  const int nmode = neq == FMMX_YUKAWA ? FMMX_G : neq;  // the tree depends on the normals only
  const bool rebuild = !tree_is_current(s.tree, mi, mj, nmode, a, n);
  if (rebuild) {
    check(fn, fmmx_set_timing(h, 1));
    check(fn, fmmx_set_points(h, mi, a[0], a[1], a[2], a[3], a[4], a[5], mj, a[6], a[7], a[8], a[9], a[10], a[11],
                              0));
  }
  for (auto& v : s.out) v.resize(std::max(mi, 1));
  const bool g = neq == FMMX_G || neq == FMMX_YUKAWA;
  check(fn, fmmx_evaluate(h, neq, vj, s.out[0].data(), g ? s.out[1].data() : nullptr, g ? s.out[2].data() : nullptr,
                          g ? s.out[3].data() : nullptr, 0));
  // fmm/fmm.cxx:46-52 zeroes vi (and gxi.. for neq 0) before accumulating
  for (int i = 0; i < mi; i++) vi[i] = (float)s.out[0][i];
  if (g)
    for (int i = 0; i < mi; i++) {
      gxi[i] = (float)s.out[1][i];
      gyi[i] = (float)s.out[2][i];
      gzi[i] = (float)s.out[3][i];
    }
  fmmx_stats st{};
  check(fn, fmmx_get_stats(h, &st));
  // seconds into the reference's timing slots (fmm/boxdatai.cxx:32 for the
  // whole tree and lists, kernel/G_p2p.cxx:39, G_p2m.cxx:57, m2m.cxx:64,
  // m2l.cxx:91, l2l.cxx:82, G_l2p.cxx:78)
  if (rebuild) add_time(10, st.ms_tree * 1e-3);
  add_time(17, st.ms_p2p * 1e-3);
  add_time(18, st.ms_p2m * 1e-3);
  add_time(19, st.ms_m2m * 1e-3);
  add_time(20, st.ms_m2l * 1e-3);
  add_time(21, st.ms_l2l * 1e-3);
  add_time(22, st.ms_l2p * 1e-3);
}

void G_dir(int n0, int n1, int n2, int n3) { direct("G_dir", FMMX_G, n0, n1, n2, n3); }
void Gni_dir(int n0, int n1, int n2, int n3) { direct("Gni_dir", FMMX_GNI, n0, n1, n2, n3); }
void Gnj_dir(int n0, int n1, int n2, int n3) { direct("Gnj_dir", FMMX_GNJ, n0, n1, n2, n3); }
void Helmholtz_G_dir(int n0, int n1, int n2, int n3) { direct("Helmholtz_G_dir", FMMX_YUKAWA, n0, n1, n2, n3); }
