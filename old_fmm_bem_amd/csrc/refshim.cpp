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
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

// include/arrays.h:38 declares ax and bx as std::complex<double> (*)[mpsym]
// with the caller's compile-time mpsym; the shim uses them as rows of
// p(p+1)/2 complex values (the same objects: p must be the caller's mp)
extern std::complex<double> (*ax)[1], (*bx)[1];

namespace {

int nmp_of(int p) { return p * (p + 1) / 2; }

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
  // operator symbols: the kernel of the sequence in progress (set by the P2P
  // and P2M symbols) and the list and expansion arrays alloc(1) made
  int plug_neq = FMMX_G;
  int plug_mi = -1, plug_mj = -1;  // particle counts of the sequence in progress
  bool lists_owned = false;
  std::vector<double> exp_buf;
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

// The engine's tree for the caller's current (sorted) arrays, built once per
// content: the operator symbols find it in place across a matvec and across
// GMRES iterations.  neq selects the normals it carries (target normals for
// Gni, source normals for Gnj); keep_lists also builds the far lists.
fmmx_handle* plug_tree(const char* fn, int mi, int mj, int neq) {
  const bool keep_lists = true;  // ifbox hands out the far lists
  Shim& s = S();
  fmmx_handle* h = handle(fn);
  const float* a[12] = {xi, yi, zi, nullptr, nullptr, nullptr, xj, yj, zj, nullptr, nullptr, nullptr};
  if (neq == FMMX_GNI) a[3] = gxi, a[4] = gyi, a[5] = gzi;
  if (neq == FMMX_GNJ) a[9] = gxj, a[10] = gyj, a[11] = gzj;
  int n[12];
  for (int k = 0; k < 12; k++) n[k] = k < 6 ? mi : mj;
  const int nmode = (neq == FMMX_YUKAWA ? FMMX_G : neq) | (keep_lists ? 0x100 : 0);
  if (!tree_is_current(s.tree, mi, mj, nmode, a, n)) {
    const double t0 = now();
    check(fn, fmmx_set_timing(h, 1));
    check(fn, fmmx_set_points(h, mi, a[0], a[1], a[2], a[3], a[4], a[5], mj, a[6], a[7], a[8], a[9], a[10], a[11],
                              keep_lists ? FMMX_KEEP_LISTS : 0));
    add_time(10, now() - t0);
  }
  return h;
}

// the tree of the sequence in progress, whatever normals it carries; built
// with the kernel of the last P2P/P2M symbol otherwise
fmmx_handle* plug_tree_any(const char* fn) {
  Shim& s = S();
  const int mi = s.plug_mi, mj = s.plug_mj;
  if (s.tree.ni == mi && s.tree.nj == mj && (s.tree.nmode & 0x100)) return plug_tree(fn, mi, mj, s.tree.nmode & 0xff);
  return plug_tree(fn, mi, mj, s.plug_neq);
}

// Particle counts of the sequence from the leaf ranges ndi/ndj (filled by
// boxdatai/boxdataj, the shim's or the reference's: fmm/boxdatai.cxx:22-29)
// for the symbols that receive leaf box counts; -1 keeps the last value.
void plug_counts(const char* fn, int lbi, int lbj) {
  Shim& s = S();
  if (lbi >= 0 && ndi) s.plug_mi = lbi > 0 ? ndi[1][lbi - 1] + 1 : 0;
  if (lbj >= 0 && ndj) s.plug_mj = lbj > 0 ? ndj[1][lbj - 1] + 1 : 0;
  if (s.plug_mi < 0 || s.plug_mj < 0) die(fn, "no box data (boxdatai/boxdataj) for this matvec");
}

int n_targets() { return S().plug_mi; }

// One engine phase for the operator symbols.  P2P and L2P add their outputs
// into the caller's float arrays (vi, and gxi.. when grad), as the reference's
// kernels accumulate into them.
void plug_phase(const char* fn, int ph, int neq, int lev, bool grad, int slot) {
  Shim& s = S();
  const double t0 = now();
  fmmx_handle* h = handle(fn);
  if (neq == FMMX_YUKAWA) check(fn, fmmx_set_yukawa(h, lambda));
  const int mi = n_targets();
  const bool writes = ph == FMMX_PHASE_P2P || ph == FMMX_PHASE_L2P;
  for (auto& v : s.out) v.resize(std::max(mi, 1));
  const bool g = grad && (neq == FMMX_G || neq == FMMX_YUKAWA);
  check(fn, fmmx_evaluate_phase(h, ph, neq, lev, vj, writes ? s.out[0].data() : nullptr,
                                writes && g ? s.out[1].data() : nullptr, writes && g ? s.out[2].data() : nullptr,
                                writes && g ? s.out[3].data() : nullptr, 0));
  if (writes) {
    for (int i = 0; i < mi; i++) vi[i] += (float)s.out[0][i];
    if (g)
      for (int i = 0; i < mi; i++) {
        gxi[i] += (float)s.out[1][i];
        gyi[i] += (float)s.out[2][i];
        gzi[i] += (float)s.out[3][i];
      }
  }
  add_time(slot, now() - t0);
}

}  // namespace

extern "C" int fmmx_ref_configure(int p, int precision, int convention, int device) {
  Shim& s = S();
  if (p < 2 || (p > 12 && p != FMMX_MAX_P) || (precision != FMMX_MIXED && precision != FMMX_FP64) ||
      (convention != FMMX_CONV_REF && convention != FMMX_CONV_EXACT) || device < 0)
    return FMMX_EINVAL;
  // mp = 18 (include/constants.h:12): FP64 in the EXACT convention only
  if (p > 12 && (precision != FMMX_FP64 || convention != FMMX_CONV_EXACT)) return FMMX_EUNSUPPORTED;
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

// fmm/alloc.cxx:32-122: the tables and the engine's own lists and
// expansions live on the device; alloc(1) makes sure the handle exists and,
// for fmm/fmm.cxx and the operator symbols, allocates the caller's box lists,
// interaction lists and expansions with the reference's sizes (ndi, ndj, nij,
// njb, neij and ax: nbne rows; nfi, nfj and bx: nbnet; nej: nbmax; nlbi,
// nlbj: lmax)
void dealloc(int ialloc);
void alloc(int ialloc) {
  double t0 = now();
  Shim& s = S();
  if (ialloc == 1) {
    (void)handle("alloc");
    if (s.lists_owned) dealloc(2);
    const int ne = std::max(nbne, 1), nt = std::max(nbnet, 1), nmp = nmp_of(s.p);
    ndi = new int*[2];
    ndj = new int*[2];
    for (int k = 0; k < 2; k++) {
      ndi[k] = new int[ne];
      ndj[k] = new int[ne];
    }
    nfi = new int[nt];
    nfj = new int[nt];
    nlbi = new int[std::max(lmax, 1)];
    nlbj = new int[std::max(lmax, 1)];
    nej = new int[std::max(nbmax, 1)];
    nij = new int[ne];
    njb = new int[ne];
    neij = new int[ne][189];
    ax = reinterpret_cast<std::complex<double>(*)[1]>(new std::complex<double>[(size_t)ne * nmp]);
    bx = reinterpret_cast<std::complex<double>(*)[1]>(new std::complex<double>[(size_t)nt * nmp]);
    s.lists_owned = true;
  }
  add_time(2, now() - t0);
}

// fmm/alloc.cxx:124-214: frees what boxpart and alloc(1) allocated (2: the
// lists and expansions only); the engine keeps its device buffers for the
// next call (as the reference's GPU staging does)
void dealloc(int ialloc) {
  Shim& s = S();
  if (ialloc == 1 && s.nek_owned) {
    delete[] nek;
    delete[] irank;
    nek = nullptr;
    irank = nullptr;
    s.nek_owned = false;
  }
  if ((ialloc == 1 || ialloc == 2) && s.lists_owned) {
    for (int k = 0; k < 2; k++) {
      delete[] ndi[k];
      delete[] ndj[k];
    }
    delete[] ndi;
    delete[] ndj;
    delete[] nfi;
    delete[] nfj;
    delete[] nlbi;
    delete[] nlbj;
    delete[] nej;
    delete[] nij;
    delete[] njb;
    delete[] neij;
    delete[] reinterpret_cast<std::complex<double>*>(ax);
    delete[] reinterpret_cast<std::complex<double>*>(bx);
    ndi = ndj = nullptr;
    nfi = nfj = nlbi = nlbj = nej = nij = njb = nullptr;
    neij = nullptr;
    ax = bx = nullptr;
    s.lists_owned = false;
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

// ---- the operator plugin: the symbols fmm/fmm.cxx:7-27 declares and calls ----
//
// Each symbol runs its phase of the matvec on the engine's device-resident
// state (fmmx_evaluate_phase).  They work under the shim's own fmm()-level
// symbols and equally under the reference's fmm.o with its own tree and list
// objects (boxdatai.o, ijbox.o, ...): every operator finds the engine's tree
// for the caller's current arrays itself, and reads only the leaf ranges
// ndi/ndj, the level offsets nlbj and the box counts it is given.  The
// reference's kernels read their interaction lists from nij/neij; the
// engine's M2L walks the same interactions by box offsets, and its level-2
// list is the one fmm/fmm.cxx:104 means (neib 1) whatever the caller built
// there.  Expansions are mirrored into the caller's bx and ax after each
// phase in the reference's layout (unscaled, complex, p(p+1)/2 per box;
// Laplace kernels only): the leaf multipoles after P2M, level lev's rows
// after M2M, the level's locals after L2L and after M2L.  The reference's
// m2l also zeroes the bx rows it has used (kernel/m2l.cxx:84-89); the mirror
// keeps them.

namespace {

// fmmx_get_multipoles_level / fmmx_get_locals rows into a caller array
void mirror(const char* fn, bool local, int lev, std::complex<double>* dst, int rows) {
  Shim& s = S();
  if (!dst || rows <= 0 || s.plug_neq == FMMX_YUKAWA) return;
  fmmx_handle* h = s.h;
  const int nb = fmmx_get_count(h, 1, local ? 0 : 1, lev);
  if (nb != rows) die(fn, "box count " + std::to_string(rows) + " differs from the engine's " + std::to_string(nb));
  const int nmp = nmp_of(s.p);
  s.exp_buf.resize((size_t)2 * rows * nmp);
  check(fn, local ? fmmx_get_locals(h, lev, s.exp_buf.data()) : fmmx_get_multipoles_level(h, lev, s.exp_buf.data()));
  std::memcpy((void*)dst, s.exp_buf.data(), sizeof(double) * s.exp_buf.size());
}

// row offset of level lev in the stacked multipoles (leaf first), as
// fmm/boxm2mj.cxx:8 stacks them in nlbj
int bx_row(int lev) {
  int r = 0;
  for (int l = lmax; l > lev; l--) r += fmmx_get_count(S().h, 1, 1, l);
  return r;
}

void check_order(const char* fn, int mp) {
  Shim& s = S();
  env_config(s);
  if (mp != s.p)
    die(fn, "expansion order mp " + std::to_string(mp) + " differs from the shim's p " + std::to_string(s.p) +
                " (fmmx_ref_configure or FMMX_REF_P)");
}

// fmm/boxdatai.cxx:10-34 and boxdataj.cxx:10-36: leaf keys of the caller's
// (sorted) arrays in array order, then one box per run of equal keys
int leaf_boxes(const char* fn, int n, const float* x, const float* y, const float* z, int* nf, int** nd,
               int* ne) {
  Shim& s = S();
  sorted_keys(fn, n, x, y, z);
  std::vector<int> key(std::max(n, 1));
  for (int k = 0; k < n; k++) key[s.perm[k]] = s.keys[k];
  if (ne) std::fill(ne, ne + nbmax, -1);
  int lb = 0;
  for (int i = 0, c = -1; i < n; i++) {
    if (key[i] == c) continue;
    c = key[i];
    if (ne) ne[c] = lb;
    nf[lb] = c;
    nd[0][lb] = i;
    if (lb > 0) nd[1][lb - 1] = i - 1;
    lb++;
  }
  if (lb > 0) nd[1][lb - 1] = n - 1;
  return lb;
}

// fmm/boxm2mi.cxx:6-24, boxm2mj.cxx:6-26: the parent level's keys (key / 8,
// deduplicated) stacked after the child level's
void parent_boxes(int& lb, int lev, int* nf, int* nl, int* ne) {
  nl[lev - 1] = nl[lev] + lb;
  const int nchild = lb;
  if (ne) std::fill(ne, ne + nbmax, -1);
  lb = 0;
  for (int ii = 0, c = -1; ii < nchild; ii++) {
    const int pk = nf[ii + nl[lev]] / 8;
    if (pk == c) continue;
    c = pk;
    nf[lb + nl[lev - 1]] = pk;
    if (ne) ne[pk] = lb;
    lb++;
  }
}

}  // namespace

// kernel/precalc.cxx:10: nmp = mp(mp+1)/2; the engine's tables and
// operators are built with its handle
void precalc(int& nmp, int mp) {
  const double t0 = now();
  check_order("precalc", mp);
  (void)handle("precalc");
  nmp = nmp_of(mp);
  add_time(3, now() - t0);
}

void boxdatai(int mi, int& lbi, double& rb) {
  const double t0 = now();
  if (!nfi || !ndi) die("boxdatai", "lists not allocated (alloc(1))");
  rb = rd / (1 << lmax);
  lbi = leaf_boxes("boxdatai", mi, xi, yi, zi, nfi, ndi, nullptr);
  S().plug_mi = mi;
  add_time(10, now() - t0);
}

void boxdataj(int mj, int& lbj, double& rb) {
  const double t0 = now();
  if (!nfj || !ndj || !nej) die("boxdataj", "lists not allocated (alloc(1))");
  rb = rd / (1 << lmax);
  lbj = leaf_boxes("boxdataj", mj, xj, yj, zj, nfj, ndj, nej);
  S().plug_mj = mj;
  add_time(11, now() - t0);
}

void boxm2mi(int& lbi, int lev) {
  const double t0 = now();
  parent_boxes(lbi, lev, nfi, nlbi, nullptr);
  add_time(10, now() - t0);
}

void boxm2mj(int& lbj, int lev) {
  const double t0 = now();
  parent_boxes(lbj, lev, nfj, nlbj, nej);
  add_time(11, now() - t0);
}

// fmm/ifbox.cxx:7-13 / ijbox.cxx:9-117: lbjr = lbj, the interaction list of
// level lev in nij/neij (neib 0: near; 1, 2: far; 3, which the reference
// never builds, gets the far list of neib 1) and njb
void ifbox(int mi, int mj, int nmp, int lbi, int lbj, int& lbjr, int lev, int neib) {
  (void)nmp;
  const double t0 = now();
  Shim& s = S();
  s.plug_mi = mi;
  s.plug_mj = mj;
  lbjr = lbj;
  fmmx_handle* h = plug_tree_any("ifbox");
  if (nij && neij && lbi > 0) {
    const int nb = fmmx_get_count(h, 1, 0, lev);
    if (nb != lbi) die("ifbox", "target box count differs from the engine's");
    check("ifbox", fmmx_get_list(h, lev, neib == 0 ? 0 : 1, nij, &neij[0][0]));
  }
  if (njb && nej && nfj && nlbj)
    for (int jj = 0; jj < lbj; jj++) {
      const int jb = jj + nlbj[lev - 1];
      njb[jj] = nej[nfj[jb]] + nlbj[lev - 1];
    }
  add_time(13, now() - t0);
}

// kernel/G_p2p.cxx:7, Gni_p2p.cxx:7, Gnj_p2p.cxx:7, Helmholtz_G_p2p.cxx: the
// near field added into vi (and gxi.. for G and Yukawa)
static void p2p_symbol(const char* fn, int neq, int lbi, int lbj) {
  Shim& s = S();
  plug_counts(fn, lbi, lbj);
  s.plug_neq = neq;
  plug_tree(fn, s.plug_mi, s.plug_mj, neq);
  plug_phase(fn, FMMX_PHASE_P2P, neq, 0, true, 17);
}
void G_p2p(int lbi, int lbj) { p2p_symbol("G_p2p", FMMX_G, lbi, lbj); }
void Gni_p2p(int lbi, int lbj) { p2p_symbol("Gni_p2p", FMMX_GNI, lbi, lbj); }
void Gnj_p2p(int lbi, int lbj) { p2p_symbol("Gnj_p2p", FMMX_GNJ, lbi, lbj); }
void Helmholtz_G_p2p(int lbi, int lbj) { p2p_symbol("Helmholtz_G_p2p", FMMX_YUKAWA, lbi, lbj); }

// kernel/G_p2m.cxx:11 (charges; the sequences of G and Gni), Gn_p2m.cxx:11
// (dipoles; Gnj), Helmholtz_G_p2m.cxx: leaf multipoles, mirrored into bx
static void p2m_symbol(const char* fn, bool dipoles, int mp, int lbj) {
  Shim& s = S();
  check_order(fn, mp);
  plug_counts(fn, -1, lbj);
  if (dipoles) s.plug_neq = FMMX_GNJ;
  else if (s.plug_neq == FMMX_GNJ) s.plug_neq = FMMX_G;
  plug_tree_any(fn);
  const int neq = s.plug_neq;
  if (dipoles) plug_tree(fn, s.plug_mi, s.plug_mj, FMMX_GNJ);
  plug_phase(fn, FMMX_PHASE_P2M, neq, 0, false, 18);
  mirror(fn, false, lmax, bx ? &bx[0][0] : nullptr, lbj);
}
void G_p2m(int nmp, int mp, int lbj) { (void)nmp; p2m_symbol("G_p2m", false, mp, lbj); }
void Gn_p2m(int nmp, int mp, int lbj) { (void)nmp; p2m_symbol("Gn_p2m", true, mp, lbj); }
void Helmholtz_G_p2m(int nmp, int mp, int lbj) {
  (void)nmp;
  S().plug_neq = FMMX_YUKAWA;
  p2m_symbol("Helmholtz_G_p2m", false, mp, lbj);
}

// kernel/m2m.cxx:13, Helmholtz_m2m.cxx: multipoles of level lev from lev+1
static void m2m_symbol(const char* fn, int mp, int lev, int lbj) {
  check_order(fn, mp);
  plug_tree_any(fn);
  plug_phase(fn, FMMX_PHASE_M2M, S().plug_neq, lev, false, 19);
  mirror(fn, false, lev, bx ? &bx[0][0] + (size_t)bx_row(lev) * nmp_of(mp) : nullptr, lbj);
}
void m2m(int nmp, int mp, int lev, int lbj, int lbjo) { (void)nmp; (void)lbjo; m2m_symbol("m2m", mp, lev, lbj); }
void Helmholtz_m2m(int nmp, int mp, int lev, int lbj, int lbjo) {
  (void)nmp;
  (void)lbjo;
  m2m_symbol("Helmholtz_m2m", mp, lev, lbj);
}

// kernel/m2l.cxx:13, Helmholtz_m2l.cxx: locals of level lev (level 2 written,
// ini != 0 as kernel/m2l.cxx:23-28 zeroes them; deeper levels added to what
// l2l wrote), mirrored into ax
static void m2l_symbol(const char* fn, int mp, int lbi, int lev) {
  check_order(fn, mp);
  plug_tree_any(fn);
  plug_phase(fn, FMMX_PHASE_M2L, S().plug_neq, lev, false, 20);
  mirror(fn, true, lev, ax ? &ax[0][0] : nullptr, lbi);
}
void m2l(int nmp, int mp, int lbi, int lbj, int lev, int ini) {
  (void)nmp;
  (void)lbj;
  (void)ini;
  m2l_symbol("m2l", mp, lbi, lev);
}
void Helmholtz_m2l(int nmp, int mp, int lbi, int lbj, int lev, int ini) {
  (void)nmp;
  (void)lbj;
  (void)ini;
  m2l_symbol("Helmholtz_m2l", mp, lbi, lev);
}

// kernel/l2l.cxx:13: locals of level lev from its parents', mirrored into ax
void l2l(int nmp, int mp, int lev, int lbi) {
  (void)nmp;
  check_order("l2l", mp);
  plug_tree_any("l2l");
  plug_phase("l2l", FMMX_PHASE_L2L, S().plug_neq, lev, false, 21);
  mirror("l2l", true, lev, ax ? &ax[0][0] : nullptr, lbi);
}

// kernel/G_l2p.cxx:11 (G, and Gnj without the gradient: the shim leaves
// gxi.. of a Gnj matvec alone), Gn_l2p.cxx:11 (Gni), Helmholtz_G_l2p.cxx:
// the far field added into vi (and gxi..)
static void l2p_symbol(const char* fn, int mp, int lbi, bool normal) {
  Shim& s = S();
  check_order(fn, mp);
  plug_counts(fn, lbi, -1);
  if (normal) s.plug_neq = FMMX_GNI;
  plug_tree_any(fn);
  if (normal) plug_tree(fn, s.plug_mi, s.plug_mj, FMMX_GNI);
  plug_phase(fn, FMMX_PHASE_L2P, s.plug_neq, 0, true, 22);
}
void G_l2p(int nmp, int mp, int lbi) { (void)nmp; l2p_symbol("G_l2p", mp, lbi, false); }
void Gn_l2p(int nmp, int mp, int lbi) { (void)nmp; l2p_symbol("Gn_l2p", mp, lbi, true); }
void Helmholtz_G_l2p(int nmp, int mp, int lbi) {
  (void)nmp;
  S().plug_neq = FMMX_YUKAWA;
  l2p_symbol("Helmholtz_G_l2p", mp, lbi, false);
}
