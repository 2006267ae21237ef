// ref_caller.cxx -- a caller written against the reference's interface, as
// test/test.cxx:98-146 and bio/main.cxx:327-336 are: it defines the global
// arrays and scalars (the reference's include/arrays.h and constants.h do
// that in the translation unit compiled with MAIN), runs the set-up chain
// nlevel -> boxpart -> sorti -> sortj -> boxallocate -> alloc(1) -> fmm ->
// unsorti -> unsortj -> dealloc(1), then the direct sum, and links against
// libfmmx_ref.so only.  Test infrastructure.
//
// With "plugin" as a third argument the matvec is not fmm() but this file's
// own sequence of the operator symbols (the order of fmm/fmm.cxx:38-150:
// precalc, box data, near list, P2P, P2M, M2M per level, M2L per level with
// L2L in between, L2P), recording the multipoles after P2M and after the
// upward pass, the locals of every level and the outputs.
//
// Built here against libfmmx_ref.so alone.  oracle/Makefile.ref also links it
// (with -DREF_GLOBALS, the reference's own include/arrays.h defining the
// globals) to the reference's fmm.o, and once more to the reference's fmm.o
// with all of its tree, list and set-up objects, so that only the kernel
// symbols come from the shim (the link-time kernel choice of
// biobem/Makefile:25-37).
//
// usage: ref_caller <input> <output> [plugin]
// input:  int ni, nj, lmax, neq, unused; float box[4] (xmin, ymin, zmin, rd);
//         float xi[3][ni], ni_normals[3][ni], xj[3][nj], nj_normals[3][nj], q[nj]
//         and for neq 10 one float, lambda
// output: records {char name[32]; int type (0 int, 1 float, 2 double); int n; data}
#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#ifdef REF_GLOBALS
#define MAIN
#include "constants.h"
#include "arrays.h"
#undef MAIN
#endif
#include "fmmx_refshim.h"

#ifndef REF_GLOBALS
float *xi, *yi, *zi, *gxi, *gyi, *gzi, *vi;
float *xj, *yj, *zj, *gxj, *gyj, *gzj, *vj;
int *nbi, *nbj, *nek, *irank;
int **ndi, *nfi, *nlbi, **ndj, *nej, *nfj, *nlbj, *nij, *njb;
int (*neij)[189];
std::complex<double> (*ax)[45], (*bx)[45];  // mpsym of the reference's mp = 9
double* tfmm;
int nbmax, nbne, nbnet, lmax;
float xmin, ymin, zmin, rd;
double lambda;
#endif

static FILE* out;
static void rec(const char* name, int type, int n, const void* data) {
  char nm[32] = {};
  std::strncpy(nm, name, 31);
  std::fwrite(nm, 1, 32, out);
  std::fwrite(&type, 4, 1, out);
  std::fwrite(&n, 4, 1, out);
  std::fwrite(data, type == 2 ? 8 : 4, (size_t)n, out);
}

// rows of complex expansions (p(p+1)/2 per box, p = 9 unless FMMX_REF_P)
static void rec_rows(const std::string& name, const std::complex<double>* a, int rows, int nmp) {
  rec(name.c_str(), 2, 2 * rows * nmp, a);
}

// One matvec through the operator symbols of order mp; rec_on: record the
// expansions, lists and near field.
static void plugin_matvec(int ni, int nj, int neq, int mp, bool rec_on) {
  const int top = 2;
  int nmp = 0, lbi = 0, lbj = 0, lbjr = 0;
  double rb = 0;
  nlbi[lmax - 1] = 0;
  nlbj[lmax - 1] = 0;
  precalc(nmp, mp);
  boxdatai(ni, lbi, rb);
  boxdataj(nj, lbj, rb);
  ifbox(ni, nj, nmp, lbi, lbj, lbjr, lmax, 0);
  for (int i = 0; i < ni; i++) vi[i] = 0;
  if (neq == 0 || neq == 10)
    for (int i = 0; i < ni; i++) gxi[i] = gyi[i] = gzi[i] = 0;
  switch (neq) {
    case 0: G_p2p(lbi, lbj); G_p2m(nmp, mp, lbj); break;
    case 1: Gni_p2p(lbi, lbj); G_p2m(nmp, mp, lbj); break;
    case 2: Gnj_p2p(lbi, lbj); Gn_p2m(nmp, mp, lbj); break;
    default: Helmholtz_G_p2p(lbi, lbj); Helmholtz_G_p2m(nmp, mp, lbj); break;
  }
  if (rec_on) {
    rec("p2p_vi", 1, ni, vi);
    rec_rows("p2m_bx", &bx[0][0], lbj, nmp);
  }
  // upward pass, one level at a time from the leaf's parents to level 2
  for (int lev = lmax - 1; lev >= top; lev--) {
    const int child = lbj;
    boxm2mi(lbi, lev);
    boxm2mj(lbj, lev);
    if (neq == 10) Helmholtz_m2m(nmp, mp, lev, lbj, child);
    else m2m(nmp, mp, lev, lbj, child);
  }
  if (lmax == top) {
    std::fill(nej, nej + nbmax, -1);
    for (int jj = 0; jj < lbj; jj++) nej[nfj[jj + nlbj[top - 1]]] = jj;
  }
  if (rec_on) {
    rec("nlbi", 0, lmax, nlbi);
    rec("nlbj", 0, lmax, nlbj);
    rec("all_nfi", 0, nlbi[top - 1] + lbi, nfi);
    rec("all_nfj", 0, nlbj[top - 1] + lbj, nfj);
    rec_rows("up_bx", &bx[0][0], nlbj[top - 1] + lbj, nmp);
  }
  // downward pass: level 2 from its full far list, then L2L + M2L per level
  for (int lev = top; lev <= lmax; lev++) {
    if (lev > top) {
      lbi = nlbi[lev - 2] - nlbi[lev - 1];
      lbj = nlbj[lev - 2] - nlbj[lev - 1];
      l2l(nmp, mp, lev, lbi);
      std::fill(nej, nej + nbmax, -1);
      for (int jj = 0; jj < lbj; jj++) nej[nfj[jj + nlbj[lev - 1]]] = jj;
    }
    ifbox(ni, nj, nmp, lbi, lbj, lbjr, lev, lev == top ? 1 : 2);
    if (rec_on) {
      std::vector<int> flat;
      for (int ii = 0; ii < lbi; ii++) flat.insert(flat.end(), neij[ii], neij[ii] + nij[ii]);
      rec(("far_nij_" + std::to_string(lev)).c_str(), 0, lbi, nij);
      rec(("far_neij_" + std::to_string(lev)).c_str(), 0, (int)flat.size(), flat.data());
    }
    if (neq == 10) Helmholtz_m2l(nmp, mp, lbi, lbjr, lev, lev == top ? lbi : 0);
    else m2l(nmp, mp, lbi, lbjr, lev, lev == top ? lbi : 0);
    if (rec_on) rec_rows("ax_" + std::to_string(lev), &ax[0][0], lbi, nmp);
  }
  if (neq == 1) Gn_l2p(nmp, mp, lbi);
  else if (neq == 10) Helmholtz_G_l2p(nmp, mp, lbi);
  else G_l2p(nmp, mp, lbi);
}

int main(int argc, char** argv) {
  if (argc != 3 && !(argc == 4 && std::string(argv[3]) == "plugin")) {
    std::fprintf(stderr, "usage: %s input output [plugin]\n", argv[0]);
    return 2;
  }
  const bool plugin = argc == 4;
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int hdr[5];
  float box[4];
  if (std::fread(hdr, 4, 5, in) != 5 || std::fread(box, 4, 4, in) != 4) return 2;
  int ni = hdr[0], nj = hdr[1], lev = hdr[2], neq = hdr[3];
  int np = ni > nj ? ni : nj;
  float** arr[14] = {&xi, &yi, &zi, &gxi, &gyi, &gzi, &vi, &xj, &yj, &zj, &gxj, &gyj, &gzj, &vj};
  for (auto a : arr) *a = new float[np]();
  nbi = new int[np];
  nbj = new int[np];
  tfmm = new double[100]();
#ifdef REF_GLOBALS
  // the work arrays of the reference's own box and sort objects, as
  // test/test.cxx:42-65 allocates them (nc holds one count per leaf box)
  {
    const int nw = std::max(np, 1 << (3 * lev));
    nfn = new int[nw];
    na = new int[nw];
    nb = new int[nw];
    nc = new int[nw];
    nd = new int[nw];
    sortd = new float[nw];
    nprocs = 1;
    myrank = 0;
    umem = 0;
  }
#endif
  float* rd_i[6] = {xi, yi, zi, gxi, gyi, gzi};
  for (auto a : rd_i)
    if (std::fread(a, 4, ni, in) != (size_t)ni) return 2;
  float* rd_j[7] = {xj, yj, zj, gxj, gyj, gzj, vj};
  for (auto a : rd_j)
    if (std::fread(a, 4, nj, in) != (size_t)nj) return 2;
  if (neq == 10) {
    float lam;
    if (std::fread(&lam, 4, 1, in) != 1) return 2;
    lambda = lam;
  }
  std::fclose(in);
  xmin = box[0];
  ymin = box[1];
  zmin = box[2];
  rd = box[3];
  out = std::fopen(argv[2], "wb");
  if (!out) return 2;

  nlevel(nj);
  lmax = lev;  // the fixture's depth instead of the switch points
  boxpart(ni, nj);
  sorti(ni);
  sortj(nj);
  boxallocate(0, ni, 0, nj);
  alloc(1);
  int counts[4] = {nbmax, nbne, nbnet, lmax};
  rec("counts", 0, 4, counts);
  int nek_boxes = 0;
  for (int k = 0; k < nbmax; k++) nek_boxes += nek[k] >= 0;
  rec("nek_boxes", 0, 1, &nek_boxes);
  const int mp = std::getenv("FMMX_REF_P") ? std::atoi(std::getenv("FMMX_REF_P")) : 9;
  bool first = true;
  auto matvec = [&]() {
    if (plugin) plugin_matvec(ni, nj, neq, mp, first);
    else fmm(ni, nj, neq);
    first = false;
  };
  matvec();
  // a second matvec on the same tree with doubled strengths (the GMRES
  // pattern of bio/main.cxx:27-53): the tree is reused, the result doubles
  std::vector<float> v1(vi, vi + ni), g1[3] = {{gxi, gxi + ni}, {gyi, gyi + ni}, {gzi, gzi + ni}};
  for (int j = 0; j < nj; j++) vj[j] *= 2;
  matvec();
  for (int j = 0; j < nj; j++) vj[j] *= 0.5f;
  rec("fmm_vi_sorted", 1, ni, v1.data());
  rec("fmm2_vi_sorted", 1, ni, vi);
  for (int i = 0; i < ni; i++) vi[i] = v1[i];
  if (neq == 0 || neq == 10) {
    // gradients of the single-strength call
    std::memcpy(gxi, g1[0].data(), sizeof(float) * ni);
    std::memcpy(gyi, g1[1].data(), sizeof(float) * ni);
    std::memcpy(gzi, g1[2].data(), sizeof(float) * ni);
  }
  rec("nbi", 0, ni, nbi);
  rec("nbj", 0, nj, nbj);
  unsorti(ni);
  unsortj(nj);
  dealloc(1);
  rec("fmm_vi", 1, ni, vi);
  if (neq == 0) {
    rec("fmm_gxi", 1, ni, gxi);
    rec("fmm_gyi", 1, ni, gyi);
    rec("fmm_gzi", 1, ni, gzi);
  }
  rec("xj_restored", 1, nj, xj);
  rec("vj_restored", 1, nj, vj);
  double tf = 0;
  for (int k = 0; k < 100; k++) tf += tfmm[k];
  float tff = (float)tf;
  rec("tfmm_total", 1, 1, &tff);

  alloc(0);
  if (neq == 0)
    G_dir(0, ni - 1, 0, nj - 1);
  else if (neq == 1)
    Gni_dir(0, ni - 1, 0, nj - 1);
  else if (neq == 2)
    Gnj_dir(0, ni - 1, 0, nj - 1);
  else
    Helmholtz_G_dir(0, ni - 1, 0, nj - 1);
  dealloc(0);
  rec("dir_vi", 1, ni, vi);
  if (neq == 0) {
    rec("dir_gxi", 1, ni, gxi);
    rec("dir_gyi", 1, ni, gyi);
    rec("dir_gzi", 1, ni, gzi);
  }
  std::fclose(out);
  for (auto a : arr) delete[] *a;
  delete[] nbi;
  delete[] nbj;
  delete[] tfmm;
  return 0;
}
