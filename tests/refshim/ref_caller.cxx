// ref_caller.cxx -- a caller written against the reference's interface, as
// test/test.cxx:98-146 and bio/main.cxx:327-336 are: it defines the global
// arrays and scalars (the reference's include/arrays.h and constants.h do
// that in the translation unit compiled with MAIN), runs the set-up chain
// nlevel -> boxpart -> sorti -> sortj -> boxallocate -> alloc(1) -> fmm ->
// unsorti -> unsortj -> dealloc(1), then the direct sum, and links against
// libfmmx_ref.so only.  Test infrastructure.
//
// usage: ref_caller <input> <output>
// input:  int ni, nj, lmax, neq, unused; float box[4] (xmin, ymin, zmin, rd);
//         float xi[3][ni], ni_normals[3][ni], xj[3][nj], nj_normals[3][nj], q[nj]
//         and for neq 10 one float, lambda
// output: records {char name[32]; int type (0 int, 1 float); int n; data}
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fmmx_refshim.h"

float *xi, *yi, *zi, *gxi, *gyi, *gzi, *vi;
float *xj, *yj, *zj, *gxj, *gyj, *gzj, *vj;
int *nbi, *nbj, *nek, *irank;
double* tfmm;
int nbmax, nbne, nbnet, lmax;
float xmin, ymin, zmin, rd;
double lambda;

static FILE* out;
static void rec(const char* name, int type, int n, const void* data) {
  char nm[32] = {};
  std::strncpy(nm, name, 31);
  std::fwrite(nm, 1, 32, out);
  std::fwrite(&type, 4, 1, out);
  std::fwrite(&n, 4, 1, out);
  std::fwrite(data, 4, (size_t)n, out);
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input output\n", argv[0]);
    return 2;
  }
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
  fmm(ni, nj, neq);
  // a second matvec on the same tree with doubled strengths (the GMRES
  // pattern of bio/main.cxx:27-53): the tree is reused, the result doubles
  std::vector<float> v1(vi, vi + ni);
  for (int j = 0; j < nj; j++) vj[j] *= 2;
  fmm(ni, nj, neq);
  for (int j = 0; j < nj; j++) vj[j] *= 0.5f;
  rec("fmm_vi_sorted", 1, ni, v1.data());
  rec("fmm2_vi_sorted", 1, ni, vi);
  for (int i = 0; i < ni; i++) vi[i] = v1[i];
  if (neq == 0 || neq == 10) fmm(ni, nj, neq);  // gradients of the single-strength call
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
