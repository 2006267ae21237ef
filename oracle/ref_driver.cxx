// ref_driver.cxx -- drives the reference's own serial CPU FMM (compiled from
// the reference sources by oracle/Makefile.ref) and dumps golden vectors.
//
// TEST INFRASTRUCTURE ONLY.  Runs in the development container, where the
// reference sources exist; its outputs become fixtures under tests/golden/.
//
// The reference objects are used unmodified.  This driver replaces only the
// top-level sequencing of fmm/fmm.cxx:29-152 so that it can record each
// phase and so that level 2 gets the interaction list fmm/fmm.cxx:104 asks
// for: the reference passes neib=3 there, which fmm/ijbox.cxx:9-117 never
// builds (the stale leaf near list would be used).  The set a level-2 M2L
// needs, all non-adjacent boxes, is neib 1.  With "asis" the driver also
// runs the shipped fmm() itself, for the record.
//
// Yukawa (neq 10) runs the shipped fmm() itself, as test/testh.cxx:134-141
// does, with lambda and scale set as test/testh.cxx:42,119-123: its M2L
// (kernel/Helmholtz_m2l.cxx) builds its own directional lists and does not
// read the level-2 list.
//
// Input (binary, little endian): int ni,nj,lmax,neq,asis; float xmin,ymin,
// zmin,rd; float xi,yi,zi,gxi,gyi,gzi [ni]; float xj,yj,zj,gxj,gyj,gzj,vj [nj];
// for neq 10 one more float, lambda.
// Output: records {char name[32]; int type(0 int,1 float,2 double); int n; data}.
#define MAIN
#include "constants.h"
#include "arrays.h"
#undef MAIN

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

extern void memoryuse();
extern void memoryfree();
extern void boxallocate(int, int, int, int);
extern void alloc(int);
extern void dealloc(int);
extern void boxpart(int, int);
extern void sorti(int&);
extern void sortj(int&);
extern void unsorti(int&);
extern void unsortj(int&);
extern void fmm(int, int, int);
extern void G_dir(int, int, int, int);
extern void Gni_dir(int, int, int, int);
extern void Gnj_dir(int, int, int, int);
extern void precalc(int&, int);
extern void boxdatai(int, int&, double&);
extern void boxdataj(int, int&, double&);
extern void boxm2mi(int&, int);
extern void boxm2mj(int&, int);
extern void ifbox(int, int, int, int, int, int&, int, int);
extern void G_p2p(int, int);
extern void Gni_p2p(int, int);
extern void Gnj_p2p(int, int);
extern void G_p2m(int, int, int);
extern void Gn_p2m(int, int, int);
extern void m2m(int, int, int, int, int);
extern void m2l(int, int, int, int, int, int);
extern void l2l(int, int, int, int);
extern void G_l2p(int, int, int);
extern void Gn_l2p(int, int, int);
extern void Helmholtz_G_dir(int, int, int, int);

static FILE* out;

static void rec(const char* name, int type, int n, const void* data) {
  char nm[32] = {0};
  snprintf(nm, sizeof nm, "%s", name);
  fwrite(nm, 1, 32, out);
  fwrite(&type, 4, 1, out);
  fwrite(&n, 4, 1, out);
  size_t es = type == 2 ? 8 : 4;
  if (n > 0) fwrite(data, es, (size_t)n, out);
}
static void reci(const char* name, int n, const int* d) { rec(name, 0, n, d); }
static void recf(const char* name, int n, const float* d) { rec(name, 1, n, d); }
static void recd(const char* name, int n, const double* d) { rec(name, 2, n, d); }
static void recc(const char* name, int n, const std::complex<double>* d) { rec(name, 2, 2 * n, d); }

static void dump_list(const char* tag, int lev, int lbi) {
  char nm[32];
  std::vector<int> flat;
  for (int ii = 0; ii < lbi; ii++)
    for (int ij = 0; ij < nij[ii]; ij++) flat.push_back(neij[ii][ij]);
  snprintf(nm, sizeof nm, "%s_nij_%d", tag, lev);
  reci(nm, lbi, nij);
  snprintf(nm, sizeof nm, "%s_neij_%d", tag, lev);
  reci(nm, (int)flat.size(), flat.data());
}

// fmm/fmm.cxx:29-152 with phase dumps and the level-2 list as neib 1
static void fmm_traced(int mi, int mj, int neq) {
  int lev = lmax, lmin = 2, nmp, lbi, lbj, lbjr, lbjo;
  double rb;
  char nm[32];
  nlbi[lev - 1] = 0;
  nlbj[lev - 1] = 0;
  precalc(nmp, mp);
  boxdatai(mi, lbi, rb);
  boxdataj(mj, lbj, rb);
  reci("leaf_nfi", lbi, nfi);
  reci("leaf_ndi0", lbi, ndi[0]);
  reci("leaf_ndi1", lbi, ndi[1]);
  reci("leaf_nfj", lbj, nfj);
  reci("leaf_ndj0", lbj, ndj[0]);
  reci("leaf_ndj1", lbj, ndj[1]);

  ifbox(mi, mj, nmp, lbi, lbj, lbjr, lev, 0);
  dump_list("near", lev, lbi);
  for (int i = 0; i < mi; i++) vi[i] = 0;
  if (neq == 0) {
    for (int i = 0; i < mi; i++) gxi[i] = gyi[i] = gzi[i] = 0;
    G_p2p(lbi, lbj);
  } else if (neq == 1) {
    Gni_p2p(lbi, lbj);
  } else {
    Gnj_p2p(lbi, lbj);
  }
  recf("p2p_vi", mi, vi);
  if (neq == 0) {
    recf("p2p_gxi", mi, gxi);
    recf("p2p_gyi", mi, gyi);
    recf("p2p_gzi", mi, gzi);
  }
  if (neq == 0 || neq == 1) G_p2m(nmp, mp, lbj);
  else Gn_p2m(nmp, mp, lbj);
  recc("p2m_bx", lbj * nmp, &bx[0][0]);

  if (lmax > lmin) {
    for (lev = lmax - 1; lev >= lmin; lev--) {
      lbjo = lbj;
      boxm2mi(lbi, lev);
      boxm2mj(lbj, lev);
      m2m(nmp, mp, lev, lbj, lbjo);
    }
    lev = lmin;
  } else {
    for (int i = 0; i < nbmax; i++) nej[i] = -1;
    for (int jj = 0; jj < lbj; jj++) nej[nfj[jj + nlbj[lev - 1]]] = jj;
  }
  reci("nlbi", lmax, nlbi);
  reci("nlbj", lmax, nlbj);
  reci("all_nfi", nlbi[lmin - 1] + lbi, nfi);
  reci("all_nfj", nlbj[lmin - 1] + lbj, nfj);
  recc("up_bx", (nlbj[lmin - 1] + lbj) * nmp, &bx[0][0]);

  ifbox(mi, mj, nmp, lbi, lbj, lbjr, lev, 1);  // fmm.cxx:104 passes 3
  dump_list("far", lev, lbi);
  m2l(nmp, mp, lbi, lbjr, lev, lbi);
  snprintf(nm, sizeof nm, "ax_%d", lev);
  recc(nm, lbi * nmp, &ax[0][0]);

  if (lmax > lmin) {
    for (lev = lmin + 1; lev <= lmax; lev++) {
      lbi = nlbi[lev - 2] - nlbi[lev - 1];
      lbj = nlbj[lev - 2] - nlbj[lev - 1];
      l2l(nmp, mp, lev, lbi);
      for (int i = 0; i < nbmax; i++) nej[i] = -1;
      for (int jj = 0; jj < lbj; jj++) nej[nfj[jj + nlbj[lev - 1]]] = jj;
      ifbox(mi, mj, nmp, lbi, lbj, lbjr, lev, 2);
      dump_list("far", lev, lbi);
      m2l(nmp, mp, lbi, lbjr, lev, 0);
      snprintf(nm, sizeof nm, "ax_%d", lev);
      recc(nm, lbi * nmp, &ax[0][0]);
    }
    lev = lmax;
  }
  if (neq == 0 || neq == 2) G_l2p(nmp, mp, lbi);
  else Gn_l2p(nmp, mp, lbi);
}

static void read_floats(FILE* f, float* d, int n) {
  if (n > 0 && fread(d, 4, (size_t)n, f) != (size_t)n) { fprintf(stderr, "short read\n"); exit(2); }
}

// The root cube of the BEM caller, bio/main.cxx:355-368 (that file needs
// PETSc and is not built here): the same statements on the same types --
// the reference's float globals xmin, ymin, zmin, rd, float locals xmax..,
// rdx.., the double literal 1+1e-5 -- compiled with the reference's flags.
// Input: int ncloud, then per cloud int n, float x[n], y[n], z[n].
// Output: one record box_<k> = {xmin, ymin, zmin, rd} per cloud.
static int box_mode(const char* fin_path, const char* fout_path) {
  FILE* fi = fopen(fin_path, "rb");
  if (!fi) return 2;
  out = fopen(fout_path, "wb");
  if (!out) return 2;
  int ncloud = 0;
  if (fread(&ncloud, 4, 1, fi) != 1) return 2;
  for (int k = 0; k < ncloud; k++) {
    int n = 0;
    if (fread(&n, 4, 1, fi) != 1) return 2;
    std::vector<float> px(n), py(n), pz(n);
    read_floats(fi, px.data(), n);
    read_floats(fi, py.data(), n);
    read_floats(fi, pz.data(), n);
    float xmax, ymax, zmax, rdx, rdy, rdz;
    xmin = 1e6;
    xmax = -1e6;
    ymin = 1e6;
    ymax = -1e6;
    zmin = 1e6;
    zmax = -1e6;
    for (int l = 0; l < n; ++l) {
      xmin = std::min(px[l], xmin);
      xmax = std::max(px[l], xmax);
      ymin = std::min(py[l], ymin);
      ymax = std::max(py[l], ymax);
      zmin = std::min(pz[l], zmin);
      zmax = std::max(pz[l], zmax);
    }
    rdx = (xmax - xmin) * (1 + 1e-5);
    rdy = (ymax - ymin) * (1 + 1e-5);
    rdz = (zmax - zmin) * (1 + 1e-5);
    rd = std::max(rdx, rdy);
    rd = std::max(rd, rdz);
    const float b[4] = {xmin, ymin, zmin, rd};
    char nm[32];
    snprintf(nm, sizeof nm, "box_%d", k);
    recf(nm, 4, b);
  }
  fclose(fi);
  fclose(out);
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 4 && std::string(argv[1]) == "box") return box_mode(argv[2], argv[3]);
  if (argc < 3) { fprintf(stderr, "usage: ref_driver in.bin out.bin | ref_driver box in.bin out.bin\n"); return 2; }
  FILE* fi = fopen(argv[1], "rb");
  if (!fi) return 2;
  int hdr[5];
  float box[4];
  if (fread(hdr, 4, 5, fi) != 5 || fread(box, 4, 4, fi) != 4) return 2;
  int ni = hdr[0], nj = hdr[1], neq = hdr[3], asis = hdr[4];
  lmax = hdr[2];
  xmin = box[0]; ymin = box[1]; zmin = box[2]; rd = box[3];
  nprocs = 1; myrank = 0; umem = 0;
  // array set-up as test/test.cxx:42-65
  xi = new float[npmax]; yi = new float[npmax]; zi = new float[npmax];
  gxi = new float[npmax]; gyi = new float[npmax]; gzi = new float[npmax]; vi = new float[npmax];
  xj = new float[npmax]; yj = new float[npmax]; zj = new float[npmax];
  gxj = new float[npmax]; gyj = new float[npmax]; gzj = new float[npmax]; vj = new float[npmax];
  nfn = new int[npmax]; na = new int[npmax]; nb = new int[npmax]; nc = new int[npmax]; nd = new int[npmax];
  nbi = new int[npmax]; nbj = new int[npmax]; sortd = new float[npmax];
  tfmm = new double[100];
  for (int i = 0; i < 100; i++) tfmm[i] = 0;
  read_floats(fi, xi, ni); read_floats(fi, yi, ni); read_floats(fi, zi, ni);
  read_floats(fi, gxi, ni); read_floats(fi, gyi, ni); read_floats(fi, gzi, ni);
  read_floats(fi, xj, nj); read_floats(fi, yj, nj); read_floats(fi, zj, nj);
  read_floats(fi, gxj, nj); read_floats(fi, gyj, nj); read_floats(fi, gzj, nj); read_floats(fi, vj, nj);
  if (neq == 10) {
    float lam;
    read_floats(fi, &lam, 1);
    lambda = lam;
    scale = rd * lambda > 1 ? 1 / rd : lambda;  // test/testh.cxx:119-123
  }
  fclose(fi);
  std::vector<float> x0(xi, xi + ni), y0(yi, yi + ni), z0(zi, zi + ni), nx0(gxi, gxi + ni),
      ny0(gyi, gyi + ni), nz0(gzi, gzi + ni);
  std::vector<float> xj0(xj, xj + nj), yj0(yj, yj + nj), zj0(zj, zj + nj), nxj0(gxj, gxj + nj),
      nyj0(gyj, gyj + nj), nzj0(gzj, gzj + nj), q0(vj, vj + nj);

  out = fopen(argv[2], "wb");
  if (!out) return 2;
  int meta[4] = {mp, lmax, neq, nebm};
  reci("meta", 4, meta);

  // setup chain as test/test.cxx:107-112 (lmax set by the caller, not nlevel)
  boxpart(ni, nj);
  sorti(ni);
  sortj(nj);
  boxallocate(0, ni, 0, nj);
  alloc(1);
  reci("nbi", ni, nbi);
  reci("nbj", nj, nbj);
  if (neq == 10) {
    double sc[2] = {lambda, scale};
    recd("yukawa", 2, sc);
    fmm(ni, nj, neq);
  } else {
    fmm_traced(ni, nj, neq);
  }
  {
    int nmp = mp * (mp + 1) / 2;
    (void)nmp;
    recf("fac", 4 * mpmax, fac);
    recf("anm", 4 * mpmax, anm);
    recf("sr", mpmax * mpmax, sr);
    recc("ynm", 4 * mpmax, ynm);
    // a sample of rotation matrices (all 1024 are 12 MB at mp=9)
    const int sample[8] = {0, 7, 100, 219, 292, 300, 511, 73};
    std::vector<std::complex<double>> dd;
    for (int s = 0; s < 8; s++)
      for (int inv = 0; inv < 2; inv++)
        for (int m = 0; m < mp; m++)
          for (int k = 0; k < mpmax; k++) dd.push_back(dnm[sample[s] + inv * nrbm][m][k]);
    recc("dnm_sample", (int)dd.size(), dd.data());
    reci("dnm_sample_idx", 8, sample);
  }
  recf("fmm_vi_sorted", ni, vi);
  if (neq == 0) {
    recf("fmm_gxi_sorted", ni, gxi);
    recf("fmm_gyi_sorted", ni, gyi);
    recf("fmm_gzi_sorted", ni, gzi);
  }
  unsorti(ni);
  unsortj(nj);
  dealloc(1);
  recf("fmm_vi", ni, vi);
  if (neq == 0) {
    recf("fmm_gxi", ni, gxi);
    recf("fmm_gyi", ni, gyi);
    recf("fmm_gzi", ni, gzi);
  }
  std::vector<float> v_fmm(vi, vi + ni);

  // direct sum (test/test.cxx:134-145) on the caller-order inputs
  for (int i = 0; i < ni; i++) {
    xi[i] = x0[i]; yi[i] = y0[i]; zi[i] = z0[i]; gxi[i] = nx0[i]; gyi[i] = ny0[i]; gzi[i] = nz0[i];
  }
  for (int j = 0; j < nj; j++) {
    xj[j] = xj0[j]; yj[j] = yj0[j]; zj[j] = zj0[j]; gxj[j] = nxj0[j]; gyj[j] = nyj0[j]; gzj[j] = nzj0[j];
    vj[j] = q0[j];
  }
  if (neq == 0) G_dir(0, ni - 1, 0, nj - 1);
  else if (neq == 1) Gni_dir(0, ni - 1, 0, nj - 1);
  else if (neq == 2) Gnj_dir(0, ni - 1, 0, nj - 1);
  else Helmholtz_G_dir(0, ni - 1, 0, nj - 1);
  recf("dir_vi", ni, vi);
  if (neq == 0) {
    recf("dir_gxi", ni, gxi);
    recf("dir_gyi", ni, gyi);
    recf("dir_gzi", ni, gzi);
  }

  fflush(out);
  if (asis) {
    // the shipped fmm() with its level-2 neib=3 call, same inputs
    for (int i = 0; i < ni; i++) {
      xi[i] = x0[i]; yi[i] = y0[i]; zi[i] = z0[i]; gxi[i] = nx0[i]; gyi[i] = ny0[i]; gzi[i] = nz0[i];
    }
    for (int j = 0; j < nj; j++) {
      xj[j] = xj0[j]; yj[j] = yj0[j]; zj[j] = zj0[j]; gxj[j] = nxj0[j]; gyj[j] = nyj0[j]; gzj[j] = nzj0[j];
      vj[j] = q0[j];
    }
    boxpart(ni, nj);
    sorti(ni);
    sortj(nj);
    boxallocate(0, ni, 0, nj);
    alloc(1);
    fmm(ni, nj, neq);
    unsorti(ni);
    unsortj(nj);
    dealloc(1);
    recf("asis_vi", ni, vi);
    if (neq == 0) {
      recf("asis_gxi", ni, gxi);
      recf("asis_gyi", ni, gyi);
      recf("asis_gzi", ni, gzi);
    }
  }
  fclose(out);
  return 0;
}
