/*
 * fmmx_refshim.h -- the reference's own C++ entry points, served by the
 * MI355X engine (libfmmx_ref.so, linked against libfmmx.so).
 *
 * A caller written against the reference (test/test.cxx, bio/main.cxx)
 * links against libfmmx_ref.so instead of the reference's fmm/, sort/ and
 * kernel/ objects and runs unchanged: the symbols below have the
 * reference's names, C++ signatures and side effects on the caller's global
 * arrays.  As in the reference, the globals themselves are DEFINED by the
 * caller (include/arrays.h and include/constants.h compiled with MAIN);
 * the shim only references the ones listed at the end of this header.
 *
 * Differences from the reference, by design:
 *  - the FMM runs on the GPU through one process-wide fmmx handle (device 0
 *    unless configured), FP64 arithmetic and the reference's conventions
 *    (softened kernel, self pair included) by default; expansion order p is
 *    9 as in include/constants.h:12 unless configured;
 *  - boxpart() fills nek/irank (used only by the MPI partitioner) but the
 *    tree lists, expansions and tables of alloc(1) stay inside the engine;
 *  - fmm(mi, mj, 2) leaves gxi, gyi, gzi untouched (the reference adds far
 *    field gradients into them, fmm/fmm.cxx:145, on top of the caller's
 *    values);
 *  - fmm(mi, mj, 10) (Yukawa, screening from the global lambda) writes the
 *    field gradient into gxi, gyi, gzi (the reference zeroes them and adds a
 *    far-field term that is not the Yukawa gradient,
 *    kernel/Helmholtz_G_l2p.cxx:60-70), and translates exactly instead of
 *    with plane waves, at any depth (the reference's path runs only at
 *    lmax 2, kernel/Helmholtz_m2l.cxx:111-122);
 *  - errors print a message and exit(1), as the reference's CUDA error path
 *    does (cutil-mp.h:776-787);
 *  - alloc(1) allocates the box lists, interaction lists and expansions
 *    fmm/fmm.cxx and the operator symbols use (the reference's sizes); the
 *    reference's tables (fac, anm, sr, ynm, dnm) are not allocated.
 */
#ifndef FMMX_REFSHIM_H
#define FMMX_REFSHIM_H

#include "fmmx.h"

/* Configuration (C linkage, no counterpart in the reference).  Must be
 * called before the first shim call to take effect; the environment
 * variables FMMX_REF_P, FMMX_REF_PRECISION (fp64|mixed),
 * FMMX_REF_CONVENTION (ref|exact) and FMMX_REF_DEVICE set the same values
 * for unchanged binaries.  p is 2..12, or 18 (a caller compiled with
 * mp = 18) with FMMX_FP64 and FMMX_CONV_EXACT.  Returns FMMX_OK,
 * FMMX_EINVAL or (p = 18 otherwise) FMMX_EUNSUPPORTED. */
extern "C" int fmmx_ref_configure(int p, int precision, int convention, int device);
/* The shim's handle (created on first use), e.g. for fmmx_get_stats. */
extern "C" fmmx_handle* fmmx_ref_handle(void);

/* fmm/nlevel.cxx:5 -- lmax from the source count (prints it as the
 * reference does) */
void nlevel(int nmax);
/* fmm/boxpart.cxx:11 -- nbmax, nek (box key -> ordinal), irank (all 0) */
void boxpart(int mi, int mj);
/* sort/sorti.cxx:12 and sort/sortj.cxx:12 -- Morton-sort the i arrays
 * (xi, yi, zi, gxi, gyi, gzi) / j arrays (xj, yj, zj, gxj, gyj, gzj, vj) in
 * place; nbi / nbj receive the permutation */
void sorti(int& mi);
void sortj(int& mj);
/* sort/sorti.cxx:38 and sort/sortj.cxx:38 -- undo it (the i side also
 * restores vi) */
void unsorti(int& mi);
void unsortj(int& mj);
/* fmm/boxallocate.cxx:10 -- nbmax, nbne, nbnet */
void boxallocate(int n0, int n1, int n2, int n3);
/* fmm/alloc.cxx:32, 124 */
void alloc(int ialloc);
void dealloc(int ialloc);
/* fmm/fmm.cxx:29 -- one FMM matvec on the sorted arrays: vi (and gxi..gzi
 * for neq 0) from vj at xj (normals gxi.. for neq 1, gxj.. for neq 2) */
void fmm(int mi, int mj, int neq);
/* ---- The operator plugin: the symbols fmm/fmm.cxx:7-27 declares ----
 * The reference picks its kernels at link time (biobem/Makefile:25-37 links
 * the kernel/..._gpu.o objects instead of kernel/....o).  These symbols serve fmm/fmm.cxx
 * compiled from the reference's sources: with the shim's set-up chain and
 * box symbols, or with the reference's own tree objects (boxdatai.o,
 * ijbox.o, ...) and only the kernels below taken from the shim.  Each kernel
 * symbol runs one phase on the engine's device-resident state
 * (fmmx_evaluate_phase), finding the engine's tree for the caller's current
 * arrays itself; P2P and L2P add into vi (and gxi.. for G), as the
 * reference's kernels do; multipoles and locals are mirrored into bx and ax
 * after each phase (Laplace kernels; the unscaled layout of
 * include/arrays.h:38).  fmm/fmm.cxx:104 asks ifbox for neib 3 at level 2,
 * a list fmm/ijbox.cxx never builds: ifbox hands out the neib-1 list there
 * and m2l at level 2 uses it. */
/* kernel/precalc.cxx:10 -- nmp = mp(mp+1)/2 (mp must be the shim's p) */
void precalc(int& nmp, int mp);
/* fmm/boxdatai.cxx:10, boxdataj.cxx:10 -- leaf boxes of the sorted arrays:
 * nfi/ndi (nfj/ndj/nej), lbi (lbj), rb = rd / 2^lmax */
void boxdatai(int mi, int& lbi, double& rb);
void boxdataj(int mj, int& lbj, double& rb);
/* fmm/boxm2mi.cxx:6, boxm2mj.cxx:6 -- the parent level's boxes stacked in
 * nfi (nfj, with nej) at nlbi[lev-1] (nlbj[lev-1]) */
void boxm2mi(int& lbi, int lev);
void boxm2mj(int& lbj, int lev);
/* fmm/ifbox.cxx:7 -- lbjr = lbj; nij/neij/njb for level lev (neib 0 near,
 * 1 and 2 far, 3 as 1) */
void ifbox(int mi, int mj, int nmp, int lbi, int lbj, int& lbjr, int lev, int neib);
/* kernel/G_p2p.cxx:7, Gni_p2p.cxx:7, Gnj_p2p.cxx:7, Helmholtz_G_p2p.cxx:7 */
void G_p2p(int lbi, int lbj);
void Gni_p2p(int lbi, int lbj);
void Gnj_p2p(int lbi, int lbj);
void Helmholtz_G_p2p(int lbi, int lbj);
/* kernel/G_p2m.cxx:11, Gn_p2m.cxx:11, Helmholtz_G_p2m.cxx:11 */
void G_p2m(int nmp, int mp, int lbj);
void Gn_p2m(int nmp, int mp, int lbj);
void Helmholtz_G_p2m(int nmp, int mp, int lbj);
/* kernel/m2m.cxx:13, Helmholtz_m2m.cxx:13 */
void m2m(int nmp, int mp, int lev, int lbj, int lbjo);
void Helmholtz_m2m(int nmp, int mp, int lev, int lbj, int lbjo);
/* kernel/m2l.cxx:13, Helmholtz_m2l.cxx:13 */
void m2l(int nmp, int mp, int lbi, int lbj, int lev, int ini);
void Helmholtz_m2l(int nmp, int mp, int lbi, int lbj, int lev, int ini);
/* kernel/l2l.cxx:13 */
void l2l(int nmp, int mp, int lev, int lbi);
/* kernel/G_l2p.cxx:11, Gn_l2p.cxx:11, Helmholtz_G_l2p.cxx:11 */
void G_l2p(int nmp, int mp, int lbi);
void Gn_l2p(int nmp, int mp, int lbi);
void Helmholtz_G_l2p(int nmp, int mp, int lbi);

/* kernel/G_dir.cxx:6, Gni_dir.cxx:6, Gnj_dir.cxx:6 -- direct sums over
 * targets n0..n1 and sources n2..n3 (inclusive), overwriting the outputs */
void G_dir(int n0, int n1, int n2, int n3);
void Gni_dir(int n0, int n1, int n2, int n3);
void Gnj_dir(int n0, int n1, int n2, int n3);
/* kernel/Helmholtz_G_dir.cxx:6 -- Yukawa direct sum, potential only */
void Helmholtz_G_dir(int n0, int n1, int n2, int n3);

/* The caller's globals the shim reads or writes (include/arrays.h:5-8,
 * 30-34, 38, 60, 64; include/constants.h:39-43, 57-61).  Also ax and bx
 * (include/arrays.h:38, std::complex<double> (*)[mpsym] with the caller's
 * mpsym), not declared here because their type depends on the caller's mp. */
extern float *xi, *yi, *zi, *gxi, *gyi, *gzi, *vi;
extern float *xj, *yj, *zj, *gxj, *gyj, *gzj, *vj;
extern int *nbi, *nbj, *nek, *irank;
extern int **ndi, *nfi, *nlbi, **ndj, *nej, *nfj, *nlbj, *nij, *njb;
extern int (*neij)[189];
extern double* tfmm;
extern int nbmax, nbne, nbnet, lmax;
extern float xmin, ymin, zmin, rd;
extern double lambda;

#endif
