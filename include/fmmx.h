/*
 * fmmx.h -- C ABI of the MI355X FMM engine (libfmmx.so).
 *
 * The engine evaluates the reference's Laplace FMM matvec (fmm/fmm.cxx:29,
 * "void fmm(int mi, int mj, int neq)") on one GPU: the tree (Morton keys,
 * stable sort, box lists, interaction lists), P2P, P2M, M2M, M2L, L2L, L2P
 * and the direct sums, with the reference's kernels G (neq 0), Gni (neq 1)
 * and Gnj (neq 2).
 *
 * The reference's interface is a set of C++ symbols over global arrays
 * (include/arrays.h, include/constants.h) that callers sequence themselves
 * (test/test.cxx:107-116, bio/main.cxx:327-336).  This ABI replaces it with
 * an explicit handle: plain pointers and sizes, no global state, integer
 * status codes instead of exit().  libfmmx_ref.so (fmmx_refshim.h) layers
 * the reference's own C++ symbols on top of it for unchanged callers.
 *
 * Threading: one handle per host thread; calls on one handle are not
 * reentrant.  All work runs on the handle's HIP stream.
 */
#ifndef FMMX_H
#define FMMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMMX_ABI_VERSION 1

/* status codes */
enum {
  FMMX_OK = 0,
  FMMX_EINVAL = 1,   /* bad argument (sizes, order, NaN coordinates, ...) */
  FMMX_ESTATE = 2,   /* call out of sequence (e.g. evaluate before set_points) */
  FMMX_EDEVICE = 3,  /* HIP runtime error */
  FMMX_ENOMEM = 4,   /* device or host allocation failed */
  FMMX_EUNSUPPORTED = 5
};

/* kernels: the reference's neq values (fmm/fmm.cxx:47-65); FMMX_YUKAWA is
 * its "Helmholtz" (neq 10): the screened kernel (pi/2) exp(-lambda r) /
 * (lambda r) (kernel/Helmholtz_G_p2p.cxx:22), lambda from fmmx_set_yukawa.
 * FMMX_HELMHOLTZ has no reference counterpart: the oscillatory kernel
 * exp(i k r) / (4 pi r) with complex charges, k from fmmx_set_helmholtz. */
enum { FMMX_G = 0, FMMX_GNI = 1, FMMX_GNJ = 2, FMMX_YUKAWA = 10, FMMX_HELMHOLTZ = 20 };

/* arithmetic of the evaluation */
enum {
  FMMX_MIXED = 0, /* FP32 pair/expansion arithmetic, FP64 tables and per-target sums */
  FMMX_FP64 = 1   /* FP64 throughout ("plain") */
};

/* conventions */
enum {
  FMMX_CONV_REF = 0,  /* reference arithmetic: r^2+eps softening (self pair
                         included, include/constants.h:34), eps added to radii
                         in expansions and rotations, float pi, float tables */
  FMMX_CONV_EXACT = 1 /* unsoftened 1/r kernel, r=0 pairs (self) excluded,
                         exact geometry and tables */
};

/* flags */
enum {
  FMMX_DEVICE_PTRS = 1, /* all array arguments of this call are device pointers */
  FMMX_KEEP_LISTS = 2   /* set_points: also build the far-field (M2L)
                           interaction lists in the reference's order for
                           fmmx_get_list (M2L itself does not need them) */
};

typedef struct fmmx_config {
  int p;          /* expansion degrees n = 0..p-1 (the reference's mp):
                     2..12 in every precision and convention, and 18 (the
                     reference's other compiled order, kernel/precalch.cxx:185)
                     with FMMX_FP64 and FMMX_CONV_EXACT */
  int lmax;       /* leaf level >= 2; 0 = automatic: FMMX_CONV_REF the
                     fmm/nlevel.cxx thresholds on nj, FMMX_CONV_EXACT the
                     measured cost model of fmmx_auto_lmax */
  int precision;  /* FMMX_MIXED or FMMX_FP64 */
  int convention; /* FMMX_CONV_REF or FMMX_CONV_EXACT */
  float xmin, ymin, zmin, rd; /* root cube (include/constants.h:56-59);
                                 rd <= 0: computed from the points as
                                 bio/main.cxx:364-380 does */
} fmmx_config;

#define FMMX_MAX_P 18

typedef struct fmmx_handle fmmx_handle;

/* Create a handle on HIP device `device`; builds the FP64 tables and the
 * dense translation operators for (p, convention).  stream may be NULL
 * (the default stream). */
int fmmx_create(const fmmx_config* cfg, int device, void* hip_stream, fmmx_handle** out);
int fmmx_destroy(fmmx_handle* h);

/* Tree set-up for targets (xi..) and sources (xj..), caller order: the
 * reference's nlevel/boxpart/sorti/sortj/boxallocate/alloc chain
 * (test/test.cxx:107-112) plus fmm.cxx steps boxdatai/boxdataj and the
 * interaction lists.  nx*: target normals (needed for FMMX_GNI) and source
 * normals (FMMX_GNJ); NULL when unused.  Reused by every evaluate until the
 * next set_points.  Passing the same x, y, z pointers for both sides (ni ==
 * nj) builds the tree once and shares it. */
int fmmx_set_points(fmmx_handle* h, int ni, const float* xi, const float* yi, const float* zi,
                    const float* nxi, const float* nyi, const float* nzi, int nj, const float* xj,
                    const float* yj, const float* zj, const float* nxj, const float* nyj,
                    const float* nzj, int flags);

/* One FMM matvec (fmm/fmm.cxx:29-152) for source strengths qj (caller
 * order).  Outputs in caller order: vi (potential, or the Gni/Gnj layer
 * potential) and, for FMMX_G only, the gradient gxi,gyi,gzi (may be NULL).
 * Unlike the reference, outputs are overwritten (not accumulated) and are
 * double.  Factor 1/(4 pi) included as in kernel/G_p2p.cxx:25.
 * FMMX_YUKAWA also returns the gradient (the reference computes the
 * potential only, kernel/Helmholtz_G_p2p.cxx).
 * FMMX_HELMHOLTZ: qj holds nj complex charges as (re, im) float pairs and vi
 * receives ni complex potentials as (re, im) double pairs; no gradient (the
 * g pointers must be NULL).  p must be one of 4, 6, 8, 10, 12 and k times
 * the level-2 box size at most 5. */
int fmmx_evaluate(fmmx_handle* h, int neq, const float* qj, double* vi, double* gxi, double* gyi,
                  double* gzi, int flags);

/* One phase of the matvec at a time, the reference's operator sequence
 * (fmm/fmm.cxx:42-150) on the tree of the last fmmx_set_points; strengths,
 * expansions and partial sums stay on the device between calls:
 *   FMMX_PHASE_P2P  near field of qj (kernel/G_p2p.cxx:7, Gni_p2p.cxx:7,
 *                   Gnj_p2p.cxx:7, Helmholtz_G_p2p.cxx) -> vi (and g)
 *   FMMX_PHASE_P2M  leaf multipoles of qj (kernel/G_p2m.cxx:11; the dipole
 *                   form Gn_p2m.cxx:11 for FMMX_GNJ)
 *   FMMX_PHASE_M2M  multipoles of level lev (2..lmax-1) from lev+1
 *                   (kernel/m2m.cxx:13)
 *   FMMX_PHASE_M2L  locals of level lev from its interaction list
 *                   (kernel/m2l.cxx:13): level 2 writes, deeper levels add to
 *                   what L2L put there
 *   FMMX_PHASE_L2L  locals of level lev (3..lmax) from lev-1, overwritten
 *                   (kernel/l2l.cxx:13)
 *   FMMX_PHASE_L2P  far field -> vi (and g) (kernel/G_l2p.cxx:11; the normal
 *                   derivative Gn_l2p.cxx:11 for FMMX_GNI)
 * neq is the kernel of the whole sequence (FMMX_G, FMMX_GNI, FMMX_GNJ or
 * FMMX_YUKAWA); qj is read by P2P and P2M only, vi and g written by P2P and
 * L2P only (caller order, overwritten; the P2P and L2P outputs add up to
 * fmmx_evaluate's).  Expansions can be read after each phase
 * (fmmx_get_multipoles_level, fmmx_get_locals).  No Helmholtz phases. */
enum {
  FMMX_PHASE_P2P = 0,
  FMMX_PHASE_P2M = 1,
  FMMX_PHASE_M2M = 2,
  FMMX_PHASE_M2L = 3,
  FMMX_PHASE_L2L = 4,
  FMMX_PHASE_L2P = 5
};
int fmmx_evaluate_phase(fmmx_handle* h, int phase, int neq, int lev, const float* qj, double* vi, double* gxi,
                        double* gyi, double* gzi, int flags);

/* ---- Adjoint of the matvec v = K q (fmmx_evaluate with kernel neq) ----
 * No reference counterpart (the reference has no derivatives); built from
 * its kernels.  Every entry recomputes the expansions it needs.
 *
 * fmmx_evaluate_transpose: u = K^T w for strengths wi on the targets (caller
 * order), one value per source (caller order).  G, Yukawa and Helmholtz are
 * symmetric kernels; the transpose of Gni (kernel/Gni_p2p.cxx:7-32, target
 * normals) is minus Gnj (kernel/Gnj_p2p.cxx:7-32) carrying the target
 * normals, and the other way round.  For G and Yukawa g*j (all three or
 * none) receive the gradient of the transposed field at the sources.  The
 * transposed problem runs on a second tree of the same root cube and depth
 * whose targets are the sources, built at the first call after set_points.
 * Helmholtz: wi and uj complex pairs as in fmmx_evaluate. */
int fmmx_evaluate_transpose(fmmx_handle* h, int neq, const float* wi, double* uj, double* gxj, double* gyj,
                            double* gzj, int flags);
/* Gradients of L = sum_i w_i (K q)_i for G or Yukawa: dq[j] = dL/dq_j =
 * (K^T w)_j; dxi = dL/dx_i = w_i grad phi(x_i) (the field gradient of
 * kernel/G_l2p.cxx:68-70); dxj = dL/dx_j = q_j grad psi(x_j) with
 * psi = K^T w.  dxi, dxj: x, y and z components as three consecutive blocks
 * of ni (nj) doubles.  Any output may be NULL. */
int fmmx_gradients(fmmx_handle* h, int neq, const float* qj, const float* wi, double* dq, double* dxi,
                   double* dxj, int flags);
/* The energy E = 1/2 q . K q of one point set (set_points with targets =
 * sources; the self pair excluded in FMMX_CONV_EXACT) for G or Yukawa, and
 * its gradients dEdq = K q and dEdx = q_k grad phi(x_k) (three blocks of n).
 * E is summed on the device in a fixed order.  Any output may be NULL. */
int fmmx_energy(fmmx_handle* h, int neq, const float* q, double* E, double* dEdq, double* dEdx, int flags);

/* Direct sums (kernel/G_dir.cxx:6, Gni_dir.cxx:6, Gnj_dir.cxx:6) on the GPU
 * for ni targets against nj sources, FP64 accumulation, in the handle's
 * convention.  Stateless apart from the handle's device and stream (and its
 * lambda or k).  FMMX_HELMHOLTZ takes and returns complex pairs as
 * fmmx_evaluate does. */
int fmmx_direct(fmmx_handle* h, int neq, int ni, const float* xi, const float* yi, const float* zi,
                const float* nxi, const float* nyi, const float* nzi, int nj, const float* xj,
                const float* yj, const float* zj, const float* nxj, const float* nyj,
                const float* nzj, const float* qj, double* vi, double* gxi, double* gyi,
                double* gzi, int flags);

/* Replace the root cube and leaf level fixed at fmmx_create (rd <= 0: from
 * the points at each set_points; lmax 0: fmm/nlevel.cxx thresholds).  The
 * reference reads them from globals at every call (include/constants.h:56-59,
 * fmm/nlevel.cxx:11); takes effect at the next set_points. */
int fmmx_set_domain(fmmx_handle* h, int lmax, float xmin, float ymin, float zmin, float rd);

/* Screening constant lambda > 0 of FMMX_YUKAWA (the reference's global
 * lambda, include/constants.h:61, test/testh.cxx:42).  Its translation
 * operators are built at the first Yukawa evaluate after a change. */
int fmmx_set_yukawa(fmmx_handle* h, double lambda);

/* M2L launch variant at level lev (2..8) for this handle: N-tiles of 16
 * targets per wave (1|2; 4 only with staging 5), operator staging (0 LDS, 1
 * register prefetch, 2 double buffer, up to p = 11; 3 and 4: the
 * error-compensated FP16 kernel on 16x16x32 MFMA with one or two operator
 * buffers, 5: the same on 32x32x16 MFMA (32 targets per wave, or 64 with 4
 * N-tiles, p <= 10); FP32 engine, Laplace, EXACT convention, p = 6..12),
 * waves per workgroup (4|8|16, 16 with staging 2), offset splits (1..16); 0
 * (-1 for staging) keeps the default.  The FP64
 * and Helmholtz kernels use the split count only.  fmmx_get_m2l_variant returns the four
 * values in effect (after set_points).  The tuner that picks them by timing
 * whole evaluations is old_fmm_bem_amd.FMM.autotune.  No reference
 * counterpart. */
int fmmx_set_m2l_variant(fmmx_handle* h, int lev, int nt, int staging, int wpg, int nsplit);
int fmmx_get_m2l_variant(fmmx_handle* h, int lev, int* v4);

/* One evaluation captured as a HIP graph: fmmx_graph_capture runs
 * fmmx_evaluate(h, neq, qj, vi, gxi, gyi, gzi, FMMX_DEVICE_PTRS) once
 * eagerly (sizing buffers, uploading operators), then records the same call
 * on the handle's stream; fmmx_graph_launch replays it asynchronously on
 * that stream, reading qj and writing vi (and g) again: write new strengths
 * into qj between replays.  Device pointers only; per-phase timing must be
 * off.  fmmx_set_points, fmmx_set_groups and a new capture drop the graph.
 * No reference counterpart: for iterative solvers that repeat one matvec. */
int fmmx_graph_capture(fmmx_handle* h, int neq, const float* qj, double* vi, double* gxi, double* gyi,
                       double* gzi);
int fmmx_graph_launch(fmmx_handle* h);

/* Independent problems in one tree: with level > 0 the points in different
 * boxes of that level (of the root cube) do not interact at all, near or
 * far; each box is its own problem and all of them run in the same launches.
 * Put each problem in its own box of the level (e.g. surface s of a batch in
 * cell s of the 2^level grid).  level 0 turns it off.  No reference
 * counterpart: for callers with many small, uneven problems (batched BEM
 * solves).  Takes effect at the next fmmx_set_points (required), which
 * fails with FMMX_EINVAL when an explicit lmax is below the level; an
 * automatic lmax (0) is raised to it. */
int fmmx_set_groups(fmmx_handle* h, int level);

/* Wavenumber k > 0 of FMMX_HELMHOLTZ.  No reference counterpart (the
 * reference has no oscillatory kernel).  Its translation operators are built
 * at the first Helmholtz evaluate after a change. */
int fmmx_set_helmholtz(fmmx_handle* h, double k);

/* Morton keys of n points at the leaf level and their stable sort
 * (fmm/boxn.cxx:5-31 then sort/csort.cxx:5-17, as sort/sorti.cxx:12-26
 * runs them): perm[k] = caller index of the k-th point in key order (the
 * reference's nbi/nbj), keys[k] = its key (may be NULL).  Needs an explicit
 * root cube and lmax (fmmx_set_domain or the config).  Synchronous. */
int fmmx_sort(fmmx_handle* h, int n, const float* x, const float* y, const float* z, int* keys, int* perm,
              int flags);

/* fmmx_sort without a handle and without allocating: no reference
 * counterpart; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The keys and stable sort of the n points (x, y, z) in the root cube
 * (xmin, ymin, zmin, rd) at leaf level lmax (1..8), as fmmx_sort computes
 * them (perm[k]: input index of the k-th point in key order, keys[k] its
 * key), asynchronously on `stream` of `device`.  Every pointer is a device
 * pointer and none may be NULL; *bad is set to 1 when a coordinate lies
 * outside the closed root cube or is not finite, else 0 (read it after the
 * stream has finished).  n is at most FMMX_SORT_NOALLOC_MAX: the
 * temporaries (ping-pong keys and values, digit counts) live in one static
 * device array of about 9 bytes per point of that maximum (18.9 MB), which
 * each device holds until the process ends.  All calls on one device share
 * that array: calls that overlap in time on one device race (serialize them
 * on one stream).  The first call on a device looks the array up; make it
 * (with n = 0 if need be) outside graph capture. */
#define FMMX_SORT_NOALLOC_MAX (1 << 21)
int fmmx_sort_noalloc(int device, void* stream, int n, const float* x, const float* y, const float* z, int lmax,
                      float xmin, float ymin, float zmin, float rd, int* keys, int* perm, int* bad);

/* Wait for the handle's stream. */
int fmmx_synchronize(fmmx_handle* h);

/* Timing and sizes of the last set_points/evaluate (device time from HIP
 * events per phase, milliseconds). */
typedef struct fmmx_stats {
  double ms_tree, ms_lists, ms_p2p, ms_p2m, ms_m2m, ms_m2l, ms_l2l, ms_l2p, ms_total;
  int lmax, nbi_leaf, nbj_leaf, nbi_total, nbj_total;
  int64_t p2p_pairs;     /* sum over near box pairs of n_i * n_j */
  int64_t m2l_pairs;     /* number of (target box, source box) M2L pairs */
  int64_t m2l_tiles;     /* M2L work tiles launched */
  int64_t m2l_pairs_leaf; /* M2L pairs at the leaf level */
  double ms_m2l_leaf;     /* duration of the M2L kernel launch that covers
                             the leaf level, alone (per-phase timing only;
                             every level's when fmmx_get_count what 2 is 1) */
} fmmx_stats;
int fmmx_get_stats(fmmx_handle* h, fmmx_stats* st);
/* Enable/disable per-phase event timing (default off: one event pair). */
int fmmx_set_timing(fmmx_handle* h, int per_phase);

/* ---- introspection of the tree for parity checks (host outputs) ---- */
/* lmax in use and box counts: what = 0 lmax, 1 boxes of side at level lev
 * (side 0 = targets i, 1 = sources j), 2: 1 if the last evaluate ran the
 * M2L of every level in one launch (FP32 Laplace and Yukawa), else 0 */
int fmmx_get_count(fmmx_handle* h, int what, int side, int lev);
/* root cube and leaf level of the tree in use (after set_points): the
 * reference's globals xmin, ymin, zmin, rd and lmax (include/constants.h:
 * 56-59), as fmmx_set_domain gave them or as computed from the points
 * (rd <= 0, bio/main.cxx:355-368; lmax 0, fmm/nlevel.cxx:8) */
int fmmx_get_domain(fmmx_handle* h, int* lmax, float* box4);
/* The depth rule of lmax 0 under FMMX_CONV_EXACT (host arithmetic only, no
 * handle): the leaf level 2..8 of least modelled cost, from the box
 * statistics of one point set at levels l = 0..8 of its root cube:
 * boxes[l] = boxes holding at least one point, sumsq[l] = sum over boxes of
 * the squared point count.  The model's far-field/near-field cost ratio was
 * fitted per order p (2..12) and precision to evaluation times measured on
 * MI355X; at p = 18 (FP64 only) it is the p = 12 ratio scaled by (p/12)^4.
 * No counterpart in the reference, whose depth is fmm/nlevel.cxx:8 (kept
 * under FMMX_CONV_REF).  The level goes to *lmax. */
int fmmx_auto_lmax(const long long* boxes, const double* sumsq, int p, int precision, int* lmax);
/* the box statistics (levels 0..8) the last set_points chose its depth from
 * (its sources, else its targets); zeros when lmax was given or the
 * convention is FMMX_CONV_REF */
int fmmx_get_occupancy(fmmx_handle* h, long long* boxes, double* sumsq);
/* sort permutation: sorted position -> caller index (sort/sorti.cxx:39-41) */
int fmmx_get_perm(fmmx_handle* h, int side, int* perm);
/* box keys at a level; particle ranges [first,last] (inclusive, as ndi/ndj
 * in fmm/boxdatai.cxx:22-29) at the leaf level, else NULL */
int fmmx_get_boxes(fmmx_handle* h, int side, int lev, int* keys, int* first, int* last);
/* interaction list: kind 0 near (leaf level, fmm/ijbox.cxx neib 0), kind 1
 * far at level lev (neib 1 at level 2, neib 2 above; needs
 * FMMX_KEEP_LISTS).  nij[nb_i(lev)], neij[nb_i(lev) * 189] row-padded as
 * neij in include/arrays.h:34. */
int fmmx_get_list(fmmx_handle* h, int lev, int kind, int* nij, int* neij);
/* Expansions of the last evaluate in the reference's layout (complex
 * interleaved, p(p+1)/2 per box, unscaled): multipoles after the upward pass
 * for all source boxes stacked leaf level first (the reference's bx rows),
 * and local expansions at level lev after its M2L. */
int fmmx_get_multipoles(fmmx_handle* h, double* bx_reim);
/* the multipoles of one level lev (its rows of fmmx_get_multipoles) */
int fmmx_get_multipoles_level(fmmx_handle* h, int lev, double* bx_reim);
int fmmx_get_locals(fmmx_handle* h, int lev, double* ax_reim);

/* Host only (no GPU needed): the dense real operator the engine applies,
 * (p*p) x (p*p) row-major, row = output slot, in the scaled basis of
 * fmmx_get_multipoles' slots: multipoles M_n / rb^n, locals L_n * rb^(n+1)
 * with real slots n(n+1)/2+m (Re) and p(p+1)/2 + n(n-1)/2 + m-1 (Im, m>0).
 * kind 0: M2M from child octant a0 (kernel/m2m.cxx:126-154); 1: L2L to
 * child octant a0 (kernel/l2l.cxx:203-237); 2: M2L for target-minus-source
 * offset (a0,a1,a2) at box size rb (kernel/m2l.cxx:43-81).
 * Yukawa (neq 10, kernel/Helmholtz_m2m.cxx, Helmholtz_m2l.cxx), where rb is
 * lambda times the box size (of the child for 3 and 4) and the bases are
 * M~_n = M_n / ((2n+1) i_n(lambda rb)), L~_n = L_n i_n(lambda rb) with
 * multipole basis k_n(lambda r) Y_n^m and local basis i_n(lambda r) Y_n^m:
 * kind 3: M2M from child octant a0; 4: L2L to child octant a0; 5: M2L for
 * offset (a0,a1,a2).  The convention does not enter kinds 3-5.
 * p = 2..FMMX_MAX_P (any order on the host; FMMX_CONV_REF up to 12: above it
 * its float-stored anm table turns subnormal or zero).
 * Helmholtz (FMMX_HELMHOLTZ), where rb is k times the box size (of the child
 * for 6 and 7): (2 p^2) x (2 p^2) operators on split expansions, the real
 * slots of the Re part then of the Im part (multipole basis
 * i h_n(kr) (kr)^(n+1) / (2n-1)!! (rb/r)^(n+1) Y_n^m, local basis
 * (2n+1)!! j_n(kr) / (kr)^n (r/rb)^n Y_n^m, both times 1/(4 pi rb)): kind 6:
 * M2M from child octant a0; 7: L2L to child octant a0; 8: M2L for offset
 * (a0,a1,a2).  A needs (2 p^2)^2 doubles.  No reference counterpart: the
 * reference has no oscillatory kernel. */
int fmmx_operator(int p, int convention, int kind, int a0, int a1, int a2, double rb, double* A);

/* Last error message of the calling thread. */
const char* fmmx_last_error(void);
int fmmx_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
