/*
 * fmm_oracle.h -- CPU FP64 restatement of the reference Laplace FMM.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (old_fmm_bem_amd/) never links, imports or calls it.
 *
 * It follows the reference serial CPU path function by function
 * (fmm/fmm.cxx, the fmm/box and sort/ files, kernel/ operators);
 * every function in fmm_oracle.c cites the reference file:line it restates.
 *
 * Two conventions:
 *   ORC_CONV_REF   -- the reference's own arithmetic: eps=1e-6 (float) added
 *                     to r^2 in P2P/direct (self pair included), eps added to
 *                     rh in every angle/radius computation, the custom
 *                     azimuth branches with float pi, fac/anm/sr tables
 *                     stored in float, outputs accumulated in float.
 *   ORC_CONV_EXACT -- the same algorithm with the quirks removed: no
 *                     softening (r=0 pairs skipped, i.e. self excluded),
 *                     atan2 azimuths, double tables, double outputs.
 */
#ifndef FMM_ORACLE_H
#define FMM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_CONV_REF = 0, ORC_CONV_EXACT = 1 };
enum { ORC_NEQ_G = 0, ORC_NEQ_GNI = 1, ORC_NEQ_GNJ = 2 };

typedef struct orc_opts {
  int mp;        /* number of expansion degrees (n = 0..mp-1) */
  int lmax;      /* leaf level, >= 2 */
  int neq;       /* 0: G, 1: Gni (target normals), 2: Gnj (source normals) */
  int conv;      /* ORC_CONV_REF or ORC_CONV_EXACT */
  int nthreads;  /* OpenMP threads (<=0: library default) */
  float xmin, ymin, zmin, rd; /* root cube, float as in include/constants.h:56-59 */
  int fp64_out;  /* 1: accumulate outputs in double even in ORC_CONV_REF */
} orc_opts;

typedef struct orc_fmm orc_fmm;

/* Build the tree: Morton keys, stable sort, box lists, all interaction
 * lists.  Inputs are in caller order and copied.  nxi.. are target normals
 * (neq 1), nxj.. source normals (neq 2); either may be NULL when unused. */
orc_fmm* orc_new(const orc_opts* o,
                 int ni, const float* xi, const float* yi, const float* zi,
                 const float* nxi, const float* nyi, const float* nzi,
                 int nj, const float* xj, const float* yj, const float* zj,
                 const float* nxj, const float* nyj, const float* nzj);
void orc_free(orc_fmm* f);

/* Run the seven FMM steps for the charges qj (caller order).  Outputs
 * (caller order): potential vi and, for neq 0, gradient gxi,gyi,gzi (may be
 * NULL).  Returns 0 on success. */
int orc_run(orc_fmm* f, const float* qj, double* vi, double* gxi, double* gyi, double* gzi);

/* Per-phase wall times of the last orc_run (seconds): 0 P2P 1 P2M 2 M2M
 * 3 M2L 4 L2L 5 L2P 6 precalc 7 tree. */
void orc_times(const orc_fmm* f, double* t8);

/* ---- introspection (for fixture comparison) ---- */
/* what: 0 nbne, 1 nbnet(j), 2 nmp, 3 lbi at level lev, 4 lbj at level lev,
 *       5 nbnet(i) */
int orc_info(const orc_fmm* f, int what, int lev);
/* sort permutations (sorted position -> caller index) */
void orc_perm(const orc_fmm* f, int* nbi, int* nbj);
/* box keys and particle ranges at a level (ranges only at lmax) */
void orc_boxes_i(const orc_fmm* f, int lev, int* keys, int* first, int* last);
void orc_boxes_j(const orc_fmm* f, int lev, int* keys, int* first, int* last);
/* interaction list at level lev: kind 0 = near (P2P, only lev==lmax),
 * kind 1 = far (M2L).  nij[lbi], neij[lbi*189] (j ordinals at that level). */
void orc_list(const orc_fmm* f, int lev, int kind, int* nij, int* neij);
/* expansions after the upward pass (bx, all levels stacked as in the
 * reference: leaf rows first), and local expansions at the leaf level after
 * the last M2L (ax).  Complex values as interleaved (re,im) doubles. */
void orc_bx(const orc_fmm* f, double* bx_reim);
void orc_ax(const orc_fmm* f, double* ax_reim);
/* local expansions at level lev after that level's M2L */
void orc_ax_level(const orc_fmm* f, int lev, double* ax_reim);

/* ---- precomputed tables (kernel/precalc.cxx) ---- */
/* fac, anm: 4*mp*mp; sr: mp^4; ynm: 4*mp*mp complex; dnm: 1024*mp*mp*mp complex */
void orc_precalc(int mp, int conv, double* fac, double* anm, double* sr,
                 double* ynm_reim, double* dnm_reim);

/* ---- direct sums (kernel/G_dir.cxx, Gni_dir.cxx, Gnj_dir.cxx) ---- */
/* Overwrite vi (and gx.. for neq 0) for targets [0,ni) against all sources. */
void orc_direct(int neq, int conv, int nthreads,
                int ni, const float* xi, const float* yi, const float* zi,
                const float* nxi, const float* nyi, const float* nzi,
                int nj, const float* xj, const float* yj, const float* zj,
                const float* nxj, const float* nyj, const float* nzj,
                const float* qj, double* vi, double* gxi, double* gyi, double* gzi);

/* Yukawa direct sum (kernel/Helmholtz_G_dir.cxx:6-20): vi = sum_j q_j
 * (pi/2) exp(-lambda r)/(lambda r).  REF: r = sqrt(r^2 + eps), self pair
 * included, float pi, float outputs, potential only (as the reference).
 * EXACT: unsoftened, r = 0 pairs excluded, double pi, plus the gradient
 * grad_i = -sum_j q_j (pi/2)/lambda (1 + lambda r) exp(-lambda r) d/r^3
 * (d = x_i - x_j; not in the reference; pinned by closed form). */
void orc_direct_yukawa(int conv, int nthreads, double lambda, int ni, const float* xi, const float* yi,
                       const float* zi, int nj, const float* xj, const float* yj, const float* zj,
                       const float* qj, double* vi, double* gxi, double* gyi, double* gzi);

/* Oscillatory Helmholtz direct sum, no reference counterpart (the
 * reference's "Helmholtz" files are the screened kernel above):
 * v_i = sum_j (qr_j + i qi_j) exp(i k r) / (4 pi r) in FP64 from the float
 * inputs.  REF: r = sqrt(r^2 + eps) with the self pair included; EXACT: r = 0
 * pairs excluded.  Pinned by closed forms in the tests (a plane-wave
 * superposition identity and the k -> 0 Laplace limit). */
void orc_direct_helmholtz(int conv, int nthreads, double k, int ni, const float* xi, const float* yi,
                          const float* zi, int nj, const float* xj, const float* yj, const float* zj,
                          const float* qr, const float* qi, double* vr, double* vim);

/* ---- single-box expansion helpers for closed-form checks ---- */
/* Multipole of point charges about centre c (kernel/G_p2m.cxx), and the
 * local-expansion evaluation at points (kernel/G_l2p.cxx) of given
 * coefficients; complex interleaved, mp*(mp+1)/2 coefficients. */
void orc_p2m_box(int mp, int conv, int n, const float* x, const float* y, const float* z,
                 const float* q, double cx, double cy, double cz, double* m_reim);
void orc_l2p_box(int mp, int conv, int n, const float* x, const float* y, const float* z,
                 double cx, double cy, double cz, const double* l_reim,
                 double* v, double* gx, double* gy, double* gz);
/* One reference-style M2L: rotate, axial translate by rh, rotate back
 * (kernel/m2l.cxx:43-81) for the lattice offset (dx,dy,dz) in boxes of size rb. */
void orc_m2l_pair(int mp, int conv, int dx, int dy, int dz, double rb,
                  const double* m_reim, double* l_reim);

#ifdef __cplusplus
}
#endif
#endif
