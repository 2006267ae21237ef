// engine.h -- device-resident state of one FMM handle and the launchers
// that operate on it.
#pragma once

#include <algorithm>
#include <array>
#include <memory>

#include "fmmx_internal.h"

namespace fmmx {

constexpr int MAXL = 8;  // deepest supported leaf level (keys of 3*8 bits)
constexpr int RB_INTS = 128;  // ints of the pinned readback buffer

#define FMMX_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(FMMX_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

// Grow-only device allocation owned by a handle, or a non-owning view of
// another DevBuf (alias); ensure() on a view drops the view and allocates.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool view = false;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p && !view) (void)hipFree(p);
  }
  void alias(const DevBuf& o) {
    if (p && !view) {
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
    p = o.p;
    bytes = o.bytes;
    view = true;
  }
  int ensure(size_t need) {
    if (view) {
      p = nullptr;
      bytes = 0;
      view = false;
    }
    if (need <= bytes && p) return FMMX_OK;
    if (p) {
      // growing: earlier asynchronous work may still read the old block
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    size_t b = need ? need : 16;
    if (hipMalloc(&p, b) != hipSuccess) return fail(FMMX_ENOMEM, "hipMalloc failed");
    bytes = b;
    return FMMX_OK;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// One particle set (targets i or sources j) in Morton order, with its boxes
// at every level 2..lmax.
struct Side {
  int n = 0;
  bool has_normals = false;
  DevBuf keys_in, idx_in, keys, perm, tmp;
  DevBuf x, y, z, nx, ny, nz;
  std::array<DevBuf, MAXL + 1> bkey, bfirst, parent, cfirst, dense;
  DevBuf blast, cnt;
  DevBuf clscnt;  // boxes per level and parity class, [8 l + class]
  std::array<int, MAXL + 1> nb{};  // host copy of box counts per level
};

struct M2LLev;  // one level of an M2L launch (kernels.hip)

struct Engine {
  fmmx_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;   // second stream: the near field beside the far field
  hipStream_t aux2 = nullptr;  // third stream: the upper levels' M2L/L2L beside the leaf M2L
  bool overlap_on = true;      // side streams in use (without per-phase timing)
  int p = 0, P2 = 0, lmax = 0;
  // multipoles are stored padded for the M2L B operand: real slot c lives at
  // (c / KS) * KSP + c % KS, KS = ceil(P2/4), KSP = KS rounded up to 4; VS = 4*KSP
  int KS = 0, KSP = 0, VS = 0;
  int mslot(int c) const { return (c / KS) * KSP + c % KS; }
  bool fp64 = false;
  float xmin = 0, ymin = 0, zmin = 0, rd = 0;
  bool have_points = false, keep_lists = false, user_box = false;
  Tables tab;

  Side si, sj;
  DevBuf rs_counts;  // digit counts per tile of the radix sort
  DevBuf box_tiles;       // boxes opened per tile of points and level (tree build)
  DevBuf xyz4;            // packed coordinates in caller order (tree build)
  // near list at the leaf level (targets): counts and rows of NNEAR
  DevBuf nij, neij, pair_ctr;
  DevBuf pair_part;  // near-pair counts per workgroup of the list build
  int near_nblk = 0;  // workgroups of the last near-list build (entries of pair_part)
  // far lists per level (only with FMMX_KEEP_LISTS): rows of NEBM
  std::array<DevBuf, MAXL + 1> fnij, fneij;
  // M2L target ordering by parity class and class counts per level
  std::array<DevBuf, MAXL + 1> cls_ord, cls_cnt;
  DevBuf cls_chunks;  // class counts per chunk of boxes
  int cls_cnt_h[MAXL + 1][8] = {};  // host copy of the class counts
  // operators (precision of the evaluation): M2M[8], L2L[8], M2L[levels][NOFF]
  DevBuf op_m2m, op_l2l, op_m2l, m2l_part, m2m_part;  // op_m2l in MFMA fragment order
  DevBuf m2l_part2;               // M2L partial sums of the levels above the leaf (third stream)
  DevBuf* part_buf = &m2l_part;   // partial-sum buffer of the M2L launches being issued
  bool m2l_per_level = false;
  // FP16 split M2L (m2l_split.hip): operator hi/lo fragments scaled by
  // 2^sa_e, split multipoles per level, largest |multipole| per level
  DevBuf op_m2l_h, op_m2l_w, mmax_buf;
  std::array<DevBuf, MAXL + 1> mexp_h;
  int sa_e = 0;
  bool split_ok = false;  // the operators are uploaded and p is supported
  bool wide_ok = false;   // the same for the 32x32x16 kernel (all 316 operators)
  bool m2l_split(int lev) const;  // the level's M2L runs the FP16 split kernel
  DevBuf offtab;    // int4 (dx,dy,dz,opidx) per (class, 189)
  DevBuf offtab_c;  // the same with w = canonical operator | reflection bits << 16
  DevBuf offtab_s;  // offtab_c with each class's offsets sorted by reflection
  bool canon = false;  // FP32 EXACT: M2L from the 56 reflection-canonical operators
  DevBuf corr_d;  // table ratio per packed (n,m) (1 in EXACT)
  DevBuf rb_dev;            // counters gathered for one readback
  int* rb_host = nullptr;   // pinned host copy of rb_dev (RB_INTS ints)
  bool tree_stats_pending = false;           // p2p_pairs and ms_tree not yet read
  int resolve_tree_stats();
  int ops_lmax = -1;
  float ops_rd = -1;
  // caller-order staging for host pointers
  DevBuf in[14], qin, out[4];
  // evaluation buffers
  DevBuf qs, vs, gxs, gys, gzs;
  std::array<DevBuf, MAXL + 1> mexp, lexp;
  // Yukawa (neq 10): screening lambda and the operator sets per level,
  // M2M by parent level 2..lmax-1, L2L by child level 3..lmax, M2L by level
  double lam = 0;
  bool yk = false;  // the evaluation in progress is Yukawa
  DevBuf yop_m2m, yop_l2l, yop_m2l;
  int yops_lmax = -1;
  float yops_rd = -1;
  double yops_lam = -1;
  // Helmholtz (FMMX_HELMHOLTZ): wavenumber, imaginary charges (key order) and
  // the operator sets per level on split expansions (two parts: the fields
  // of the real and of the imaginary charges, locals of Re and Im of the
  // potential): M2M (2 VS)^2 by parent level 2..lmax-1, L2L (2 P2)^2 by child
  // level 3..lmax, M2L (2 P2)^2 k-major by level
  double hk = 0;
  bool hh = false;  // the evaluation in progress is Helmholtz
  DevBuf qs2, hop_m2m, hop_l2l, hop_m2l;
  int hops_lmax = -1;
  float hops_rd = -1;
  double hops_k = -1;
  int parts() const { return hh ? 2 : 1; }
  // Groups: with group_level > 0 the points in different boxes of that level
  // are independent problems (no near or far interaction between them); the
  // far field then starts at the group level.
  int group_level = 0;
  int group_shift(int lev) const { return group_level > 0 ? lev - group_level : -1; }
  int first_m2l_level() const { return std::max(2, group_level); }
  int set_groups(int level);
  // one evaluation captured as a HIP graph (device pointers), replayed on demand
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipStream_t gstream = nullptr;  // capture and replay stream
  hipEvent_t gev[2] = {};         // orders replays after / before the caller's stream
  bool capturing = false;
  void drop_graph();
  int graph_capture(int neq, const float* qj, double* vi, double* gxi, double* gyi, double* gzi);
  int graph_launch();
  // stand-alone Morton sort (fmmx_sort): staging, keys, permutation
  DevBuf srt_in[3], srt_keys_in, srt_idx_in, srt_keys, srt_perm, srt_bad;
  DevBuf occ_bits;  // points per box of levels 8..0 and their statistics (automatic depth)
  // box statistics the last automatic depth was chosen from
  long occ_boxes[MAXL + 1] = {};
  double occ_sumsq[MAXL + 1] = {};
  DevBuf bad_flag;  // set by the key kernel for coordinates outside the root cube
  int last_neq = -1;
  // The transposed problem (targets: this handle's sources, sources: its
  // targets) in a second engine on the same root cube and depth, built at
  // the first transposed evaluation after set_points (adjoint.cpp)
  std::unique_ptr<Engine> tr;
  bool tr_valid = false;
  DevBuf adj_w, adj_wk, adj_q, adj_w2, adj_t[19];  // adjoint staging and temporaries
  int transposed(Engine*& t);
  int evaluate_transpose(int neq, const float* wi, double* uj, double* gxj, double* gyj, double* gzj, int flags);
  int gradients(int neq, const float* qj, const float* wi, double* dq, double* dxi, double* dxj, int flags);
  int energy(int neq, const float* qj, double* E, double* dEdq, double* dEdx, int flags);
  // timing
  bool per_phase = false;
  hipEvent_t ev[16] = {};
  fmmx_stats stats{};
  int64_t p2p_pairs = 0, m2l_pairs = 0;

  ~Engine();
  int init(const fmmx_config* c, int dev, void* stream);
  int set_points(int ni, const float* xi, const float* yi, const float* zi, const float* nxi, const float* nyi,
                 const float* nzi, int nj, const float* xj, const float* yj, const float* zj, const float* nxj,
                 const float* nyj, const float* nzj, int flags);
  int evaluate(int neq, const float* qj, double* vi, double* gxi, double* gyi, double* gzi, int flags);
  int phase(int ph, int neq, int lev, const float* qj, double* vi, double* gxi, double* gyi, double* gzi, int flags);
  int check_eval(int neq, const float* qj, const double* vi, const double* gxi, const double* gyi, const double* gzi);
  int prepare(int neq);
  int direct(int neq, int ni, const float* xi, const float* yi, const float* zi, const float* nxi,
             const float* nyi, const float* nzi, int nj, const float* xj, const float* yj, const float* zj,
             const float* nxj, const float* nyj, const float* nzj, const float* qj, double* vi, double* gxi,
             double* gyi, double* gzi, int flags);

  int set_domain(int lmax, float xmin, float ymin, float zmin, float rd);
  int set_yukawa(double lambda);
  int upload_yukawa_operators();
  int set_helmholtz(double k);
  int upload_helmholtz_operators();

  // tree.hip
  int sort_points(int n, const float* x, const float* y, const float* z, int* keys, int* perm, int flags);
  // boxes holding a point and sums of squared points per box, levels 0..8,
  // for n points in the current root cube
  int occupancy(int n, const float* x, const float* y, const float* z, long boxes[MAXL + 1],
                double sumsq[MAXL + 1]);
  int build_side(Side& s, int n, const float* x, const float* y, const float* z, const float* nx,
                 const float* ny, const float* nz, bool dense, const Side* same = nullptr);
  int build_lists();
  int build_far_lists();
  int build_m2l_order();
  // kernels.hip
  int run_p2p(int neq);
  int run_p2m(int neq);
  int run_m2m(int lev);
  // M2L launch variant per level: MFMA kernel N-tiles per wave, operator
  // staging (0 LDS, 1 register prefetch, 2 double buffer; 3 and 4: the
  // error-compensated FP16 kernel on 16x16x32 MFMA with one or two operator
  // buffers; 5: the same on 32x32x16 MFMA), waves per workgroup, offset
  // splits.  tune[] holds the autotuned choice (0 fields:
  // the defaults); FMMX_M2L_* environment variables override both.
  struct M2LVariant {
    int nt = 0, pa = -1, wpg = 0, nsplit = 0;
  };
  M2LVariant tune[MAXL + 1];
  M2LVariant m2l_variant(int lev) const;
  int m2l_splits(int lev) const { return m2l_variant(lev).nsplit; }
  int run_m2l(int lev, bool init, int phase = 0);
  int run_m2l_all(int l0);  // FP32 MFMA: every level in one launch
  int run_msplit(int l0, int l1, bool wide);  // split multipoles for the FP16 kernels
  bool last_m2l_all = false;  // the last evaluation took that form
  bool m2l_timed = false;     // the last evaluation recorded ev[8], ev[9] around its M2L launch
  bool m2l_all_levels() const {
    if (fp64 || hh) return false;
    const char* e = getenv("FMMX_M2L_FUSED");
    return !e || atoi(e) != 0;
  }
  M2LLev m2l_desc(int lev, int nt, int wpg, int nsplit, float* part, int init, int wg0);
  int run_l2l(int lev, bool add = false);
  // dst: outputs in the caller's order (v, gx, gy, gz; NULL entries skipped),
  // written by L2P itself instead of a separate scatter
  int run_l2p(int neq, double* const* dst = nullptr);
  int run_direct(int neq, int ni, const float* xi, const float* yi, const float* zi, const float* nxi,
                 const float* nyi, const float* nzi, int nj, const float* xj, const float* yj,
                 const float* zj, const float* nxj, const float* nyj, const float* nzj, const float* qj,
                 double* vi, double* gxi, double* gyi, double* gzi);
  int upload_operators();
  double c4() const { return cfg.convention == FMMX_CONV_REF ? 0.25 / (double)PIF : 0.25 / M_PI; }
  // (pi/2)/lambda of kernel/Helmholtz_G_p2p.cxx:22 (float pi in REF)
  double c_yukawa() const { return (cfg.convention == FMMX_CONV_REF ? 0.5 * (double)PIF : 0.5 * M_PI) / lam; }
  // translation operators of the evaluation in progress
  // (FP32 Laplace and Yukawa: MFMA fragment order, FR floats per octant)
  size_t frag_size() const { return (size_t)((P2 + 3) / 4) * ((P2 + 15) / 16) * 64; }
  template <class T>
  const T* m2m_ops(int parent_lev) const {
    if (hh) return hop_m2m.as<T>() + (size_t)(parent_lev - 2) * 8 * 4 * VS * VS;
    const size_t per = fp64 ? (size_t)VS * VS : frag_size();
    return yk ? yop_m2m.as<T>() + (size_t)(parent_lev - 2) * 8 * per : op_m2m.as<T>();
  }
  template <class T>
  const T* l2l_ops(int child_lev) const {
    if (hh) return hop_l2l.as<T>() + (size_t)(child_lev - 3) * 8 * 4 * P2 * P2;
    const size_t per = fp64 ? (size_t)P2 * P2 : frag_size();
    return yk ? yop_l2l.as<T>() + (size_t)(child_lev - 3) * 8 * per : op_l2l.as<T>();
  }
  double rb(int lev) const { return (double)rd / (double)(1 << lev); }
};

// fmmx_sort_noalloc (tree.hip)
int sort_noalloc(int device, void* stream, int n, const float* x, const float* y, const float* z, int lmax,
                 float xmin, float ymin, float zmin, float rd, int* keys, int* perm, int* bad);

}  // namespace fmmx
