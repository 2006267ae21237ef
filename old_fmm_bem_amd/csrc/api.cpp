// api.cpp -- handle lifetime, set-up, the matvec sequence of fmm/fmm.cxx:29-152
// on device-resident data, and the extern "C" entry points of fmmx.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>

#include "engine.h"
#include "helmholtz.h"
#include "m2l.h"
#include "yukawa.h"
#include "tree_util.h"

namespace fmmx {

namespace {

constexpr int TPB = 256;
inline int nblk(long n) { return (int)((n + TPB - 1) / TPB); }

__global__ void k_gather_q(int n, const uint32_t* __restrict__ perm, const float* __restrict__ q,
                           float* __restrict__ qs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) qs[i] = q[perm[i]];
}

// device counters the host needs after the tree build, gathered into one
// small buffer so that a single copy brings them back:
// [0, 9) target boxes per level, [9, 18) source boxes per level, [18] the
// out-of-cube flag, [19 + 8 l + c) target boxes of level l in parity class c
struct CountSrc {
  const int* cnt[2];  // box counts of both sides
  const int* cls;     // target boxes per level and class
  const int* bad;     // coordinates outside the root cube
};
constexpr int RB_CLS = 2 * (MAXL + 1) + 1;
static_assert(RB_CLS + 8 * (MAXL + 1) <= RB_INTS, "readback buffer");
__global__ void k_pack_counts(CountSrc src, int lmax, int* __restrict__ out) {
  const int t = threadIdx.x;
  if (t < 2 * (MAXL + 1)) {
    const int* c = src.cnt[t / (MAXL + 1)];
    out[t] = c ? c[t % (MAXL + 1)] : 0;
  } else if (t == 2 * (MAXL + 1)) {
    out[t] = src.bad ? *src.bad : 0;
  }
  for (int e = t; e < 8 * (MAXL + 1); e += blockDim.x) {
    const int l = e / 8;
    out[RB_CLS + e] = (l >= 2 && l <= lmax && src.cls) ? src.cls[e] : 0;
  }
}

// up to four key-ordered outputs back to caller order in one pass (NULL: skip)
__global__ void k_scatter4(int n, const uint32_t* __restrict__ perm, const double* __restrict__ s0,
                           const double* __restrict__ s1, const double* __restrict__ s2, const double* __restrict__ s3,
                           double* __restrict__ d0, double* __restrict__ d1, double* __restrict__ d2,
                           double* __restrict__ d3) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  if (d0) d0[j] = s0[i];
  if (d1) d1[j] = s1[i];
  if (d2) d2[j] = s2[i];
  if (d3) d3[j] = s3[i];
}

// complex charges given as (re, im) pairs -> key-ordered real and imaginary parts
__global__ void k_gather_qc(int n, const uint32_t* __restrict__ perm, const float* __restrict__ q,
                            float* __restrict__ qr, float* __restrict__ qi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint32_t j = perm ? perm[i] : (uint32_t)i;
    qr[i] = q[2 * (size_t)j];
    qi[i] = q[2 * (size_t)j + 1];
  }
}

// key-ordered real and imaginary parts -> (re, im) pairs in caller order
__global__ void k_scatter_c(int n, const uint32_t* __restrict__ perm, const double* __restrict__ re,
                            const double* __restrict__ im, double* __restrict__ d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint32_t j = perm ? perm[i] : (uint32_t)i;
    d[2 * (size_t)j] = re[i];
    d[2 * (size_t)j + 1] = im[i];
  }
}

__global__ void k_zero_i32(int n, int* __restrict__ a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_zero_f64(long n, double* __restrict__ a) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// order-preserving float <-> int map for atomic min/max
__device__ inline int f2o(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__global__ void k_init_mm(int* mm) {
  if (threadIdx.x < 3) mm[threadIdx.x] = 0x7fffffff;
  else if (threadIdx.x < 6) mm[threadIdx.x] = (int)0x80000000;
}
__global__ void k_minmax(int n, const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                         int* mm) {
  __shared__ int s[6];
  if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? 0x7fffffff : (int)0x80000000;
  __syncthreads();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    int a = f2o(x[i]), b = f2o(y[i]), c = f2o(z[i]);
    atomicMin(&s[0], a); atomicMin(&s[1], b); atomicMin(&s[2], c);
    atomicMax(&s[3], a); atomicMax(&s[4], b); atomicMax(&s[5], c);
  }
  __syncthreads();
  if (threadIdx.x < 3) atomicMin(&mm[threadIdx.x], s[threadIdx.x]);
  else if (threadIdx.x < 6) atomicMax(&mm[threadIdx.x], s[threadIdx.x]);
}
inline float o2f(int o) {
  int i = o >= 0 ? o : o ^ 0x7fffffff;
  float f;
  std::memcpy(&f, &i, 4);
  return f;
}

// fmm/nlevel.cxx:8-26 ("gpu p=5" switch points)
int nlevel(int nmax) {
  const float sw[6] = {1e5f, 5e5f, 3e6f, 2e7f, 1.5e8f, 1.3e9f};
  int l = 1;
  for (int k = 0; k < 6; k++)
    if ((float)nmax < sw[k]) return l + k + 1;
  return l + 7;
}

// Leaf level of the EXACT convention when the caller leaves lmax to the
// engine: the level 2..8 of least modelled cost
//   near field  3^d S(l)            (S: sum over boxes of squared counts)
//   far field   r (6^d - 3^d) T(l)  (T: boxes of levels 2..l)
// where d = log2(B(l) / B(l-1)) in [0, 3] is the dimension the point set
// shows at that scale (B: boxes holding a point; 3 for a volume, 2 for a
// surface), so 3^d and 6^d - 3^d count the near and interaction-list boxes
// that hold points.  r, the cost of one far-field box against one near pair,
// was fitted per order and precision to evaluation times measured on MI355X
// (uniform cube, sphere surface and Plummer inputs, N 1e4..3e6, every leaf
// level with 3..2000 points per box): the choice is within 5% of the fastest
// level on average over those 158 cases.  The model holds where boxes hold
// several points (a volume's boxes hold one point each once B stops growing,
// d drops to 0 and the far-field term with it), so the levels considered
// are those with 3..2000 points per occupied box on average (level 2 when
// none has 3, the deepest level when none has at most 2000).
int auto_depth(const long boxes[MAXL + 1], const double sumsq[MAXL + 1], int p, bool fp64) {
  // r at p = 4, 6, 8, 10, 12 (non-decreasing in p), geometric in between
  static const double r_mixed[5] = {112, 112, 316, 501, 501};
  static const double r_fp64[5] = {316, 501, 501, 1000, 2820};
  const double* rt = fp64 ? r_fp64 : r_mixed;
  const double u = std::min(std::max((p - 4) / 2.0, 0.0), 4.0);
  const int k = std::min((int)u, 3);
  double r = rt[k] * std::pow(rt[k + 1] / rt[k], u - k);
  // order 18 (FP64): not fitted; the far field's cost per box grows as p^4
  // (M2L) while a near pair's does not
  if (p > 12) r *= std::pow(p / 12.0, 4);
  const double n = (double)boxes[0] > 0 ? std::sqrt(sumsq[0]) : 0.0;  // one box holds all points
  int best = -1, lo = 2, hi = MAXL;
  while (lo < MAXL && boxes[lo] > 0 && n / (double)boxes[lo] > 2000) lo++;
  while (hi > lo && (boxes[hi] <= 0 || n / (double)boxes[hi] < 3)) hi--;
  double cbest = 0, tot = 0;
  for (int l = 2; l <= hi; l++) {
    tot += (double)boxes[l];
    if (l < lo) continue;
    const double d = boxes[l - 1] > 0 ? std::min(std::max(std::log2((double)boxes[l] / (double)boxes[l - 1]), 0.0), 3.0)
                                       : 3.0;
    const double c = std::pow(3.0, d) * sumsq[l] + r * (std::pow(6.0, d) - std::pow(3.0, d)) * tot;
    if (best < 0 || c < cbest) best = l, cbest = c;
  }
  return best;
}

// process-wide cache of host operators keyed by (p, conv, lmax, rd)
struct OpSet {
  std::vector<double> m2m, l2l;  // transposed [op][k][row]
  std::vector<double> m2m_frag, l2l_frag;  // MFMA fragment order (FP32 engine)
  std::vector<double> m2l_frag;  // MFMA fragment order (FP32 engine)
  std::vector<double> m2l_canon; // fragment order, the 56 offsets with dx, dy, dz >= 0 (EXACT)
  std::vector<double> m2l_kmaj;  // transposed [op][k][row] (FP64 engine)
  // FP16 split M2L (EXACT, p 6..12): hi/lo fragments of the 56 canonical
  // operators scaled by 2^split_e
  std::vector<uint4> m2l_split;
  int split_e = 0;
  // the same split for the 32x32x16 kernel: all 316 operators as they act
  std::vector<uint4> m2l_wide;
  int nlev = 1;
};
std::mutex g_cache_mu;
std::map<std::tuple<int, int, int, float>, std::shared_ptr<OpSet>> g_cache;
std::map<std::pair<int, int>, std::shared_ptr<Tables>> g_tables;

}  // namespace

std::shared_ptr<Tables> get_tables(int p, int conv) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto& t = g_tables[{p, conv}];
  if (!t) {
    t = std::make_shared<Tables>();
    build_tables(*t, p, conv);
  }
  return t;
}

namespace {

// MFMA 16x16x4 fragment order of a row-major P2 x P2 operator: [ks][mt][lane]
// holds A[16 mt + lane%16][(lane/16) * KS + ks] (zero outside)
void fragment_into(const double* A, int P2, double* dst) {
  const int MT = (P2 + 15) / 16, KS = (P2 + 3) / 4;
  for (int ks = 0; ks < KS; ks++)
    for (int mt = 0; mt < MT; mt++)
      for (int l = 0; l < 64; l++) {
        const int i = 16 * mt + (l & 15), k = (l >> 4) * KS + ks;
        dst[((size_t)ks * MT + mt) * 64 + l] = (i < P2 && k < P2) ? A[(size_t)i * P2 + k] : 0.0;
      }
}

void transpose_into(const double* A, int P2, double* dst) {
  for (int r = 0; r < P2; r++)
    for (int k = 0; k < P2; k++) dst[(size_t)k * P2 + r] = A[(size_t)r * P2 + k];
}

std::shared_ptr<OpSet> get_ops(const Tables& T, int lmax, float rd) {
  const int P2 = T.p * T.p;
  const bool per_level = T.conv == FMMX_CONV_REF;
  auto key = std::make_tuple(T.p, T.conv, per_level ? lmax : 0, per_level ? rd : 0.f);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) return it->second;
  }
  auto S = std::make_shared<OpSet>();
  const size_t sq = (size_t)P2 * P2;
  S->l2l.resize(8 * sq);
  std::vector<double> A(sq);
  const int KS = (P2 + 3) / 4, KSP = (KS + 3) / 4 * 4, VS = 4 * KSP;
  const size_t vq = (size_t)VS * VS;
  S->m2m.assign(8 * vq, 0.0);
  // MFMA fragment order for the FP32 engine (orders up to P_ALL)
  const bool frag = T.p <= P_ALL;
  const size_t fr = frag ? (size_t)((P2 + 3) / 4) * ((P2 + 15) / 16) * 64 : 0;
  S->m2m_frag.resize(8 * fr);
  S->l2l_frag.resize(8 * fr);
  for (int o = 0; o < 8; o++) {
    op_m2m(T, o, A.data());
    if (frag) fragment_into(A.data(), P2, &S->m2m_frag[o * fr]);
    // padded multipole layout, transposed [k][row]
    for (int r = 0; r < P2; r++)
      for (int k = 0; k < P2; k++)
        S->m2m[o * vq + (size_t)((k / KS) * KSP + k % KS) * VS + (r / KS) * KSP + r % KS] = A[(size_t)r * P2 + k];
    op_l2l(T, o, A.data());
    if (frag) fragment_into(A.data(), P2, &S->l2l_frag[o * fr]);
    transpose_into(A.data(), P2, &S->l2l[o * sq]);
  }
  S->nlev = per_level ? lmax - 1 : 1;
  S->m2l_frag.resize((size_t)S->nlev * NOFF * fr);
  S->m2l_kmaj.resize((size_t)S->nlev * NOFF * sq);
  const int njobs = S->nlev * NOFF;
  // parallel over (level, offset)
  std::vector<std::thread> th;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      std::vector<double> B(sq);
      for (int j = (int)t; j < njobs; j += (int)nt) {
        int li = j / NOFF, oi = j % NOFF, d[3];
        offset_vector(oi, d);
        double rbv = per_level ? (double)rd / (double)(1 << (li + 2)) : 1.0;
        op_m2l(T, d[0], d[1], d[2], rbv, B.data());
        if (frag) fragment_into(B.data(), P2, &S->m2l_frag[((size_t)li * NOFF + oi) * fr]);
        transpose_into(B.data(), P2, &S->m2l_kmaj[((size_t)li * NOFF + oi) * sq]);
      }
    });
  for (auto& x : th) x.join();
  if (!per_level && frag) {
    // EXACT: every offset is a reflection of one with non-negative components
    S->m2l_canon.resize((size_t)NCANON * fr);
    for (int ci = 0; ci < NCANON; ci++) {
      int d[3];
      canon_vector(ci, d);
      const int oi = offset_index(d[0], d[1], d[2]);
      std::copy(&S->m2l_frag[(size_t)oi * fr], &S->m2l_frag[(size_t)oi * fr] + fr, &S->m2l_canon[(size_t)ci * fr]);
    }
    if (m2l_split_supported(T.p)) {
      std::vector<double> ck((size_t)NCANON * sq);
      for (int ci = 0; ci < NCANON; ci++) {
        int d[3];
        canon_vector(ci, d);
        const int oi = offset_index(d[0], d[1], d[2]);
        std::copy(&S->m2l_kmaj[(size_t)oi * sq], &S->m2l_kmaj[(size_t)oi * sq] + sq, &ck[(size_t)ci * sq]);
      }
      S->split_e = split_scale_exp(ck.data(), ck.size());
      split_fragments(ck.data(), NCANON, T.p, S->split_e, S->m2l_split);
      if (split_wide_fits(T.p)) split_fragments_wide(S->m2l_kmaj.data(), NOFF, T.p, S->split_e, S->m2l_wide);
    }
  }
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_cache[key] = S;
  return S;
}

// Yukawa operator sets keyed by (p, lambda, rd, lmax): M2M per parent level
// 2..lmax-1 and L2L per child level 3..lmax (8 octants each, layouts as the
// Laplace sets), M2L per level 2..lmax (316 offsets)
std::map<std::tuple<int, double, float, int>, std::shared_ptr<OpSet>> g_yuk_cache;

std::shared_ptr<OpSet> get_yuk_ops(int p, double lam, float rd, int lmax) {
  const int P2 = p * p;
  auto key = std::make_tuple(p, lam, rd, lmax);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_yuk_cache.find(key);
    if (it != g_yuk_cache.end()) return it->second;
  }
  auto S = std::make_shared<OpSet>();
  const size_t sq = (size_t)P2 * P2;
  const int KS = (P2 + 3) / 4, KSP = (KS + 3) / 4 * 4, VS = 4 * KSP;
  const size_t vq = (size_t)VS * VS;
  const int nt_lev = std::max(1, lmax - 2);  // translation levels
  S->m2m.assign((size_t)nt_lev * 8 * vq, 0.0);
  S->l2l.assign((size_t)nt_lev * 8 * sq, 0.0);
  S->nlev = lmax - 1;
  const bool frag = p <= P_ALL;  // MFMA fragment order: the FP32 engine's orders
  const size_t fr = frag ? (size_t)((P2 + 3) / 4) * ((P2 + 15) / 16) * 64 : 0;
  S->m2m_frag.assign((size_t)nt_lev * 8 * fr, 0.0);
  S->l2l_frag.assign((size_t)nt_lev * 8 * fr, 0.0);
  S->m2l_frag.resize((size_t)S->nlev * NOFF * fr);
  S->m2l_kmaj.resize((size_t)S->nlev * NOFF * sq);
  const int ntr = lmax > 2 ? (lmax - 2) * 8 : 0, njobs = ntr + S->nlev * NOFF;
  std::vector<std::thread> th;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      std::vector<double> A(sq), B(sq);
      for (int j = (int)t; j < njobs; j += (int)nt) {
        if (j < ntr) {
          // child level lc = 3 + j/8 (parent lc-1), octant j%8
          const int li = j / 8, o = j % 8, lc = 3 + li;
          yuk_op_m2m_l2l(p, lam * (double)rd / (double)(1 << lc), o, A.data(), B.data());
          double* M = &S->m2m[((size_t)li * 8 + o) * vq];
          for (int r = 0; r < P2; r++)
            for (int k = 0; k < P2; k++)
              M[(size_t)((k / KS) * KSP + k % KS) * VS + (r / KS) * KSP + r % KS] = A[(size_t)r * P2 + k];
          transpose_into(B.data(), P2, &S->l2l[((size_t)li * 8 + o) * sq]);
          if (frag) {
            fragment_into(A.data(), P2, &S->m2m_frag[((size_t)li * 8 + o) * fr]);
            fragment_into(B.data(), P2, &S->l2l_frag[((size_t)li * 8 + o) * fr]);
          }
        } else {
          const int jj = j - ntr, li = jj / NOFF, oi = jj % NOFF;
          int d[3];
          offset_vector(oi, d);
          yuk_op_m2l(p, lam * (double)rd / (double)(1 << (li + 2)), d[0], d[1], d[2], A.data());
          if (frag) fragment_into(A.data(), P2, &S->m2l_frag[((size_t)li * NOFF + oi) * fr]);
          transpose_into(A.data(), P2, &S->m2l_kmaj[((size_t)li * NOFF + oi) * sq]);
        }
      }
    });
  for (auto& x : th) x.join();
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_yuk_cache[key] = S;
  return S;
}

// Helmholtz operator sets keyed by (p, k, rd, lmax), on split expansions:
// M2M per parent level 2..lmax-1 (padded multipole layout of both parts,
// (2 VS)^2, k-major), L2L per child level 3..lmax ((2 P2)^2, k-major), M2L per
// level 2..lmax (316 offsets, (2 P2)^2, k-major)
std::map<std::tuple<int, double, float, int>, std::shared_ptr<OpSet>> g_hh_cache;

std::shared_ptr<OpSet> get_hh_ops(int p, double k, float rd, int lmax) {
  const int P2 = p * p, D = 2 * P2;
  auto key = std::make_tuple(p, k, rd, lmax);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_hh_cache.find(key);
    if (it != g_hh_cache.end()) return it->second;
  }
  auto S = std::make_shared<OpSet>();
  const size_t dq = (size_t)D * D;
  const int KS = (P2 + 3) / 4, KSP = (KS + 3) / 4 * 4, VS = 4 * KSP, VH = 2 * VS;
  const size_t vq = (size_t)VH * VH;
  auto pos = [&](int c) {  // split slot -> padded position of its part
    const int t = c / P2, r = c % P2;
    return t * VS + (r / KS) * KSP + r % KS;
  };
  const int nt_lev = std::max(1, lmax - 2);
  S->m2m.assign((size_t)nt_lev * 8 * vq, 0.0);
  S->l2l.assign((size_t)nt_lev * 8 * dq, 0.0);
  S->nlev = lmax - 1;
  S->m2l_kmaj.resize((size_t)S->nlev * NOFF * dq);
  // FP32 engine: MFMA fragment order of the D x D operators
  const size_t frh = (size_t)(D / 4) * ((D + 15) / 16) * 64;
  S->m2l_frag.resize((size_t)S->nlev * NOFF * frh);
  const int ntr = lmax > 2 ? (lmax - 2) * 8 : 0, njobs = ntr + S->nlev * NOFF;
  std::vector<std::thread> th;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      std::vector<double> A(dq), B(dq);
      for (int j = (int)t; j < njobs; j += (int)nt) {
        if (j < ntr) {
          const int li = j / 8, o = j % 8, lc = 3 + li;
          hh_op_m2m_l2l(p, k * (double)rd / (double)(1 << lc), o, A.data(), B.data());
          double* M = &S->m2m[((size_t)li * 8 + o) * vq];
          for (int r = 0; r < D; r++)
            for (int c = 0; c < D; c++) M[(size_t)pos(c) * VH + pos(r)] = A[(size_t)r * D + c];
          transpose_into(B.data(), D, &S->l2l[((size_t)li * 8 + o) * dq]);
        } else {
          const int jj = j - ntr, li = jj / NOFF, oi = jj % NOFF;
          int d[3];
          offset_vector(oi, d);
          hh_op_m2l(p, k * (double)rd / (double)(1 << (li + 2)), d[0], d[1], d[2], A.data());
          transpose_into(A.data(), D, &S->m2l_kmaj[((size_t)li * NOFF + oi) * dq]);
          fragment_into(A.data(), D, &S->m2l_frag[((size_t)li * NOFF + oi) * frh]);
        }
      }
    });
  for (auto& x : th) x.join();
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_hh_cache[key] = S;
  return S;
}

template <typename T>
int upload(DevBuf& dst, const std::vector<double>& src, hipStream_t s) {
  int rc = dst.ensure(src.size() * sizeof(T));
  if (rc) return rc;
  std::vector<T> tmp(src.begin(), src.end());
  FMMX_HIP(hipMemcpyAsync(dst.p, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice, s));
  FMMX_HIP(hipStreamSynchronize(s));
  return FMMX_OK;
}

}  // namespace

void Engine::drop_graph() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  gexec = nullptr;
  graph = nullptr;
}

int Engine::graph_capture(int neq, const float* qj, double* vi, double* gxi, double* gyi, double* gzi) {
  if (per_phase) return fail(FMMX_ESTATE, "graph capture needs per-phase timing off (fmmx_set_timing 0)");
  drop_graph();
  // eager first: sizes every buffer and uploads the operators, so that the
  // captured evaluation allocates and copies nothing
  int rc = evaluate(neq, qj, vi, gxi, gyi, gzi, FMMX_DEVICE_PTRS);
  if (rc) return rc;
  FMMX_HIP(hipStreamSynchronize(stream));
  // capture on a stream of the handle's own (the caller's stream may be the
  // null stream, which cannot be captured); replays are ordered with the
  // caller's stream by events
  if (!gstream) {
    FMMX_HIP(hipStreamCreateWithFlags(&gstream, hipStreamNonBlocking));
    FMMX_HIP(hipEventCreateWithFlags(&gev[0], hipEventDisableTiming));
    FMMX_HIP(hipEventCreateWithFlags(&gev[1], hipEventDisableTiming));
  }
  hipStream_t caller = stream;
  stream = gstream;
  hipError_t eb = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
  if (eb != hipSuccess) {
    stream = caller;
    return fail(FMMX_EDEVICE, std::string("hipStreamBeginCapture: ") + hipGetErrorString(eb));
  }
  capturing = true;
  rc = evaluate(neq, qj, vi, gxi, gyi, gzi, FMMX_DEVICE_PTRS);
  capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(stream, &g);
  stream = caller;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess || !g) return fail(FMMX_EDEVICE, std::string("stream capture failed: ") + hipGetErrorString(ec));
  graph = g;
  if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
    drop_graph();
    return fail(FMMX_EDEVICE, "hipGraphInstantiate failed");
  }
  return FMMX_OK;
}

int Engine::graph_launch() {
  if (!gexec) return fail(FMMX_ESTATE, "no captured evaluation (fmmx_graph_capture)");
  FMMX_HIP(hipSetDevice(device));
  FMMX_HIP(hipEventRecord(gev[0], stream));
  FMMX_HIP(hipStreamWaitEvent(gstream, gev[0], 0));
  FMMX_HIP(hipGraphLaunch(gexec, gstream));
  FMMX_HIP(hipEventRecord(gev[1], gstream));
  FMMX_HIP(hipStreamWaitEvent(stream, gev[1], 0));
  return FMMX_OK;
}

Engine::~Engine() {
  drop_graph();
  if (rb_host) (void)hipHostFree(rb_host);
  if (gstream) (void)hipStreamDestroy(gstream);
  for (auto& e : gev)
    if (e) (void)hipEventDestroy(e);
  if (aux) (void)hipStreamDestroy(aux);
  if (aux2) (void)hipStreamDestroy(aux2);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
}

int Engine::init(const fmmx_config* c, int dev, void* strm) {
  if (!c) return fail(FMMX_EINVAL, "config is NULL");
  if (c->p < 2 || (c->p > P_ALL && c->p != P_HI)) return fail(FMMX_EINVAL, "p must be 2..12 or 18");
  if (c->lmax != 0 && (c->lmax < 2 || c->lmax > MAXL)) return fail(FMMX_EINVAL, "lmax out of range [2, 8]");
  if (c->precision != FMMX_MIXED && c->precision != FMMX_FP64) return fail(FMMX_EINVAL, "bad precision");
  if (c->convention != FMMX_CONV_REF && c->convention != FMMX_CONV_EXACT) return fail(FMMX_EINVAL, "bad convention");
  // order 18: the vector-unit FP64 kernels only; the REF convention's
  // float-stored anm table turns subnormal or zero at the degrees its
  // operators need (up to 2p - 1 = 35)
  if (c->p > P_ALL && (c->precision != FMMX_FP64 || c->convention != FMMX_CONV_EXACT))
    return fail(FMMX_EUNSUPPORTED, "p = 18 needs FMMX_FP64 and FMMX_CONV_EXACT");
  if (c->rd > 0 && !(std::isfinite(c->xmin) && std::isfinite(c->ymin) && std::isfinite(c->zmin) && std::isfinite(c->rd)))
    return fail(FMMX_EINVAL, "root cube not finite");
  cfg = *c;
  device = dev;
  FMMX_HIP(hipSetDevice(dev));
  stream = (hipStream_t)strm;
  p = c->p;
  P2 = p * p;
  KS = (P2 + 3) / 4;
  KSP = (KS + 3) / 4 * 4;
  VS = 4 * KSP;
  fp64 = c->precision == FMMX_FP64;
  user_box = c->rd > 0;
  for (auto& e : ev) FMMX_HIP(hipEventCreate(&e));
  FMMX_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
  FMMX_HIP(hipStreamCreateWithFlags(&aux2, hipStreamNonBlocking));
  // FMMX_OVERLAP=0 runs every phase on the caller's stream (measurement)
  if (const char* e = getenv("FMMX_OVERLAP")) overlap_on = atoi(e) != 0;
  tab = *get_tables(p, c->convention);
  // device tables: table-ratio corrections and the M2L offset table
  std::vector<double> corr(tab.corr);
  int rc = corr_d.ensure(corr.size() * 8);
  if (rc) return rc;
  FMMX_HIP(hipMemcpy(corr_d.p, corr.data(), corr.size() * 8, hipMemcpyHostToDevice));
  std::vector<int> ot(8 * NEBM * 4);
  for (int cls = 0; cls < 8; cls++) {
    int px = cls & 1, py = (cls >> 1) & 1, pz = (cls >> 2) & 1, o = 0;
    for (int dx = px - 3; dx <= px + 2; dx++)
      for (int dy = py - 3; dy <= py + 2; dy++)
        for (int dz = pz - 3; dz <= pz + 2; dz++) {
          if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1) continue;
          int* e = &ot[(cls * NEBM + o) * 4];
          e[0] = dx; e[1] = dy; e[2] = dz; e[3] = offset_index(dx, dy, dz);
          o++;
        }
  }
  if ((rc = offtab.ensure(ot.size() * 4))) return rc;
  FMMX_HIP(hipMemcpy(offtab.p, ot.data(), ot.size() * 4, hipMemcpyHostToDevice));
  // the same offsets against the reflection-canonical operators: w = canonical
  // index | reflection bits (x, y, z) << 16
  canon = !fp64 && c->convention == FMMX_CONV_EXACT;
  if (const char* e = getenv("FMMX_M2L_CANON")) canon = canon && atoi(e) != 0;
  for (size_t e = 0; e < ot.size(); e += 4) {
    const int dx = ot[e], dy = ot[e + 1], dz = ot[e + 2];
    const int r = (dx < 0 ? 1 : 0) | (dy < 0 ? 2 : 0) | (dz < 0 ? 4 : 0);
    const int ci = canon_index(std::abs(dx), std::abs(dy), std::abs(dz));
    if (ci < 0) return fail(FMMX_ESTATE, "M2L offset outside the canonical table");
    ot[e + 3] = ci | (r << 16);
  }
  if ((rc = offtab_c.ensure(ot.size() * 4))) return rc;
  FMMX_HIP(hipMemcpy(offtab_c.p, ot.data(), ot.size() * 4, hipMemcpyHostToDevice));
  // the same entries of each class in order of reflection (stable): the FP16
  // kernel with DMA-staged operators re-signs its accumulator once per group
  for (int cls = 0; cls < 8; cls++) {
    int* row = &ot[(size_t)cls * NEBM * 4];
    std::vector<std::array<int, 4>> v(NEBM);
    for (int o = 0; o < NEBM; o++) v[o] = {row[4 * o], row[4 * o + 1], row[4 * o + 2], row[4 * o + 3]};
    std::stable_sort(v.begin(), v.end(), [](const std::array<int, 4>& a, const std::array<int, 4>& b) {
      return (a[3] >> 16) < (b[3] >> 16);
    });
    for (int o = 0; o < NEBM; o++)
      for (int k = 0; k < 4; k++) row[4 * o + k] = v[o][k];
  }
  if ((rc = offtab_s.ensure(ot.size() * 4))) return rc;
  FMMX_HIP(hipMemcpy(offtab_s.p, ot.data(), ot.size() * 4, hipMemcpyHostToDevice));
  return FMMX_OK;
}

int Engine::set_domain(int lm, float x0, float y0, float z0, float r) {
  if (lm != 0 && (lm < 2 || lm > MAXL)) return fail(FMMX_EINVAL, "lmax out of range [2, 8]");
  if (r > 0 && !(std::isfinite(x0) && std::isfinite(y0) && std::isfinite(z0) && std::isfinite(r)))
    return fail(FMMX_EINVAL, "root cube not finite");
  if (cfg.lmax == lm && cfg.xmin == x0 && cfg.ymin == y0 && cfg.zmin == z0 && cfg.rd == r) return FMMX_OK;
  cfg.lmax = lm;
  cfg.xmin = x0;
  cfg.ymin = y0;
  cfg.zmin = z0;
  cfg.rd = r;
  user_box = r > 0;
  have_points = false;  // the tree of the previous domain no longer applies
  tr_valid = false;
  return FMMX_OK;
}

int Engine::set_yukawa(double lambda) {
  if (!(lambda > 0) || !std::isfinite(lambda)) return fail(FMMX_EINVAL, "lambda must be positive and finite");
  lam = lambda;
  return FMMX_OK;
}

int Engine::set_groups(int level) {
  if (level < 0 || level > MAXL) return fail(FMMX_EINVAL, "group level out of range [0, 8]");
  drop_graph();
  group_level = level;
  have_points = false;  // the near lists depend on it
  tr_valid = false;
  return FMMX_OK;
}

int Engine::set_helmholtz(double k) {
  if (!(k > 0) || !std::isfinite(k)) return fail(FMMX_EINVAL, "k must be positive and finite");
  hk = k;
  return FMMX_OK;
}

int Engine::upload_helmholtz_operators() {
  if (hops_lmax == lmax && hops_rd == rd && hops_k == hk) return FMMX_OK;
  auto S = get_hh_ops(p, hk, rd, lmax);
  int rc;
  if (fp64) {
    if ((rc = upload<double>(hop_m2m, S->m2m, stream)) || (rc = upload<double>(hop_l2l, S->l2l, stream)) ||
        (rc = upload<double>(hop_m2l, S->m2l_kmaj, stream)))
      return rc;
  } else {
    if ((rc = upload<float>(hop_m2m, S->m2m, stream)) || (rc = upload<float>(hop_l2l, S->l2l, stream)) ||
        (rc = upload<float>(hop_m2l, S->m2l_frag, stream)))
      return rc;
  }
  hops_lmax = lmax;
  hops_rd = rd;
  hops_k = hk;
  return FMMX_OK;
}

int Engine::upload_yukawa_operators() {
  if (yops_lmax == lmax && yops_rd == rd && yops_lam == lam) return FMMX_OK;
  auto S = get_yuk_ops(p, lam, rd, lmax);
  int rc;
  if (fp64) {
    if ((rc = upload<double>(yop_m2m, S->m2m, stream)) || (rc = upload<double>(yop_l2l, S->l2l, stream)) ||
        (rc = upload<double>(yop_m2l, S->m2l_kmaj, stream)))
      return rc;
  } else {
    if ((rc = upload<float>(yop_m2m, S->m2m_frag, stream)) || (rc = upload<float>(yop_l2l, S->l2l_frag, stream)) ||
        (rc = upload<float>(yop_m2l, S->m2l_frag, stream)))
      return rc;
  }
  yops_lmax = lmax;
  yops_rd = rd;
  yops_lam = lam;
  return FMMX_OK;
}

int Engine::upload_operators() {
  if (ops_lmax == lmax && ops_rd == rd) return FMMX_OK;
  auto S = get_ops(tab, lmax, rd);
  m2l_per_level = S->nlev > 1;
  int rc;
  if (fp64) {
    if ((rc = upload<double>(op_m2m, S->m2m, stream)) || (rc = upload<double>(op_l2l, S->l2l, stream)) ||
        (rc = upload<double>(op_m2l, S->m2l_kmaj, stream)))
      return rc;
  } else {
    if ((rc = upload<float>(op_m2m, S->m2m_frag, stream)) || (rc = upload<float>(op_l2l, S->l2l_frag, stream)) ||
        (rc = upload<float>(op_m2l, canon ? S->m2l_canon : S->m2l_frag, stream)))
      return rc;
    split_ok = false;
    if (canon && !S->m2l_split.empty()) {
      if ((rc = op_m2l_h.ensure(S->m2l_split.size() * 16))) return rc;
      FMMX_HIP(hipMemcpyAsync(op_m2l_h.p, S->m2l_split.data(), S->m2l_split.size() * 16, hipMemcpyHostToDevice,
                              stream));
      FMMX_HIP(hipStreamSynchronize(stream));
      sa_e = S->split_e;
      split_ok = true;
    }
    wide_ok = false;
    if (split_ok && !S->m2l_wide.empty()) {
      if ((rc = op_m2l_w.ensure(S->m2l_wide.size() * 16))) return rc;
      FMMX_HIP(hipMemcpyAsync(op_m2l_w.p, S->m2l_wide.data(), S->m2l_wide.size() * 16, hipMemcpyHostToDevice,
                              stream));
      FMMX_HIP(hipStreamSynchronize(stream));
      wide_ok = true;
    }
  }
  ops_lmax = lmax;
  ops_rd = rd;
  return FMMX_OK;
}

static const float* stage(DevBuf& b, const float* src, int n, bool dev, hipStream_t s, int& rc) {
  if (!src || dev) return src;
  rc = b.ensure((size_t)(n > 0 ? n : 1) * 4);
  if (rc) return nullptr;
  if (n > 0 && hipMemcpyAsync(b.p, src, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess) {
    rc = fail(FMMX_EDEVICE, "hipMemcpyAsync H2D failed");
    return nullptr;
  }
  return b.as<float>();
}

int Engine::set_points(int ni, const float* xi, const float* yi, const float* zi, const float* nxi, const float* nyi,
                       const float* nzi, int nj, const float* xj, const float* yj, const float* zj, const float* nxj,
                       const float* nyj, const float* nzj, int flags) {
  if (ni < 0 || nj < 0) return fail(FMMX_EINVAL, "negative particle count");
  drop_graph();  // a captured evaluation refers to this tree's buffers
  tree_stats_pending = false;
  if ((ni > 0 && (!xi || !yi || !zi)) || (nj > 0 && (!xj || !yj || !zj)))
    return fail(FMMX_EINVAL, "coordinate pointer is NULL");
  FMMX_HIP(hipSetDevice(device));
  const bool dev = flags & FMMX_DEVICE_PTRS;
  keep_lists = flags & FMMX_KEEP_LISTS;
  int rc = FMMX_OK;
  const float* a[14] = {xi, yi, zi, nxi, nyi, nzi, xj, yj, zj, nxj, nyj, nzj, nullptr, nullptr};
  const float* d[14];
  // one tree for both sides when the caller passes the same coordinate arrays
  const bool same = ni == nj && ni > 0 && xi == xj && yi == yj && zi == zj;
  for (int k = 0; k < 12; k++) {
    if (same && k >= 6 && k < 9) {
      d[k] = d[k - 6];
      continue;
    }
    d[k] = stage(in[k], a[k], k < 6 ? ni : nj, dev, stream, rc);
    if (rc) return rc;
  }
  have_points = false;
  tr_valid = false;
  FMMX_HIP(hipEventRecord(ev[0], stream));
  // root cube
  if (user_box) {
    xmin = cfg.xmin; ymin = cfg.ymin; zmin = cfg.zmin; rd = cfg.rd;
  } else {
    DevBuf& mm = si.tmp;
    if ((rc = mm.ensure(64))) return rc;
    k_init_mm<<<1, 64, 0, stream>>>(mm.as<int>());
    if (ni > 0) k_minmax<<<nblk(ni), TPB, 0, stream>>>(ni, d[0], d[1], d[2], mm.as<int>());
    if (nj > 0) k_minmax<<<nblk(nj), TPB, 0, stream>>>(nj, d[6], d[7], d[8], mm.as<int>());
    if (!rb_host) FMMX_HIP(hipHostMalloc((void**)&rb_host, RB_INTS * 4, hipHostMallocDefault));
    int* h = rb_host;
    FMMX_HIP(hipMemcpyAsync(h, mm.p, 24, hipMemcpyDeviceToHost, stream));
    FMMX_HIP(hipStreamSynchronize(stream));
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      lo[k] = o2f(h[k]);
      hi[k] = o2f(h[3 + k]);
    }
    if (ni + nj == 0) lo[0] = lo[1] = lo[2] = 0, hi[0] = hi[1] = hi[2] = 1;
    // bio/main.cxx:364-368: the extent is a float difference, multiplied by
    // the double 1+1e-5 and rounded once to the float rdx; rd is the largest
    float r = 0;
    for (int k = 0; k < 3; k++) {
      if (!std::isfinite(lo[k]) || !std::isfinite(hi[k])) return fail(FMMX_EINVAL, "non-finite coordinates");
      const float ext = hi[k] - lo[k];
      r = std::max(r, (float)((double)ext * (1 + 1e-5)));
    }
    if (!(r > 0)) r = 1.0f;
    xmin = lo[0]; ymin = lo[1]; zmin = lo[2]; rd = r;
  }
  for (int l = 0; l <= MAXL; l++) occ_boxes[l] = 0, occ_sumsq[l] = 0;
  if (cfg.lmax > 0) {
    lmax = cfg.lmax;
  } else if (cfg.convention == FMMX_CONV_REF) {
    lmax = std::min(MAXL, nlevel(nj));
  } else {
    // modelled cost on the box statistics of the sources (targets when
    // there are none)
    const int o = nj > 0 ? 6 : 0;
    if ((rc = occupancy(nj > 0 ? nj : ni, d[o], d[o + 1], d[o + 2], occ_boxes, occ_sumsq))) return rc;
    lmax = auto_depth(occ_boxes, occ_sumsq, p, fp64);
  }
  // an automatic depth reaches at least the group level (independent
  // problems need their own boxes)
  if (cfg.lmax <= 0) lmax = std::max(lmax, std::min(group_level, MAXL));
  if (group_level > lmax)
    return fail(FMMX_EINVAL, "group level " + std::to_string(group_level) + " is deeper than lmax " +
                                 std::to_string(lmax) + " (fmmx_set_groups)");
  if ((rc = upload_operators())) return rc;
  if ((rc = bad_flag.ensure(4))) return rc;
  k_zero_i32<<<1, 32, 0, stream>>>(1, bad_flag.as<int>());
  // tree of both sides (keys, sort, gathers, per-level boxes)
  // targets = sources: one tree, whose dense tables the source side uses
  if ((rc = build_side(si, ni, d[0], d[1], d[2], d[3], d[4], d[5], same))) return rc;
  if ((rc = build_side(sj, nj, d[6], d[7], d[8], d[9], d[10], d[11], true, same ? &si : nullptr))) return rc;
  // near lists and parity-class order on upper-bound grids (the counts are
  // read on the device), queued ahead of the readback so that the GPU runs
  // them while the host waits
  if ((rc = build_lists())) return rc;
  if ((rc = build_m2l_order())) return rc;
  // the one readback of the tree build: box counts of every level (buffer
  // sizes and grids of the phases), the target boxes per parity class (M2L
  // launch shapes) and the out-of-cube flag; pinned, since small pageable
  // copies cost tens of us each
  if ((rc = rb_dev.ensure(RB_INTS * 4))) return rc;
  if (!rb_host) FMMX_HIP(hipHostMalloc((void**)&rb_host, RB_INTS * 4, hipHostMallocDefault));
  CountSrc cs{};
  cs.cnt[0] = ni > 0 ? si.cnt.as<int>() : nullptr;
  cs.cnt[1] = nj > 0 ? sj.cnt.as<int>() : nullptr;
  cs.cls = ni > 0 ? si.clscnt.as<int>() : nullptr;
  cs.bad = bad_flag.as<int>();
  k_pack_counts<<<1, 64, 0, stream>>>(cs, lmax, rb_dev.as<int>());
  FMMX_HIP(hipMemcpyAsync(rb_host, rb_dev.p, RB_INTS * 4, hipMemcpyDeviceToHost, stream));
  FMMX_HIP(hipStreamSynchronize(stream));
  if (rb_host[2 * (MAXL + 1)]) return fail(FMMX_EINVAL, "coordinates outside the root cube or not finite");
  for (int l = 2; l <= lmax; l++) {
    si.nb[l] = rb_host[l];
    sj.nb[l] = rb_host[MAXL + 1 + l];
  }
  for (int l = 2; l <= MAXL; l++)
    for (int c = 0; c < 8; c++) cls_cnt_h[l][c] = l <= lmax ? rb_host[RB_CLS + 8 * l + c] : 0;
  if (keep_lists && (rc = build_far_lists())) return rc;
  FMMX_HIP(hipEventRecord(ev[1], stream));
  // statistics read when asked for (fmmx_get_stats), without a wait here:
  // the near pair count and the tree time
  tree_stats_pending = true;
  have_points = true;
  last_neq = -1;
  return FMMX_OK;
}

// The statistics of the last set_points (near pairs, tree time), once its
// work has finished.
int Engine::resolve_tree_stats() {
  if (!tree_stats_pending) return FMMX_OK;
  FMMX_HIP(hipEventSynchronize(ev[1]));
  p2p_pairs = 0;
  if (near_nblk > 0) {
    std::vector<unsigned long long> part((size_t)near_nblk);
    FMMX_HIP(hipMemcpyAsync(part.data(), pair_part.p, part.size() * 8, hipMemcpyDeviceToHost, stream));
    FMMX_HIP(hipStreamSynchronize(stream));
    for (unsigned long long v : part) p2p_pairs += (int64_t)v;
  }
  float ms = 0;
  FMMX_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  stats.ms_tree = ms;
  tree_stats_pending = false;
  return FMMX_OK;
}

// Arguments of an evaluation (whole or one phase) against the tree in use.
int Engine::check_eval(int neq, const float* qj, const double* vi, const double* gxi, const double* gyi,
                       const double* gzi) {
  if (!have_points) return fail(FMMX_ESTATE, "evaluate before set_points");
  if (neq != FMMX_G && neq != FMMX_GNI && neq != FMMX_GNJ && neq != FMMX_YUKAWA && neq != FMMX_HELMHOLTZ)
    return fail(FMMX_EINVAL, "neq must be 0, 1, 2, 10 or 20");
  if (neq == FMMX_YUKAWA && !(lam > 0)) return fail(FMMX_ESTATE, "Yukawa needs lambda > 0 (fmmx_set_yukawa)");
  if (neq == FMMX_HELMHOLTZ) {
    if (!(hk > 0)) return fail(FMMX_ESTATE, "Helmholtz needs k > 0 (fmmx_set_helmholtz)");
    if (gxi || gyi || gzi) return fail(FMMX_EUNSUPPORTED, "Helmholtz returns the potential only");
    if (p < 4 || p > 12 || p % 2) return fail(FMMX_EUNSUPPORTED, "Helmholtz needs p in {4, 6, 8, 10, 12}");
    if (hk * rb(2) > HH_KRB_MAX) return fail(FMMX_EUNSUPPORTED, "k times the level-2 box size exceeds 5");
  }
  if (neq == FMMX_GNI && si.n > 0 && !si.has_normals) return fail(FMMX_EINVAL, "Gni needs target normals");
  if (neq == FMMX_GNJ && sj.n > 0 && !sj.has_normals) return fail(FMMX_EINVAL, "Gnj needs source normals");
  if ((sj.n > 0 && !qj) || (si.n > 0 && !vi)) return fail(FMMX_EINVAL, "q or v pointer is NULL");
  return FMMX_OK;
}

// Evaluation buffers and the kernel's operators (allocates only when a
// buffer grows, so a repeated evaluation on one tree allocates nothing).
int Engine::prepare(int neq) {
  FMMX_HIP(hipSetDevice(device));
  const int ni = si.n, nj = sj.n;
  const size_t es = fp64 ? 8 : 4;
  int rc = FMMX_OK;
  const size_t ni1 = (size_t)(ni > 0 ? ni : 1);
  yk = neq == FMMX_YUKAWA;
  hh = neq == FMMX_HELMHOLTZ;
  const int np = parts();
  if ((rc = qs.ensure((size_t)(nj > 0 ? nj : 1) * 4)) || (rc = vs.ensure(ni1 * 8)) || (rc = gxs.ensure(ni1 * 8)) ||
      (rc = gys.ensure(ni1 * 8)) || (rc = gzs.ensure(ni1 * 8)) ||
      (hh && (rc = qs2.ensure((size_t)(nj > 0 ? nj : 1) * 4))))
    return rc;
  for (int l = 2; l <= lmax; l++) {
    if ((rc = mexp[l].ensure((size_t)std::max(1, sj.nb[l]) * VS * np * es)) ||
        (rc = lexp[l].ensure((size_t)std::max(1, si.nb[l]) * P2 * np * es)))
      return rc;
    if (split_ok && !hh && !fp64 && (rc = mexp_h[l].ensure((size_t)(sj.nb[l] + 1) * split_box_halfs(p) * 2)))
      return rc;
  }
  if (yk && (rc = upload_yukawa_operators())) return rc;
  if (hh && (rc = upload_helmholtz_operators())) return rc;
  return FMMX_OK;
}

int Engine::evaluate(int neq, const float* qj, double* vi, double* gxi, double* gyi, double* gzi, int flags) {
  int rc = check_eval(neq, qj, vi, gxi, gyi, gzi);
  if (rc || (rc = prepare(neq))) return rc;
  const bool dev = flags & FMMX_DEVICE_PTRS;
  const int ni = si.n, nj = sj.n;
  const size_t ni1 = (size_t)(ni > 0 ? ni : 1);
  const float* q = stage(qin, qj, hh ? 2 * nj : nj, dev, stream, rc);
  if (rc) return rc;
  // Yukawa: no far field where lambda * (leaf box) makes it vanish in FP64
  const bool far = !yk || lam * rb(lmax) <= YUK_LRB_MAX;
  // outputs in the caller's order (device pointers, or staging for host ones)
  const bool g = neq == FMMX_G || neq == FMMX_YUKAWA;
  double* outs[4] = {vi, g ? gxi : nullptr, g ? gyi : nullptr, g ? gzi : nullptr};
  double* dsts[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int k = 0; !hh && k < 4; k++) {
    if (!outs[k] || ni == 0) continue;
    dsts[k] = outs[k];
    if (!dev) {
      if ((rc = out[k].ensure(ni1 * 8))) return rc;
      dsts[k] = out[k].as<double>();
    }
  }
  int e = 2;
  auto mark = [&]() -> int {
    if (per_phase) FMMX_HIP(hipEventRecord(ev[e++], stream));
    return FMMX_OK;
  };
  if (!capturing) FMMX_HIP(hipEventRecord(ev[2], stream));
  m2l_timed = false;
  e = 3;
  k_zero_f64<<<1, 64, 0, stream>>>(14, pair_ctr.as<double>() + 2);  // M2L pair counters per level
  if (nj > 0) {
    if (hh) k_gather_qc<<<nblk(nj), TPB, 0, stream>>>(nj, sj.perm.as<uint32_t>(), q, qs.as<float>(), qs2.as<float>());
    else k_gather_q<<<nblk(nj), TPB, 0, stream>>>(nj, sj.perm.as<uint32_t>(), q, qs.as<float>());
  }
  // Without per-phase timing the near field (vector ALU) runs on the second
  // stream beside the far field (MFMA); L2P, which adds to its results, waits.
  const bool overlap = overlap_on && !per_phase && aux && ni > 0;
  if (ni > 0) {
    if (overlap) {
      FMMX_HIP(hipEventRecord(ev[10], stream));
      FMMX_HIP(hipStreamWaitEvent(aux, ev[10], 0));
      hipStream_t main = stream;
      stream = aux;
      rc = run_p2p(neq);
      stream = main;
      if (rc) return rc;
      FMMX_HIP(hipEventRecord(ev[11], aux));
    } else if ((rc = run_p2p(neq))) {
      return rc;
    }
    if ((rc = mark())) return rc;  // ev3: after P2P
    if (far && (rc = run_p2m(neq))) return rc;
    if ((rc = mark())) return rc;  // ev4: after P2M
    for (int l = lmax - 1; far && l >= 2; l--)
      if ((rc = run_m2m(l))) return rc;
    if ((rc = mark())) return rc;  // ev5: after M2M
    const int l0 = first_m2l_level();
    last_m2l_all = far && m2l_all_levels();
    if (last_m2l_all) {
      // FP32 MFMA: every level's M2L in one launch (the levels above the leaf
      // fill the leaf launch's tail), then the L2L chain adds each parent's
      // locals; the same order with and without per-phase timing
      if ((rc = run_m2l_all(l0))) return rc;
      for (int l = l0 + 1; l <= lmax; l++)
        if ((rc = run_l2l(l, true))) return rc;
    } else if (far && overlap && !hh && lmax > l0 && m2l_splits(lmax) > 1) {
      // levels 2..lmax-1 (M2L, L2L) on the third stream while the leaf M2L
      // computes its partial sums; then the leaf L2L and the reduction that
      // adds the partials (the serialized order below, bit for bit)
      FMMX_HIP(hipEventRecord(ev[12], stream));
      FMMX_HIP(hipStreamWaitEvent(aux2, ev[12], 0));
      hipStream_t main = stream;
      stream = aux2;
      part_buf = &m2l_part2;
      rc = run_m2l(l0, true);
      for (int l = l0 + 1; !rc && l < lmax; l++) {
        if (!(rc = run_l2l(l))) rc = run_m2l(l, false);
      }
      stream = main;
      part_buf = &m2l_part;
      if (rc) return rc;
      FMMX_HIP(hipEventRecord(ev[13], aux2));
      if ((rc = run_m2l(lmax, false, 1))) return rc;
      FMMX_HIP(hipStreamWaitEvent(stream, ev[13], 0));
      if ((rc = run_l2l(lmax))) return rc;
      if ((rc = run_m2l(lmax, false, 2))) return rc;
    } else {
      if (far && (rc = run_m2l(l0, true))) return rc;
      for (int l = l0 + 1; far && l <= lmax; l++) {
        if ((rc = run_l2l(l))) return rc;
        if ((rc = run_m2l(l, false))) return rc;
      }
    }
    if ((rc = mark())) return rc;  // ev6: after M2L+L2L
    if (overlap) FMMX_HIP(hipStreamWaitEvent(stream, ev[11], 0));
    // Laplace and Yukawa: L2P writes the outputs in the caller's order itself
    if (far && (rc = run_l2p(neq, hh ? nullptr : dsts))) return rc;
    if ((rc = mark())) return rc;  // ev7: after L2P
  }
  const bool scattered = far && ni > 0 && !hh;
  // outputs in caller order
  if (hh) {
    if (ni > 0) {
      double* dst = vi;
      if (!dev) {
        if ((rc = out[0].ensure(2 * ni1 * 8))) return rc;
        dst = out[0].as<double>();
      }
      k_scatter_c<<<nblk(ni), TPB, 0, stream>>>(ni, si.perm.as<uint32_t>(), vs.as<double>(), gxs.as<double>(), dst);
      if (!dev) FMMX_HIP(hipMemcpyAsync(vi, dst, (size_t)ni * 16, hipMemcpyDeviceToHost, stream));
    }
    if (!capturing) FMMX_HIP(hipEventRecord(ev[15], stream));
    if (!dev) FMMX_HIP(hipStreamSynchronize(stream));
    last_neq = neq;
    return FMMX_OK;
  }
  double* srcs[4] = {vs.as<double>(), gxs.as<double>(), gys.as<double>(), gzs.as<double>()};
  if (ni > 0 && !scattered)
    k_scatter4<<<nblk(ni), TPB, 0, stream>>>(ni, si.perm.as<uint32_t>(), srcs[0], srcs[1], srcs[2], srcs[3], dsts[0],
                                             dsts[1], dsts[2], dsts[3]);
  for (int k = 0; k < 4; k++)
    if (!dev && dsts[k]) FMMX_HIP(hipMemcpyAsync(outs[k], dsts[k], (size_t)ni * 8, hipMemcpyDeviceToHost, stream));
  if (!capturing) FMMX_HIP(hipEventRecord(ev[15], stream));
  if (!dev) FMMX_HIP(hipStreamSynchronize(stream));
  last_neq = neq;
  return FMMX_OK;
}

// One phase of the matvec (the reference's operator symbols one at a time,
// fmm/fmm.cxx:42-150) on the device-resident state of the last set_points:
// strengths, expansions and near-field sums persist between phases.  The
// serialized order of evaluate() without the one-launch M2L.
int Engine::phase(int ph, int neq, int lev, const float* qj, double* vi, double* gxi, double* gyi, double* gzi,
                  int flags) {
  if (neq == FMMX_HELMHOLTZ) return fail(FMMX_EUNSUPPORTED, "phases of the Helmholtz kernel are not provided");
  const bool needs_q = ph == FMMX_PHASE_P2P || ph == FMMX_PHASE_P2M;
  const bool writes = ph == FMMX_PHASE_P2P || ph == FMMX_PHASE_L2P;
  if (ph < FMMX_PHASE_P2P || ph > FMMX_PHASE_L2P) return fail(FMMX_EINVAL, "phase out of range [0, 5]");
  int rc = check_eval(neq, needs_q ? qj : (const float*)1, writes ? vi : (const double*)1, gxi, gyi, gzi);
  if (rc) return rc;
  const int l0 = first_m2l_level();
  if ((ph == FMMX_PHASE_M2M && (lev < 2 || lev >= lmax)) || (ph == FMMX_PHASE_M2L && (lev < l0 || lev > lmax)) ||
      (ph == FMMX_PHASE_L2L && (lev <= l0 || lev > lmax)))
    return fail(FMMX_EINVAL, "level out of range for this phase");
  if ((rc = prepare(neq))) return rc;
  const bool dev = flags & FMMX_DEVICE_PTRS;
  const int ni = si.n, nj = sj.n;
  const size_t ni1 = (size_t)(ni > 0 ? ni : 1);
  // Yukawa: no far field where lambda * (leaf box) makes it vanish in FP64
  const bool far = !yk || lam * rb(lmax) <= YUK_LRB_MAX;
  if (needs_q && nj > 0) {
    const float* q = stage(qin, qj, nj, dev, stream, rc);
    if (rc) return rc;
    k_gather_q<<<nblk(nj), TPB, 0, stream>>>(nj, sj.perm.as<uint32_t>(), q, qs.as<float>());
  }
  const bool g = neq == FMMX_G || neq == FMMX_YUKAWA;
  double* outs[4] = {vi, g ? gxi : nullptr, g ? gyi : nullptr, g ? gzi : nullptr};
  double* dsts[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int k = 0; writes && k < 4; k++) {
    if (!outs[k] || ni == 0) continue;
    dsts[k] = outs[k];
    if (!dev) {
      if ((rc = out[k].ensure(ni1 * 8))) return rc;
      dsts[k] = out[k].as<double>();
    }
  }
  double* sums[4] = {vs.as<double>(), gxs.as<double>(), gys.as<double>(), gzs.as<double>()};
  switch (ph) {
    case FMMX_PHASE_P2P:
      if (ni > 0) {
        // the near field alone: P2P writes the per-target sums, which go out
        // in caller order and are then cleared for the far field
        for (int k = 0; k < 4; k++) k_zero_f64<<<nblk(ni), TPB, 0, stream>>>(ni, sums[k]);
        if ((rc = run_p2p(neq))) return rc;
        k_scatter4<<<nblk(ni), TPB, 0, stream>>>(ni, si.perm.as<uint32_t>(), sums[0], sums[1], sums[2], sums[3],
                                                 dsts[0], dsts[1], dsts[2], dsts[3]);
        for (int k = 0; k < 4; k++) k_zero_f64<<<nblk(ni), TPB, 0, stream>>>(ni, sums[k]);
      }
      break;
    case FMMX_PHASE_P2M:
      if (far && (rc = run_p2m(neq))) return rc;
      break;
    case FMMX_PHASE_M2M:
      if (far && (rc = run_m2m(lev))) return rc;
      break;
    case FMMX_PHASE_M2L:
      if (far && (rc = run_m2l(lev, lev == l0))) return rc;
      break;
    case FMMX_PHASE_L2L:
      if (far && (rc = run_l2l(lev))) return rc;
      break;
    case FMMX_PHASE_L2P:
      if (ni > 0) {
        // L2P adds the far field to the per-target sums: cleared first, so
        // that the outputs hold the far field alone
        for (int k = 0; k < 4; k++) k_zero_f64<<<nblk(ni), TPB, 0, stream>>>(ni, sums[k]);
        if (far) {
          if ((rc = run_l2p(neq, dsts))) return rc;
        } else {
          k_scatter4<<<nblk(ni), TPB, 0, stream>>>(ni, si.perm.as<uint32_t>(), sums[0], sums[1], sums[2], sums[3],
                                                   dsts[0], dsts[1], dsts[2], dsts[3]);
        }
      }
      break;
  }
  FMMX_HIP(hipGetLastError());
  for (int k = 0; writes && k < 4; k++)
    if (!dev && dsts[k]) FMMX_HIP(hipMemcpyAsync(outs[k], dsts[k], (size_t)ni * 8, hipMemcpyDeviceToHost, stream));
  if (!dev) FMMX_HIP(hipStreamSynchronize(stream));
  last_neq = neq;
  return FMMX_OK;
}

int Engine::direct(int neq, int ni, const float* xi, const float* yi, const float* zi, const float* nxi,
                   const float* nyi, const float* nzi, int nj, const float* xj, const float* yj, const float* zj,
                   const float* nxj, const float* nyj, const float* nzj, const float* qj, double* vi, double* gxi,
                   double* gyi, double* gzi, int flags) {
  if (ni < 0 || nj < 0) return fail(FMMX_EINVAL, "negative particle count");
  if (neq != FMMX_G && neq != FMMX_GNI && neq != FMMX_GNJ && neq != FMMX_YUKAWA && neq != FMMX_HELMHOLTZ)
    return fail(FMMX_EINVAL, "neq must be 0, 1, 2, 10 or 20");
  if (neq == FMMX_YUKAWA && !(lam > 0)) return fail(FMMX_ESTATE, "Yukawa needs lambda > 0 (fmmx_set_yukawa)");
  if (neq == FMMX_HELMHOLTZ) {
    if (!(hk > 0)) return fail(FMMX_ESTATE, "Helmholtz needs k > 0 (fmmx_set_helmholtz)");
    if (gxi || gyi || gzi) return fail(FMMX_EUNSUPPORTED, "Helmholtz returns the potential only");
    if (ni > 0 && (!xi || !yi || !zi || !vi)) return fail(FMMX_EINVAL, "NULL target pointer");
    if (nj > 0 && (!xj || !yj || !zj || !qj)) return fail(FMMX_EINVAL, "NULL source pointer");
    FMMX_HIP(hipSetDevice(device));
    const bool dev = flags & FMMX_DEVICE_PTRS;
    int rc = FMMX_OK;
    const float* a[7] = {xi, yi, zi, xj, yj, zj, qj};
    const int na[7] = {ni, ni, ni, nj, nj, nj, 2 * nj};
    const int slot[7] = {0, 1, 2, 6, 7, 8, 12};
    const float* d[7];
    for (int k = 0; k < 7; k++) {
      d[k] = stage(k < 6 ? in[slot[k]] : qin, a[k], na[k], dev, stream, rc);
      if (rc) return rc;
    }
    const size_t nj1 = (size_t)std::max(1, nj), ni1 = (size_t)std::max(1, ni);
    if ((rc = in[3].ensure(nj1 * 4)) || (rc = in[4].ensure(nj1 * 4)) || (rc = out[0].ensure(ni1 * 8)) ||
        (rc = out[1].ensure(ni1 * 8)) || (rc = out[2].ensure(2 * ni1 * 8)))
      return rc;
    if (nj > 0) k_gather_qc<<<nblk(nj), TPB, 0, stream>>>(nj, nullptr, d[6], in[3].as<float>(), in[4].as<float>());
    if ((rc = run_direct(neq, ni, d[0], d[1], d[2], nullptr, nullptr, nullptr, nj, d[3], d[4], d[5], in[4].as<float>(),
                         nullptr, nullptr, in[3].as<float>(), out[0].as<double>(), out[1].as<double>(), nullptr,
                         nullptr)))
      return rc;
    double* dst = dev ? vi : out[2].as<double>();
    if (ni > 0) k_scatter_c<<<nblk(ni), TPB, 0, stream>>>(ni, nullptr, out[0].as<double>(), out[1].as<double>(), dst);
    if (!dev && ni > 0) FMMX_HIP(hipMemcpyAsync(vi, dst, (size_t)ni * 16, hipMemcpyDeviceToHost, stream));
    if (!dev) FMMX_HIP(hipStreamSynchronize(stream));
    return FMMX_OK;
  }
  if (neq == FMMX_GNI && ni > 0 && (!nxi || !nyi || !nzi)) return fail(FMMX_EINVAL, "Gni needs target normals");
  if (neq == FMMX_GNJ && nj > 0 && (!nxj || !nyj || !nzj)) return fail(FMMX_EINVAL, "Gnj needs source normals");
  if (ni > 0 && (!xi || !yi || !zi || !vi)) return fail(FMMX_EINVAL, "NULL target pointer");
  if (nj > 0 && (!xj || !yj || !zj || !qj)) return fail(FMMX_EINVAL, "NULL source pointer");
  if (neq == FMMX_G && ni > 0 && (!gxi || !gyi || !gzi)) return fail(FMMX_EINVAL, "G needs gradient outputs");
  FMMX_HIP(hipSetDevice(device));
  const bool dev = flags & FMMX_DEVICE_PTRS;
  int rc = FMMX_OK;
  const float* a[13] = {xi, yi, zi, nxi, nyi, nzi, xj, yj, zj, nxj, nyj, nzj, qj};
  const float* d[13];
  for (int k = 0; k < 13; k++) {
    d[k] = stage(k < 12 ? in[k] : qin, a[k], k < 6 ? ni : nj, dev, stream, rc);
    if (rc) return rc;
  }
  double* o[4] = {vi, gxi, gyi, gzi};
  double* od[4];
  const bool ygrad = neq == FMMX_YUKAWA && gxi && gyi && gzi;  // Yukawa gradients are optional
  if (neq == FMMX_YUKAWA && !ygrad) o[1] = o[2] = o[3] = nullptr;
  for (int k = 0; k < 4; k++) {
    od[k] = o[k];
    if ((!dev && o[k]) || (k > 0 && neq == FMMX_YUKAWA && !ygrad)) {
      if ((rc = out[k].ensure((size_t)std::max(1, ni) * 8))) return rc;
      od[k] = out[k].as<double>();
    }
  }
  if ((rc = run_direct(neq, ni, d[0], d[1], d[2], d[3], d[4], d[5], nj, d[6], d[7], d[8], d[9], d[10], d[11], d[12],
                       od[0], od[1], od[2], od[3])))
    return rc;
  if (!dev) {
    for (int k = 0; k < 4; k++)
      if (o[k] && ni > 0 && (k == 0 || neq == FMMX_G || neq == FMMX_YUKAWA))
        FMMX_HIP(hipMemcpyAsync(o[k], od[k], (size_t)ni * 8, hipMemcpyDeviceToHost, stream));
    FMMX_HIP(hipStreamSynchronize(stream));
  }
  return FMMX_OK;
}

}  // namespace fmmx

using fmmx::Engine;

struct fmmx_handle {
  Engine e;
};

#define FMMX_GUARD(body)                                              \
  try {                                                               \
    body                                                              \
  } catch (const std::bad_alloc&) {                                   \
    return fmmx::fail(FMMX_ENOMEM, "host allocation failed");         \
  } catch (const std::exception& ex) {                                \
    return fmmx::fail(FMMX_EDEVICE, ex.what());                       \
  }

extern "C" {

int fmmx_abi_version(void) { return FMMX_ABI_VERSION; }

int fmmx_operator(int p, int convention, int kind, int a0, int a1, int a2, double rb, double* A) {
  FMMX_GUARD({
    if (p < 2 || p > fmmx::MAXP || !A) return fmmx::fail(FMMX_EINVAL, "bad p or NULL output");
    if (convention != FMMX_CONV_REF && convention != FMMX_CONV_EXACT) return fmmx::fail(FMMX_EINVAL, "bad convention");
    if (convention == FMMX_CONV_REF && p > fmmx::P_ALL)
      return fmmx::fail(FMMX_EUNSUPPORTED, "the REF convention's float tables turn subnormal above p = 12");
    auto T = fmmx::get_tables(p, convention);
    if (kind == 0 || kind == 1) {
      if (a0 < 0 || a0 > 7) return fmmx::fail(FMMX_EINVAL, "octant out of range");
      if (kind == 0) fmmx::op_m2m(*T, a0, A);
      else fmmx::op_l2l(*T, a0, A);
      return FMMX_OK;
    }
    if (kind == 2) {
      if (fmmx::offset_index(a0, a1, a2) < 0 || !(rb > 0)) return fmmx::fail(FMMX_EINVAL, "bad M2L offset or rb");
      fmmx::op_m2l(*T, a0, a1, a2, rb, A);
      return FMMX_OK;
    }
    if (kind == 3 || kind == 4) {
      if (a0 < 0 || a0 > 7 || !(rb > 0)) return fmmx::fail(FMMX_EINVAL, "octant out of range or lambda*rb <= 0");
      std::vector<double> other((size_t)p * p * p * p);
      if (kind == 3) fmmx::yuk_op_m2m_l2l(p, rb, a0, A, other.data());
      else fmmx::yuk_op_m2m_l2l(p, rb, a0, other.data(), A);
      return FMMX_OK;
    }
    if (kind == 5) {
      if (fmmx::offset_index(a0, a1, a2) < 0 || !(rb > 0)) return fmmx::fail(FMMX_EINVAL, "bad M2L offset or lambda*rb");
      fmmx::yuk_op_m2l(p, rb, a0, a1, a2, A);
      return FMMX_OK;
    }
    if (kind == 6 || kind == 7) {
      if (a0 < 0 || a0 > 7 || !(rb > 0) || rb > 0.5 * fmmx::HH_KRB_MAX)
        return fmmx::fail(FMMX_EINVAL, "octant out of range or k*rb outside (0, HH_KRB_MAX/2]");
      std::vector<double> other((size_t)4 * p * p * p * p);
      if (kind == 6) fmmx::hh_op_m2m_l2l(p, rb, a0, A, other.data());
      else fmmx::hh_op_m2m_l2l(p, rb, a0, other.data(), A);
      return FMMX_OK;
    }
    if (kind == 8) {
      if (fmmx::offset_index(a0, a1, a2) < 0 || !(rb > 0) || rb > fmmx::HH_KRB_MAX)
        return fmmx::fail(FMMX_EINVAL, "bad M2L offset or k*rb outside (0, HH_KRB_MAX]");
      fmmx::hh_op_m2l(p, rb, a0, a1, a2, A);
      return FMMX_OK;
    }
    return fmmx::fail(FMMX_EINVAL, "bad kind");
  })
}

const char* fmmx_last_error(void) { return fmmx::last_error().c_str(); }

int fmmx_create(const fmmx_config* cfg, int device, void* stream, fmmx_handle** out) {
  FMMX_GUARD({
    if (!out) return fmmx::fail(FMMX_EINVAL, "out is NULL");
    *out = nullptr;
    auto* h = new fmmx_handle();
    int rc = h->e.init(cfg, device, stream);
    if (rc) {
      delete h;
      return rc;
    }
    *out = h;
    return FMMX_OK;
  })
}

int fmmx_destroy(fmmx_handle* h) {
  FMMX_GUARD({
    if (!h) return FMMX_OK;
    (void)hipSetDevice(h->e.device);
    (void)hipStreamSynchronize(h->e.stream);
    delete h;
    return FMMX_OK;
  })
}

int fmmx_set_points(fmmx_handle* h, int ni, const float* xi, const float* yi, const float* zi, const float* nxi,
                    const float* nyi, const float* nzi, int nj, const float* xj, const float* yj, const float* zj,
                    const float* nxj, const float* nyj, const float* nzj, int flags) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.set_points(ni, xi, yi, zi, nxi, nyi, nzi, nj, xj, yj, zj, nxj, nyj, nzj, flags);
  })
}

int fmmx_evaluate(fmmx_handle* h, int neq, const float* qj, double* vi, double* gxi, double* gyi, double* gzi,
                  int flags) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.evaluate(neq, qj, vi, gxi, gyi, gzi, flags);
  })
}

int fmmx_direct(fmmx_handle* h, int neq, int ni, const float* xi, const float* yi, const float* zi, const float* nxi,
                const float* nyi, const float* nzi, int nj, const float* xj, const float* yj, const float* zj,
                const float* nxj, const float* nyj, const float* nzj, const float* qj, double* vi, double* gxi,
                double* gyi, double* gzi, int flags) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.direct(neq, ni, xi, yi, zi, nxi, nyi, nzi, nj, xj, yj, zj, nxj, nyj, nzj, qj, vi, gxi, gyi, gzi,
                       flags);
  })
}

int fmmx_set_domain(fmmx_handle* h, int lmax, float xmin, float ymin, float zmin, float rd) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.set_domain(lmax, xmin, ymin, zmin, rd);
  })
}

int fmmx_graph_capture(fmmx_handle* h, int neq, const float* qj, double* vi, double* gxi, double* gyi,
                       double* gzi) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.graph_capture(neq, qj, vi, gxi, gyi, gzi);
  })
}

int fmmx_set_m2l_variant(fmmx_handle* h, int lev, int nt, int staging, int wpg, int nsplit) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    if (lev < 2 || lev > fmmx::MAXL) return fmmx::fail(FMMX_EINVAL, "level out of range [2, 8]");
    if (nt < 0 || nt > 4 || nt == 3 || (nt == 4 && staging != 5) || staging < -1 || staging > 5 || (wpg != 0 && wpg != 4 && wpg != 8 && wpg != 16) ||
        nsplit < 0 || nsplit > 16)
      return fmmx::fail(FMMX_EINVAL, "bad M2L variant");
    auto& t = h->e.tune[lev];
    t.nt = nt;
    t.pa = staging;
    t.wpg = wpg;
    t.nsplit = nsplit;
    return FMMX_OK;
  })
}

int fmmx_get_m2l_variant(fmmx_handle* h, int lev, int* v4) {
  FMMX_GUARD({
    if (!h || !v4) return fmmx::fail(FMMX_EINVAL, "NULL argument");
    if (lev < 2 || lev > fmmx::MAXL) return fmmx::fail(FMMX_EINVAL, "level out of range [2, 8]");
    if (!h->e.have_points) return fmmx::fail(FMMX_ESTATE, "the split count depends on the points (set_points)");
    const auto v = h->e.m2l_variant(lev);
    v4[0] = v.nt;
    v4[1] = v.pa;
    v4[2] = v.wpg;
    v4[3] = v.nsplit;
    return FMMX_OK;
  })
}

int fmmx_graph_launch(fmmx_handle* h) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.graph_launch();
  })
}

int fmmx_set_groups(fmmx_handle* h, int level) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.set_groups(level);
  })
}

int fmmx_set_helmholtz(fmmx_handle* h, double k) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.set_helmholtz(k);
  })
}

int fmmx_set_yukawa(fmmx_handle* h, double lambda) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.set_yukawa(lambda);
  })
}

int fmmx_sort(fmmx_handle* h, int n, const float* x, const float* y, const float* z, int* keys, int* perm,
              int flags) {
  FMMX_GUARD({
    if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
    return h->e.sort_points(n, x, y, z, keys, perm, flags);
  })
}

int fmmx_sort_noalloc(int device, void* stream, int n, const float* x, const float* y, const float* z, int lmax,
                      float xmin, float ymin, float zmin, float rd, int* keys, int* perm, int* bad) {
  FMMX_GUARD({ return fmmx::sort_noalloc(device, stream, n, x, y, z, lmax, xmin, ymin, zmin, rd, keys, perm, bad); })
}

int fmmx_synchronize(fmmx_handle* h) {
  if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
  if (hipStreamSynchronize(h->e.stream) != hipSuccess) return fmmx::fail(FMMX_EDEVICE, "stream synchronize failed");
  return FMMX_OK;
}

int fmmx_set_timing(fmmx_handle* h, int per_phase) {
  if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
  h->e.per_phase = per_phase != 0;
  return FMMX_OK;
}

int fmmx_get_stats(fmmx_handle* h, fmmx_stats* st) {
  if (!h || !st) return fmmx::fail(FMMX_EINVAL, "NULL argument");
  Engine& e = h->e;
  if (int rc = e.resolve_tree_stats()) return rc;
  if (hipEventSynchronize(e.ev[15]) != hipSuccess) return fmmx::fail(FMMX_EDEVICE, "event synchronize failed");
  fmmx_stats s = e.stats;
  float ms = 0;
  if (e.last_neq >= 0 && hipEventElapsedTime(&ms, e.ev[2], e.ev[15]) == hipSuccess) s.ms_total = ms;
  if (e.per_phase && e.last_neq >= 0 && e.si.n > 0) {
    double* dst[5] = {&s.ms_p2p, &s.ms_p2m, &s.ms_m2m, &s.ms_m2l, &s.ms_l2p};
    for (int k = 0; k < 5; k++)
      if (hipEventElapsedTime(&ms, e.ev[2 + k], e.ev[3 + k]) == hipSuccess) *dst[k] = ms;
    s.ms_l2l = 0;  // interleaved with M2L, counted there
  }
  s.lmax = e.lmax;
  s.nbi_leaf = e.si.nb[e.lmax];
  s.nbj_leaf = e.sj.nb[e.lmax];
  s.nbi_total = s.nbj_total = 0;
  for (int l = 2; l <= e.lmax; l++) {
    s.nbi_total += e.si.nb[l];
    s.nbj_total += e.sj.nb[l];
  }
  s.p2p_pairs = e.p2p_pairs;
  unsigned long long mp[16] = {};
  // the M2L pair counters belong to the last evaluation on this tree
  if (e.last_neq >= 0 && e.pair_ctr.p && hipMemcpy(mp, e.pair_ctr.p, sizeof(mp), hipMemcpyDeviceToHost) != hipSuccess)
    return fmmx::fail(FMMX_EDEVICE, "hipMemcpy D2H failed");
  s.m2l_pairs = 0;
  for (int l = 2; l <= e.lmax; l++) s.m2l_pairs += (int64_t)mp[2 + l];
  s.m2l_pairs_leaf = (int64_t)mp[2 + e.lmax];
  s.ms_m2l_leaf = 0;
  if (e.per_phase && e.m2l_timed && e.last_neq >= 0 && hipEventElapsedTime(&ms, e.ev[8], e.ev[9]) == hipSuccess)
    s.ms_m2l_leaf = ms;
  // a query that failed above (events of a phase that did not run) must not
  // leave the runtime's last error set for the caller's next launch
  (void)hipGetLastError();
  *st = s;
  return FMMX_OK;
}

int fmmx_evaluate_phase(fmmx_handle* h, int phase, int neq, int lev, const float* qj, double* vi, double* gxi,
                        double* gyi, double* gzi, int flags) {
  if (!h) return fmmx::fail(FMMX_EINVAL, "NULL handle");
  FMMX_GUARD({ return h->e.phase(phase, neq, lev, qj, vi, gxi, gyi, gzi, flags); })
}

int fmmx_evaluate_transpose(fmmx_handle* h, int neq, const float* wi, double* uj, double* gxj, double* gyj,
                            double* gzj, int flags) {
  if (!h) return fmmx::fail(FMMX_EINVAL, "NULL handle");
  FMMX_GUARD({ return h->e.evaluate_transpose(neq, wi, uj, gxj, gyj, gzj, flags); })
}

int fmmx_gradients(fmmx_handle* h, int neq, const float* qj, const float* wi, double* dq, double* dxi, double* dxj,
                   int flags) {
  if (!h) return fmmx::fail(FMMX_EINVAL, "NULL handle");
  FMMX_GUARD({ return h->e.gradients(neq, qj, wi, dq, dxi, dxj, flags); })
}

int fmmx_energy(fmmx_handle* h, int neq, const float* q, double* E, double* dEdq, double* dEdx, int flags) {
  if (!h) return fmmx::fail(FMMX_EINVAL, "NULL handle");
  FMMX_GUARD({ return h->e.energy(neq, q, E, dEdq, dEdx, flags); })
}

int fmmx_auto_lmax(const long long* boxes, const double* sumsq, int p, int precision, int* lmax) {
  if (!boxes || !sumsq || !lmax) return fmmx::fail(FMMX_EINVAL, "NULL argument");
  if (p < 2 || (p > fmmx::P_ALL && p != fmmx::P_HI)) return fmmx::fail(FMMX_EINVAL, "p must be 2..12 or 18");
  if (precision != FMMX_MIXED && precision != FMMX_FP64) return fmmx::fail(FMMX_EINVAL, "unknown precision");
  if (p > fmmx::P_ALL && precision != FMMX_FP64) return fmmx::fail(FMMX_EUNSUPPORTED, "p = 18 needs FMMX_FP64");
  long b[fmmx::MAXL + 1];
  double q[fmmx::MAXL + 1];
  for (int l = 0; l <= fmmx::MAXL; l++) {
    if (boxes[l] < 0 || !(sumsq[l] >= 0)) return fmmx::fail(FMMX_EINVAL, "negative box statistics");
    b[l] = (long)boxes[l];
    q[l] = sumsq[l];
  }
  *lmax = fmmx::auto_depth(b, q, p, precision == FMMX_FP64);
  return FMMX_OK;
}

int fmmx_get_occupancy(fmmx_handle* h, long long* boxes, double* sumsq) {
  if (!h || !boxes || !sumsq) return fmmx::fail(FMMX_EINVAL, "NULL argument");
  Engine& e = h->e;
  if (!e.have_points) return fmmx::fail(FMMX_ESTATE, "no tree (fmmx_set_points)");
  for (int l = 0; l <= fmmx::MAXL; l++) {
    boxes[l] = e.occ_boxes[l];
    sumsq[l] = e.occ_sumsq[l];
  }
  return FMMX_OK;
}

int fmmx_get_domain(fmmx_handle* h, int* lmax, float* box4) {
  if (!h) return fmmx::fail(FMMX_EINVAL, "NULL handle");
  Engine& e = h->e;
  if (!e.have_points) return fmmx::fail(FMMX_ESTATE, "no tree (fmmx_set_points)");
  if (lmax) *lmax = e.lmax;
  if (box4) {
    box4[0] = e.xmin;
    box4[1] = e.ymin;
    box4[2] = e.zmin;
    box4[3] = e.rd;
  }
  return FMMX_OK;
}

int fmmx_get_count(fmmx_handle* h, int what, int side, int lev) {
  if (!h) return -1;
  Engine& e = h->e;
  if (what == 0) return e.lmax;
  if (what == 2) return e.last_m2l_all ? 1 : 0;
  if (lev < 2 || lev > e.lmax) return -1;
  return side == 0 ? e.si.nb[lev] : e.sj.nb[lev];
}

int fmmx_get_perm(fmmx_handle* h, int side, int* perm) {
  if (!h || !perm) return fmmx::fail(FMMX_EINVAL, "NULL argument");
  fmmx::Side& s = side == 0 ? h->e.si : h->e.sj;
  if (s.n > 0 && hipMemcpy(perm, s.perm.p, (size_t)s.n * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return fmmx::fail(FMMX_EDEVICE, "copy failed");
  return FMMX_OK;
}

int fmmx_get_boxes(fmmx_handle* h, int side, int lev, int* keys, int* first, int* last) {
  if (!h) return fmmx::fail(FMMX_EINVAL, "handle is NULL");
  Engine& e = h->e;
  if (lev < 2 || lev > e.lmax) return fmmx::fail(FMMX_EINVAL, "level out of range");
  fmmx::Side& s = side == 0 ? e.si : e.sj;
  size_t nb = (size_t)s.nb[lev];
  if (nb == 0) return FMMX_OK;
  if (keys && hipMemcpy(keys, s.bkey[lev].p, nb * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return fmmx::fail(FMMX_EDEVICE, "copy failed");
  if (lev == e.lmax) {
    if (first && hipMemcpy(first, s.bfirst[lev].p, nb * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return fmmx::fail(FMMX_EDEVICE, "copy failed");
    if (last && hipMemcpy(last, s.blast.p, nb * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return fmmx::fail(FMMX_EDEVICE, "copy failed");
  }
  return FMMX_OK;
}

int fmmx_get_list(fmmx_handle* h, int lev, int kind, int* nij_out, int* neij_out) {
  if (!h || !nij_out || !neij_out) return fmmx::fail(FMMX_EINVAL, "NULL argument");
  Engine& e = h->e;
  if (lev < 2 || lev > e.lmax) return fmmx::fail(FMMX_EINVAL, "level out of range");
  size_t nb = (size_t)e.si.nb[lev];
  if (kind == 0) {
    if (lev != e.lmax) return fmmx::fail(FMMX_EINVAL, "near list exists at the leaf level only");
    if (nb == 0) return FMMX_OK;
    std::vector<int> rows(nb * fmmx::NNEAR);
    if (hipMemcpy(nij_out, e.nij.p, nb * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(rows.data(), e.neij.p, rows.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return fmmx::fail(FMMX_EDEVICE, "copy failed");
    for (size_t b = 0; b < nb; b++)
      for (int c = 0; c < fmmx::NEBM; c++) neij_out[b * fmmx::NEBM + c] = c < nij_out[b] ? rows[b * fmmx::NNEAR + c] : -1;
    return FMMX_OK;
  }
  if (!e.keep_lists) return fmmx::fail(FMMX_ESTATE, "far lists need FMMX_KEEP_LISTS at set_points");
  if (nb == 0) return FMMX_OK;
  if (hipMemcpy(nij_out, e.fnij[lev].p, nb * 4, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(neij_out, e.fneij[lev].p, nb * fmmx::NEBM * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return fmmx::fail(FMMX_EDEVICE, "copy failed");
  return FMMX_OK;
}

static int copy_expansion(Engine& e, fmmx::DevBuf& b, int nb, int lev, bool local, double* out) {
  const int p = e.p, nr = fmmx::nre(p);
  const int P2 = local ? e.P2 : e.VS;  // stored vector stride
  auto at = [&](int c) { return local ? c : e.mslot(c); };
  std::vector<double> v((size_t)nb * P2);
  if (nb == 0) return FMMX_OK;
  if (e.fp64) {
    if (hipMemcpy(v.data(), b.p, v.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
      return fmmx::fail(FMMX_EDEVICE, "copy failed");
  } else {
    std::vector<float> f(v.size());
    if (hipMemcpy(f.data(), b.p, f.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
      return fmmx::fail(FMMX_EDEVICE, "copy failed");
    std::copy(f.begin(), f.end(), v.begin());
  }
  const double rb = e.rb(lev);
  for (int i = 0; i < nb; i++)
    for (int n = 0; n < p; n++) {
      const double s = local ? 1.0 / std::pow(rb, n + 1) : std::pow(rb, n);
      for (int m = 0; m <= n; m++) {
        const double re = v[(size_t)i * P2 + at(fmmx::slot_re(n, m))];
        const double im = m ? v[(size_t)i * P2 + at(fmmx::slot_im(p, n, m))] : 0.0;
        out[2 * ((size_t)i * nr + fmmx::slot_re(n, m))] = s * re;
        out[2 * ((size_t)i * nr + fmmx::slot_re(n, m)) + 1] = s * im;
      }
    }
  return FMMX_OK;
}

int fmmx_get_multipoles(fmmx_handle* h, double* bx) {
  if (!h || !bx) return fmmx::fail(FMMX_EINVAL, "NULL argument");
  Engine& e = h->e;
  if (e.last_neq < 0) return fmmx::fail(FMMX_ESTATE, "no evaluation yet");
  if (hipStreamSynchronize(e.stream) != hipSuccess) return fmmx::fail(FMMX_EDEVICE, "sync failed");
  size_t row = 0;
  const int nr = fmmx::nre(e.p);
  for (int l = e.lmax; l >= 2; l--) {
    int rc = copy_expansion(e, e.mexp[l], e.sj.nb[l], l, false, bx + 2 * row * nr);
    if (rc) return rc;
    row += e.sj.nb[l];
  }
  return FMMX_OK;
}

int fmmx_get_multipoles_level(fmmx_handle* h, int lev, double* bx) {
  if (!h || !bx) return fmmx::fail(FMMX_EINVAL, "NULL argument");
  Engine& e = h->e;
  if (e.last_neq < 0) return fmmx::fail(FMMX_ESTATE, "no evaluation yet");
  if (lev < 2 || lev > e.lmax) return fmmx::fail(FMMX_EINVAL, "level out of range");
  if (hipStreamSynchronize(e.stream) != hipSuccess) return fmmx::fail(FMMX_EDEVICE, "sync failed");
  return copy_expansion(e, e.mexp[lev], e.sj.nb[lev], lev, false, bx);
}

int fmmx_get_locals(fmmx_handle* h, int lev, double* ax) {
  if (!h || !ax) return fmmx::fail(FMMX_EINVAL, "NULL argument");
  Engine& e = h->e;
  if (e.last_neq < 0) return fmmx::fail(FMMX_ESTATE, "no evaluation yet");
  if (lev < 2 || lev > e.lmax) return fmmx::fail(FMMX_EINVAL, "level out of range");
  if (hipStreamSynchronize(e.stream) != hipSuccess) return fmmx::fail(FMMX_EDEVICE, "sync failed");
  return copy_expansion(e, e.lexp[lev], e.si.nb[lev], lev, true, ax);
}

}  // extern "C"
