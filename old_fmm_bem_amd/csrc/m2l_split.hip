// m2l_split.hip -- M2L of the mixed-precision engine on the FP16 matrix cores
// (v_mfma_f32_16x16x32_f16, CDNA4) with error compensation: operands split into
// FP16 hi + lo pairs, three products per pair of operands, FP32 accumulation
// (the hi/lo splitting of Ootomo and Yokota, "Recovering single precision
// accuracy from Tensor Cores while surpassing the FP32 theoretical peak
// performance", IJHPCA 2022).  The tiling, class order, offset splits and XCD
// mapping are those of the FP32 MFMA kernel (kernels.hip, k_m2l): a workgroup
// owns a tile of targets of one parity class at a level, all sharing the 189
// offsets of kernel/m2l.cxx's far list (fmm/ijbox.cxx:366-401); per offset the
// split operator is staged in LDS and every wave multiplies it with its
// targets' gathered source multipoles.
#include "m2l.h"
#include "tree_util.h"

namespace fmmx {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// sign masks of a reflection in the split operator's fragment order:
// byte g of col[r][ks] flags the 8 k-positions 8 ks .. 8 ks + 7 of lane group g
// whose input slot T_r negates; bit j of row[r][mt] flags output row 16 mt + j
template <int P>
struct HSign {
  using H = HLayout<P>;
  uint32_t col[8][H::KH];
  uint32_t row[8][H::MT];
  constexpr HSign() : col{}, row{} {
    constexpr ReflMask<P> RM{};
    for (int r = 0; r < 8; r++) {
      for (int ks = 0; ks < H::KH; ks++)
        for (int g = 0; g < 4; g++)
          for (int t = 0; t < 8; t++) {
            const int c = H::slot(g, 8 * ks + t);
            if (c >= 0 && RM.neg(r, c)) col[r][ks] |= 1u << (8 * g + t);
          }
      for (int mt = 0; mt < H::MT; mt++)
        for (int j = 0; j < 16; j++)
          if (16 * mt + j < H::P2 && RM.neg(r, 16 * mt + j)) row[r][mt] |= 1u << j;
    }
  }
};

template <int P>
__constant__ HSign<P> c_hsign = HSign<P>();

// 2^e that brings the level's largest magnitude just below 2^SPLIT_TOP
__device__ __forceinline__ int split_exp(const unsigned* mmax) {
  const float m = __uint_as_float(*mmax);
  if (!(m > 0.f) || !(m < 3.0e38f)) return 0;
  int e;
  (void)frexpf(m, &e);
  return min(100, max(-100, SPLIT_TOP - e));
}

// bits of v (0..3) at positions 0 and 3 (one axis of a Morton key)
__device__ __forceinline__ uint32_t dil3(int v) { return (uint32_t)(v & 1) | ((uint32_t)(v & 2) << 2); }

__device__ __forceinline__ uint32_t pack_h2(float a, float b) {
  const _Float16 ha = (_Float16)a, hb = (_Float16)b;
  return (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}

constexpr int MS_T = 256;

__global__ void k_msplit_zero(MSplitLevels S) {
  if ((int)threadIdx.x < S.nl) *S.mmax[threadIdx.x] = 0u;
}

// largest |multipole| of every level (valid slots only), as float bits;
// mmax must hold 0 before the launch
template <int P>
__global__ __launch_bounds__(MS_T) void k_msplit_max(MSplitLevels S) {
  using ML = MLayout<P>;
  const long t = (long)blockIdx.x * MS_T + threadIdx.x;  // one thread per (box, lane group)
  __shared__ unsigned smax[MAXL + 1];
  if (threadIdx.x <= MAXL) smax[threadIdx.x] = 0u;
  __syncthreads();
  const long b = t >> 2;
  int k = 0;
  while (k + 1 < S.nl && b >= S.off[k + 1]) k++;
  float mx = 0.f;
  if (b < S.off[S.nl] && b != S.off[k + 1] - 1) {  // not the zero box
    const int g = (int)(t & 3);
    const float* m = S.M[k] + (size_t)(b - S.off[k]) * ML::VS + g * ML::KSP;
#pragma unroll
    for (int u = 0; u < ML::KS; u++)
      if (g * ML::KS + u < ML::P2) mx = fmaxf(mx, fabsf(m[u]));
  }
  // non-negative floats order as their bit patterns; a wave within one
  // level reduces first (one LDS atomic instead of 64 on one address)
  const int k0 = __builtin_amdgcn_readfirstlane(k);
  if (__ballot(k != k0) == 0) {
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0 && mx > 0.f) atomicMax(&smax[k0], __float_as_uint(mx));
  } else if (mx > 0.f) {
    atomicMax(&smax[k], __float_as_uint(mx));
  }
  __syncthreads();
  if ((int)threadIdx.x < S.nl && smax[threadIdx.x]) atomicMax(S.mmax[threadIdx.x], smax[threadIdx.x]);
}

// the split pairs of every level: lane group g (of G) of box b holds
// hi[u], lo[u] of input slot g L::KS + u scaled by 2^split_exp, in the
// layout L (HLayout, 4 groups: 16x16x32; WLayout, 2 groups: 32x32x16).  One
// thread per 32-bit word of the output (hi pairs, lo pairs, zero padding),
// so that both the multipole reads and the writes are contiguous across a
// wave.
template <int P, class L, int G>
__global__ __launch_bounds__(MS_T) void k_msplit(MSplitLevels S) {
  using ML = MLayout<P>;
  constexpr int JW = L::GH / 2;  // words per lane group
  // 32-bit index arithmetic (the division by JW is by a constant)
  const unsigned t = blockIdx.x * MS_T + threadIdx.x;
  const unsigned bg = t / JW;
  const int j = (int)(t % JW);
  const long b = bg / G;
  const int g = (int)(bg % G);
  if (b >= S.off[S.nl]) return;
  int k = 0;
  while (k + 1 < S.nl && b >= S.off[k + 1]) k++;
  // the last box of a level's range is the zero box (read where a target
  // has no source box)
  const bool zero = b == S.off[k + 1] - 1;
  uint32_t w = 0u;
  if (!zero && j < L::SEG) {
    const int u = 2 * (j % (L::SEG / 2));
    const float sc = ldexpf(1.f, split_exp(S.mmax[k]));
    const float* m = S.M[k] + (size_t)(b - S.off[k]) * ML::VS;
    const int c0 = g * L::KS + u, c1 = c0 + 1;
    const float x0 = (u < L::KS && c0 < ML::P2) ? m[ML::at(c0)] * sc : 0.f;
    const float x1 = (u + 1 < L::KS && c1 < ML::P2) ? m[ML::at(c1)] * sc : 0.f;
    const float h0 = (float)(_Float16)x0, h1 = (float)(_Float16)x1;
    w = j < L::SEG / 2 ? pack_h2(h0, h1) : pack_h2(x0 - h0, x1 - h1);
  }
  uint32_t* dst = reinterpret_cast<uint32_t*>(S.Mh[k]);
  dst[((size_t)(b - S.off[k]) * L::VSH + g * L::GH) / 2 + j] = w;
}

// B' fragment of k-step ks from a lane's loaded words: k-positions
// 8 ks + 2 d, 8 ks + 2 d + 1 (a fixed selection; aligned steps are the loaded
// registers themselves)
template <class H>
__device__ __forceinline__ h8 bfrag(const uint32_t* bw, int ks) {
  const int w0 = H::word(8 * ks), w1 = H::word(8 * ks + 2), w2 = H::word(8 * ks + 4), w3 = H::word(8 * ks + 6);
  return __builtin_bit_cast(h8, u32x4{w0 < 0 ? 0u : bw[w0], w1 < 0 ? 0u : bw[w1], w2 < 0 ? 0u : bw[w2],
                                      w3 < 0 ? 0u : bw[w3]});
}

template <int P, int NT, int WPG, int DB>
__global__ __launch_bounds__(64 * WPG) void k_m2l_h(M2LLaunch A) {
  M2LLev Q = A.lv[0];
#pragma unroll
  for (int k = 1; k < MAXL + 1; k++)
    if (k < A.nlv && (int)blockIdx.x >= A.lv[k].wg0) Q = A.lv[k];
  const int lev = Q.lev, nb = Q.nb, nsplit = Q.nsplit, init = Q.init, region = Q.region, gsh = Q.gsh;
  const int* __restrict__ cls_cnt = Q.cls_cnt;
  const uint32_t* __restrict__ ord = Q.ord;
  const uint32_t* __restrict__ ikey = Q.ikey;
  const int* __restrict__ dense_j = Q.dense_j;
  const uint4* __restrict__ Mh = Q.Mh;
  const uint4* __restrict__ opH = Q.opH;
  float* __restrict__ L = Q.L;
  float* __restrict__ part = Q.part;
  const int4* __restrict__ offtab = Q.offtab;
  const int bid = (int)blockIdx.x - Q.wg0, nwg = Q.nwg;
  using H = HLayout<P>;
  constexpr int P2 = H::P2, MT = H::MT, KH = H::KH, FRH = H::FRH, NB4 = H::NB4;
  constexpr int NTH = 64 * WPG, TN = 16 * NT * WPG, NVT = (FRH + NTH - 1) / NTH;
  __shared__ uint4 As[(DB ? 2 : 1) * FRH];
  __shared__ int s_class, s_start, s_n;
  const int tid = threadIdx.x;
  // tile, class and split of this workgroup: the FP32 kernel's mapping
  // (workgroup b runs on XCD b % 8; with region set, XCD x owns a block of
  // tiles of every class, so siblings share source multipoles in one L2)
  int cls0, split, tile;
  if (region) {
    const int xcd = bid & 7, item = bid >> 3, tpr = (nwg >> 6) / nsplit;
    cls0 = item & 7;
    split = (item >> 3) % nsplit;
    tile = xcd * tpr + (item >> 3) / nsplit;
  } else {
    const int item = bid >> 3, tpc = (nwg >> 3) / nsplit;
    cls0 = bid & 7;
    split = item / tpc;
    tile = item % tpc;
  }
  if (tid == 0) {
    int start = 0;
    for (int c = 0; c < cls0; c++) start += cls_cnt[c];
    const int cnt = cls_cnt[cls0], nt = (cnt + TN - 1) / TN;
    s_class = tile < nt ? cls0 : -1;
    s_start = start + tile * TN;
    s_n = tile < nt ? min(TN, cnt - tile * TN) : 0;
  }
  __syncthreads();
  const int cls = s_class, ntile = s_n;
  if (cls < 0) return;
  const int w = tid >> 6, l = tid & 63, j = l & 15, kq = l >> 4;
  const int g = 1 << lev;
  int tord[NT], tx[NT], ty[NT], tz[NT];
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const int t = (w * NT + n) * 16 + j;
    tord[n] = -1;
    tx[n] = ty[n] = tz[n] = 0;
    if (t < ntile) {
      tord[n] = (int)ord[s_start + t];
      morton_decode(ikey[tord[n]], lev, tx[n], ty[n], tz[n]);
    }
  }
  const bool wave_on = w * NT * 16 < ntile;
  const int per = (NEBM + nsplit - 1) / nsplit, o0 = split * per, o1 = min(NEBM, o0 + per);
  const float oscale = ldexpf(1.f, -(split_exp(Q.mmax) + Q.sa_e));
  f32x4 acc[NT][MT];
#pragma unroll
  for (int n = 0; n < NT; n++)
#pragma unroll
    for (int mt = 0; mt < MT; mt++) acc[n][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
  unsigned long long mypairs = 0;
  const HSign<P>& HS = c_hsign<P>;
  uint32_t bcur[NT][4 * NB4];
  bool cur_any[NT] = {};
  // source multipoles of offset o_ for the wave's targets (zero where the
  // target has no source box); any_[n]: some target of N-tile n has one
#define FMMX_LOAD_BH(o_, dst_, any_)                                                          \
  {                                                                                           \
    const int4 off_ = offtab[cls * NEBM + (o_)];                                              \
    _Pragma("unroll") for (int n = 0; n < NT; n++) {                                          \
      int s_ = -1;                                                                            \
      if (tord[n] >= 0) {                                                                     \
        const int jx = tx[n] - off_.x, jy = ty[n] - off_.y, jz = tz[n] - off_.z;              \
        if (jx >= 0 && jx < g && jy >= 0 && jy < g && jz >= 0 && jz < g &&                    \
            (gsh < 0 || !((jx ^ tx[n]) >> gsh | (jy ^ ty[n]) >> gsh | (jz ^ tz[n]) >> gsh)))   \
          s_ = dense_j[morton_encode(jx, jy, jz, lev)];                                       \
      }                                                                                       \
      if (s_ >= 0 && kq == 0) mypairs++;                                                      \
      any_[n] = __ballot(s_ >= 0) != 0;                                                       \
      const uint4* ms_ = Mh + ((size_t)(s_ >= 0 ? s_ : Q.nbj) * H::VSH + kq * H::GH) / 8;     \
      _Pragma("unroll") for (int v_ = 0; v_ < NB4; v_++) {                                    \
        const uint4 m4_ = ms_[v_];                                                            \
        dst_[n][4 * v_] = m4_.x;                                                              \
        dst_[n][4 * v_ + 1] = m4_.y;                                                          \
        dst_[n][4 * v_ + 2] = m4_.z;                                                          \
        dst_[n][4 * v_ + 3] = m4_.w;                                                          \
      }                                                                                       \
    }                                                                                         \
  }
  // offset o_'s split operator (fragment e of this thread) into registers,
  // reflected: element (row i, k-position s) negated where
  // T_r(i) T_r(slot(s)) = -1.  Fragment e holds row 16 mt + lane % 16 of lane
  // group lane / 16 at k-step ks, with e / 64 = ks MT + mt the same for a
  // whole wave
#define FMMX_LOAD_AH(o_)                                                                      \
  {                                                                                           \
    const int w_ = offtab[cls * NEBM + (o_)].w, refl_ = w_ >> 16;                             \
    const uint4* src_ = opH + (size_t)(w_ & 0xffff) * FRH;                                    \
    _Pragma("unroll") for (int v = 0; v < NVT; v++) {                                         \
      const int e = tid + NTH * v;                                                            \
      if (e < FRH) {                                                                          \
        uint4 x = src_[e];                                                                    \
        if (refl_) {                                                                          \
          const int f = e >> 6, ks = f / MT, mt = f % MT;                                     \
          uint32_t m = (HS.col[refl_][ks] >> (8 * kq)) & 0xffu;                               \
          if ((HS.row[refl_][mt] >> j) & 1u) m ^= 0xffu;                                      \
          x.x ^= ((m & 1u) << 15) | ((m & 2u) << 30);                                         \
          x.y ^= ((m & 4u) << 13) | ((m & 8u) << 28);                                         \
          x.z ^= ((m & 16u) << 11) | ((m & 32u) << 26);                                       \
          x.w ^= ((m & 64u) << 9) | ((m & 128u) << 24);                                       \
        }                                                                                     \
        areg[v] = x;                                                                          \
      }                                                                                       \
    }                                                                                         \
  }
#define FMMX_STORE_AH(buf_)                                                                   \
  {                                                                                           \
    _Pragma("unroll") for (int v = 0; v < NVT; v++) {                                         \
      const int e = tid + NTH * v;                                                            \
      if (e < FRH) As[(buf_) * FRH + e] = areg[v];                                            \
    }                                                                                         \
  }
  uint4 areg[NVT];
  if (DB && o0 < o1) {
    FMMX_LOAD_AH(o0);
    FMMX_STORE_AH(0);
  }
  if (o0 < o1) FMMX_LOAD_BH(o0, bcur, cur_any);
  if (DB) __syncthreads();
  for (int o = o0; o < o1; o++) {
    const int buf = DB ? (o - o0) & 1 : 0;
    if (!DB) {
      // one buffer: every wave is done with the previous operator
      __syncthreads();
      FMMX_LOAD_AH(o);
      FMMX_STORE_AH(0);
      __syncthreads();
    } else if (o + 1 < o1) {
      // two buffers: offset o+1's operator is in flight during o's products
      FMMX_LOAD_AH(o + 1);
    }
    uint32_t bnext[NT][4 * NB4];
    bool next_any[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) next_any[n] = false;
    if (o + 1 < o1) FMMX_LOAD_BH(o + 1, bnext, next_any);
    bool run = false;
#pragma unroll
    for (int n = 0; n < NT; n++) run = run || cur_any[n];
    if (wave_on && run) {
      // the operator fragments of step ks+1 are read from LDS while the
      // MFMAs of step ks run
      const uint4* Ab = As + buf * FRH;
      h8 af[2][MT];
#pragma unroll
      for (int mt = 0; mt < MT; mt++) af[0][mt] = __builtin_bit_cast(h8, Ab[mt * 64 + l]);
#pragma unroll
      for (int ks = 0; ks < KH; ks++) {
        if (ks + 1 < KH) {
#pragma unroll
          for (int mt = 0; mt < MT; mt++) af[(ks + 1) & 1][mt] = __builtin_bit_cast(h8, Ab[((ks + 1) * MT + mt) * 64 + l]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
          for (int n = 0; n < NT; n++)
            acc[n][mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(af[ks & 1][mt], bfrag<H>(bcur[n], ks), acc[n][mt],
                                                                0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (DB) {
      // buffer buf ^ 1 was last read before the previous barrier
      if (o + 1 < o1) FMMX_STORE_AH(buf ^ 1);
      __syncthreads();
    }
    if (o + 1 < o1) {
#pragma unroll
      for (int n = 0; n < NT; n++) {
        cur_any[n] = next_any[n];
#pragma unroll
        for (int v = 0; v < 4 * NB4; v++) bcur[n][v] = bnext[n][v];
      }
    }
  }
#undef FMMX_LOAD_AH
#undef FMMX_STORE_AH
#undef FMMX_LOAD_BH
#pragma unroll
  for (int n = 0; n < NT; n++) {
    if (tord[n] < 0) continue;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * mt + 4 * kq + r;
        if (i < P2) {
          const float v = acc[n][mt][r] * oscale;
          if (nsplit == 1) {
            float* dst = L + (size_t)tord[n] * P2 + i;
            *dst = init ? v : *dst + v;
          } else {
            part[((size_t)split * nb + tord[n]) * P2 + i] = v;
          }
        }
      }
  }
  if (mypairs) atomicAdd(Q.npairs, mypairs);
}

// Signs of a reflection for the direct-staging kernel, which keeps the
// canonical operator as stored and moves T_r to the operands:
// T_r A T_r B = T_r (A (T_r B)).  bneg[r][g]: bit u set where T_r negates
// input slot g KS + u; rneg[r]: bit i set where T_r negates output row i.
template <int P>
struct HSign2 {
  static constexpr int W = (P * P + 31) / 32;
  uint64_t bneg[8][4];
  uint32_t rneg[8][W];
  constexpr HSign2() : bneg{}, rneg{} {
    constexpr ReflMask<P> RM{};
    using H = HLayout<P>;
    for (int r = 0; r < 8; r++) {
      for (int g = 0; g < 4; g++)
        for (int u = 0; u < H::KS; u++)
          if (g * H::KS + u < H::P2 && RM.neg(r, g * H::KS + u)) bneg[r][g] |= 1ull << u;
      for (int i = 0; i < H::P2; i++)
        if (RM.neg(r, i)) rneg[r][i >> 5] |= 1u << (i & 31);
    }
  }
};
template <int P>
__constant__ HSign2<P> c_hsign2 = HSign2<P>();

// Split M2L with the operators copied global -> LDS by the DMA path
// (global_load_lds_dwordx4: no registers, no ALU), two LDS buffers: offset
// o+1's operator streams in while o's products run.  The offsets of a class
// come sorted by reflection (offtab_s), the operator is the canonical one as
// stored, and the reflection moves to the operands: the B words are negated
// by T_r (one mask per reflection group), and the accumulator holds
// T_r (partial sum) for the current group r, re-signed by T_r T_r' where the
// group changes and by T_r after the last offset.
template <int P, int NT, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_m2l_hd(M2LLaunch A) {
  M2LLev Q = A.lv[0];
#pragma unroll
  for (int k = 1; k < MAXL + 1; k++)
    if (k < A.nlv && (int)blockIdx.x >= A.lv[k].wg0) Q = A.lv[k];
  const int lev = Q.lev, nb = Q.nb, nbj = Q.nbj, nsplit = Q.nsplit, init = Q.init, region = Q.region, gsh = Q.gsh;
  const int* __restrict__ cls_cnt = Q.cls_cnt;
  const uint32_t* __restrict__ ord = Q.ord;
  const uint32_t* __restrict__ ikey = Q.ikey;
  const int* __restrict__ dense_j = Q.dense_j;
  const uint4* __restrict__ Mh = Q.Mh;
  const uint4* __restrict__ opH = Q.opH;
  float* __restrict__ L = Q.L;
  float* __restrict__ part = Q.part;
  const int4* __restrict__ offtab = Q.offtab;
  const int bid = (int)blockIdx.x - Q.wg0, nwg = Q.nwg;
  using H = HLayout<P>;
  constexpr int P2 = H::P2, MT = H::MT, KH = H::KH, FRH = H::FRH, NB4 = H::NB4, NCH = KH * MT;
  constexpr int TN = 16 * NT * WPG;
  __shared__ uint4 As[2 * FRH];
  __shared__ int s_class, s_start, s_n;
  const int tid = threadIdx.x;
  int cls0, split, tile;
  if (region) {
    const int xcd = bid & 7, item = bid >> 3, tpr = (nwg >> 6) / nsplit;
    cls0 = item & 7;
    split = (item >> 3) % nsplit;
    tile = xcd * tpr + (item >> 3) / nsplit;
  } else {
    const int item = bid >> 3, tpc = (nwg >> 3) / nsplit;
    cls0 = bid & 7;
    split = item / tpc;
    tile = item % tpc;
  }
  if (tid == 0) {
    int start = 0;
    for (int c = 0; c < cls0; c++) start += cls_cnt[c];
    const int cnt = cls_cnt[cls0], nt = (cnt + TN - 1) / TN;
    s_class = tile < nt ? cls0 : -1;
    s_start = start + tile * TN;
    s_n = tile < nt ? min(TN, cnt - tile * TN) : 0;
  }
  __syncthreads();
  const int cls = s_class, ntile = s_n;
  if (cls < 0) return;
  // the wave index, known to the compiler as uniform (scalar addressing of
  // the operator copies)
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, j = l & 15, kq = l >> 4;
  const int g = 1 << lev;
  int tord[NT], tx[NT], ty[NT], tz[NT];
  uint32_t tk[NT];  // target keys (Morton: x bits 3b+1, y 3b, z 3b+2)
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const int t = (w * NT + n) * 16 + j;
    tord[n] = -1;
    tx[n] = ty[n] = tz[n] = 0;
    tk[n] = 0u;
    if (t < ntile) {
      tord[n] = (int)ord[s_start + t];
      tk[n] = ikey[tord[n]];
      morton_decode(tk[n], lev, tx[n], ty[n], tz[n]);
    }
  }
  const uint32_t MY = 0x49249249u & ((1u << (3 * lev)) - 1u), MX = MY << 1, MZ = MY << 2;
  const bool wave_on = w * NT * 16 < ntile;
  const int per = (NEBM + nsplit - 1) / nsplit, o0 = split * per, o1 = min(NEBM, o0 + per);
  const float oscale = ldexpf(1.f, -(split_exp(Q.mmax) + Q.sa_e));
  f32x4 acc[NT][MT];
#pragma unroll
  for (int n = 0; n < NT; n++)
#pragma unroll
    for (int mt = 0; mt < MT; mt++) acc[n][mt] = f32x4{0.f, 0.f, 0.f, 0.f};
  unsigned long long mypairs = 0;
  const HSign2<P>& HS = c_hsign2<P>;
  uint32_t bcur[NT][4 * NB4];
  bool cur_any[NT] = {};
#define FMMX_LOAD_BH(o_, dst_, any_)                                                          \
  {                                                                                           \
    const int4 off_ = offtab[cls * NEBM + (o_)];                                              \
    /* the offset's key digits (wave-uniform): the source key is the target  */                \
    /* key minus the offset in dilated-integer arithmetic, per axis          */                \
    const uint32_t dx_ = dil3(abs(off_.x)) << 1, dy_ = dil3(abs(off_.y)), dz_ = dil3(abs(off_.z)) << 2; \
    _Pragma("unroll") for (int n = 0; n < NT; n++) {                                          \
      int s_ = -1;                                                                            \
      if (tord[n] >= 0) {                                                                     \
        const int jx = tx[n] - off_.x, jy = ty[n] - off_.y, jz = tz[n] - off_.z;              \
        if ((unsigned)jx < (unsigned)g && (unsigned)jy < (unsigned)g && (unsigned)jz < (unsigned)g && \
            (gsh < 0 || !((jx ^ tx[n]) >> gsh | (jy ^ ty[n]) >> gsh | (jz ^ tz[n]) >> gsh))) { \
          const uint32_t kx_ = off_.x >= 0 ? ((tk[n] & MX) - dx_) & MX : ((tk[n] | ~MX) + dx_) & MX; \
          const uint32_t ky_ = off_.y >= 0 ? ((tk[n] & MY) - dy_) & MY : ((tk[n] | ~MY) + dy_) & MY; \
          const uint32_t kz_ = off_.z >= 0 ? ((tk[n] & MZ) - dz_) & MZ : ((tk[n] | ~MZ) + dz_) & MZ; \
          s_ = dense_j[kx_ | ky_ | kz_];                                                      \
        }                                                                                     \
      }                                                                                       \
      if (s_ >= 0 && kq == 0) mypairs++;                                                      \
      any_[n] = __ballot(s_ >= 0) != 0;                                                       \
      /* no source box: the level's zero box after the last one */                            \
      const uint4* ms_ = Mh + ((size_t)(s_ >= 0 ? s_ : nbj) * H::VSH + kq * H::GH) / 8;       \
      _Pragma("unroll") for (int v_ = 0; v_ < NB4; v_++) {                                    \
        const uint4 m4_ = ms_[v_];                                                            \
        dst_[n][4 * v_] = m4_.x;                                                              \
        dst_[n][4 * v_ + 1] = m4_.y;                                                          \
        dst_[n][4 * v_ + 2] = m4_.z;                                                          \
        dst_[n][4 * v_ + 3] = m4_.w;                                                          \
      }                                                                                       \
    }                                                                                         \
  }
  // offset o_'s canonical operator into LDS buffer buf_: wave w copies the
  // 1 KB chunks w, w + WPG, ... (lane l: 16 bytes at chunk offset 16 l)
#define FMMX_DMA_A(o_, buf_)                                                                  \
  {                                                                                           \
    const uint4* src_ = opH + (size_t)(offtab[cls * NEBM + (o_)].w & 0xffff) * FRH + l;       \
    for (int c_ = w; c_ < NCH; c_ += WPG)                                                     \
      __builtin_amdgcn_global_load_lds((const void*)(src_ + c_ * 64),                         \
                                       (__attribute__((address_space(3))) void*)(As + (buf_) * FRH + c_ * 64), \
                                       16, 0, 0);                                             \
  }
  if (o0 < o1) {
    FMMX_DMA_A(o0, 0);
    FMMX_LOAD_BH(o0, bcur, cur_any);
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  int rcur = -1;
  uint32_t bm[4 * NB4];
#pragma unroll
  for (int q = 0; q < 4 * NB4; q++) bm[q] = 0u;
  uint32_t bnext[NT][4 * NB4];
  bool next_any[NT] = {};
  // one offset: DMA of o+1's operator, o+1's multipoles into BN_, o's
  // products from BC_ (the two register sets alternate, no copies)
#define FMMX_STEP(o_, BC_, AC_, BN_, AN_)                                                     \
  {                                                                                           \
    const int buf = ((o_) - o0) & 1;                                                          \
    if ((o_) + 1 < o1) {                                                                      \
      FMMX_DMA_A((o_) + 1, buf ^ 1);                                                          \
      FMMX_LOAD_BH((o_) + 1, BN_, AN_);                                                       \
    }                                                                                         \
    const int r = offtab[cls * NEBM + (o_)].w >> 16;                                          \
    if (r != rcur) {                                                                          \
      if (rcur >= 0) {                                                                        \
        _Pragma("unroll") for (int mt = 0; mt < MT; mt++)                                     \
        _Pragma("unroll") for (int q = 0; q < 4; q++) {                                       \
          const int i = 16 * mt + 4 * kq + q;                                                 \
          if (i < P2) {                                                                       \
            const uint32_t f = ((HS.rneg[rcur][i >> 5] ^ HS.rneg[r][i >> 5]) >> (i & 31)) & 1u; \
            _Pragma("unroll") for (int n = 0; n < NT; n++)                                    \
              acc[n][mt][q] = __uint_as_float(__float_as_uint(acc[n][mt][q]) ^ (f << 31));    \
          }                                                                                   \
        }                                                                                     \
      }                                                                                       \
      rcur = r;                                                                               \
      const uint64_t bn = HS.bneg[r][kq];                                                     \
      _Pragma("unroll") for (int q = 0; q < 4 * NB4; q++) {                                   \
        const int u0 = q < H::SEG / 2 ? 2 * q : q < H::SEG ? 2 * (q - H::SEG / 2) : -1;       \
        bm[q] = u0 < 0 ? 0u : (uint32_t)((bn >> u0) & 1u) << 15 | (uint32_t)((bn >> (u0 + 1)) & 1u) << 31; \
      }                                                                                       \
    }                                                                                         \
    bool run_ = false;                                                                        \
    _Pragma("unroll") for (int n = 0; n < NT; n++) run_ = run_ || AC_[n];                     \
    if (wave_on && run_) {                                                                    \
      _Pragma("unroll") for (int n = 0; n < NT; n++)                                          \
      _Pragma("unroll") for (int q = 0; q < 4 * NB4; q++) BC_[n][q] ^= bm[q];                 \
      const uint4* Ab = As + buf * FRH;                                                       \
      h8 af[2][MT];                                                                           \
      _Pragma("unroll") for (int mt = 0; mt < MT; mt++) af[0][mt] = __builtin_bit_cast(h8, Ab[mt * 64 + l]); \
      _Pragma("unroll") for (int ks = 0; ks < KH; ks++) {                                     \
        if (ks + 1 < KH) {                                                                    \
          _Pragma("unroll") for (int mt = 0; mt < MT; mt++)                                   \
            af[(ks + 1) & 1][mt] = __builtin_bit_cast(h8, Ab[((ks + 1) * MT + mt) * 64 + l]); \
        }                                                                                     \
        __builtin_amdgcn_sched_barrier(0);                                                    \
        _Pragma("unroll") for (int mt = 0; mt < MT; mt++)                                     \
        _Pragma("unroll") for (int n = 0; n < NT; n++)                                        \
          acc[n][mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(af[ks & 1][mt], bfrag<H>(BC_[n], ks), acc[n][mt], 0, 0, 0); \
        __builtin_amdgcn_sched_barrier(0);                                                    \
      }                                                                                       \
    }                                                                                         \
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");                                          \
    __syncthreads();                                                                          \
  }
  for (int o = o0; o < o1; o += 2) {
    FMMX_STEP(o, bcur, cur_any, bnext, next_any);
    if (o + 1 < o1) FMMX_STEP(o + 1, bnext, next_any, bcur, cur_any);
  }
#undef FMMX_STEP
#undef FMMX_LOAD_BH
#undef FMMX_DMA_A
#pragma unroll
  for (int n = 0; n < NT; n++) {
    if (tord[n] < 0) continue;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = 16 * mt + 4 * kq + q;
        if (i < P2) {
          float v = acc[n][mt][q] * oscale;
          if (rcur >= 0 && ((HS.rneg[rcur][i >> 5] >> (i & 31)) & 1u)) v = -v;
          if (nsplit == 1) {
            float* dst = L + (size_t)tord[n] * P2 + i;
            *dst = init ? v : *dst + v;
          } else {
            part[((size_t)split * nb + tord[n]) * P2 + i] = v;
          }
        }
      }
  }
  if (mypairs) atomicAdd(Q.npairs, mypairs);
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

// B' fragment of k-step ks in the 32x32x16 layout
template <class W>
__device__ __forceinline__ h8 bfrag_w(const uint32_t* bw, int ks) {
  const int w0 = W::word(8 * ks), w1 = W::word(8 * ks + 2), w2 = W::word(8 * ks + 4), w3 = W::word(8 * ks + 6);
  return __builtin_bit_cast(h8, u32x4{w0 < 0 ? 0u : bw[w0], w1 < 0 ? 0u : bw[w1], w2 < 0 ? 0u : bw[w2],
                                      w3 < 0 ? 0u : bw[w3]});
}

// Split M2L on v_mfma_f32_32x32x16_f16: a wave owns NT groups of 32
// targets (one per lane of a lane group; the two lane groups hold the two
// halves of the input slots) and MT tiles of 32 output rows per group.  The
// operator of every one of the 316 offsets is stored as it acts (no
// reflections to apply), copied global -> LDS by the DMA path into two
// buffers: offset o+1's operator streams in while o's products run.  Each
// operator fragment read from LDS feeds NT products, so NT = 2 halves the
// LDS reads per product (4 waves of 64 targets, one wave per SIMD).  Output
// lane layout: element e of row tile mt is row 32 mt + 8 (e / 4) +
// 4 (lane / 32) + e % 4, column lane % 32.
template <int P, int WPG, int DB, int NT>
__global__ __launch_bounds__(64 * WPG) void k_m2l_w(M2LLaunch A) {
  M2LLev Q = A.lv[0];
#pragma unroll
  for (int k = 1; k < MAXL + 1; k++)
    if (k < A.nlv && (int)blockIdx.x >= A.lv[k].wg0) Q = A.lv[k];
  const int lev = Q.lev, nb = Q.nb, nbj = Q.nbj, nsplit = Q.nsplit, init = Q.init, region = Q.region, gsh = Q.gsh;
  const int* __restrict__ cls_cnt = Q.cls_cnt;
  const uint32_t* __restrict__ ord = Q.ord;
  const uint32_t* __restrict__ ikey = Q.ikey;
  const int* __restrict__ dense_j = Q.dense_j;
  const uint4* __restrict__ Mh = Q.Mh;
  const uint4* __restrict__ opH = Q.opH;
  float* __restrict__ L = Q.L;
  float* __restrict__ part = Q.part;
  const int4* __restrict__ offtab = Q.offtab;
  const int bid = (int)blockIdx.x - Q.wg0, nwg = Q.nwg;
  using W = WLayout<P>;
  constexpr int P2 = W::P2, MT = W::MT, KH = W::KH, FRW = W::FRW, NB4 = W::NB4, NCH = KH * MT;
  constexpr int TN = 32 * NT * WPG;
  // two operator buffers as distinct objects, so that the compiler sees the
  // fragment reads of one and the copy into the other as independent (no
  // wait for the copy before the products)
  __shared__ uint4 As0[FRW], As1[DB ? FRW : 1];
  __shared__ int s_class, s_start, s_n;
  const int tid = threadIdx.x;
  int cls0, split, tile;
  if (region) {
    const int xcd = bid & 7, item = bid >> 3, tpr = (nwg >> 6) / nsplit;
    cls0 = item & 7;
    split = (item >> 3) % nsplit;
    tile = xcd * tpr + (item >> 3) / nsplit;
  } else {
    const int item = bid >> 3, tpc = (nwg >> 3) / nsplit;
    cls0 = bid & 7;
    split = item / tpc;
    tile = item % tpc;
  }
  if (tid == 0) {
    int start = 0;
    for (int c = 0; c < cls0; c++) start += cls_cnt[c];
    const int cnt = cls_cnt[cls0], nt = (cnt + TN - 1) / TN;
    s_class = tile < nt ? cls0 : -1;
    s_start = start + tile * TN;
    s_n = tile < nt ? min(TN, cnt - tile * TN) : 0;
  }
  __syncthreads();
  const int cls = s_class, ntile = s_n;
  if (cls < 0) return;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, j = l & 31, kq = l >> 5;
  const int g = 1 << lev;
  int tord[NT], tx[NT], ty[NT], tz[NT];
  uint32_t tk[NT];
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const int t = (w * NT + n) * 32 + j;
    tord[n] = -1;
    tx[n] = ty[n] = tz[n] = 0;
    tk[n] = 0u;
    if (t < ntile) {
      tord[n] = (int)ord[s_start + t];
      tk[n] = ikey[tord[n]];
      morton_decode(tk[n], lev, tx[n], ty[n], tz[n]);
    }
  }
  const uint32_t MY = 0x49249249u & ((1u << (3 * lev)) - 1u), MX = MY << 1, MZ = MY << 2;
  const bool wave_on = w * NT * 32 < ntile;
  const int per = (NEBM + nsplit - 1) / nsplit, o0 = split * per, o1 = min(NEBM, o0 + per);
  const float oscale = ldexpf(1.f, -(split_exp(Q.mmax) + Q.sa_e));
  f32x16 acc[NT][MT];
#pragma unroll
  for (int n = 0; n < NT; n++)
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[n][mt][e] = 0.f;
  unsigned long long mypairs = 0;
  uint32_t bcur[NT][4 * NB4], bnext[NT][4 * NB4];
  bool cur_any = false, next_any = false;
#define FMMX_LOAD_BW(o_, dst_, any_)                                                            \
  {                                                                                             \
    const int4 off_ = offtab[cls * NEBM + (o_)];                                                \
    const uint32_t dx_ = dil3(abs(off_.x)) << 1, dy_ = dil3(abs(off_.y)), dz_ = dil3(abs(off_.z)) << 2; \
    bool a_ = false;                                                                            \
    _Pragma("unroll") for (int n = 0; n < NT; n++) {                                            \
      int s_ = -1;                                                                              \
      if (tord[n] >= 0) {                                                                       \
        const int jx = tx[n] - off_.x, jy = ty[n] - off_.y, jz = tz[n] - off_.z;                \
        if ((unsigned)jx < (unsigned)g && (unsigned)jy < (unsigned)g && (unsigned)jz < (unsigned)g && \
            (gsh < 0 || !((jx ^ tx[n]) >> gsh | (jy ^ ty[n]) >> gsh | (jz ^ tz[n]) >> gsh))) {  \
          const uint32_t t_ = tk[n];                                                            \
          const uint32_t kx_ = off_.x >= 0 ? ((t_ & MX) - dx_) & MX : ((t_ | ~MX) + dx_) & MX;  \
          const uint32_t ky_ = off_.y >= 0 ? ((t_ & MY) - dy_) & MY : ((t_ | ~MY) + dy_) & MY;  \
          const uint32_t kz_ = off_.z >= 0 ? ((t_ & MZ) - dz_) & MZ : ((t_ | ~MZ) + dz_) & MZ;  \
          s_ = dense_j[kx_ | ky_ | kz_];                                                        \
        }                                                                                       \
      }                                                                                         \
      if (s_ >= 0 && kq == 0) mypairs++;                                                        \
      a_ = a_ || s_ >= 0;                                                                       \
      /* no source box: the level's zero box after the last one */                              \
      const uint4* ms_ = Mh + ((size_t)(s_ >= 0 ? s_ : nbj) * W::VSH + kq * W::GH) / 8;         \
      _Pragma("unroll") for (int v_ = 0; v_ < NB4; v_++) {                                      \
        const uint4 m4_ = ms_[v_];                                                              \
        dst_[n][4 * v_] = m4_.x;                                                                \
        dst_[n][4 * v_ + 1] = m4_.y;                                                            \
        dst_[n][4 * v_ + 2] = m4_.z;                                                            \
        dst_[n][4 * v_ + 3] = m4_.w;                                                            \
      }                                                                                         \
    }                                                                                           \
    any_ = __ballot(a_) != 0;                                                                   \
  }
#define FMMX_DMA_W(o_, buf_)                                                                    \
  {                                                                                             \
    const uint4* src_ = opH + (size_t)(offtab[cls * NEBM + (o_)].w & 0xffff) * FRW + l;         \
    for (int c_ = w; c_ < NCH; c_ += WPG)                                                       \
      __builtin_amdgcn_global_load_lds((const void*)(src_ + c_ * 64),                           \
                                       (__attribute__((address_space(3))) void*)((buf_) + c_ * 64), 16, 0, 0); \
  }
#define FMMX_MFMA_W(Ab_, ks_, BC_)                                                              \
  _Pragma("unroll") for (int n = 0; n < NT; n++) {                                              \
    const h8 bf_ = bfrag_w<W>(BC_[n], ks_);                                                     \
    _Pragma("unroll") for (int mt = 0; mt < MT; mt++)                                           \
      acc[n][mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ab_[mt], bf_, acc[n][mt], 0, 0, 0);   \
  }
#define FMMX_STEP_W(o_, BC_, AC_, BN_, AN_, ACUR_, ANEXT_)                                       \
  {                                                                                             \
    /* the next multipoles are plain loads, in flight during the products; */                    \
    /* the next operator's copy is issued after this step's last read of  */                    \
    /* LDS (a copy into LDS orders every later LDS read behind it)        */                    \
    if ((o_) + 1 < o1) FMMX_LOAD_BW((o_) + 1, BN_, AN_);                                        \
    bool dma_ = (o_) + 1 < o1;                                                                  \
    if (wave_on && AC_) {                                                                       \
      const uint4* Ab = ACUR_;                                                                  \
      h8 af[2][MT];                                                                             \
      _Pragma("unroll") for (int mt = 0; mt < MT; mt++) af[0][mt] = __builtin_bit_cast(h8, Ab[mt * 64 + l]); \
      _Pragma("unroll") for (int ks = 0; ks < KH; ks++) {                                       \
        if (ks + 1 < KH) {                                                                      \
          _Pragma("unroll") for (int mt = 0; mt < MT; mt++)                                     \
            af[(ks + 1) & 1][mt] = __builtin_bit_cast(h8, Ab[((ks + 1) * MT + mt) * 64 + l]);   \
        }                                                                                       \
        if (ks + 2 == KH && dma_) {                                                             \
          FMMX_DMA_W((o_) + 1, ANEXT_);                                                         \
          dma_ = false;                                                                         \
        }                                                                                       \
        __builtin_amdgcn_sched_barrier(0);                                                      \
        FMMX_MFMA_W(af[ks & 1], ks, BC_);                                                       \
        __builtin_amdgcn_sched_barrier(0);                                                      \
      }                                                                                         \
    }                                                                                           \
    if (dma_) FMMX_DMA_W((o_) + 1, ANEXT_);                                                     \
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");                                            \
    __syncthreads();                                                                            \
  }
  if constexpr (DB) {
    if (o0 < o1) {
      FMMX_DMA_W(o0, As0);
      FMMX_LOAD_BW(o0, bcur, cur_any);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    for (int o = o0; o < o1; o += 2) {
      FMMX_STEP_W(o, bcur, cur_any, bnext, next_any, As0, As1);
      if (o + 1 < o1) FMMX_STEP_W(o + 1, bnext, next_any, bcur, cur_any, As1, As0);
    }
  } else {
    // one buffer, two workgroups per CU: one computes while the other waits
    // for its operator copy and multipoles
    for (int o = o0; o < o1; o++) {
      __syncthreads();  // every wave is done with the previous operator
      FMMX_DMA_W(o, As0);
      FMMX_LOAD_BW(o, bcur, cur_any);
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      if (wave_on && cur_any) {
        h8 af[2][MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++) af[0][mt] = __builtin_bit_cast(h8, As0[mt * 64 + l]);
#pragma unroll
        for (int ks = 0; ks < KH; ks++) {
          if (ks + 1 < KH) {
#pragma unroll
            for (int mt = 0; mt < MT; mt++) af[(ks + 1) & 1][mt] = __builtin_bit_cast(h8, As0[((ks + 1) * MT + mt) * 64 + l]);
          }
          __builtin_amdgcn_sched_barrier(0);
          FMMX_MFMA_W(af[ks & 1], ks, bcur);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    (void)bnext;
    (void)next_any;
  }
#undef FMMX_STEP_W
#undef FMMX_MFMA_W
#undef FMMX_DMA_W
#undef FMMX_LOAD_BW
#pragma unroll
  for (int n = 0; n < NT; n++) {
    if (tord[n] < 0) continue;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int i = 32 * mt + 8 * (e / 4) + 4 * kq + e % 4;
        if (i < P2) {
          const float v = acc[n][mt][e] * oscale;
          if (nsplit == 1) {
            float* dst = L + (size_t)tord[n] * P2 + i;
            *dst = init ? v : *dst + v;
          } else {
            part[((size_t)split * nb + tord[n]) * P2 + i] = v;
          }
        }
      }
  }
  if (mypairs) atomicAdd(Q.npairs, mypairs);
}

template <typename F>
int dispatch_split_p(int p, F&& f) {
  switch (p) {
#define FMMX_CASE(P) \
  case P:            \
    return f(std::integral_constant<int, P>{});
    FMMX_CASE(6) FMMX_CASE(7) FMMX_CASE(8) FMMX_CASE(9) FMMX_CASE(10) FMMX_CASE(11) FMMX_CASE(12)
#undef FMMX_CASE
  }
  return fail(FMMX_EUNSUPPORTED, "the split FP16 M2L covers p = 6..12");
}

}  // namespace

bool m2l_split_supported(int p) { return p >= 6 && p <= 12; }

int launch_msplit(int p, const MSplitLevels& S, hipStream_t stream) {
  const long nt = 4 * S.off[S.nl];
  if (nt == 0) return FMMX_OK;
  const int nblk = (int)((nt + MS_T - 1) / MS_T);
  return dispatch_split_p(p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    k_msplit_zero<<<1, 32, 0, stream>>>(S);
    k_msplit_max<P><<<nblk, MS_T, 0, stream>>>(S);
    if (S.off[S.nl] * 4 * (HLayout<P>::GH / 2) >= (1L << 32))
      return fail(FMMX_EUNSUPPORTED, "too many boxes for the multipole split");
    if (S.wide) {
      const long nw = S.off[S.nl] * 2 * (WLayout<P>::GH / 2);
      k_msplit<P, WLayout<P>, 2><<<(int)((nw + MS_T - 1) / MS_T), MS_T, 0, stream>>>(S);
    } else {
      const long nw = S.off[S.nl] * 4 * (HLayout<P>::GH / 2);
      k_msplit<P, HLayout<P>, 4><<<(int)((nw + MS_T - 1) / MS_T), MS_T, 0, stream>>>(S);
    }
    return hipGetLastError() == hipSuccess ? FMMX_OK : fail(FMMX_EDEVICE, "multipole split launch failed");
  });
}

int launch_m2l_split(int p, const M2LLaunch& A, int nwg, int nt, int wpg, int db, hipStream_t stream) {
  return dispatch_split_p(p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    if (db && split_db_fits(P)) {
      // operators by DMA into two LDS buffers, reflections on the operands
      if constexpr ((size_t)2 * HLayout<P>::FRH * 16 <= 160 * 1024 - 64) {
        if (nt == 2) k_m2l_hd<P, 2, 4><<<nwg, 256, 0, stream>>>(A);
        else k_m2l_hd<P, 1, 8><<<nwg, 512, 0, stream>>>(A);
      }
    } else if (nt == 2) {
      k_m2l_h<P, 2, 4, 0><<<nwg, 256, 0, stream>>>(A);
    } else if (wpg == 4) {
      k_m2l_h<P, 1, 4, 0><<<nwg, 256, 0, stream>>>(A);
    } else {
      k_m2l_h<P, 1, 8, 0><<<nwg, 512, 0, stream>>>(A);
    }
    return hipGetLastError() == hipSuccess ? FMMX_OK : fail(FMMX_EDEVICE, "split M2L launch failed");
  });
}

int launch_m2l_wide(int p, const M2LLaunch& A, int nwg, int wpg, int nt, hipStream_t stream) {
  return dispatch_split_p(p, [&](auto pc) -> int {
    constexpr int P = decltype(pc)::value;
    if constexpr ((size_t)2 * WLayout<P>::FRW * 16 <= 160 * 1024 - 64) {
      if (nt == 4) k_m2l_w<P, 4, 1, 2><<<nwg, 256, 0, stream>>>(A);
      else if (wpg == 4) k_m2l_w<P, 4, 0, 1><<<nwg, 256, 0, stream>>>(A);
      else k_m2l_w<P, 8, 1, 1><<<nwg, 512, 0, stream>>>(A);
      return hipGetLastError() == hipSuccess ? FMMX_OK : fail(FMMX_EDEVICE, "split M2L launch failed");
    }
    return fail(FMMX_EUNSUPPORTED, "the 32x32x16 split M2L needs p <= 10");
  });
}

}  // namespace fmmx
