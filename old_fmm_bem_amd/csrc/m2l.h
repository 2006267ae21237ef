// m2l.h -- expansion layouts, reflection signs and launch descriptors shared
// by the M2L kernels (kernels.hip: FP32 and FP64 forms; m2l_split.hip: the
// error-compensated FP16 form).
#pragma once

#include "engine.h"

namespace fmmx {

// storage of a multipole slot c in the padded layout read by the M2L B
// operand: four segments of KSP = roundup4(ceil(P2/4)) floats
template <int P>
struct MLayout {
  static constexpr int P2 = P * P, KS = (P2 + 3) / 4, KSP = (KS + 3) / 4 * 4, VS = 4 * KSP;
  static __host__ __device__ constexpr int at(int c) { return (c / KS) * KSP + c % KS; }
};

// Reflection of an M2L offset through the coordinate planes: with reflection
// bits r = (x, y, z), M2L(reflected offset) = T_r M2L(offset) T_r for the
// diagonal +-1 matrix T_r on the real slots: x flips Re by (-1)^m and Im by
// (-1)^(m+1); y flips Im; z flips both by (-1)^(n+m).  Bit c of mask[r] is set
// where T_r has -1.
template <int P>
struct ReflMask {
  static constexpr int W = (P * P + 31) / 32;
  uint32_t mask[8][W];
  constexpr ReflMask() : mask{} {
    for (int r = 0; r < 8; r++) {
      int c = 0;
      for (int im = 0; im < 2; im++)
        for (int n = 0; n < P; n++)
          for (int m = im; m <= n; m++, c++) {
            int neg = 0;
            if (r & 1) neg ^= (m & 1) ^ im;
            if (r & 2) neg ^= im;
            if (r & 4) neg ^= (n + m) & 1;
            if (neg) mask[r][c >> 5] |= 1u << (c & 31);
          }
    }
  }
  __host__ __device__ constexpr bool neg(int r, int c) const { return (mask[r][c >> 5] >> (c & 31)) & 1u; }
};

// Error-compensated FP16 form of the M2L product (v_mfma_f32_16x16x32_f16):
// an FP32 value x is carried as hi = fp16(x) and lo = fp16(x - hi), and
// A B = Ahi Bhi + Alo Bhi + Ahi Blo up to the dropped Alo Blo term (relative
// size 2^-22), with FP32 accumulation.  The three products are one GEMM over
// a K dimension three times as long: lane group g (lanes 16g..16g+15) of an
// MFMA owns the KS input slots g KS .. g KS + KS - 1 and runs them through 8 KH
// k-positions as three segments of SEG = KS rounded up to even:
//   k-position s = seg SEG + u:  A' = (Ahi, Alo, Ahi)[seg],  B' = (hi, hi, lo)[seg]
// at input slot g KS + u (zero for u >= KS, slots >= P2 and seg >= 3).  The
// multipoles of a box are stored per lane group as hi[SEG], lo[SEG], padding
// (GH halfs), so a lane loads its B' words with NB4 16-byte loads and the
// even SEG keeps every pair of k-positions inside one 32-bit word.
// Operators are scaled by 2^sa_e and multipoles by 2^sb_e (powers of two,
// exact) so that their largest magnitude sits just below 2^14: hi stays in
// range and lo above the FP16 subnormals for all but negligible entries.
template <int P>
struct HLayout {
  static constexpr int P2 = P * P, MT = (P2 + 15) / 16, KS = (P2 + 3) / 4, SEG = KS + (KS & 1);
  static constexpr int KH = (3 * SEG + 7) / 8;         // k-steps of 32 (8 per lane group)
  static constexpr int GH = (2 * SEG + 7) / 8 * 8;     // halfs per lane group in a stored box
  static constexpr int VSH = 4 * GH;                   // halfs per stored box
  static constexpr int NB4 = GH / 8;                   // 16-byte loads per lane group
  static constexpr int FRH = KH * MT * 64;             // 16-byte operator fragments
  // 32-bit word of a lane's loaded box holding k-positions s, s+1 (s even):
  // hi, hi, lo segments, then the zero padding words, else -1 (constant zero)
  static __host__ __device__ constexpr int word(int s) {
    return s / SEG < 2 ? (s % SEG) / 2
                       : s / SEG == 2 ? SEG / 2 + (s % SEG) / 2 : (SEG + (s - 3 * SEG) / 2 < GH / 2 ? SEG + (s - 3 * SEG) / 2 : -1);
  }
  // input slot of k-position s of lane group g and its segment (-1: none)
  static __host__ __device__ constexpr int slot(int g, int s) {
    return (s / SEG < 3 && s % SEG < KS && g * KS + s % SEG < P2) ? g * KS + s % SEG : -1;
  }
};
// The same split on v_mfma_f32_32x32x16_f16 (two lane groups of 32 lanes,
// 8 k-positions each per step; output tiles of 32 rows x 32 targets): lane
// group g owns the KS = ceil(P2/2) input slots g KS .. g KS + KS - 1.  Stored
// split multipoles: per box two lane groups of GH halfs (hi[SEG], lo[SEG],
// zero padding).
template <int P>
struct WLayout {
  static constexpr int P2 = P * P, MT = (P2 + 31) / 32, KS = (P2 + 1) / 2, SEG = KS + (KS & 1);
  static constexpr int KH = (3 * SEG + 7) / 8;      // k-steps of 16 (8 per lane group)
  static constexpr int GH = (2 * SEG + 7) / 8 * 8;  // halfs per lane group in a stored box
  static constexpr int VSH = 2 * GH;                // halfs per stored box
  static constexpr int NB4 = GH / 8;                // 16-byte loads per lane group
  static constexpr int FRW = KH * MT * 64;          // 16-byte operator fragments
  // word of a lane's loaded box holding k-positions s, s+1 (s even): hi,
  // hi, lo segments, then the zero padding words, else -1 (constant zero)
  static __host__ __device__ constexpr int word(int s) {
    return s / SEG < 2 ? (s % SEG) / 2
                       : s / SEG == 2 ? SEG / 2 + (s % SEG) / 2 : (SEG + (s - 3 * SEG) / 2 < GH / 2 ? SEG + (s - 3 * SEG) / 2 : -1);
  }
};

// largest magnitude of the scaled split operands (2^14)
constexpr int SPLIT_TOP = 14;

// One M2L launch can cover several levels: workgroups [wg0, wg0 + nwg) of
// the grid run level `lev` with that level's tiles, splits and buffers.
// FP16 form: Mh (split multipoles), opH (split operators, fragment order),
// mmax (the level's largest |multipole| as float bits) and sa_e.
struct M2LLev {
  const int* cls_cnt;
  const uint32_t* ord;
  const uint32_t* ikey;
  const int* dense_j;
  const float* M;
  float* L;
  float* part;
  const float* opF;
  const int4* offtab;
  unsigned long long* npairs;
  const uint4* Mh;
  const uint4* opH;
  const unsigned* mmax;
  int lev, nb, nbj, nsplit, init, region, gsh, wg0, nwg, sa_e;
};
struct M2LLaunch {
  M2LLev lv[MAXL + 1];  // ascending wg0
  int nlv;
};

// levels of the multipole split (largest magnitude, then the FP16 pairs):
// boxes [off[k], off[k+1]) of the flattened range belong to level k, the
// last of them a zero box (the split multipoles hold nb + 1 boxes)
struct MSplitLevels {
  const float* M[MAXL + 1];
  uint4* Mh[MAXL + 1];
  unsigned* mmax[MAXL + 1];
  long off[MAXL + 2];
  int nl;
  int wide;  // layout of the 32x32x16 kernel (WLayout), else HLayout
};

// m2l_split.hip
bool m2l_split_supported(int p);
int launch_msplit(int p, const MSplitLevels& S, hipStream_t stream);
// launch shapes: N-tiles per wave nt (1, 2), waves per workgroup wpg (4, 8),
// db: two operator buffers (8 waves, one N-tile; p with split_db_fits)
int launch_m2l_split(int p, const M2LLaunch& A, int nwg, int nt, int wpg, int db, hipStream_t stream);
bool split_db_fits(int p);
// the 32x32x16 form, all 316 operators: nt 2 (32 targets per wave) with 8
// waves and two LDS buffers or 4 waves and one (two workgroups per CU); nt 4
// (64 targets per wave), 4 waves and two buffers; p with split_wide_fits
int launch_m2l_wide(int p, const M2LLaunch& A, int nwg, int wpg, int nt, hipStream_t stream);
bool split_wide_fits(int p);
void split_fragments_wide(const double* kmaj, int count, int p, int sa_e, std::vector<uint4>& out);
// 16-byte fragments of one split operator; halfs of one split box
inline size_t split_frag_count(int p) {
  const int P2 = p * p, MT = (P2 + 15) / 16, KS = (P2 + 3) / 4, SEG = KS + (KS & 1), KH = (3 * SEG + 7) / 8;
  return (size_t)KH * MT * 64;
}
inline size_t split_box_halfs(int p) {
  const int KS = (p * p + 3) / 4, SEG = KS + (KS & 1);
  const int KS2 = (p * p + 1) / 2, SEG2 = KS2 + (KS2 & 1);
  return std::max<size_t>((size_t)4 * ((2 * SEG + 7) / 8 * 8), (size_t)2 * ((2 * SEG2 + 7) / 8 * 8));
}
// operator fragments of the split form from k-major FP64 operators (count
// of them, P2 x P2 each, element (row i, column k) at [k P2 + i]), scaled by
// 2^sa_e; sa_e from the largest magnitude of n values
void split_fragments(const double* ops, int count, int p, int sa_e, std::vector<uint4>& out);
int split_scale_exp(const double* ops, size_t n);

}  // namespace fmmx
