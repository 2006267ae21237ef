// tree.hip -- GPU tree build: Morton keys (fmm/boxn.cxx), stable radix sort
// (the reference's counting sort, sort/csort.cxx, is stable; so is an LSD
// radix sort), SoA gathers (sort/sortvar.cxx), per-level box lists
// (fmm/boxdatai.cxx, boxdataj.cxx, boxm2mi.cxx, boxm2mj.cxx), and the
// interaction lists of fmm/ijbox.cxx in the reference's loop order.
#include <cmath>
#include <mutex>

#include "engine.h"
#include "tree_util.h"

namespace fmmx {

namespace {

constexpr int TPB = 256;
inline int nblk(long n) { return (int)((n + TPB - 1) / TPB); }

// fmm/boxn.cxx:12-20: (x - xmin) in float, times the scale kl/rd (a float
// quotient widened to double, fmm/boxn.cxx:9), truncate.
// The reference clamps only the upper face.  A coordinate outside the closed
// cube [min, min + rd] (or not finite) sets *bad, which the host turns into
// FMMX_EINVAL; its key is clamped to the nearest box so that every index
// stays in range.
__global__ void k_keys(int n, const float* __restrict__ x, const float* __restrict__ y,
                       const float* __restrict__ z, float xmin, float ymin, float zmin, float rd, double rbinv,
                       int lmax, uint32_t* __restrict__ key, uint32_t* __restrict__ idx,
                       float4* __restrict__ xyz, int* __restrict__ bad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = x[i], py = y[i], pz = z[i];
  // the coordinates packed as well (when xyz is set): the gather into key
  // order then reads one 16-byte record per point instead of three words
  if (xyz) xyz[i] = make_float4(px, py, pz, 0.f);
  const float d0 = px - xmin, d1 = py - ymin, d2 = pz - zmin;
  // NaN fails every comparison
  if (!(d0 >= 0.f && d0 <= rd && d1 >= 0.f && d1 <= rd && d2 >= 0.f && d2 <= rd)) atomicOr(bad, 1);
  int kl = 1 << lmax;
  int f0 = (int)((double)d0 * rbinv);
  int f1 = (int)((double)d1 * rbinv);
  int f2 = (int)((double)d2 * rbinv);
  f0 = min(max(f0, 0), kl - 1);
  f1 = min(max(f1, 0), kl - 1);
  f2 = min(max(f2, 0), kl - 1);
  key[i] = morton_encode(f0, f1, f2, lmax);
  if (idx) idx[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------
// Stable LSD radix sort of (key, index) pairs, 8-bit digits: the GPU form of
// the reference's counting sort (sort/csort.cxx:5-17: counts per key, prefix
// sums, placement in input order).  Per pass: digit counts per tile of
// RS_TILE keys (digit-major, tile-minor), the prefix over tiles of every
// digit (one workgroup per digit, also writing the digit's total), then a
// stable scatter: within a tile each wave ranks its keys chunk by chunk (64
// keys, same-digit lanes found by 8 ballots), the waves' counts are
// prefix-summed in wave order, the digit totals give each digit's start,
// and tile, wave and lane order give the positions, so equal keys keep
// their input order.
constexpr int RS_T = 256, RS_CH = 4, RS_TILE = RS_T * RS_CH;  // keys per tile
constexpr int RS_SCAN_T = 1024;

inline int rs_tiles(long n) { return (int)((n + RS_TILE - 1) / RS_TILE); }
inline int rs_passes(int bits) { return bits <= 0 ? 1 : (bits + 7) / 8; }
inline long rs_count_words(long n) { return 256L * (rs_tiles(n) + 1); }

// digit counts of one tile: counts[d * ntile + tile]
__global__ __launch_bounds__(RS_T) void k_rs_hist(int n, const uint32_t* __restrict__ keys, int shift, int ntile,
                                                  uint32_t* __restrict__ counts) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const long base = (long)blockIdx.x * RS_TILE;
#pragma unroll
  for (int c = 0; c < RS_CH; c++) {
    const long i = base + c * RS_T + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[(size_t)threadIdx.x * ntile + blockIdx.x] = h[threadIdx.x];
}

// block-wide exclusive scan of one value per thread (RS_SCAN_T threads);
// *all receives the block total
__device__ __forceinline__ uint32_t rs_block_scan(uint32_t v, uint32_t* wsum, uint32_t* all) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  uint32_t before = 0u, tot = 0u;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
    if (w < wv) before += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *all = tot;
  return before + inc - v;
}

// prefix over tiles of digit blockIdx.x, in place; total[d] = the digit's count
__global__ __launch_bounds__(RS_SCAN_T) void k_rs_rowscan(int ntile, uint32_t* __restrict__ counts,
                                                          uint32_t* __restrict__ total) {
  __shared__ uint32_t wsum[RS_SCAN_T / 64];
  uint32_t* row = counts + (size_t)blockIdx.x * ntile;
  uint32_t carry = 0u;
  for (int r0 = 0; r0 < ntile; r0 += RS_SCAN_T) {
    const int i = r0 + threadIdx.x;
    const uint32_t v = i < ntile ? row[i] : 0u;
    uint32_t all;
    const uint32_t ex = rs_block_scan(v, wsum, &all);
    if (i < ntile) row[i] = carry + ex;
    carry += all;
  }
  if (threadIdx.x == 0) total[blockIdx.x] = carry;
}

// stable scatter of one pass (vin NULL: the values are the input positions)
__global__ __launch_bounds__(RS_T) void k_rs_scatter(int n, const uint32_t* __restrict__ kin,
                                                     const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
                                                     uint32_t* __restrict__ vout, int shift, int ntile,
                                                     const uint32_t* __restrict__ offs,
                                                     const uint32_t* __restrict__ total) {
  __shared__ uint32_t wc[RS_T / 64][256];
  __shared__ uint32_t wsum[RS_T / 64];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  for (int e = t; e < (RS_T / 64) * 256; e += RS_T) (&wc[0][0])[e] = 0u;
  // start of each digit: exclusive scan of the digit totals (RS_T == 256)
  uint32_t dall;
  const uint32_t dstart = rs_block_scan(total[t], wsum, &dall);
  const long base = (long)blockIdx.x * RS_TILE + (long)w * (RS_CH * 64);
  const uint64_t below = (1ull << l) - 1ull;
  uint32_t k[RS_CH], v[RS_CH], rk[RS_CH];
#pragma unroll
  for (int c = 0; c < RS_CH; c++) {
    const long i = base + c * 64 + l;
    const bool ok = i < n;
    k[c] = ok ? kin[i] : 0u;
    v[c] = ok ? (vin ? vin[i] : (uint32_t)i) : 0u;
    const uint32_t d = (k[c] >> shift) & 255u;
    // lanes of this chunk holding the same digit
    uint64_t m = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t bb = __ballot((d >> b) & 1u);
      m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t prev = wc[w][d];
    rk[c] = prev + (uint32_t)__popcll(m & below);
    // the highest such lane advances the wave's count of the digit
    if (ok && (m >> l) == 1ull) wc[w][d] = prev + (uint32_t)__popcll(m);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // wave offsets within the tile plus the tile's and the digit's starts
  {
    const int d = t;
    uint32_t run = dstart + offs[(size_t)d * ntile + blockIdx.x];
#pragma unroll
    for (int ww = 0; ww < RS_T / 64; ww++) {
      const uint32_t c = wc[ww][d];
      wc[ww][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < RS_CH; c++) {
    const long i = base + c * 64 + l;
    if (i < n) {
      const uint32_t pos = wc[w][(k[c] >> shift) & 255u] + rk[c];
      kout[pos] = k[c];
      vout[pos] = v[c];
    }
  }
}

// packed coordinates into key order, as three columns
__global__ void k_gather_xyz(int n, const uint32_t* __restrict__ perm, const float4* __restrict__ xyz,
                             float* __restrict__ b0, float* __restrict__ b1, float* __restrict__ b2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = xyz[perm[i]];
  b0[i] = v.x;
  b1[i] = v.y;
  b2[i] = v.z;
}

// three coordinate (or normal) columns into key order in one pass
__global__ void k_gather3(int n, const uint32_t* __restrict__ perm, const float* __restrict__ a0,
                          const float* __restrict__ a1, const float* __restrict__ a2, float* __restrict__ b0,
                          float* __restrict__ b1, float* __restrict__ b2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  b0[i] = a0[j];
  b1[i] = a1[j];
  b2[i] = a2[j];
}

// Boxes of every level in two passes over the sorted keys (fmm/boxdatai.cxx:
// 10-34, boxm2mi.cxx:6-24: a box of level l starts where the leaf keys
// shifted by 3 (lmax - l) change).  Keys are sorted, so a point that opens a
// box at level l opens one at every deeper level too: per point the coarsest
// such level lf(i) says everything.  Tiles of BX_TILE points: the first pass
// counts the boxes each tile opens per level, the second sums the counts of
// the earlier tiles, scans its own and writes per box: key, first particle,
// parent (the level l-1 box holding that particle), first child of the
// parent (children of a parent are consecutive, fmm/boxm2mj.cxx:267-275),
// leaf last particle, the dense key -> box table (when given), and the
// level totals with the first-child sentinels.
constexpr int BX_T = 256, BX_PT = 8, BX_TILE = BX_T * BX_PT;

struct BoxOut {
  uint32_t* bkey[MAXL + 1];
  int *bfirst[MAXL + 1], *parent[MAXL + 1], *cfirst[MAXL + 1], *dense[MAXL + 1];
  int* blast;
  int* cnt;
  int* clscnt;  // boxes per level and parity class, [8 l + class] (nullable)
};
// per tile of points: [0, 7) boxes opened at levels 2..8, [8 + 8 (l - 2) + c)
// of them in parity class c
constexpr int BX_REC = 64;
constexpr int BX_CG = 16;  // tiles per block in the class totals

// parity class of a box, (x & 1) | (y & 1) << 1 | (z & 1) << 2, from its key
__device__ __forceinline__ int box_class(uint32_t k) { return ((k >> 1) & 1) | ((k & 1) << 1) | (k & 4); }

// coarsest level >= 2 at which point i opens a box (lmax + 1: none)
__device__ __forceinline__ int open_level(const uint32_t* __restrict__ keys, int i, int lmax) {
  if (i == 0) return 2;
  const uint32_t d = keys[i] ^ keys[i - 1];
  if (!d) return lmax + 1;
  // highest differing bit -> level: bits 3 (lmax - l) .. 3 (lmax - l) + 2 belong to level l
  const int hb = 31 - __clz(d);
  return max(2, lmax - hb / 3);
}

// block-wide sums (red = true) or exclusive scans of the per-thread counts
// v[0..6] (levels 2..8) over BX_T threads; tot receives the block totals
__device__ __forceinline__ void bx_block(int v[7], int tot[7], bool red) {
  __shared__ int ws[BX_T / 64][7];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int inc[7];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    inc[k] = v[k];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int u = __shfl_up(inc[k], d, 64);
      if (lane >= d) inc[k] += u;
    }
  }
  __syncthreads();  // ws may still be read by an earlier call
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < 7; k++) ws[wv][k] = inc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 7; k++) {
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < BX_T / 64; w++) {
      if (w < wv) before += ws[w][k];
      all += ws[w][k];
    }
    tot[k] = all;
    if (!red) v[k] = before + inc[k] - v[k];
  }
}

__global__ __launch_bounds__(BX_T) void k_box_count(int n, const uint32_t* __restrict__ keys, int lmax,
                                                    int* __restrict__ tcnt, int* __restrict__ clscnt) {
  __shared__ int scl[7 * 8];
  if (threadIdx.x < 7 * 8) scl[threadIdx.x] = 0;
  // the class totals, which the next kernel accumulates
  if (clscnt && blockIdx.x == 0 && threadIdx.x < 8 * (MAXL + 1)) clscnt[threadIdx.x] = 0;
  __syncthreads();
  int c[7] = {0, 0, 0, 0, 0, 0, 0}, tot[7];
  const long i0 = (long)blockIdx.x * BX_TILE + (long)threadIdx.x * BX_PT;
#pragma unroll
  for (int k = 0; k < BX_PT; k++) {
    const long i = i0 + k;
    if (i < n) {
      const int lf = open_level(keys, (int)i, lmax);
      const uint32_t key = keys[i];
#pragma unroll
      for (int l = 2; l <= MAXL; l++)
        if (l >= lf && l <= lmax) {
          c[l - 2]++;
          atomicAdd(&scl[(l - 2) * 8 + box_class(key >> (3 * (lmax - l)))], 1);
        }
    }
  }
  bx_block(c, tot, true);  // synchronises the block: scl is complete
  int* rec = tcnt + (size_t)blockIdx.x * BX_REC;
  if (threadIdx.x < 7) rec[threadIdx.x] = tot[threadIdx.x];
  if (threadIdx.x < 7 * 8) rec[8 + threadIdx.x] = scl[threadIdx.x];
}

__global__ __launch_bounds__(BX_T) void k_box_place(int n, const uint32_t* __restrict__ keys, int lmax,
                                                    const int* __restrict__ tcnt, BoxOut o) {
  // class totals: block k adds the class counts of tiles BX_CG k ..
  // BX_CG k + BX_CG - 1 (written by the counting pass) to the totals
  if (o.clscnt && (long)blockIdx.x * BX_CG < gridDim.x) {
    const int e = threadIdx.x % BX_REC, r = threadIdx.x / BX_REC;  // 4 rows of 64 columns
    int sum = 0;
#pragma unroll
    for (int u = 0; u < BX_CG / (BX_T / BX_REC); u++) {
      const long t = (long)blockIdx.x * BX_CG + u * (BX_T / BX_REC) + r;
      if (t < gridDim.x) sum += tcnt[t * BX_REC + e];
    }
    const int l = 2 + (e - 8) / 8;
    if (e >= 8 && l <= lmax && sum) atomicAdd(&o.clscnt[8 * l + (e - 8) % 8], sum);
  }
  // boxes opened by the earlier tiles, per level
  int pre[7] = {0, 0, 0, 0, 0, 0, 0}, tot[7];
  for (int t = threadIdx.x; t < (int)blockIdx.x; t += BX_T)
#pragma unroll
    for (int k = 0; k < 7; k++) pre[k] += tcnt[(size_t)t * BX_REC + k];
  bx_block(pre, tot, true);
  int run[7], c[7] = {0, 0, 0, 0, 0, 0, 0}, ctot[7];
  const long i0 = (long)blockIdx.x * BX_TILE + (long)threadIdx.x * BX_PT;
  int lfs[BX_PT];
#pragma unroll
  for (int k = 0; k < BX_PT; k++) {
    const long i = i0 + k;
    lfs[k] = i < n ? open_level(keys, (int)i, lmax) : MAXL + 2;
#pragma unroll
    for (int l = 2; l <= MAXL; l++) c[l - 2] += (l >= lfs[k] && l <= lmax) ? 1 : 0;
  }
  bx_block(c, ctot, false);
#pragma unroll
  for (int k = 0; k < 7; k++) run[k] = tot[k] + c[k];  // boxes before this thread's first point
  // levels unrolled (run[] stays in registers); level l is live when
  // lf <= l <= lmax
#pragma unroll
  for (int k = 0; k < BX_PT; k++) {
    const long i = i0 + k;
    if (i >= n) break;
    const int lf = lfs[k];
    const uint32_t key = keys[i];
#pragma unroll
    for (int l = 2; l <= MAXL; l++) {
      if (l < lf || l > lmax) continue;
      const int b = run[l - 2]++;
      const uint32_t bk = key >> (3 * (lmax - l));
      o.bkey[l][b] = bk;
      o.bfirst[l][b] = (int)i;
      if (o.dense[l]) o.dense[l][bk] = b;
      if (l >= 3) {
        const int par = run[l - 3] - 1;  // the level l-1 box holding point i
        o.parent[l][b] = par;
        if (l > lf) o.cfirst[l - 1][par] = b;  // point i opens the parent too: first child
      }
      if (l == lmax && b > 0) o.blast[b - 1] = (int)i - 1;
    }
    if (i == n - 1) {
#pragma unroll
      for (int l = 2; l <= MAXL; l++) {
        if (l > lmax) continue;
        o.cnt[l] = run[l - 2];
        if (l >= 3) o.cfirst[l - 1][run[l - 3]] = run[l - 2];
        if (l == lmax) o.blast[run[l - 2] - 1] = n - 1;
      }
    }
  }
}

// dense key -> box tables of every level filled with -1 in one launch (the
// box kernel then enters the boxes)
struct DenseLevels {
  int* dense[MAXL + 1];
  long off[MAXL + 2];  // flattened entries of levels 2.. (8^l each)
  int lmax;
};

__global__ void k_dense_fill(DenseLevels D) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= D.off[D.lmax + 1]) return;
  int l = 2;
  while (t >= D.off[l + 1]) l++;
  D.dense[l][t - D.off[l]] = -1;
}

__global__ void k_fill_int(long n, int v, int* __restrict__ a) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// fmm/ijbox.cxx:322-345 (neib 0): jx -> jy -> jz over the 3x3x3 block, empty
// and out-of-grid boxes skipped.  The reference clips to the bounding box of
// the source boxes (:312-321); positions outside it hold no source box, so
// the lists are identical.  Also counts particle pairs for the statistics.
// 32 lanes per target box, lane k < 27 for the neighbour position k in the
// reference's loop order (k = 9 (dx + 1) + 3 (dy + 1) + dz + 1); the boxes
// found are compacted in lane order.
__global__ void k_near_list(const int* __restrict__ nbi_d, const uint32_t* __restrict__ bkey, int lev,
                            const int* __restrict__ dense_j, const int* __restrict__ ifirst,
                            const int* __restrict__ ilast, const int* __restrict__ jfirst,
                            const int* __restrict__ jlast, int* __restrict__ nij, int* __restrict__ neij,
                            unsigned long long* __restrict__ pairs, int gsh) {
  // the grid covers an upper bound of the target boxes; their count is on the device
  const int nbi = *nbi_d;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = (int)(t >> 5), k = (int)(t & 31), half = (int)(threadIdx.x & 32);
  int jj = -1;
  long ni = 0;
  if (b < nbi && k < 27) {
    int ix, iy, iz;
    const int g = 1 << lev;
    morton_decode(bkey[b], lev, ix, iy, iz);
    const int jx = ix + k / 9 - 1, jy = iy + (k / 3) % 3 - 1, jz = iz + k % 3 - 1;
    const bool in = (unsigned)jx < (unsigned)g && (unsigned)jy < (unsigned)g && (unsigned)jz < (unsigned)g &&
                    // groups: boxes in different cells of the group level do not interact
                    !(gsh >= 0 && ((jx ^ ix) >> gsh | (jy ^ iy) >> gsh | (jz ^ iz) >> gsh));
    if (in) jj = dense_j[morton_encode(jx, jy, jz, lev)];
    if (jj >= 0) ni = ilast[b] - ifirst[b] + 1;
  }
  const unsigned long long found = __ballot(jj >= 0);
  const uint32_t mine = (uint32_t)(found >> half);  // this box's 32 lanes
  const int pos = __popc(mine & ((1u << k) - 1u));
  if (jj >= 0) neij[(size_t)b * NNEAR + pos] = jj;
  if (k == 0 && b < nbi) nij[b] = __popc(mine);
  // pair count: one partial sum per workgroup (pairs[blockIdx.x]; a single
  // counter would take same-address atomics from every wave, which serialise)
  __shared__ long wsum[4];
  long npair = jj >= 0 ? ni * (jlast[jj] - jfirst[jj] + 1) : 0;
  for (int o = 32; o > 0; o >>= 1) npair += __shfl_xor(npair, o);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = npair;
  __syncthreads();
  if (threadIdx.x == 0) pairs[blockIdx.x] = (unsigned long long)(wsum[0] + wsum[1] + wsum[2] + wsum[3]);
}

// fmm/ijbox.cxx:346-365 (neib 1, used at level 2) and :366-401 (neib 2)
__global__ void k_far_list(int nbi, const uint32_t* __restrict__ ikey, int nbj, const uint32_t* __restrict__ jkey,
                           int lev, const int* __restrict__ dense_j, int* __restrict__ nij,
                           int* __restrict__ neij) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbi) return;
  int ix, iy, iz, g = 1 << lev, c = 0;
  morton_decode(ikey[b], lev, ix, iy, iz);
  int* row = neij + (long)b * NEBM;
  if (lev == 2) {
    for (int jj = 0; jj < nbj; jj++) {
      int jx, jy, jz;
      morton_decode(jkey[jj], lev, jx, jy, jz);
      if (jx < ix - 1 || ix + 1 < jx || jy < iy - 1 || iy + 1 < jy || jz < iz - 1 || iz + 1 < jz) row[c++] = jj;
    }
  } else {
    int ixp = (ix + 2) / 2, iyp = (iy + 2) / 2, izp = (iz + 2) / 2;
    for (int jxp = ixp - 1; jxp <= ixp + 1; jxp++)
      for (int jyp = iyp - 1; jyp <= iyp + 1; jyp++)
        for (int jzp = izp - 1; jzp <= izp + 1; jzp++)
          for (int jx = max(2 * jxp - 2, 0); jx <= min(2 * jxp - 1, g - 1); jx++)
            for (int jy = max(2 * jyp - 2, 0); jy <= min(2 * jyp - 1, g - 1); jy++)
              for (int jz = max(2 * jzp - 2, 0); jz <= min(2 * jzp - 1, g - 1); jz++)
                if (jx < ix - 1 || ix + 1 < jx || jy < iy - 1 || iy + 1 < jy || jz < iz - 1 || iz + 1 < jz) {
                  int jj = dense_j[morton_encode(jx, jy, jz, lev)];
                  if (jj >= 0) row[c++] = jj;
                }
  }
  nij[b] = c;
}

// Parity class (x & 1 | (y & 1) << 1 | (z & 1) << 2) of each target box:
// M2L and L2L work is grouped by class, so that every target of a tile shares
// the same 189 offsets (and one L2L octant).  The class order is a stable
// partition of the boxes of each level (box order kept within a class), built
// for all levels in two launches over chunks of CL_CH boxes: class counts per
// chunk, then each box's position from the counts of the chunks before it and
// a block scan inside its chunk.  Counts of a thread's CL_PER boxes are packed
// as 16-bit fields, two classes per 32-bit word.
constexpr int CL_T = 256, CL_PER = 4, CL_CH = CL_T * CL_PER;
struct ClassLevels {
  const uint32_t* key[MAXL + 1];
  uint32_t* ord[MAXL + 1];
  int* cnt[MAXL + 1];
  const int* nb;         // boxes per level (device; the chunks cover an upper bound)
  int chunk0[MAXL + 2];  // first chunk of each level
  int lmin, lmax;
  int* ccnt;  // 8 class counts per chunk
};

__device__ __forceinline__ void class_chunk(const ClassLevels& L, int& lev, int& ch) {
  lev = L.lmin;
  while (lev < L.lmax && (int)blockIdx.x >= L.chunk0[lev + 1]) lev++;
  ch = (int)blockIdx.x - L.chunk0[lev];
}

// the classes of a thread's boxes (-1 past the end) and their packed counts
__device__ __forceinline__ void class_load(const ClassLevels& L, int lev, int ch, int* c, uint32_t* w) {
  const int b0 = ch * CL_CH + (int)threadIdx.x * CL_PER, nb = L.nb[lev];
  w[0] = w[1] = w[2] = w[3] = 0;
#pragma unroll
  for (int k = 0; k < CL_PER; k++) {
    c[k] = -1;
    if (b0 + k < nb) {
      int x, y, z;
      morton_decode(L.key[lev][b0 + k], lev, x, y, z);
      c[k] = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
      w[c[k] >> 1] += 1u << (16 * (c[k] & 1));
    }
  }
}

// block-wide exclusive scan of the packed words; tot receives the block totals
__device__ __forceinline__ void class_scan(uint32_t* w, uint32_t* tot) {
  __shared__ uint32_t wsum[CL_T / 64][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t inc[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    inc[q] = w[q];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t v = __shfl_up(inc[q], d, 64);
      if (lane >= d) inc[q] += v;
    }
  }
  if (lane == 63)
#pragma unroll
    for (int q = 0; q < 4; q++) wsum[wv][q] = inc[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; q++) {
    uint32_t before = 0, all = 0;
    for (int v = 0; v < CL_T / 64; v++) {
      if (v < wv) before += wsum[v][q];
      all += wsum[v][q];
    }
    w[q] = before + inc[q] - w[q];
    tot[q] = all;
  }
}

__global__ __launch_bounds__(CL_T) void k_class_count(ClassLevels L) {
  int lev, ch, c[CL_PER];
  uint32_t w[4], tot[4];
  class_chunk(L, lev, ch);
  class_load(L, lev, ch, c, w);
  class_scan(w, tot);
  if (threadIdx.x < 8) L.ccnt[blockIdx.x * 8 + threadIdx.x] = (int)((tot[threadIdx.x >> 1] >> (16 * (threadIdx.x & 1))) & 0xffffu);
}

__global__ __launch_bounds__(CL_T) void k_class_place(ClassLevels L) {
  int lev, ch, c[CL_PER];
  uint32_t w[4], tot[4];
  class_chunk(L, lev, ch);
  class_load(L, lev, ch, c, w);
  class_scan(w, tot);
  // counts of the level's earlier chunks and of all its chunks, per class
  __shared__ int pre[8], all[8];
  if (threadIdx.x < 8) pre[threadIdx.x] = all[threadIdx.x] = 0;
  __syncthreads();
  const int nch = L.chunk0[lev + 1] - L.chunk0[lev];
  int p8[8] = {0, 0, 0, 0, 0, 0, 0, 0}, a8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = threadIdx.x; j < nch; j += CL_T) {
    const int* cc = L.ccnt + (size_t)(L.chunk0[lev] + j) * 8;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      a8[q] += cc[q];
      if (j < ch) p8[q] += cc[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 8; q++) {
    if (a8[q]) atomicAdd(&all[q], a8[q]);
    if (p8[q]) atomicAdd(&pre[q], p8[q]);
  }
  __syncthreads();
  if (ch == 0 && threadIdx.x < 8) L.cnt[lev][threadIdx.x] = all[threadIdx.x];
  int base[8], s = 0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    base[q] = s + pre[q] + (int)((w[q >> 1] >> (16 * (q & 1))) & 0xffffu);
    s += all[q];
  }
  const int b0 = ch * CL_CH + (int)threadIdx.x * CL_PER;
#pragma unroll
  for (int k = 0; k < CL_PER; k++) {
    if (c[k] < 0) continue;
    int pos = 0;
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (q == c[k]) pos = base[q]++;
    L.ord[lev][pos] = (uint32_t)(b0 + k);
  }
}

// ---------------------------------------------------------------------
// Box statistics of every level (automatic depth): the points per level-8
// cell, counted in Morton order so that the 8 children of a box are 8
// consecutive counts; each level's counts are the sums of its children's.
// Per level l: boxes[l] = boxes holding a point, sumsq[l] = sum of the
// squared counts (the near-field pair count of a box grows with it).
constexpr int OCC_L = 8;
constexpr long OCC_CELLS = 1L << (3 * OCC_L);  // 2^24 counts, 64 MB

__global__ void k_occ_mark(int n, const float* __restrict__ x, const float* __restrict__ y,
                           const float* __restrict__ z, float xmin, float ymin, float zmin, double rbinv,
                           unsigned* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t k = 0xffffffffu;
  if (i < n) {
    const int kl = 1 << OCC_L;
    const int f0 = min(max((int)((double)(x[i] - xmin) * rbinv), 0), kl - 1);
    const int f1 = min(max((int)((double)(y[i] - ymin) * rbinv), 0), kl - 1);
    const int f2 = min(max((int)((double)(z[i] - zmin) * rbinv), 0), kl - 1);
    k = morton_encode(f0, f1, f2, OCC_L);
  }
  // a wave whose points share one cell (clustered or sorted input) adds
  // once: same-address atomics serialise
  const uint32_t k0 = __builtin_amdgcn_readfirstlane(k);
  const unsigned long long same = __ballot(k == k0 && i < n);
  const unsigned long long act = __ballot(i < n);
  if (same == act) {
    if ((threadIdx.x & 63) == 0 && act) atomicAdd(&cnt[k0], (unsigned)__popcll(act));
  } else if (i < n) {
    atomicAdd(&cnt[k], 1u);
  }
}

// one level up: out[b] = sum of in[8b .. 8b+7] for nout boxes, and the
// statistics of the input level, st[0] += occupied boxes, st[1] += sum of
// squared counts (one atomic pair per workgroup)
__global__ __launch_bounds__(256) void k_occ_up(long nout, const unsigned* __restrict__ in,
                                                unsigned* __restrict__ out, unsigned long long* __restrict__ st) {
  __shared__ unsigned long long red[2][4];
  unsigned long long nb = 0, sq = 0;
  for (long b = (long)blockIdx.x * 256 + threadIdx.x; b < nout; b += (long)gridDim.x * 256) {
    const uint4 a = *(const uint4*)(in + 8 * b), c = *(const uint4*)(in + 8 * b + 4);
    const unsigned v[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
    unsigned t = 0;
    for (int k = 0; k < 8; k++) {
      t += v[k];
      nb += v[k] != 0;
      sq += (unsigned long long)v[k] * v[k];
    }
    out[b] = t;
  }
  for (int o = 32; o > 0; o >>= 1) {
    nb += __shfl_xor(nb, o);
    sq += __shfl_xor(sq, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = nb;
    red[1][threadIdx.x >> 6] = sq;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const unsigned long long v = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
    if (v) atomicAdd(&st[threadIdx.x], v);
  }
}

}  // namespace

// Stable radix sort of the keys in kbuf[0] by their low `bits` bits, the
// input positions as values, on stream s.  kbuf[0], kbuf[1] (keys) and
// vbuf[0], vbuf[1] (values) ping-pong, n words each.  The sorted keys and
// values end in
// kbuf[rs_out_slot(bits)] and vbuf[rs_out_slot(bits)].  counts holds
// rs_count_words(n) words: digit counts per tile and the 256 digit totals.
int rs_out_slot(int bits) { return rs_passes(bits) & 1; }

void radix_sort(int n, int bits, uint32_t* const kbuf[2], uint32_t* const vbuf[2], uint32_t* counts,
                hipStream_t s) {
  if (n <= 0) return;
  const int nt = rs_tiles(n), passes = rs_passes(bits);
  uint32_t* total = counts + 256L * nt;
  for (int ps = 0; ps < passes; ps++) {
    const int src = ps & 1, dst = src ^ 1;
    k_rs_hist<<<nt, RS_T, 0, s>>>(n, kbuf[src], 8 * ps, nt, counts);
    k_rs_rowscan<<<256, RS_SCAN_T, 0, s>>>(nt, counts, total);
    k_rs_scatter<<<nt, RS_T, 0, s>>>(n, kbuf[src], ps == 0 ? nullptr : vbuf[src], kbuf[dst], vbuf[dst], 8 * ps, nt,
                                     counts, total);
  }
}

// ---------------------------------------------------------------------
// fmmx_sort_noalloc: the keys and stable sort of fmmx_sort with every
// temporary in `scratch`, a static device array of fixed size (each device
// holds its own copy until the process ends): the ping-pong keys and values
// of the radix passes, the digit counts per tile.  Host code never uses the
// array's name as an address: hipGetSymbolAddress gives each device's copy,
// once per device, kept in g_scratch_ptr.
namespace {
constexpr long NA_MAX = FMMX_SORT_NOALLOC_MAX;
constexpr long NA_COUNTS = 256L * ((NA_MAX + RS_TILE - 1) / RS_TILE + 1);
constexpr long NA_WORDS = 2 * NA_MAX + NA_COUNTS;
static __device__ __align__(256) uint32_t scratch[NA_WORDS];
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

int scratch_ptr(int device, uint32_t** out) {
  if (device < 0 || device >= 64) return fail(FMMX_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FMMX_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = static_cast<uint32_t*>(g_scratch_ptr[device]);
  return FMMX_OK;
}
}  // namespace

int sort_noalloc(int device, void* stream, int n, const float* x, const float* y, const float* z, int lmax,
                 float xmin, float ymin, float zmin, float rd, int* keys, int* perm, int* bad) {
  if (n < 0 || n > NA_MAX)
    return fail(FMMX_EINVAL, "n = " + std::to_string(n) + " outside [0, FMMX_SORT_NOALLOC_MAX = " +
                                 std::to_string(NA_MAX) + "]");
  if (lmax < 1 || lmax > MAXL) return fail(FMMX_EINVAL, "lmax out of range [1, 8]");
  if (!(rd > 0) || !std::isfinite(rd) || !std::isfinite(xmin) || !std::isfinite(ymin) || !std::isfinite(zmin))
    return fail(FMMX_EINVAL, "root cube not finite or rd <= 0");
  if (n > 0 && (!x || !y || !z || !keys || !perm || !bad)) return fail(FMMX_EINVAL, "NULL pointer");
  int prev = 0;
  FMMX_HIP(hipGetDevice(&prev));
  if (prev != device) FMMX_HIP(hipSetDevice(device));
  uint32_t* sp = nullptr;
  int rc = scratch_ptr(device, &sp);
  if (!rc && n > 0) {
    hipStream_t s = (hipStream_t)stream;
    const int slot = rs_out_slot(3 * lmax);
    uint32_t* kb[2];
    uint32_t* vb[2];
    kb[slot] = reinterpret_cast<uint32_t*>(keys);
    kb[slot ^ 1] = sp;
    vb[slot] = reinterpret_cast<uint32_t*>(perm);
    vb[slot ^ 1] = sp + NA_MAX;
    const double rbinv = (double)((float)(1 << lmax) / rd);
    k_fill_int<<<1, 32, 0, s>>>(1, 0, bad);
    k_keys<<<nblk(n), TPB, 0, s>>>(n, x, y, z, xmin, ymin, zmin, rd, rbinv, lmax, kb[0], nullptr, nullptr, bad);
    radix_sort(n, 3 * lmax, kb, vb, sp + 2 * NA_MAX, s);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail(FMMX_EDEVICE, std::string("fmmx_sort_noalloc launch: ") + hipGetErrorString(e));
  }
  if (prev != device) (void)hipSetDevice(prev);
  return rc;
}

// Keys of one point set at the leaf level and their stable sort, as
// sort/sorti.cxx:12-26 (boxn, then the counting sort of sort/csort.cxx) leaves
// them in nfn/na (sorted keys) and nbi (sorted position -> caller index).
int Engine::sort_points(int n, const float* x, const float* y, const float* z, int* keys_out, int* perm_out,
                        int flags) {
  if (n < 0) return fail(FMMX_EINVAL, "negative particle count");
  if (n > 0 && (!x || !y || !z || !perm_out)) return fail(FMMX_EINVAL, "NULL pointer");
  if (!user_box || cfg.lmax <= 0) return fail(FMMX_ESTATE, "fmmx_sort needs an explicit root cube and lmax");
  if (n == 0) return FMMX_OK;
  FMMX_HIP(hipSetDevice(device));
  const bool dev = flags & FMMX_DEVICE_PTRS;
  const size_t nb = (size_t)n * 4;
  int rc;
  if ((rc = srt_keys_in.ensure(nb)) || (rc = srt_idx_in.ensure(nb)) || (rc = srt_keys.ensure(nb)) ||
      (rc = srt_perm.ensure(nb)))
    return rc;
  const float* d[3] = {x, y, z};
  if (!dev) {
    for (int k = 0; k < 3; k++) {
      if ((rc = srt_in[k].ensure(nb))) return rc;
      FMMX_HIP(hipMemcpyAsync(srt_in[k].p, d[k], nb, hipMemcpyHostToDevice, stream));
      d[k] = srt_in[k].as<float>();
    }
  }
  const int lm = cfg.lmax;
  double rbinv = (double)((float)(1 << lm) / cfg.rd);
  if ((rc = srt_bad.ensure(4)) || (rc = rs_counts.ensure((size_t)rs_count_words(n) * 4))) return rc;
  const int slot = rs_out_slot(3 * lm);
  uint32_t* kb[2];
  uint32_t* vb[2];
  kb[slot] = srt_keys.as<uint32_t>();
  kb[slot ^ 1] = srt_keys_in.as<uint32_t>();
  vb[slot] = srt_perm.as<uint32_t>();
  vb[slot ^ 1] = srt_idx_in.as<uint32_t>();
  k_fill_int<<<1, 32, 0, stream>>>(1, 0, srt_bad.as<int>());
  k_keys<<<nblk(n), TPB, 0, stream>>>(n, d[0], d[1], d[2], cfg.xmin, cfg.ymin, cfg.zmin, cfg.rd, rbinv, lm, kb[0],
                                      nullptr, nullptr, srt_bad.as<int>());
  radix_sort(n, 3 * lm, kb, vb, rs_counts.as<uint32_t>(), stream);
  FMMX_HIP(hipGetLastError());
  const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  FMMX_HIP(hipMemcpyAsync(perm_out, srt_perm.p, nb, kind, stream));
  if (keys_out) FMMX_HIP(hipMemcpyAsync(keys_out, srt_keys.p, nb, kind, stream));
  int bad = 0;
  FMMX_HIP(hipMemcpyAsync(&bad, srt_bad.p, 4, hipMemcpyDeviceToHost, stream));
  FMMX_HIP(hipStreamSynchronize(stream));
  if (bad) return fail(FMMX_EINVAL, "coordinates outside the root cube or not finite");
  return FMMX_OK;
}

int Engine::build_side(Side& s, int n, const float* x, const float* y, const float* z, const float* nx,
                       const float* ny, const float* nz, bool dense, const Side* same) {
  s.n = n;
  s.has_normals = nx && ny && nz;
  size_t nn = (size_t)(n > 0 ? n : 1);
  int rc;
  // dense key -> box tables of every level (sizes 8^l)
  DenseLevels D{};
  if (dense && !(same && n > 0)) {
    D.lmax = lmax;
    D.off[2] = 0;
    for (int lev = 2; lev <= lmax; lev++) {
      const long nd = 1L << (3 * lev);
      if ((rc = s.dense[lev].ensure(nd * 4))) return rc;
      D.dense[lev] = s.dense[lev].as<int>();
      D.off[lev + 1] = D.off[lev] + nd;
    }
  }
  if (same && n > 0) {
    // the same points as `same` (targets = sources): its keys, order and
    // boxes are this side's too; only normals and dense tables are built
    // (aliased when `same` built them)
    s.keys.alias(same->keys);
    s.perm.alias(same->perm);
    s.x.alias(same->x);
    s.y.alias(same->y);
    s.z.alias(same->z);
    s.blast.alias(same->blast);
    s.cnt.alias(same->cnt);
    s.clscnt.alias(same->clscnt);
    for (int lev = 2; lev <= lmax; lev++) {
      s.bkey[lev].alias(same->bkey[lev]);
      s.bfirst[lev].alias(same->bfirst[lev]);
      s.parent[lev].alias(same->parent[lev]);
      s.cfirst[lev].alias(same->cfirst[lev]);
    }
    s.nb.fill(0);
    if (s.has_normals) {
      if ((rc = s.nx.ensure(nn * 4)) || (rc = s.ny.ensure(nn * 4)) || (rc = s.nz.ensure(nn * 4))) return rc;
      k_gather3<<<nblk(n), TPB, 0, stream>>>(n, s.perm.as<uint32_t>(), nx, ny, nz, s.nx.as<float>(),
                                             s.ny.as<float>(), s.nz.as<float>());
    }
    // `same` was built with its dense tables in this call (set_points)
    if (dense)
      for (int lev = 2; lev <= lmax; lev++) s.dense[lev].alias(same->dense[lev]);
    return FMMX_OK;
  }
  if ((rc = s.keys_in.ensure(nn * 4)) || (rc = s.idx_in.ensure(nn * 4)) || (rc = s.keys.ensure(nn * 4)) ||
      (rc = s.perm.ensure(nn * 4)) || (rc = s.x.ensure(nn * 4)) || (rc = s.y.ensure(nn * 4)) ||
      (rc = s.z.ensure(nn * 4)) || (rc = s.cnt.ensure((MAXL + 1) * 4)) || (rc = s.blast.ensure(nn * 4)))
    return rc;
  if (s.has_normals &&
      ((rc = s.nx.ensure(nn * 4)) || (rc = s.ny.ensure(nn * 4)) || (rc = s.nz.ensure(nn * 4))))
    return rc;
  s.nb.fill(0);
  if (n == 0) return FMMX_OK;
  for (int lev = lmax; lev >= 2; lev--) {
    if ((rc = s.bkey[lev].ensure(nn * 4)) || (rc = s.bfirst[lev].ensure(nn * 4)) ||
        (rc = s.parent[lev].ensure(nn * 4)) || (rc = s.cfirst[lev].ensure((nn + 1) * 4)))
      return rc;
  }
  const int ntile = (int)((n + BX_TILE - 1) / BX_TILE);
  if ((rc = xyz4.ensure(nn * 16)) || (rc = box_tiles.ensure((size_t)ntile * BX_REC * 4)) ||
      (rc = s.clscnt.ensure((MAXL + 1) * 8 * 4)))
    return rc;
  if (dense) k_dense_fill<<<(int)((D.off[lmax + 1] + TPB - 1) / TPB), TPB, 0, stream>>>(D);
  double rbinv = (double)((float)(1 << lmax) / rd);
  if ((rc = rs_counts.ensure((size_t)rs_count_words(n) * 4))) return rc;
  const int slot = rs_out_slot(3 * lmax);
  uint32_t* kb[2];
  uint32_t* vb[2];
  kb[slot] = s.keys.as<uint32_t>();
  kb[slot ^ 1] = s.keys_in.as<uint32_t>();
  vb[slot] = s.perm.as<uint32_t>();
  vb[slot ^ 1] = s.idx_in.as<uint32_t>();
  k_keys<<<nblk(n), TPB, 0, stream>>>(n, x, y, z, xmin, ymin, zmin, rd, rbinv, lmax, kb[0], nullptr,
                                      xyz4.as<float4>(), bad_flag.as<int>());
  radix_sort(n, 3 * lmax, kb, vb, rs_counts.as<uint32_t>(), stream);
  const uint32_t* perm = s.perm.as<uint32_t>();
  k_gather_xyz<<<nblk(n), TPB, 0, stream>>>(n, perm, xyz4.as<float4>(), s.x.as<float>(), s.y.as<float>(),
                                            s.z.as<float>());
  if (s.has_normals)
    k_gather3<<<nblk(n), TPB, 0, stream>>>(n, perm, nx, ny, nz, s.nx.as<float>(), s.ny.as<float>(),
                                           s.nz.as<float>());
  // boxes of every level: keys, first particles, parents, first children,
  // leaf last particles, dense tables and counts (two passes)
  BoxOut o{};
  for (int lev = 2; lev <= lmax; lev++) {
    o.bkey[lev] = s.bkey[lev].as<uint32_t>();
    o.bfirst[lev] = s.bfirst[lev].as<int>();
    o.parent[lev] = s.parent[lev].as<int>();
    o.cfirst[lev] = s.cfirst[lev].as<int>();
    o.dense[lev] = dense ? s.dense[lev].as<int>() : nullptr;
  }
  o.blast = s.blast.as<int>();
  o.cnt = s.cnt.as<int>();
  o.clscnt = s.clscnt.as<int>();
  k_box_count<<<ntile, BX_T, 0, stream>>>(n, s.keys.as<uint32_t>(), lmax, box_tiles.as<int>(), o.clscnt);
  k_box_place<<<ntile, BX_T, 0, stream>>>(n, s.keys.as<uint32_t>(), lmax, box_tiles.as<int>(), o);
  return FMMX_OK;
}

// The near lists and the parity-class order are launched before the box
// counts reach the host: grids and buffers cover min(N, 8^lev) boxes per
// level and the kernels read the counts from the device.
int Engine::build_lists() {
  int rc;
  const int ni = si.n, nj = sj.n;
  const int ub = (int)std::min<long>(ni, 1L << (3 * lmax));
  size_t nb1 = (size_t)(ub > 0 ? ub : 1);
  if ((rc = nij.ensure(nb1 * 4)) || (rc = neij.ensure(nb1 * NNEAR * 4))) return rc;
  // [2 + lev]: M2L pairs at level lev (zeroed by each evaluation)
  if ((rc = pair_ctr.ensure(16 * 8))) return rc;
  // the near-pair count stays as one partial sum per workgroup, summed on
  // the host when statistics are asked for
  near_nblk = 0;
  if (ub > 0 && nj > 0) {
    near_nblk = nblk(32L * ub);
    if ((rc = pair_part.ensure((size_t)near_nblk * 8))) return rc;
    k_near_list<<<near_nblk, TPB, 0, stream>>>(si.cnt.as<int>() + lmax, si.bkey[lmax].as<uint32_t>(), lmax,
                                               sj.dense[lmax].as<int>(), si.bfirst[lmax].as<int>(),
                                               si.blast.as<int>(), sj.bfirst[lmax].as<int>(), sj.blast.as<int>(),
                                               nij.as<int>(), neij.as<int>(), pair_part.as<unsigned long long>(),
                                               group_shift(lmax));
  } else if (ub > 0)
    k_fill_int<<<nblk(ub), TPB, 0, stream>>>(ub, 0, nij.as<int>());
  return FMMX_OK;
}

int Engine::build_far_lists() {
  int rc;
  for (int lev = 2; lev <= lmax; lev++) {
    int nbi = si.nb[lev];
    size_t nb1 = (size_t)(nbi > 0 ? nbi : 1);
    if ((rc = fnij[lev].ensure(nb1 * 4)) || (rc = fneij[lev].ensure(nb1 * NEBM * 4))) return rc;
    if (nbi == 0) continue;
    if (sj.nb[lev] == 0) {
      k_fill_int<<<nblk(nbi), TPB, 0, stream>>>(nbi, 0, fnij[lev].as<int>());
      continue;
    }
    k_far_list<<<nblk(nbi), TPB, 0, stream>>>(nbi, si.bkey[lev].as<uint32_t>(), sj.nb[lev],
                                              sj.bkey[lev].as<uint32_t>(), lev, sj.dense[lev].as<int>(),
                                              fnij[lev].as<int>(), fneij[lev].as<int>());
  }
  return FMMX_OK;
}

int Engine::build_m2l_order() {
  int rc;
  ClassLevels L{};
  L.lmin = 2;
  L.lmax = lmax;
  L.nb = si.n > 0 ? si.cnt.as<int>() : nullptr;
  int nch = 0;
  for (int lev = 2; lev <= lmax; lev++) {
    const int ub = (int)std::min<long>(si.n, 1L << (3 * lev));
    size_t nb1 = (size_t)(ub > 0 ? ub : 1);
    if ((rc = cls_ord[lev].ensure(nb1 * 4)) || (rc = cls_cnt[lev].ensure(8 * 4))) return rc;
    if (ub == 0) k_fill_int<<<1, 32, 0, stream>>>(8, 0, cls_cnt[lev].as<int>());
    L.key[lev] = si.bkey[lev].as<uint32_t>();
    L.ord[lev] = cls_ord[lev].as<uint32_t>();
    L.cnt[lev] = cls_cnt[lev].as<int>();
    L.chunk0[lev] = nch;
    nch += (ub + CL_CH - 1) / CL_CH;
  }
  L.chunk0[lmax + 1] = nch;
  if (nch == 0) return FMMX_OK;
  if ((rc = cls_chunks.ensure((size_t)nch * 8 * 4))) return rc;
  L.ccnt = cls_chunks.as<int>();
  k_class_count<<<nch, CL_T, 0, stream>>>(L);
  k_class_place<<<nch, CL_T, 0, stream>>>(L);
  FMMX_HIP(hipGetLastError());
  return FMMX_OK;
}

// Box statistics of levels 0..8 for the n points (x, y, z) in the current
// root cube: boxes[l] holding a point and sumsq[l], the sum of the squared
// points per box (synchronous: one readback).
int Engine::occupancy(int n, const float* x, const float* y, const float* z, long boxes[MAXL + 1],
                      double sumsq[MAXL + 1]) {
  for (int l = 0; l <= MAXL; l++) boxes[l] = 0, sumsq[l] = 0;
  if (n <= 0) return FMMX_OK;
  static_assert(OCC_L == MAXL, "statistics from the deepest level");
  int rc;
  // counts of levels 8..0 (level l: 8^l words) then 2 words of statistics
  // per level
  long words = 0;
  for (int l = 0; l <= OCC_L; l++) words += 1L << (3 * l);
  const long st_off = (words + 3) / 4 * 4;  // 16-byte aligned
  if ((rc = occ_bits.ensure((size_t)(st_off + 4 * (OCC_L + 1)) * 4))) return rc;
  unsigned* base = occ_bits.as<unsigned>();
  unsigned* lvl[OCC_L + 1];
  long o = 0;
  for (int l = OCC_L; l >= 0; l--) {
    lvl[l] = base + o;  // level l starts at a multiple of 8 words (16-byte aligned for l >= 1)
    o += 1L << (3 * l);
  }
  unsigned long long* st = (unsigned long long*)(base + st_off);
  k_fill_int<<<(int)((st_off + 4 * (OCC_L + 1) + TPB - 1) / TPB), TPB, 0, stream>>>(st_off + 4 * (OCC_L + 1), 0,
                                                                                    (int*)base);
  const double rbinv = (double)((float)(1 << OCC_L) / rd);
  k_occ_mark<<<nblk(n), TPB, 0, stream>>>(n, x, y, z, xmin, ymin, zmin, rbinv, lvl[OCC_L]);
  for (int l = OCC_L; l >= 1; l--) {
    const long nout = 1L << (3 * (l - 1));
    const int g = (int)std::min<long>(1024, (nout + 255) / 256);
    k_occ_up<<<g, 256, 0, stream>>>(nout, lvl[l], lvl[l - 1], st + 2 * l);
  }
  if (!rb_host) FMMX_HIP(hipHostMalloc((void**)&rb_host, RB_INTS * 4, hipHostMallocDefault));
  static_assert(4 * (MAXL + 1) <= RB_INTS, "readback buffer");
  FMMX_HIP(hipMemcpyAsync(rb_host, st, (size_t)2 * (OCC_L + 1) * 8, hipMemcpyDeviceToHost, stream));
  FMMX_HIP(hipStreamSynchronize(stream));
  const unsigned long long* h = (const unsigned long long*)rb_host;
  boxes[0] = 1;
  sumsq[0] = (double)n * (double)n;
  for (int l = 1; l <= OCC_L; l++) {
    boxes[l] = (long)h[2 * l];
    sumsq[l] = (double)h[2 * l + 1];
  }
  return FMMX_OK;
}

}  // namespace fmmx
