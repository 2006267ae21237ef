// tree_util.h -- Morton key helpers shared by the tree and phase kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace fmmx {

// ---- Morton helpers: x -> bit 3b+1, y -> 3b, z -> 3b+2 (fmm/boxn.cxx:21-30)
__device__ __host__ inline uint32_t morton_encode(int x, int y, int z, int lev) {
  uint32_t k = 0;
  for (int b = 0; b < lev; b++) {
    k |= (uint32_t)((y >> b) & 1) << (3 * b);
    k |= (uint32_t)((x >> b) & 1) << (3 * b + 1);
    k |= (uint32_t)((z >> b) & 1) << (3 * b + 2);
  }
  return k;
}
__device__ __host__ inline void morton_decode(uint32_t k, int lev, int& x, int& y, int& z) {
  x = y = z = 0;
  for (int b = 0; b < lev; b++) {
    y |= (int)((k >> (3 * b)) & 1) << b;
    x |= (int)((k >> (3 * b + 1)) & 1) << b;
    z |= (int)((k >> (3 * b + 2)) & 1) << b;
  }
}

}  // namespace fmmx
