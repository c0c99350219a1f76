// mt_morton.h -- Morton keys and the device key lookup shared by the sources.
//
// A reference key interleaves the cell's lower corner at finest resolution:
// x bit i -> key bit 3i, y -> 3i+1, z -> 3i+2 (Morton_Assist.h:229-289,
// Lat_Voxelize.cpp:245).  A level's cells live in 4x4x4 bricks of 64 slots;
// find() resolves a key to its slot by a binary search over the level's sorted
// brick codes (the device form of D_map_define<T>::find, mortonTrieMap.cpp:116-126).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mt {

__host__ __device__ inline uint64_t spread3(uint64_t v) {
  v &= 0x1FFFFF;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}
__host__ __device__ inline uint64_t compact3(uint64_t v) {
  v &= 0x1249249249249249ull;
  v = (v ^ (v >> 2)) & 0x10C30C30C30C30C3ull;
  v = (v ^ (v >> 4)) & 0x100F00F00F00F00Full;
  v = (v ^ (v >> 8)) & 0x1F0000FF0000FFull;
  v = (v ^ (v >> 16)) & 0x1F00000000FFFFull;
  v = (v ^ (v >> 32)) & 0x1FFFFFull;
  return v;
}
__host__ __device__ inline uint64_t morton3(uint64_t x, uint64_t y, uint64_t z) {
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

__device__ inline int64_t bsearch_u64(const uint64_t* a, int64_t n, uint64_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return (lo < n && a[lo] == x) ? lo : -1;
}

struct LevelView {
  const uint64_t* bcode;
  const uint16_t* kind;
  int64_t n_bricks;
  int shift;  // 3*(Lmax-L)
};
// find(): slot of a reference key on a level, -1 if absent
__device__ inline int32_t find_slot(const LevelView& v, uint64_t key) {
  if (v.shift && (key & ((1ull << v.shift) - 1))) return -1;
  uint64_t code = key >> v.shift;
  int64_t b = bsearch_u64(v.bcode, v.n_bricks, code >> 6);
  if (b < 0) return -1;
  int32_t s = (int32_t)((b << 6) | (code & 63));
  return v.kind[s] ? s : -1;
}

}  // namespace mt
