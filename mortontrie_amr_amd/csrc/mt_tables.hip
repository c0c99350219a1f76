// mt_tables.hip -- device construction of the per-level cell tables.
//
// The reference keeps each level's cells in associative containers keyed by
// Morton code (header/Grid_Class.h:116-137, mortonTrieMap.cpp) and resolves
// every neighbour with find() (lbm_kernel/Stream_Manager.cpp:53-62).  Here the
// uploaded key sets are sorted on the GPU (LSD radix sort, 8-bit digits, stable
// tile ranking in LDS), grouped into 4x4x4 Morton bricks, and every neighbour /
// parent / child relation is resolved once into index tables by a binary
// descent over the sorted brick codes.  find() semantics are preserved: a key
// is present iff its brick exists and its slot is occupied.
#include <algorithm>
#include <climits>
#include <cstring>

#include "mt_case.h"
#include "mt_device.h"
#include "mt_morton.h"

namespace mt {

// ------------------------------------------------------------------ scan
constexpr int kScanTile = 1024;  // 256 threads x 4

__device__ int block_exclusive_scan(int v, int* lds, int& total) {
  int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int x = (t >= off) ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += x;
    __syncthreads();
  }
  int incl = lds[t];
  total = lds[255];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan_blocksum(const int* in, int64_t n, int* bsum) {
  __shared__ int lds[256];
  int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
  int s = 0;
  for (int j = 0; j < 4; ++j)
    if (base + j < n) s += in[base + j];
  int tot;
  block_exclusive_scan(s, lds, tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
// single block: exclusive scan of nb block sums, chunked
__global__ void k_scan_bsum(int* bsum, int64_t nb, int* total) {
  __shared__ int lds[256];
  int64_t chunk = (nb + 255) / 256;
  int64_t lo = threadIdx.x * chunk, hi = lo + chunk < nb ? lo + chunk : nb;
  int s = 0;
  for (int64_t i = lo; i < hi; ++i) s += bsum[i];
  int tot;
  int ex = block_exclusive_scan(s, lds, tot);
  int run = ex;
  for (int64_t i = lo; i < hi; ++i) {
    int v = bsum[i];
    bsum[i] = run;
    run += v;
  }
  if (threadIdx.x == 0 && total) *total = tot;
}
__global__ void k_scan_apply(const int* in, int64_t n, const int* bsum, int* out) {
  __shared__ int lds[256];
  int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
  int v[4], s = 0;
  for (int j = 0; j < 4; ++j) {
    v[j] = (base + j < n) ? in[base + j] : 0;
    s += v[j];
  }
  int tot;
  int ex = block_exclusive_scan(s, lds, tot) + bsum[blockIdx.x];
  for (int j = 0; j < 4; ++j) {
    if (base + j < n) out[base + j] = ex;
    ex += v[j];
  }
}

// exclusive scan of n ints; *h_total receives the sum
static int scan_exclusive(Case& c, const int* d_in, int* d_out, int64_t n, int64_t* h_total) {
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb < 1) nb = 1;
  int* bsum = nullptr;
  int* tot = nullptr;
  MT_HIP(dev_malloc((void**)&bsum, sizeof(int) * (nb + 1), c.stream));
  tot = bsum + nb;
  k_scan_blocksum<<<nb, 256, 0, c.stream>>>(d_in, n, bsum);
  k_scan_bsum<<<1, 256, 0, c.stream>>>(bsum, nb, tot);
  k_scan_apply<<<nb, 256, 0, c.stream>>>(d_in, n, bsum, d_out);
  int t = 0;
  MT_HIP(hipMemcpyAsync(&t, tot, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipStreamSynchronize(c.stream));
  dev_free(bsum, c.stream);
  *h_total = t;
  return MT_OK;
}

// ------------------------------------------------------------------ radix sort
constexpr int kRsTile = 256;

__global__ void k_rs_hist(const uint64_t* keys, int64_t n, int shift, int64_t n_tiles, int* hist) {
  __shared__ int h[256];
  int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * kRsTile + t;
  if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
  __syncthreads();
  hist[(int64_t)t * n_tiles + blockIdx.x] = h[t];
}
__global__ void k_rs_scatter(const uint64_t* kin, const int32_t* vin, uint64_t* kout, int32_t* vout, int64_t n,
                             int shift, int64_t n_tiles, const int* offs) {
  __shared__ int dg[kRsTile];
  int t = threadIdx.x;
  int64_t i = (int64_t)blockIdx.x * kRsTile + t;
  uint64_t k = 0;
  int d = 256;
  if (i < n) {
    k = kin[i];
    d = (int)((k >> shift) & 255);
  }
  dg[t] = d;
  __syncthreads();
  if (i < n) {
    int rank = 0;
    for (int j = 0; j < t; ++j) rank += (dg[j] == d);
    int64_t pos = offs[(int64_t)d * n_tiles + blockIdx.x] + rank;
    kout[pos] = k;
    vout[pos] = vin[i];
  }
}
__global__ void k_iota(int32_t* v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int32_t)i;
}

// sorts (keys, payload = original index) ascending; results in *ks, *vs
static int radix_sort(Case& c, const uint64_t* d_keys, int64_t n, uint64_t max_key, uint64_t** ks, int32_t** vs) {
  uint64_t *ka = nullptr, *kb = nullptr;
  int32_t *va = nullptr, *vb = nullptr;
  int* hist = nullptr;
  int64_t n_tiles = (n + kRsTile - 1) / kRsTile;
  if (n_tiles < 1) n_tiles = 1;
  MT_HIP(dev_malloc((void**)&ka, 8 * (n + 1), c.stream));
  MT_HIP(dev_malloc((void**)&kb, 8 * (n + 1), c.stream));
  MT_HIP(dev_malloc((void**)&va, 4 * (n + 1), c.stream));
  MT_HIP(dev_malloc((void**)&vb, 4 * (n + 1), c.stream));
  MT_HIP(dev_malloc((void**)&hist, sizeof(int) * 256 * n_tiles, c.stream));
  MT_HIP(hipMemcpyAsync(ka, d_keys, 8 * n, hipMemcpyDeviceToDevice, c.stream));
  k_iota<<<(n + 255) / 256 + 1, 256, 0, c.stream>>>(va, n);
  int bits = 0;
  while (bits < 64 && (max_key >> bits)) ++bits;
  int passes = (bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * p;
    k_rs_hist<<<n_tiles, kRsTile, 0, c.stream>>>(ka, n, shift, n_tiles, hist);
    int64_t tot;
    int r = scan_exclusive(c, hist, hist, 256 * n_tiles, &tot);
    if (r) return r;
    k_rs_scatter<<<n_tiles, kRsTile, 0, c.stream>>>(ka, va, kb, vb, n, shift, n_tiles, hist);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  MT_HIP(hipStreamSynchronize(c.stream));
  dev_free(kb, c.stream);
  dev_free(vb, c.stream);
  dev_free(hist, c.stream);
  *ks = ka;
  *vs = va;
  return MT_OK;
}

// ------------------------------------------------------------------ build

__global__ void k_codes(const uint64_t* ks, int64_t n, int shift, int* head, int* err) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = ks[i];
  if (shift && (k & ((1ull << shift) - 1))) atomicOr(err, 1);  // not a cell of this level
  if (i > 0 && ks[i - 1] == k) atomicOr(err, 2);                // duplicate key
  uint64_t code = k >> shift;
  head[i] = (i == 0) || ((ks[i - 1] >> shift) >> 6) != (code >> 6);
}
__global__ void k_fill_empty(uint64_t* key, uint16_t* kind, int64_t ns) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  key[s] = ~0ull;
  kind[s] = 0;
}
__global__ void k_place(const uint64_t* ks, const int32_t* vs, const uint16_t* kinds_in, const int* head,
                        const int* bidx, int64_t n, int shift, uint64_t* bcode, uint64_t* key, uint16_t* kind) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = ks[i];
  uint64_t code = k >> shift;
  int64_t b = bidx[i] + head[i] - 1;
  int64_t s = (b << 6) | (int64_t)(code & 63);
  key[s] = k;
  kind[s] = kinds_in[vs[i]];
  if (head[i]) bcode[b] = code >> 6;
}
// brick -> the member whose level-L box holds it (boxes are brick-aligned)
__global__ void k_bmember(const uint64_t* bcode, int64_t nb, const MemberParams* mp, int nm, int L, int32_t* bm,
                          int* err) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  uint64_t c = bcode[b];
  const int64_t x = 4 * (int64_t)compact3(c), y = 4 * (int64_t)compact3(c >> 1), z = 4 * (int64_t)compact3(c >> 2);
  int found = -1;
  for (int m = 0; m < nm && found < 0; ++m) {
    const MemberParams& P = mp[m];
    if (x >= P.lo[L][0] && x < P.hi[L][0] && y >= P.lo[L][1] && y < P.hi[L][1] && z >= P.lo[L][2] &&
        z < P.hi[L][2])
      found = m;
  }
  if (found < 0) {
    atomicOr(err, 8);
    found = 0;
  }
  bm[b] = found;
}
// 27 neighbour bricks of each brick within its member's box; a periodic axis of
// the member wraps inside the box, any other step out of the box is no brick
__global__ void k_bnbr(const uint64_t* bcode, int64_t nb, const int32_t* bm, const MemberParams* mp, int L,
                       int32_t* bnbr) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  uint64_t c = bcode[b];
  const int64_t p[3] = {(int64_t)compact3(c), (int64_t)compact3(c >> 1), (int64_t)compact3(c >> 2)};
  const MemberParams& P = mp[bm[b]];
  int64_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = P.lo[L][a] >> 2;
    hi[a] = (P.hi[L][a] + 3) >> 2;
  }
  for (int o = 0; o < 27; ++o) {
    const int d[3] = {o % 3 - 1, (o / 3) % 3 - 1, o / 9 - 1};
    int64_t V[3];
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
      V[a] = p[a] + d[a];
      if (V[a] < lo[a] || V[a] >= hi[a]) {
        if (P.periodic[a]) {
          const int64_t n = hi[a] - lo[a];
          V[a] = lo[a] + ((V[a] - lo[a]) % n + n) % n;
        } else {
          ok = false;
        }
      }
    }
    bnbr[b * 27 + o] = ok ? (int32_t)bsearch_u64(bcode, nb, morton3(V[0], V[1], V[2])) : -1;
  }
}
__global__ void k_pmask(const int32_t* bnbr, const uint16_t* kind, int64_t ns, int Q, uint32_t* pmask) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  if (!kind[s]) {
    pmask[s] = 0;
    return;
  }
  int b = (int)(s >> 6);
  int lx, ly, lz;
  brick_coords((int)(s & 63), lx, ly, lz);
  const int* nb27 = bnbr + (int64_t)b * 27;
  uint32_t m = 0;
  for (int q = 1; q < Q; ++q) {
    int t = slot_at(nb27, lx, ly, lz, kEX[q], kEY[q], kEZ[q]);
    if (t >= 0 && kind[t]) m |= 1u << q;
  }
  pmask[s] = m;
}
__global__ void k_flag_kind(const uint16_t* kind, int64_t ns, uint16_t mask, int* flag) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  flag[s] = (kind[s] & mask) != 0;
}
__global__ void k_flag_ib(const uint16_t* kind, const int32_t* dis_row, int64_t ns, int* flag) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  flag[s] = (kind[s] & MT_KIND_DIS) && (kind[s] & MT_KIND_FLUID) && dis_row[s] >= 0;
}
// Coalescence inside the fine level's fused update: the 8 children of an F2C
// parent form one aligned group of 8 consecutive fine slots, child j at offset j
// (the child order j = x + 2y + 4z of LBM_Manager.cpp:1074-1082 is the low
// Morton bits), so the fine block holding them can average them.  A parent
// qualifies when its children are C2F cells that the fused update streams
// (not lat_f, not treated solid-boundary cells); cpar[group] = parent slot.
__global__ void k_cpar(const int32_t* f2c, const int32_t* child, int64_t n, const uint16_t* fkind,
                       const int32_t* fdis_row, int32_t* cpar, int* bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t c0 = child[i * 8];
  bool ok = c0 >= 0 && (c0 & 7) == 0;
  for (int j = 0; j < 8 && ok; ++j) {
    const int32_t cj = child[i * 8 + j];
    ok = cj == c0 + j;
    if (!ok) break;
    const uint16_t k = fkind[cj];
    ok = (k & MT_KIND_C2F) && !(k & MT_KIND_FLUID) && !((k & MT_KIND_DIS) && fdis_row[cj] >= 0);
  }
  if (ok) cpar[c0 >> 3] = f2c[i];
  else *bad = 1;
}
// Explosions inside the fused updates: bit q of tmask[s] is set when the pull
// source of population q of a streamed cell is an explosion target (C2F /
// ghost_C2F not in dis), whose value the pull then reads from the pair's
// target buffer.  The source is x + e_q, except on a wall link of a treated
// solid-boundary cell (wmask, finest level), whose pull loads x - e_q instead.
__global__ void k_tmask(const int32_t* bnbr, const uint16_t* kind, const uint32_t* pmask, const uint32_t* wmask,
                        int64_t ns, int Q, uint16_t targets, uint32_t* tmask) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  uint32_t m = 0;
  if (kind[s] & (MT_KIND_FLUID | MT_KIND_C2F)) {
    int lx, ly, lz;
    brick_coords((int)(s & 63), lx, ly, lz);
    const int* nb27 = bnbr + (s >> 6) * 27;
    const uint32_t pm = pmask[s];
    const bool treated = wmask && (kind[s] & MT_KIND_FLUID) && (kind[s] & MT_KIND_DIS);
    const uint32_t wm = treated ? wmask[s] : 0u;
    for (int q = 1; q < Q; ++q) {
      const bool wall = (wm >> q) & 1;
      if (!((pm >> (wall ? qinv(q) : q)) & 1)) continue;
      const int t = wall ? slot_at(nb27, lx, ly, lz, -kEX[q], -kEY[q], -kEZ[q])
                         : slot_at(nb27, lx, ly, lz, kEX[q], kEY[q], kEZ[q]);
      if (t >= 0 && (kind[t] & targets) && !(kind[t] & MT_KIND_DIS)) m |= 1u << q;
    }
    // bit 0: the cell is a target itself, so its rest population (the pull
    // from its own slot) is in the target buffer too
    if ((kind[s] & MT_KIND_C2F & targets) && !(kind[s] & MT_KIND_DIS)) m |= 1u;
  }
  tmask[s] = m;
}
__global__ void k_gxrow(const int32_t* gf2c, int64_t n, int32_t* gxrow) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gxrow[gf2c[i]] = (int32_t)i;
}
__global__ void k_flag_target(const uint16_t* kind, int64_t ns, uint16_t targets, int* flag) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  flag[s] = (kind[s] & targets) && !(kind[s] & MT_KIND_DIS);
}
// The classes order_active sorts the active bricks by, one 64-lane block per
// entry of the active list: bit 0 a treated solid-boundary cell, bit 1 a
// coalescence group (its first slot has a parent), bit 2 a ghost_F2C parent,
// bit 3 a cell that a sub-step without the C2F stream still updates (a lat_f
// cell, or a C2F cell in dis), bit 4 a cell of the domain boundary that the
// update reconstructs (wet faces, edges and corners: face bits 0x3F00).
__global__ __launch_bounds__(64) void k_brick_order_flags(const int32_t* act, const uint16_t* kind,
                                                          const int32_t* dis_row, const int32_t* cpar,
                                                          const int32_t* gxrow, uint8_t* out) {
  const int i = threadIdx.x;
  const int64_t s = (int64_t)act[blockIdx.x] * 64 + i;
  const uint16_t k = kind[s];
  const bool t = (k & MT_KIND_DIS) && (k & MT_KIND_FLUID) && dis_row[s] >= 0;
  const bool g = cpar && (i & 7) == 0 && cpar[s >> 3] >= 0;
  const bool x = gxrow && gxrow[s] >= 0;
  const bool n = (k & MT_KIND_FLUID) || ((k & MT_KIND_C2F) && (k & MT_KIND_DIS));
  const bool w = (k & 0x3F00) != 0;
  const uint64_t bt = __ballot(t), bg = __ballot(g), bx = __ballot(x), bn = __ballot(n), bw = __ballot(w);
  if (i == 0) out[blockIdx.x] = (uint8_t)((bt ? 1 : 0) | (bg ? 2 : 0) | (bx ? 4 : 0) | (bn ? 8 : 0) | (bw ? 16 : 0));
}
__global__ void k_flag_brick(const uint16_t* kind, int64_t nb, uint16_t mask, int* flag) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int any = 0;
  for (int l = 0; l < 64; ++l) any |= (kind[b * 64 + l] & mask) != 0;
  flag[b] = any;
}
__global__ void k_emit(const int* flag, const int* pos, int64_t n, int32_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) out[pos[i]] = (int32_t)i;
}
__global__ void k_find(LevelView v, const uint64_t* keys, int64_t n, int32_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = find_slot(v, keys[i]);
}
// children of F2C / ghost_F2C parents (child order of LBM_Manager.cpp:1074-1082)
__global__ void k_children(const int32_t* parents, int64_t n, const uint64_t* pkey, int pshift, LevelView fine,
                           const uint16_t* fkind, const uint64_t* dis_keys, int64_t n_dis, int ghost_only,
                           int32_t* child_all, int32_t* child_write, int* err) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = pkey[parents[i]];
  uint64_t px = compact3(k) >> pshift, py = compact3(k >> 1) >> pshift, pz = compact3(k >> 2) >> pshift;
  for (int j = 0; j < 8; ++j) {
    uint64_t cx = 2 * px + (j & 1), cy = 2 * py + ((j >> 1) & 1), cz = 2 * pz + ((j >> 2) & 1);
    uint64_t ck = morton3(cx << fine.shift / 3, cy << fine.shift / 3, cz << fine.shift / 3);
    int32_t s = find_slot(fine, ck);
    bool in_dis = n_dis > 0 && bsearch_u64(dis_keys, n_dis, ck) >= 0;
    if (!ghost_only) {
      if (s < 0) atomicOr(err, 4);  // coalescence reads every child (LBM_Manager.cpp:1155)
      child_all[i * 8 + j] = s;
      child_write[i * 8 + j] = (s >= 0 && !in_dis) ? s : -1;
    } else {
      bool g = s >= 0 && (fkind[s] & MT_KIND_GHOST_C2F);
      child_write[i * 8 + j] = (g && !in_dis) ? s : -1;
    }
  }
}
__global__ void k_dis_rows(const uint64_t* ks, const int32_t* vs, int64_t n, LevelView v, const double* vals_in,
                           int m, int32_t* dis_row, double* vals_out, uint16_t* kind) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t s = find_slot(v, ks[i]);
  for (int j = 0; j < m; ++j) vals_out[i * m + j] = vals_in[(int64_t)vs[i] * m + j];
  if (s >= 0) {
    dis_row[s] = (int32_t)i;
    kind[s] |= MT_KIND_DIS;
  }
}
// wall-link mask of the dis cells: bit q set when dis[inv(q) - 1] != -1
// (SinglePointInterpolation::treat, ComplexBoundary_Manager.cpp:77-80)
__global__ void k_wmask(const int32_t* dis_row, const double* dis, int64_t ns, int Q, uint32_t* wmask) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  uint32_t m = 0;
  const int row = dis_row[s];
  if (row >= 0)
    for (int q = 1; q < Q; ++q)
      if (dis[(int64_t)row * (Q - 1) + qinv(q) - 1] != -1.0) m |= 1u << q;
  wmask[s] = m;
}
__global__ void k_fill_i32(int32_t* a, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

static inline int64_t nblk(int64_t n, int b = 256) { return n <= 0 ? 1 : (n + b - 1) / b; }

// Copy one member's state between two cases, brick by brick: every brick of
// member sm in the source, moved by (dx,dy,dz) bricks, is a brick of the
// destination; rows[r] = (source row, destination row) of n_slots elements.
template <typename T>
__global__ void k_copy_member(const uint64_t* sbcode, const int32_t* sbm, int32_t sm, const uint64_t* dbcode,
                              int64_t dnb, int64_t dx, int64_t dy, int64_t dz, const T* const* srows,
                              T* const* drows, int nrows, int* err) {
  const int64_t b = blockIdx.x;
  if (sbm[b] != sm) return;  // uniform over the block
  const uint64_t c = sbcode[b];
  const uint64_t code = morton3(compact3(c) + dx, compact3(c >> 1) + dy, compact3(c >> 2) + dz);
  const int64_t db = bsearch_u64(dbcode, dnb, code);
  if (db < 0) {
    if (threadIdx.x == 0) atomicOr(err, 1);
    return;
  }
  for (int r = 0; r < nrows; ++r) drows[r][db * 64 + threadIdx.x] = srows[r][b * 64 + threadIdx.x];
}

template <typename T>
static int copy_rows(Case& dst, DevLevel& ld, Case& src, DevLevel& ls, int32_t sm, int64_t d[3],
                     const std::vector<std::pair<void*, void*>>& rows, int* d_err) {
  std::vector<const T*> hs;
  std::vector<T*> hd;
  for (auto& r : rows) {
    hs.push_back((const T*)r.first);
    hd.push_back((T*)r.second);
  }
  T** dev = nullptr;
  MT_HIP(dev_malloc((void**)&dev, 2 * sizeof(T*) * rows.size(), dst.stream));
  MT_HIP(hipMemcpyAsync(dev, hs.data(), sizeof(T*) * rows.size(), hipMemcpyHostToDevice, dst.stream));
  MT_HIP(hipMemcpyAsync(dev + rows.size(), hd.data(), sizeof(T*) * rows.size(), hipMemcpyHostToDevice, dst.stream));
  k_copy_member<T><<<ls.n_bricks, 64, 0, dst.stream>>>(ls.bcode, ls.bmember, sm, ld.bcode, ld.n_bricks, d[0], d[1], d[2],
                                                       (const T* const*)dev, dev + rows.size(), (int)rows.size(),
                                                       d_err);
  MT_HIP(hipGetLastError());
  MT_HIP(hipStreamSynchronize(dst.stream));
  dev_free(dev, dst.stream);
  return MT_OK;
}

// Both cases settled (reference-state buffers); member geometries equal up to
// a whole-brick shift.  Copies f, both fcol buffers and both rho / u buffers.
int copy_member(Case& dst, int32_t dm, Case& src, int32_t sm) {
  if (dst.Q != src.Q || dst.prec != src.prec) {
    set_error("members differ in Q or precision");
    return MT_ERR_ARG;
  }
  const MemberParams& PS = src.h_mem[sm];
  const MemberParams& PD = dst.h_mem[dm];
  int* d_err = nullptr;
  MT_HIP(dev_malloc((void**)&d_err, 4, dst.stream));
  MT_HIP(hipMemsetAsync(d_err, 0, 4, dst.stream));
  MT_HIP(hipStreamSynchronize(src.stream));
  int r = MT_OK;
  for (int L = 0; L <= src.Lmax && r == MT_OK; ++L) {
    if (L > dst.Lmax) {
      set_error("destination has fewer levels");
      r = MT_ERR_ARG;
      break;
    }
    DevLevel& ls = src.lv[L];
    DevLevel& ld = dst.lv[L];
    int64_t d[3];
    for (int a = 0; a < 3; ++a) {
      const int64_t sh = (int64_t)PD.lo[L][a] - PS.lo[L][a];
      if (sh % 4) {
        set_error("member origins differ by a partial brick");
        r = MT_ERR_ARG;
      }
      d[a] = sh / 4;
    }
    if (r) break;
    const size_t pb = src.pop_bytes(), rb = src.real_bytes();
    std::vector<std::pair<void*, void*>> pops, reals;
    auto prow = [&](void* base, int64_t ns, int q) { return (void*)((char*)base + pb * (size_t)(q * ns)); };
    auto rrow = [&](void* base, int64_t ns, int j) { return (void*)((char*)base + rb * (size_t)(j * ns)); };
    for (int q = 0; q < src.Q; ++q) {
      pops.push_back({prow(ls.f, ls.n_slots, q), prow(ld.f, ld.n_slots, q)});
      for (int k = 0; k < 2; ++k)
        pops.push_back({prow(ls.fcol[k ^ ls.pf], ls.n_slots, q), prow(ld.fcol[k ^ ld.pf], ld.n_slots, q)});
    }
    for (int k = 0; k < 2; ++k) {
      reals.push_back({rrow(ls.rho[k ^ ls.pr], ls.n_slots, 0), rrow(ld.rho[k ^ ld.pr], ld.n_slots, 0)});
      for (int j = 0; j < 3; ++j)
        reals.push_back({rrow(ls.u[k ^ ls.pr], ls.n_slots, j), rrow(ld.u[k ^ ld.pr], ld.n_slots, j)});
    }
    if (pb == 8) r = copy_rows<uint64_t>(dst, ld, src, ls, sm, d, pops, d_err);
    else if (pb == 4) r = copy_rows<uint32_t>(dst, ld, src, ls, sm, d, pops, d_err);
    else r = copy_rows<uint16_t>(dst, ld, src, ls, sm, d, pops, d_err);
    if (r == MT_OK) {
      if (rb == 8) r = copy_rows<uint64_t>(dst, ld, src, ls, sm, d, reals, d_err);
      else r = copy_rows<uint32_t>(dst, ld, src, ls, sm, d, reals, d_err);
    }
  }
  int herr = 0;
  if (r == MT_OK) {
    MT_HIP(hipMemcpy(&herr, d_err, 4, hipMemcpyDeviceToHost));
    if (herr) {
      set_error("a source brick of the member has no destination brick (geometries differ)");
      r = MT_ERR_ARG;
    }
  }
  dev_free(d_err, dst.stream);
  return r;
}

// tau_L of a member and the transfer factors that follow from it.
void set_member_tau(const Case& c, MemberParams& P, const double* tau) {
  const bool conv = (c.d.options & MT_OPT_CONVECTIVE_XFER) != 0;
  for (int L = 0; L <= c.Lmax; ++L) P.tau[L] = tau[L];
  for (int L = 1; L <= c.Lmax; ++L) P.xscale[L] = conv ? P.tau[L] / (2.0 * P.tau[L - 1]) : (2.0 * P.tau[L]) / P.tau[L - 1];
  for (int L = 0; L < c.Lmax; ++L) P.cscale[L] = conv ? (2.0 * P.tau[L]) / P.tau[L + 1] : P.tau[L] / (2.0 * P.tau[L + 1]);
}

// Per-member parameters on the device: tau_L = 2^L (tau0 - 1/2) + 1/2
// (LBM_Manager.cpp:814-818), the transfer factors 2 tau_f / tau_c and
// tau_c / (2 tau_f) (LBM_Manager.cpp:1065, 1137; swapped for the convective
// option), the wet-node face configuration and the member's box per level.
int build_members(Case& c) {
  c.h_mem.clear();
  c.wet_any = c.res_any = c.nee_code = false;
  for (const mt_member_desc& m : c.h_mdesc) {
    MemberParams P = {};
    double tau[kMaxLevels];
    for (int L = 0; L <= c.Lmax; ++L) tau[L] = double(1 << L) * (m.tau0 - 0.5) + 0.5;
    set_member_tau(c, P, tau);
    P.les_C = m.les_C;
    P.rho0 = m.rho0;
    for (int a = 0; a < 3; ++a) P.u0[a] = m.u0[a];
    P.bc.wet_mask = 0;
    for (int f = 0; f < 6; ++f) {
      P.bc.kind[f] = m.face_kind[f];
      for (int a = 0; a < 3; ++a) P.bc.u[f][a] = m.face_u[f][a];
      if (m.face_kind[f] == MT_BC_ZOUHE || m.face_kind[f] == MT_BC_OUTFLOW || m.face_kind[f] == MT_BC_NEE)
        P.bc.wet_mask |= 1 << f;
      if (m.face_kind[f] == MT_BC_RESERVOIR) c.res_any = true;
    }
    if (P.bc.wet_mask) c.wet_any = true;
    // the non-equilibrium extrapolation runs on an NEE face and on the edges and
    // corners where wet faces of two axes meet
    int axes = 0;
    for (int f = 0; f < 6; ++f) {
      if (m.face_kind[f] == MT_BC_NEE) c.nee_code = true;
      if ((P.bc.wet_mask >> f) & 1) axes |= 1 << (f >> 1);
    }
    if (__builtin_popcount(axes) >= 2) c.nee_code = true;
    for (int L = 0; L <= c.Lmax; ++L)
      for (int a = 0; a < 3; ++a) {
        P.lo[L][a] = m.origin0[a] << L;
        P.hi[L][a] = (m.origin0[a] + m.dims0[a]) << L;
      }
    for (int a = 0; a < 3; ++a) P.periodic[a] = m.face_kind[2 * a] == MT_BC_PERIODIC;
    c.h_mem.push_back(P);
  }
  c.d_mem = c.alloc<MemberParams>((int64_t)c.h_mem.size());
  if (!c.d_mem) {
    set_error("out of device memory");
    return MT_ERR_HIP;
  }
  MT_HIP(hipMemcpyAsync(c.d_mem, c.h_mem.data(), sizeof(MemberParams) * c.h_mem.size(), hipMemcpyHostToDevice,
                        c.stream));
  return MT_OK;
}

// The fused kernel gives XCD x the active-list range [x*per, (x+1)*per), per =
// ceil(n_active/8), and dispatches each range in list order.  Bricks whose
// update has a longer dependent tail go first within their range, so they do
// not finish last: treated solid-boundary cells, then wet domain-boundary cells
// (the reconstruction's neighbour loads; the 64^3 cavity's FP32 step runs 11%
// faster with its wall bricks first), then coalescence groups, then ghost_F2C
// parents (explosion stores), then the rest; Morton order otherwise.
// A level that coalesces into the coarser one gets a second order for the
// sub-step that does (coalescence groups first), over the same XCD ranges.
static int order_active(Case& c, int L) {
  DevLevel& l = c.lv[L];
  const int64_t na = l.n_active;
  if (na <= 1) return MT_OK;
  // the classes are found on the device, one byte per active brick, so only
  // the active list and these bytes cross to the host
  uint8_t* d_fl = nullptr;
  MT_HIP(dev_malloc((void**)&d_fl, na, c.stream));
  k_brick_order_flags<<<(unsigned)na, 64, 0, c.stream>>>(l.active, l.kind, l.dis_row, l.cpar, l.gxrow, d_fl);
  MT_HIP(hipGetLastError());
  std::vector<int32_t> act(na);
  std::vector<uint8_t> fl(na);
  MT_HIP(hipMemcpyAsync(act.data(), l.active, 4 * na, hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipMemcpyAsync(fl.data(), d_fl, na, hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipStreamSynchronize(c.stream));
  dev_free(d_fl, c.stream);
  // bit 0: treated cells, bit 1: coalescence groups, bit 2: ghost_F2C parents,
  // bit 4: wet domain-boundary cells
  auto rank_a = [](int k) { return (k & 1) ? 0 : (k & 16) ? 1 : (k & 2) ? 2 : (k & 4) ? 3 : 4; };
  // in a sub-step that coalesces, a coalescing brick's tail (the parents'
  // rescale, collision and explosion) is the longest
  auto rank_b = [](int k) { return (k & 2) ? 0 : (k & 1) ? 1 : (k & 4) ? 2 : 3; };
  const int64_t per = ((na + 7) / 8 * 8) / 8;
  std::vector<int32_t> act_b(na), slot_b(na);
  for (int64_t x = 0; x * per < na; ++x) {
    const int64_t lo = x * per, hi = std::min<int64_t>(na, (x + 1) * per);
    std::vector<std::pair<int, int32_t>> v;
    for (int64_t i = lo; i < hi; ++i) v.push_back({fl[i], act[i]});
    std::stable_sort(v.begin(), v.end(), [&](auto& p, auto& q) { return rank_a(p.first) < rank_a(q.first); });
    std::vector<int64_t> ib(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
      act[lo + i] = v[i].second;
      fl[lo + i] = (uint8_t)v[i].first;
      ib[i] = (int64_t)i;
    }
    std::stable_sort(ib.begin(), ib.end(), [&](int64_t p, int64_t q) { return rank_b(v[p].first) < rank_b(v[q].first); });
    for (size_t i = 0; i < v.size(); ++i) {
      act_b[lo + i] = v[ib[i]].second;
      slot_b[lo + i] = (int32_t)(lo + ib[i]);
    }
  }
  MT_HIP(hipMemcpyAsync(l.active, act.data(), 4 * na, hipMemcpyHostToDevice, c.stream));
  // the bricks with some lat_f cell or C2F cell in dis (a brick of C2F cells
  // only has nothing to do in a sub-step that drops their stream)
  std::vector<int32_t> act_a, slot_a;
  for (int64_t i = 0; i < na; ++i) {
    if (fl[i] & 8) {
      act_a.push_back(act[i]);
      slot_a.push_back((int32_t)i);
    }
  }
  l.n_active_a = (int64_t)act_a.size();
  if (l.n_active_a) {
    l.active_a = c.alloc<int32_t>(l.n_active_a);
    l.slot_a = c.alloc<int32_t>(l.n_active_a);
    if (!l.active_a || !l.slot_a) {
      set_error("out of device memory");
      return MT_ERR_HIP;
    }
    MT_HIP(hipMemcpyAsync(l.active_a, act_a.data(), 4 * l.n_active_a, hipMemcpyHostToDevice, c.stream));
    MT_HIP(hipMemcpyAsync(l.slot_a, slot_a.data(), 4 * l.n_active_a, hipMemcpyHostToDevice, c.stream));
  }
  if (l.cpar) {
    l.active_b = c.alloc<int32_t>(na);
    l.slot_b = c.alloc<int32_t>(na);
    if (!l.active_b || !l.slot_b) {
      set_error("out of device memory");
      return MT_ERR_HIP;
    }
    MT_HIP(hipMemcpyAsync(l.active_b, act_b.data(), 4 * na, hipMemcpyHostToDevice, c.stream));
    MT_HIP(hipMemcpyAsync(l.slot_b, slot_b.data(), 4 * na, hipMemcpyHostToDevice, c.stream));
  }
  MT_HIP(hipStreamSynchronize(c.stream));
  return MT_OK;
}

static int compact_list(Case& c, const int* d_flag, int64_t n, int32_t** out, int64_t* count) {
  int* pos = nullptr;
  MT_HIP(dev_malloc((void**)&pos, sizeof(int) * (n + 1), c.stream));
  int64_t tot;
  int r = scan_exclusive(c, d_flag, pos, n, &tot);
  if (r) return r;
  *out = c.alloc<int32_t>(tot);
  if (!*out) { set_error("hipMalloc failed"); return MT_ERR_HIP; }
  k_emit<<<nblk(n), 256, 0, c.stream>>>(d_flag, pos, n, *out);
  MT_HIP(hipStreamSynchronize(c.stream));
  dev_free(pos, c.stream);
  *count = tot;
  return MT_OK;
}

static LevelView view(const Case& c, int L) {
  LevelView v;
  v.bcode = c.lv[L].bcode;
  v.kind = c.lv[L].kind;
  v.n_bricks = c.lv[L].n_bricks;
  v.shift = 3 * (c.Lmax - L);
  return v;
}

int find_slots(Case& c, int L, const uint64_t* d_keys, int64_t n, int32_t* d_out) {
  if (n <= 0) return MT_OK;
  k_find<<<nblk(n), 256, 0, c.stream>>>(view(c, L), d_keys, n, d_out);
  MT_HIP(hipGetLastError());
  return MT_OK;
}

int build_tables(Case& c) {
  int* d_err = c.alloc<int>(1);
  MT_HIP(hipMemsetAsync(d_err, 0, sizeof(int), c.stream));
  for (int L = 0; L <= c.Lmax; ++L) {
    DevLevel& l = c.lv[L];
    const int64_t n = l.n_up;
    if (n <= 0) { set_error("level " + std::to_string(L) + " has no cells"); return MT_ERR_ARG; }
    for (int a = 0; a < 3; ++a) l.dims[a] = c.d.dims0[a] << L;
    // the uploaded keys and kinds, staged on the device by mt_case_upload_level
    const uint64_t* d_keys = l.up_keys;
    const uint16_t* d_kinds = l.up_kinds;
    uint64_t* ks = nullptr;
    int32_t* vs = nullptr;
    int r = radix_sort(c, d_keys, n, l.up_max, &ks, &vs);
    if (r) return r;
    int shift = 3 * (c.Lmax - L);
    int *head = nullptr, *bidx = nullptr;
    MT_HIP(dev_malloc((void**)&head, 4 * n, c.stream));
    MT_HIP(dev_malloc((void**)&bidx, 4 * n, c.stream));
    k_codes<<<nblk(n), 256, 0, c.stream>>>(ks, n, shift, head, d_err);
    int64_t nb;
    r = scan_exclusive(c, head, bidx, n, &nb);
    if (r) return r;
    int herr = 0;
    MT_HIP(hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, c.stream));
    MT_HIP(hipStreamSynchronize(c.stream));
    if (herr & 3) {
      set_error("level " + std::to_string(L) + ": " +
                ((herr & 2) ? "duplicate key in the uploaded set" : "key is not a cell of this level"));
      return MT_ERR_ARG;
    }
    l.n_cells = n;
    l.n_bricks = nb;
    l.n_slots = nb * 64;
    l.bcode = c.alloc<uint64_t>(nb);
    l.bnbr = c.alloc<int32_t>(nb * 27);
    l.key = c.alloc<uint64_t>(l.n_slots);
    l.kind = c.alloc<uint16_t>(l.n_slots);
    l.pmask = c.alloc<uint32_t>(l.n_slots);
    l.dis_row = c.alloc<int32_t>(l.n_slots);
    if (!l.bcode || !l.bnbr || !l.key || !l.kind || !l.pmask || !l.dis_row) {
      set_error("hipMalloc failed for level tables");
      return MT_ERR_HIP;
    }
    k_fill_empty<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.key, l.kind, l.n_slots);
    k_fill_i32<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.dis_row, l.n_slots, -1);
    k_place<<<nblk(n), 256, 0, c.stream>>>(ks, vs, d_kinds, head, bidx, n, shift, l.bcode, l.key, l.kind);
    l.bmember = c.alloc<int32_t>(nb);
    if (!l.bmember) {
      set_error("hipMalloc failed for level tables");
      return MT_ERR_HIP;
    }
    k_bmember<<<nblk(nb), 256, 0, c.stream>>>(l.bcode, nb, c.d_mem, (int)c.h_mem.size(), L, l.bmember, d_err);
    k_bnbr<<<nblk(nb), 256, 0, c.stream>>>(l.bcode, nb, l.bmember, c.d_mem, L, l.bnbr);
    k_pmask<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.bnbr, l.kind, l.n_slots, c.Q, l.pmask);
    MT_HIP(hipGetLastError());
    MT_HIP(hipStreamSynchronize(c.stream));
    drop_level_upload(l, c.stream);
    dev_free(ks, c.stream);
    dev_free(vs, c.stream);
    dev_free(head, c.stream);
    dev_free(bidx, c.stream);
  }
  // dis (finest level)
  int Lm = c.Lmax;
  DevLevel& lm = c.lv[Lm];
  uint64_t* d_dis_sorted = nullptr;
  const int64_t nd = c.n_up_dis;
  if (nd > 0) {
    int m = c.Q - 1;
    uint64_t* ks = nullptr;
    int32_t* vs = nullptr;
    // the uploaded dis, staged on the device by mt_case_upload_dis
    const double* d_v = c.up_dis_vals;
    int r = radix_sort(c, c.up_dis_keys, nd, c.up_dis_max, &ks, &vs);
    if (r) return r;
    lm.dis = c.alloc<double>(nd * m);
    lm.n_dis = nd;
    k_dis_rows<<<nblk(nd), 256, 0, c.stream>>>(ks, vs, nd, view(c, Lm), d_v, m, lm.dis_row, lm.dis, lm.kind);
    lm.wmask = c.alloc<uint32_t>(lm.n_slots);
    if (!lm.wmask) {
      set_error("out of device memory");
      return MT_ERR_HIP;
    }
    k_wmask<<<nblk(lm.n_slots), 256, 0, c.stream>>>(lm.dis_row, lm.dis, lm.n_slots, c.Q, lm.wmask);
    d_dis_sorted = c.alloc<uint64_t>(nd);
    MT_HIP(hipMemcpyAsync(d_dis_sorted, ks, 8 * nd, hipMemcpyDeviceToDevice, c.stream));
    MT_HIP(hipStreamSynchronize(c.stream));
    drop_dis_upload(c, c.stream);
    dev_free(ks, c.stream);
    dev_free(vs, c.stream);
  }
  // lists
  for (int L = 0; L <= c.Lmax; ++L) {
    DevLevel& l = c.lv[L];
    int* flag = nullptr;
    int64_t nmax = std::max(l.n_slots, l.n_bricks);
    MT_HIP(dev_malloc((void**)&flag, 4 * (nmax + 1), c.stream));
    k_flag_brick<<<nblk(l.n_bricks), 256, 0, c.stream>>>(l.kind, l.n_bricks, MT_KIND_FLUID | MT_KIND_C2F, flag);
    int r = compact_list(c, flag, l.n_bricks, &l.active, &l.n_active);
    if (r) return r;
    k_flag_kind<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.kind, l.n_slots, MT_KIND_F2C, flag);
    r = compact_list(c, flag, l.n_slots, &l.f2c, &l.n_f2c);
    if (r) return r;
    k_flag_kind<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.kind, l.n_slots, MT_KIND_GHOST_F2C, flag);
    r = compact_list(c, flag, l.n_slots, &l.gf2c, &l.n_gf2c);
    if (r) return r;
    k_flag_kind<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.kind, l.n_slots, MT_KIND_BCRES, flag);
    r = compact_list(c, flag, l.n_slots, &l.bcres, &l.n_bcres);
    if (r) return r;
    if (L == c.Lmax) {
      k_flag_ib<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.kind, l.dis_row, l.n_slots, flag);
      r = compact_list(c, flag, l.n_slots, &l.ibl, &l.n_ibl);
      if (r) return r;
    }
    MT_HIP(hipStreamSynchronize(c.stream));
    dev_free(flag, c.stream);
    if (L < c.Lmax && (l.n_f2c || l.n_gf2c)) {
      l.f2c_child = c.alloc<int32_t>(l.n_f2c * 8);
      l.f2c_child_write = c.alloc<int32_t>(l.n_f2c * 8);
      l.gf2c_child = c.alloc<int32_t>(l.n_gf2c * 8);
      LevelView fine = view(c, L + 1);
      int pshift = c.Lmax - L;
      if (l.n_f2c)
        k_children<<<nblk(l.n_f2c), 256, 0, c.stream>>>(l.f2c, l.n_f2c, l.key, pshift, fine, c.lv[L + 1].kind,
                                                        d_dis_sorted, nd, 0, l.f2c_child, l.f2c_child_write, d_err);
      if (l.n_gf2c)
        k_children<<<nblk(l.n_gf2c), 256, 0, c.stream>>>(l.gf2c, l.n_gf2c, l.key, pshift, fine,
                                                         c.lv[L + 1].kind, d_dis_sorted, nd, 1, nullptr,
                                                         l.gf2c_child, d_err);
      l.n_xp = l.n_f2c + l.n_gf2c;
      l.xp_list = c.alloc<int32_t>(l.n_xp);
      l.xp_child = c.alloc<int32_t>(l.n_xp * 8);
      if (l.n_f2c) {
        MT_HIP(hipMemcpyAsync(l.xp_list, l.f2c, 4 * l.n_f2c, hipMemcpyDeviceToDevice, c.stream));
        MT_HIP(hipMemcpyAsync(l.xp_child, l.f2c_child_write, 32 * l.n_f2c, hipMemcpyDeviceToDevice, c.stream));
      }
      if (l.n_gf2c) {
        MT_HIP(hipMemcpyAsync(l.xp_list + l.n_f2c, l.gf2c, 4 * l.n_gf2c, hipMemcpyDeviceToDevice, c.stream));
        MT_HIP(hipMemcpyAsync(l.xp_child + 8 * l.n_f2c, l.gf2c_child, 32 * l.n_gf2c, hipMemcpyDeviceToDevice,
                              c.stream));
      }
    } else if (L == c.Lmax && l.n_f2c) {
      set_error("F2C cells on the finest level");
      return MT_ERR_ARG;
    }
  }
  int herr = 0;
  MT_HIP(hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipStreamSynchronize(c.stream));
  if (!herr) {
    for (int L = 0; L < c.Lmax; ++L) {
      DevLevel& l = c.lv[L];
      DevLevel& lf = c.lv[L + 1];
      if (!l.n_f2c) continue;
      lf.cpar = c.alloc<int32_t>(lf.n_slots / 8);
      if (!lf.cpar) {
        set_error("out of device memory");
        return MT_ERR_HIP;
      }
      int* d_bad = nullptr;
      MT_HIP(dev_malloc((void**)&d_bad, 4, c.stream));
      MT_HIP(hipMemsetAsync(d_bad, 0, 4, c.stream));
      k_fill_i32<<<nblk(lf.n_slots / 8), 256, 0, c.stream>>>(lf.cpar, lf.n_slots / 8, -1);
      k_cpar<<<nblk(l.n_f2c), 256, 0, c.stream>>>(l.f2c, l.f2c_child, l.n_f2c, lf.kind, lf.dis_row, lf.cpar, d_bad);
      MT_HIP(hipGetLastError());
      int hbad = 1;
      MT_HIP(hipMemcpyAsync(&hbad, d_bad, 4, hipMemcpyDeviceToHost, c.stream));
      MT_HIP(hipStreamSynchronize(c.stream));
      dev_free(d_bad, c.stream);
      l.coal_in_fused = hbad == 0;
    }
    // explosions inside the fused updates: needs the in-kernel coalescence on
    // every interface
    const bool both = (c.d.options & MT_OPT_EXPLODE_BOTH) != 0;
    bool xin = c.Lmax >= 1;
    for (int L = 0; L < c.Lmax && xin; ++L) xin = c.lv[L].coal_in_fused;
    for (int L = 0; L <= c.Lmax && xin; ++L) {
      DevLevel& l = c.lv[L];
      if (L >= 1) {
        const uint32_t* wm = L == c.Lmax ? l.wmask : nullptr;
        // every target, and (reference explosion) the ghost_C2F targets alone
        for (int g = 0; g < (both ? 1 : 2); ++g) {
          const uint16_t targets = g ? MT_KIND_GHOST_C2F : (MT_KIND_C2F | MT_KIND_GHOST_C2F);
          uint32_t*& mask = g ? l.tmask_g : l.tmask;
          mask = c.alloc<uint32_t>(l.n_slots);
          int* flag = nullptr;
          MT_HIP(dev_malloc((void**)&flag, 4 * (l.n_slots + 1), c.stream));
          k_flag_target<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.kind, l.n_slots, targets, flag);
          int r = g ? compact_list(c, flag, l.n_slots, &l.xtg_list, &l.n_xtg)
                    : compact_list(c, flag, l.n_slots, &l.xt_list, &l.n_xt);
          if (r) return r;
          MT_HIP(hipStreamSynchronize(c.stream));
          dev_free(flag, c.stream);
          if (!mask) {
            set_error("out of device memory");
            return MT_ERR_HIP;
          }
          k_tmask<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.bnbr, l.kind, l.pmask, wm, l.n_slots, c.Q, targets, mask);
        }
        for (int j = 0; j < 2; ++j) {
          l.xg[j] = c.alloc<char>((l.n_slots / 8) * c.Q * (int64_t)c.pop_bytes());
          if (!l.xg[j]) {
            set_error("out of device memory");
            return MT_ERR_HIP;
          }
        }
      }
      if (L < c.Lmax && l.n_gf2c) {
        l.gxrow = c.alloc<int32_t>(l.n_slots);
        if (!l.gxrow) {
          set_error("out of device memory");
          return MT_ERR_HIP;
        }
        k_fill_i32<<<nblk(l.n_slots), 256, 0, c.stream>>>(l.gxrow, l.n_slots, -1);
        k_gxrow<<<nblk(l.n_gf2c), 256, 0, c.stream>>>(l.gf2c, l.n_gf2c, l.gxrow);
      }
      MT_HIP(hipGetLastError());
    }
    MT_HIP(hipStreamSynchronize(c.stream));
    c.xin = xin;
    for (int L = 0; L <= c.Lmax; ++L) {
      int r = order_active(c, L);
      if (r) return r;
    }
  }
  if (herr & 4) {
    set_error("an F2C cell has a child missing on the finer level");
    return MT_ERR_LOOKUP;
  }
  if (herr & 8) {
    set_error("a cell lies outside every member's box");
    return MT_ERR_ARG;
  }
  return MT_OK;
}


// ------------------------------------------------------------------ moving markers
// Marker positions are in finest-level cell units (cell i spans [i, i + 1)).  A
// marker's kernel cells are the nodes i with |i + 1/2 - x| < 3/2 per axis, in
// the candidate order z, y, x outermost to innermost, and its weights the
// product of the three-point kernel of Roma, Peskin & Berger (J. Comput. Phys.
// 153, 1999) over the axes; the arithmetic is the host builder's (no
// contraction, the same operation order), so both give the same bits.

namespace {

__device__ inline double mv_roma3(double r) {
#pragma clang fp contract(off)
  r = fabs(r);
  if (r <= 0.5) return (1.0 + sqrt(1.0 - 3.0 * (r * r))) / 3.0;
  if (r < 1.5) {
    const double t = 1.0 - r;
    return (5.0 - 3.0 * r - sqrt(fmax(0.0, 1.0 - 3.0 * (t * t)))) / 6.0;
  }
  return 0.0;
}

// rigid motion over dt: x' = center + vel dt + Rot(omega dt)(x - center)
// (Rodrigues), and the velocity of the moved point, vel + omega x (x' - center
// - vel dt), stored in the reported frame (negated)
__global__ void k_mv_rigid(const double* x, int64_t n, double dt, double cx, double cy, double cz, double vx,
                           double vy, double vz, double wx, double wy, double wz, double* xo, double* vo) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double rx = x[3 * m] - cx, ry = x[3 * m + 1] - cy, rz = x[3 * m + 2] - cz;
  const double wn = sqrt(wx * wx + wy * wy + wz * wz);
  double px = rx, py = ry, pz = rz;
  if (wn > 0.0) {
    const double kx = wx / wn, ky = wy / wn, kz = wz / wn, th = wn * dt;
    const double cs = cos(th), sn = sin(th), kr = kx * rx + ky * ry + kz * rz;
    px = rx * cs + (ky * rz - kz * ry) * sn + kx * kr * (1.0 - cs);
    py = ry * cs + (kz * rx - kx * rz) * sn + ky * kr * (1.0 - cs);
    pz = rz * cs + (kx * ry - ky * rx) * sn + kz * kr * (1.0 - cs);
  }
  xo[3 * m] = cx + vx * dt + px;
  xo[3 * m + 1] = cy + vy * dt + py;
  xo[3 * m + 2] = cz + vz * dt + pz;
  vo[3 * m] = -(vx + (wy * pz - wz * py));
  vo[3 * m + 1] = -(vy + (wz * px - wx * pz));
  vo[3 * m + 2] = -(vz + (wx * py - wy * px));
}

// flags: [0] markers whose support is invalid, [1] the first of them, [2] some
// marker's kernel cells changed, [3] used entries
__global__ void k_mv_stencil(const double* x, int64_t n, LevelView v, const int32_t* cur_slot, const double* cur_w,
                             int32_t* st_slot, double* st_w, int32_t* flags) {
#pragma clang fp contract(off)
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double cx = x[3 * m] - 0.5, cy = x[3 * m + 1] - 0.5, cz = x[3 * m + 2] - 0.5;
  const double hi = (double)(1 << 20);
  bool ok = cx >= 1.0 && cy >= 1.0 && cz >= 1.0 && cx < hi && cy < hi && cz < hi;  // false for NaN
  int cnt = 0;
  int32_t* sl_out = st_slot + 27 * m;
  double* w_out = st_w + 27 * m;
  if (ok) {
    const int64_t bx = (int64_t)floor(cx) - 1, by = (int64_t)floor(cy) - 1, bz = (int64_t)floor(cz) - 1;
    for (int k = 0; k < 4 && ok; ++k)
      for (int j = 0; j < 4 && ok; ++j)
        for (int i = 0; i < 4 && ok; ++i) {
          const double w = mv_roma3((double)(bx + i) - cx) * mv_roma3((double)(by + j) - cy) *
                           mv_roma3((double)(bz + k) - cz);
          if (!(w > 0.0)) continue;
          const int32_t sl = cnt < 27 ? find_slot(v, morton3(bx + i, by + j, bz + k)) : -1;
          const uint16_t kd = sl >= 0 ? v.kind[sl] : 0;
          if (sl < 0 || !(kd & MT_KIND_FLUID) || (kd & (MT_KIND_C2F | MT_KIND_DIS | MT_KIND_SOLID))) {
            ok = false;
            break;
          }
          sl_out[cnt] = sl;
          w_out[cnt] = w;
          ++cnt;
        }
  }
  if (!ok || cnt == 0) {
    atomicAdd(&flags[0], 1);
    atomicMin(&flags[1], (int32_t)m);
    return;
  }
  for (int e = cnt; e < 27; ++e) {
    sl_out[e] = sl_out[0];
    w_out[e] = 0.0;
  }
  bool changed = false;
  for (int e = 0; e < 27; ++e)
    changed |= sl_out[e] != cur_slot[27 * m + e] || (w_out[e] != 0.0) != (cur_w[27 * m + e] != 0.0);
  if (changed) atomicOr(&flags[2], 1);
  atomicAdd(&flags[3], cnt);
}

// the same kernel cells: new weights in place
__global__ void k_mv_reweight(const double* w, int64_t ne, const int32_t* ent, double* mk_w, double* band_w) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  mk_w[e] = w[e];
  if (ent[e] >= 0) band_w[ent[e]] = w[e];
}

// sort keys (slot, entry); unused entries sort last
__global__ void k_mv_keys(const int32_t* slot, const double* w, int64_t ne, uint64_t last, uint64_t* keys) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  keys[e] = ((w[e] != 0.0 ? (uint64_t)slot[e] : last) << 32) | (uint64_t)e;
}

__global__ void k_mv_heads(const uint64_t* ks, int64_t nv, int64_t ne, int* head) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  head[i] = i < nv && (i == 0 || (ks[i] >> 32) != (ks[i - 1] >> 32));
}

__global__ void k_mv_emit(const uint64_t* ks, const int* head, const int* pos, int64_t nv, const double* w,
                          int32_t* band_slot, int32_t* band_ptr, int32_t* band_mk, double* band_w, int32_t* mk_band,
                          int32_t* mk_ent) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int64_t e = (int64_t)(ks[i] & 0xFFFFFFFFull);
  const int32_t b = pos[i] + head[i] - 1;
  if (head[i]) {
    band_slot[b] = (int32_t)(ks[i] >> 32);
    band_ptr[b] = (int32_t)i;
  }
  if (i == nv - 1) band_ptr[b + 1] = (int32_t)nv;
  band_mk[i] = (int32_t)(e / 27);
  band_w[i] = w[e];
  mk_band[e] = b;
  mk_ent[e] = (int32_t)i;
}

// unused entries point at the band cell of the marker's first used entry
__global__ void k_mv_unused(const double* w, int64_t n, int32_t* mk_band, int32_t* mk_ent) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  int32_t first = -1;
  for (int j = 0; j < 27; ++j) {
    const int64_t e = 27 * m + j;
    if (w[e] != 0.0) {
      if (first < 0) first = mk_band[e];
    } else {
      mk_band[e] = first;  // entry 0 is always used (a stencil lists its used cells first)
      mk_ent[e] = -1;
    }
  }
}

__global__ void k_mv_rows(const int32_t* band_slot, const int32_t* n_band, int32_t* row, int clear) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= *n_band) return;
  row[band_slot[b]] = clear ? -1 : (int32_t)b;
}

}  // namespace

int move_markers(Case& c, bool rigid, double dt, const double* center, const double* vel, const double* omega) {
  const int64_t n = c.n_mk, ne = 27 * n;
  if (rigid)
    k_mv_rigid<<<nblk(n), 256, 0, c.stream>>>(c.mk_x, n, dt, center[0], center[1], center[2], vel[0], vel[1], vel[2],
                                              omega[0], omega[1], omega[2], c.mv_x, c.mv_v);
  const int32_t init[4] = {0, INT32_MAX, 0, 0};
  int32_t fl[4];
  MT_HIP(hipMemcpyAsync(c.mv_flags, init, sizeof init, hipMemcpyHostToDevice, c.stream));
  k_mv_stencil<<<nblk(n), 256, 0, c.stream>>>(c.mv_x, n, view(c, c.Lmax), c.mk_slot, c.mk_w, c.mv_slot, c.mv_w,
                                              c.mv_flags);
  MT_HIP(hipGetLastError());
  MT_HIP(hipMemcpyAsync(fl, c.mv_flags, sizeof fl, hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipStreamSynchronize(c.stream));
  if (fl[0]) {
    set_error("marker " + std::to_string(fl[1]) +
              ": a kernel cell is not a streamed fluid cell of the finest level (markers not moved)");
    return MT_ERR_ARG;
  }
  if (!fl[2]) {
    k_mv_reweight<<<nblk(ne), 256, 0, c.stream>>>(c.mv_w, ne, c.mk_ent, c.mk_w, c.band_w);
  } else {
    // rebuild the band: entries sorted by (slot, entry), one band cell per run
    const int64_t nv = fl[3];
    uint64_t* keys = nullptr;
    int* head = nullptr;
    int* pos = nullptr;
    MT_HIP(dev_malloc((void**)&keys, 8 * ne, c.stream));
    MT_HIP(dev_malloc((void**)&head, sizeof(int) * (ne + 1), c.stream));
    MT_HIP(dev_malloc((void**)&pos, sizeof(int) * (ne + 1), c.stream));
    const uint64_t last = (uint64_t)c.lv[c.Lmax].n_slots;
    k_mv_keys<<<nblk(ne), 256, 0, c.stream>>>(c.mv_slot, c.mv_w, ne, last, keys);
    uint64_t* ks = nullptr;
    int32_t* vs = nullptr;
    int r = radix_sort(c, keys, ne, (last << 32) | (uint64_t)ne, &ks, &vs);
    if (r) return r;
    k_mv_heads<<<nblk(ne), 256, 0, c.stream>>>(ks, nv, ne, head);
    int64_t nb = 0;
    r = scan_exclusive(c, head, pos, ne, &nb);
    if (r) return r;
    k_mv_rows<<<nblk(c.band_cap), 256, 0, c.stream>>>(c.band_slot, c.d_nband, c.band_row, 1);
    k_mv_emit<<<nblk(nv), 256, 0, c.stream>>>(ks, head, pos, nv, c.mv_w, c.band_slot, c.band_ptr, c.band_mk,
                                              c.band_w, c.mk_band, c.mk_ent);
    k_mv_unused<<<nblk(n), 256, 0, c.stream>>>(c.mv_w, n, c.mk_band, c.mk_ent);
    k_fill_i32<<<1, 256, 0, c.stream>>>(c.d_nband, 1, (int32_t)nb);
    k_mv_rows<<<nblk(c.band_cap), 256, 0, c.stream>>>(c.band_slot, c.d_nband, c.band_row, 0);
    MT_HIP(hipMemcpyAsync(c.mk_w, c.mv_w, 8 * ne, hipMemcpyDeviceToDevice, c.stream));
    MT_HIP(hipGetLastError());
    MT_HIP(hipStreamSynchronize(c.stream));
    dev_free(ks, c.stream);
    dev_free(vs, c.stream);
    dev_free(keys, c.stream);
    dev_free(head, c.stream);
    dev_free(pos, c.stream);
    c.n_band = nb;
  }
  MT_HIP(hipMemcpyAsync(c.mk_slot, c.mv_slot, 4 * ne, hipMemcpyDeviceToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.mk_x, c.mv_x, 8 * 3 * n, hipMemcpyDeviceToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.mk_v, c.mv_v, 8 * 3 * n, hipMemcpyDeviceToDevice, c.stream));
  MT_HIP(hipGetLastError());
  MT_HIP(hipStreamSynchronize(c.stream));
  c.mk_x_set = true;
  return MT_OK;
}

}  // namespace mt
