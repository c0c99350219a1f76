// mt_voxel.hip -- the per-level cell sets built on the GPU from a geometry.
//
// The role of Grid_Manager::initial + Lat_Manager::voxelize
// (Grid_Manager.cpp:40-134, Lat_Voxelize.cpp:17-1285): from a solid body, the
// domain and the refinement shells, produce every level's cells with their kinds
// (lat_f, lat_overlap_F2C / C2F, ghost_overlap_C2F / F2C, lat_sf, lat_bc;
// header/Lat_Manager.hpp:245-316) and the finest level's link fractions `dis`.
// The body is an analytic sphere or a closed triangle mesh (an STL surface):
//   * inside:  r < R, or the mesh's generalised winding number above 1/2
//              (Van Oosterom-Strackee solid angles);
//   * refine:  level L+1 covers the level-L cells inside the body or whose
//              centre lies within shells[L] (level-0 units) of its surface
//              (for the sphere r <= R + shells[L]);
//   * links:   the fraction t in [0, 1] along x -> x + h e_q of the first
//              crossing of the surface from a cell centre outside the body
//              (the sphere's closed form, or Moller-Trumbore over the
//              triangles as Shape.cpp:809-843 does), -1 where there is none.
// Every level lives on the device as a dense grid of flags during the build;
// the rules are set operations (26- or Q-neighbourhood dilations, parents,
// faces), one kernel each, and the sets leave as Morton keys.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mt_case.h"
#include "mt_device.h"
#include "mt_morton.h"

// the geometry's arithmetic is evaluated as written (no contraction into fused
// multiply-adds), so the analytic sphere gives the host builder's bits
#pragma clang fp contract(off)

#define VX_TRY(x)                 \
  do {                            \
    int r_ = (x);                 \
    if (r_ != MT_OK) return r_;   \
  } while (0)

namespace mt {
namespace {

// working flags of a dense level grid
enum : uint32_t {
  W_D = 1u << 0,         // the level has this cell
  W_REF = 1u << 1,       // the geometry asks for refinement here
  W_R = 1u << 2,         // covered by the finer level
  W_BAND = 1u << 3,      // F2C: refined cell next to an unrefined one
  W_INT = 1u << 4,       // refined, not F2C
  W_C2F = 1u << 5,       // child of a coarser F2C cell
  W_IN = 1u << 6,        // centre inside the body
  W_SHELL = 1u << 7,     // level-0 reservoir shell
  W_SOL = 1u << 8,       // solid (finest level)
  W_FLUID = 1u << 9,
  W_GHOST = 1u << 10,
};

struct Grid {
  int n[3];
  int L;
  __host__ __device__ int64_t cells() const { return (int64_t)n[0] * n[1] * n[2]; }
};

__device__ __forceinline__ void unlin(int64_t i, const Grid& g, int& x, int& y, int& z) {
  z = (int)(i % g.n[2]);
  const int64_t t = i / g.n[2];
  y = (int)(t % g.n[1]);
  x = (int)(t / g.n[1]);
}
__device__ __forceinline__ int64_t lin(int x, int y, int z, const Grid& g) {
  return ((int64_t)x * g.n[1] + y) * g.n[2] + z;
}
// neighbour (x+dx, y+dy, z+dz) with wrap on periodic axes, -1 outside
__device__ __forceinline__ int64_t nbr(int x, int y, int z, int dx, int dy, int dz, const Grid& g, int per) {
  int c[3] = {x + dx, y + dy, z + dz};
  for (int a = 0; a < 3; ++a) {
    if (c[a] < 0 || c[a] >= g.n[a]) {
      if (!((per >> a) & 1)) return -1;
      c[a] = (c[a] + g.n[a]) % g.n[a];
    }
  }
  return lin(c[0], c[1], c[2], g);
}

// ---- geometry
struct Geom {
  int shape;  // MT_SHAPE_*
  double c[3], r;
  const double* tri;  // [n][9]
  int64_t n_tri;
};

__device__ double tri_solid_angle(const double* t, const double* p) {
  double a[3], b[3], c[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = t[k] - p[k];
    b[k] = t[3 + k] - p[k];
    c[k] = t[6 + k] - p[k];
  }
  const double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  const double lb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double lc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                     a[2] * (b[0] * c[1] - b[1] * c[0]);
  const double den = la * lb * lc + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * lc +
                     (a[0] * c[0] + a[1] * c[1] + a[2] * c[2]) * lb + (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) * la;
  return 2.0 * atan2(det, den);
}

// squared distance from p to the triangle (closest point by regions, Ericson 5.1.5)
__device__ double tri_dist2(const double* t, const double* p) {
  const double* A = t;
  const double* B = t + 3;
  const double* C = t + 6;
  double ab[3], ac[3], ap[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = B[k] - A[k];
    ac[k] = C[k] - A[k];
    ap[k] = p[k] - A[k];
  }
  auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  auto d2 = [&](const double* q) {
    const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return dx * dx + dy * dy + dz * dz;
  };
  const double d1 = dot(ab, ap), d2v = dot(ac, ap);
  if (d1 <= 0 && d2v <= 0) return d2(A);
  double bp[3];
  for (int k = 0; k < 3; ++k) bp[k] = p[k] - B[k];
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0 && d4 <= d3) return d2(B);
  double q[3];
  const double vc = d1 * d4 - d3 * d2v;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    const double v = d1 / (d1 - d3);
    for (int k = 0; k < 3; ++k) q[k] = A[k] + v * ab[k];
    return d2(q);
  }
  double cp[3];
  for (int k = 0; k < 3; ++k) cp[k] = p[k] - C[k];
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0 && d5 <= d6) return d2(C);
  const double vb = d5 * d2v - d1 * d6;
  if (vb <= 0 && d2v >= 0 && d6 <= 0) {
    const double w = d2v / (d2v - d6);
    for (int k = 0; k < 3; ++k) q[k] = A[k] + w * ac[k];
    return d2(q);
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int k = 0; k < 3; ++k) q[k] = B[k] + w * (C[k] - B[k]);
    return d2(q);
  }
  const double den = 1.0 / (va + vb + vc);
  const double v = vb * den, w = vc * den;
  for (int k = 0; k < 3; ++k) q[k] = A[k] + ab[k] * v + ac[k] * w;
  return d2(q);
}

// Moller-Trumbore: parameter t of the crossing of p + t d with the triangle, or -1
__device__ double tri_ray(const double* t, const double* p, const double* d) {
  double e1[3], e2[3], s[3];
  for (int k = 0; k < 3; ++k) {
    e1[k] = t[3 + k] - t[k];
    e2[k] = t[6 + k] - t[k];
    s[k] = p[k] - t[k];
  }
  const double h[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
  const double a = e1[0] * h[0] + e1[1] * h[1] + e1[2] * h[2];
  if (fabs(a) < 1e-300) return -1.0;
  const double f = 1.0 / a;
  const double u = f * (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]);
  if (u < 0.0 || u > 1.0) return -1.0;
  const double qv[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
  const double v = f * (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]);
  if (v < 0.0 || u + v > 1.0) return -1.0;
  return f * (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]);
}

__device__ bool geom_inside(const Geom& G, const double* p) {
  if (G.shape == MT_SHAPE_SPHERE) {
    const double dx = p[0] - G.c[0], dy = p[1] - G.c[1], dz = p[2] - G.c[2];
    return sqrt(dx * dx + dy * dy + dz * dz) < G.r;
  }
  double w = 0;
  for (int64_t i = 0; i < G.n_tri; ++i) w += tri_solid_angle(G.tri + 9 * i, p);
  return fabs(w) > 2.0 * M_PI;  // winding number above 1/2
}

// inside the body, or within `shell` of its surface
__device__ bool geom_refine(const Geom& G, const double* p, double shell, bool inside) {
  if (G.shape == MT_SHAPE_SPHERE) {
    const double dx = p[0] - G.c[0], dy = p[1] - G.c[1], dz = p[2] - G.c[2];
    return sqrt(dx * dx + dy * dy + dz * dz) <= G.r + shell;
  }
  if (inside) return true;
  const double s2 = shell * shell;
  for (int64_t i = 0; i < G.n_tri; ++i)
    if (tri_dist2(G.tri + 9 * i, p) <= s2) return true;
  return false;
}

// link fraction from an outside centre p along h e, or -1
__device__ double geom_link(const Geom& G, const double* p, int ex, int ey, int ez, double h) {
  if (G.shape == MT_SHAPE_SPHERE) {
    // the same arithmetic as the analytic host builder (cases.sphere_ib_dist)
    const double dx = p[0] - G.c[0], dy = p[1] - G.c[1], dz = p[2] - G.c[2];
    const double a = h * h * (double)(ex * ex + ey * ey + ez * ez);
    const double b = 2.0 * h * (ex * dx + ey * dy + ez * dz);
    const double c = dx * dx + dy * dy + dz * dz - G.r * G.r;
    const double disc = b * b - 4.0 * a * c;
    if (!(disc >= 0) || !(c > 0)) return -1.0;
    const double t = (-b - sqrt(disc)) / (2.0 * a);
    return (t >= 0.0 && t <= 1.0) ? t : -1.0;
  }
  const double d[3] = {h * ex, h * ey, h * ez};
  double best = 2.0;
  for (int64_t i = 0; i < G.n_tri; ++i) {
    const double t = tri_ray(G.tri + 9 * i, p, d);
    if (t >= 0.0 && t < best) best = t;
  }
  return best <= 1.0 ? best : -1.0;
}

__device__ __forceinline__ void centre(int x, int y, int z, int L, double* p) {
  const double h = 1.0 / (double)(1 << L);
  p[0] = (x + 0.5) * h;
  p[1] = (y + 0.5) * h;
  p[2] = (z + 0.5) * h;
}

// ---- kernels (one thread per cell of the dense level grid)
#define CELL_LOOP(g)                                        \
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; \
  if (i >= (g).cells()) return;                             \
  int x, y, z;                                              \
  unlin(i, (g), x, y, z)

// D, C2F from the coarser level; the reservoir shell; inside / refine
__global__ void k_vx_init(Grid g, uint32_t* W, const uint32_t* Wc, Grid gc, int Lmax, int reservoir, Geom G,
                          double shell, int solid) {
  CELL_LOOP(g);
  uint32_t w = 0;
  if (g.L == 0) {
    w |= W_D;
    if (reservoir && (x == 0 || y == 0 || z == 0 || x == g.n[0] - 1 || y == g.n[1] - 1 || z == g.n[2] - 1))
      w |= W_SHELL;
  } else {
    const uint32_t p = Wc[lin(x >> 1, y >> 1, z >> 1, gc)];
    if (p & W_R) w |= W_D;
    if (p & W_BAND) w |= W_C2F;
  }
  if (w & W_D) {
    double c[3];
    centre(x, y, z, g.L, c);
    const bool in = solid ? geom_inside(G, c) : false;
    if (in) w |= W_IN;
    if (g.L < Lmax && geom_refine(G, c, shell, in)) w |= W_REF;
  }
  W[i] = w;
}

// R = D & refine
__global__ void k_vx_r(Grid g, uint32_t* W) {
  CELL_LOOP(g);
  const uint32_t w = W[i];
  if ((w & W_D) && (w & W_REF)) W[i] = w | W_R;
}

// band = R & (some 26-neighbour not R); interior = R & ~band; error bits
__global__ void k_vx_band(Grid g, uint32_t* W, int per, int reservoir, int* err) {
  CELL_LOOP(g);
  const uint32_t w = W[i];
  if (!(w & W_R)) return;
  bool nr = false, shell = false, c2f = false;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        if (!dx && !dy && !dz) continue;
        // the band looks across periodic faces (level 0); the checks do not
        const int64_t jp = nbr(x, y, z, dx, dy, dz, g, per);
        if (jp >= 0 && !(W[jp] & W_R)) nr = true;
        const int64_t j = nbr(x, y, z, dx, dy, dz, g, 0);
        if (j < 0) continue;
        const uint32_t v = W[j];
        if (v & W_SHELL) shell = true;
        if (v & W_C2F) c2f = true;
      }
  if (reservoir && (shell || (w & W_SHELL))) atomicOr(err, 1);  // refinement touches the boundary layer
  // the domain edge (a refined cell within one cell of it)
  if (x <= 1 || y <= 1 || z <= 1 || x >= g.n[0] - 2 || y >= g.n[1] - 2 || z >= g.n[2] - 2) atomicOr(err, 2);
  if (c2f || (w & W_C2F)) atomicOr(err, 4);  // refinement too close to the C2F band
  W[i] = w | (nr ? W_BAND : W_INT);
}

// solid, reservoir, fluid, and the level's kinds
__global__ void k_vx_kinds(Grid g, uint32_t* W, uint16_t* K, int finest, int solid, int reservoir, int* err) {
  CELL_LOOP(g);
  uint32_t w = W[i];
  const bool D = w & W_D, in = w & W_IN;
  if (finest && solid && D && in && !(w & W_C2F)) w |= W_SOL;
  if (!finest && solid && D && !(w & W_INT) && in) atomicOr(err, 8);  // solid not fully inside the finest level
  const bool bcres = reservoir && (w & W_SHELL);
  const bool fluid = D && !(w & (W_INT | W_BAND | W_C2F | W_SOL)) && !bcres;
  if (fluid) w |= W_FLUID;
  uint16_t k = 0;
  if (fluid) k |= MT_KIND_FLUID;
  if (w & W_BAND) k |= MT_KIND_F2C;
  if (w & W_C2F) k |= MT_KIND_C2F;
  if (bcres) k |= MT_KIND_BCRES;
  W[i] = w;
  K[i] = k;
}

// Q-neighbourhood of `bits` (zero outside): the velocity offsets of D3Q19/D3Q27
__device__ __forceinline__ bool qnbr_any(int x, int y, int z, const Grid& g, const uint32_t* W, uint32_t bits,
                                        int Q) {
  for (int q = 1; q < Q; ++q) {
    const int64_t j = nbr(x, y, z, kEX[q], kEY[q], kEZ[q], g, 0);
    if (j >= 0 && (W[j] & bits)) return true;
  }
  return false;
}

// ghost_C2F: Q-neighbours of C2F cells outside D; their parents must be fluid
__global__ void k_vx_ghost(Grid g, uint32_t* W, uint16_t* K, const uint16_t* Kc, Grid gc, int Q, int* err) {
  CELL_LOOP(g);
  const uint32_t w = W[i];
  if (w & W_D) return;
  if (!qnbr_any(x, y, z, g, W, W_C2F, Q)) return;
  if (!(Kc[lin(x >> 1, y >> 1, z >> 1, gc)] & MT_KIND_FLUID)) atomicOr(err, 16);
  K[i] |= MT_KIND_GHOST_C2F;
}

// ghost_F2C on a coarser level: fluid cells Q-adjacent to F2C cells
__global__ void k_vx_ghost_f2c(Grid g, const uint32_t* W, uint16_t* K, int Q) {
  CELL_LOOP(g);
  if (!(K[i] & MT_KIND_FLUID)) return;
  if ((W[i] & W_BAND) || qnbr_any(x, y, z, g, W, W_BAND, Q)) K[i] |= MT_KIND_GHOST_F2C;
}

// solid cells some fluid or C2F cell pulls from stay (zero populations); the others go
__global__ void k_vx_solid(Grid g, const uint32_t* W, uint16_t* K, int Q) {
  CELL_LOOP(g);
  const uint32_t w = W[i];
  if (!(w & W_SOL)) return;
  const bool live = (w & (W_FLUID | W_C2F)) || qnbr_any(x, y, z, g, W, W_FLUID | W_C2F, Q);
  K[i] = live ? (uint16_t)(K[i] | MT_KIND_SOLID) : (uint16_t)0;
}

// level-0 face bits (W,E,S,N,B,T = 0..5) on the reservoir shell or the fluid cells
// of every non-periodic face
__global__ void k_vx_faces(Grid g, const uint32_t* W, uint16_t* K, const int* face_kind, int reservoir) {
  CELL_LOOP(g);
  const uint32_t w = W[i];
  const bool on = reservoir ? (w & W_SHELL) != 0 : (w & W_FLUID) != 0;
  if (!on) return;
  const int c[3] = {x, y, z};
  uint16_t k = K[i];
  for (int f = 0; f < 6; ++f) {
    if (!reservoir && face_kind[f] == MT_BC_PERIODIC) continue;
    const int a = f >> 1;
    if ((f & 1) ? c[a] == g.n[a] - 1 : c[a] == 0) k |= (uint16_t)(0x0100u << f);
  }
  K[i] = k;
}

// the finest level's link fractions: fluid and C2F cells with a crossing
__global__ void k_vx_dis(Grid g, const uint32_t* W, uint16_t* K, double* vals, int Q, Geom G, int Lmax) {
  CELL_LOOP(g);
  if (!(K[i] & (MT_KIND_FLUID | MT_KIND_C2F))) return;
  if (W[i] & W_IN) {
    // links start from centres outside the body
    for (int q = 1; q < Q; ++q) vals[i * (Q - 1) + (q - 1)] = -1.0;
    return;
  }
  double c[3];
  centre(x, y, z, g.L, c);
  const double h = 1.0 / (double)(1 << Lmax);
  bool any = false;
  if (G.shape == MT_SHAPE_MESH) {
    // no link (at most h sqrt 3 long) reaches a surface farther away
    double d2 = 1e300;
    for (int64_t t = 0; t < G.n_tri; ++t) d2 = fmin(d2, tri_dist2(G.tri + 9 * t, c));
    if (d2 > 3.0 * h * h * (1.0 + 1e-9)) {
      for (int q = 1; q < Q; ++q) vals[i * (Q - 1) + (q - 1)] = -1.0;
      return;
    }
  }
  for (int q = 1; q < Q; ++q) {
    const double t = geom_link(G, c, kEX[q], kEY[q], kEZ[q], h);
    vals[i * (Q - 1) + (q - 1)] = t;
    any |= t >= 0.0;
  }
  if (any) K[i] |= MT_KIND_DIS;
}

int64_t nblk(int64_t n) { return (n + 255) / 256; }

}  // namespace

struct LevelSet {
  int Q = 19, Lmax = 0;
  std::vector<std::vector<uint64_t>> keys;
  std::vector<std::vector<uint16_t>> kinds;
  std::vector<uint64_t> dis_keys;
  std::vector<double> dis_vals;
};

}  // namespace mt

using namespace mt;
struct mt_levelset : public mt::LevelSet {};

namespace {
const char* const kVoxErr[] = {"refinement touches the boundary layer", "refinement touches the domain edge",
                               "refinement too close to the C2F band", "solid is not fully inside the finest level",
                               "ghost_C2F parent is not a fluid cell"};

int build(const mt_levelset_desc* d, mt_levelset& out) {
  const int Lmax = d->Lmax, Q = d->Q;
  int per = 0;
  for (int a = 0; a < 3; ++a)
    if (d->face_kind[2 * a] == MT_BC_PERIODIC) per |= 1 << a;
  hipStream_t st = nullptr;
  MT_HIP(stream_acquire(&st));
  struct Guard {
    std::vector<void*> p;
    hipStream_t s;
    ~Guard() {
      (void)hipStreamSynchronize(s);
      for (void* q : p) dev_free(q, s);
      stream_release(s);
    }
  } guard{{}, st};
  auto alloc = [&](void** p, size_t bytes) -> int {
    MT_HIP(dev_malloc(p, bytes, st));
    guard.p.push_back(*p);
    return MT_OK;
  };
  Geom G{};
  G.shape = d->shape;
  for (int a = 0; a < 3; ++a) G.c[a] = d->center[a];
  G.r = d->radius;
  if (d->shape == MT_SHAPE_MESH) {
    void* t = nullptr;
    VX_TRY(alloc(&t, sizeof(double) * 9 * std::max<int64_t>(d->n_tri, 1)));
    MT_HIP(hipMemcpyAsync(t, d->tri, sizeof(double) * 9 * d->n_tri, hipMemcpyHostToDevice, st));
    G.tri = (const double*)t;
    G.n_tri = d->n_tri;
  }
  int* err = nullptr;
  int* fk = nullptr;
  VX_TRY(alloc((void**)&err, sizeof(int)));
  VX_TRY(alloc((void**)&fk, sizeof(int) * 6));
  MT_HIP(hipMemsetAsync(err, 0, sizeof(int), st));
  MT_HIP(hipMemcpyAsync(fk, d->face_kind, sizeof(int) * 6, hipMemcpyHostToDevice, st));
  std::vector<Grid> g(Lmax + 1);
  std::vector<uint32_t*> W(Lmax + 1);
  std::vector<uint16_t*> K(Lmax + 1);
  double* dis = nullptr;
  auto check = [&](int L) -> int {
    int h = 0;
    MT_HIP(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, st));
    MT_HIP(hipStreamSynchronize(st));
    for (int b = 0; b < 5; ++b)
      if (h & (1 << b)) {
        set_error((b == 0 ? std::string() : "level " + std::to_string(L) + ": ") + kVoxErr[b]);
        return MT_ERR_ARG;
      }
    return MT_OK;
  };
  for (int L = 0; L <= Lmax; ++L) {
    for (int a = 0; a < 3; ++a) g[L].n[a] = d->dims0[a] << L;
    g[L].L = L;
    const int64_t n = g[L].cells();
    VX_TRY(alloc((void**)&W[L], sizeof(uint32_t) * n));
    VX_TRY(alloc((void**)&K[L], sizeof(uint16_t) * n));
    const Grid gc = L ? g[L - 1] : g[L];
    const int res = (L == 0 && d->reservoir) ? 1 : 0;
    k_vx_init<<<nblk(n), 256, 0, st>>>(g[L], W[L], L ? W[L - 1] : nullptr, gc, Lmax, res, G,
                                        L < Lmax ? d->shells[L] : 0.0, d->solid);
    if (L < Lmax) {
      k_vx_r<<<nblk(n), 256, 0, st>>>(g[L], W[L]);
      k_vx_band<<<nblk(n), 256, 0, st>>>(g[L], W[L], L == 0 ? per : 0, res, err);
    }
    k_vx_kinds<<<nblk(n), 256, 0, st>>>(g[L], W[L], K[L], L == Lmax, d->solid, res, err);
    MT_HIP(hipGetLastError());
    VX_TRY(check(L));
    if (L > 0) {
      k_vx_ghost<<<nblk(n), 256, 0, st>>>(g[L], W[L], K[L], K[L - 1], g[L - 1], Q, err);
      k_vx_ghost_f2c<<<nblk(g[L - 1].cells()), 256, 0, st>>>(g[L - 1], W[L - 1], K[L - 1], Q);
      MT_HIP(hipGetLastError());
      VX_TRY(check(L));
    }
    if (L == Lmax && d->solid) k_vx_solid<<<nblk(n), 256, 0, st>>>(g[L], W[L], K[L], Q);
    if (L == 0) k_vx_faces<<<nblk(n), 256, 0, st>>>(g[L], W[L], K[L], fk, d->reservoir);
    MT_HIP(hipGetLastError());
  }
  if (d->solid) {
    const int64_t n = g[Lmax].cells();
    VX_TRY(alloc((void**)&dis, sizeof(double) * (Q - 1) * n));
    k_vx_dis<<<nblk(n), 256, 0, st>>>(g[Lmax], W[Lmax], K[Lmax], dis, Q, G, Lmax);
    MT_HIP(hipGetLastError());
  }
  // the sets leave as Morton keys in ascending order
  out.Q = Q;
  out.Lmax = Lmax;
  out.keys.assign(Lmax + 1, {});
  out.kinds.assign(Lmax + 1, {});
  for (int L = 0; L <= Lmax; ++L) {
    const int64_t n = g[L].cells();
    std::vector<uint16_t> hk(n);
    MT_HIP(hipMemcpyAsync(hk.data(), K[L], 2 * n, hipMemcpyDeviceToHost, st));
    std::vector<double> hd;
    if (L == Lmax && dis) {
      hd.resize((size_t)(Q - 1) * n);
      MT_HIP(hipMemcpyAsync(hd.data(), dis, sizeof(double) * hd.size(), hipMemcpyDeviceToHost, st));
    }
    MT_HIP(hipStreamSynchronize(st));
    const int sh = Lmax - L;
    std::vector<std::pair<uint64_t, int64_t>> ord;
    for (int64_t i = 0; i < n; ++i) {
      if (!hk[i]) continue;
      const int64_t z = i % g[L].n[2], t = i / g[L].n[2], y = t % g[L].n[1], x = t / g[L].n[1];
      const uint64_t key = spread3((uint64_t)x << sh) | (spread3((uint64_t)y << sh) << 1) |
                           (spread3((uint64_t)z << sh) << 2);
      ord.emplace_back(key, i);
    }
    std::sort(ord.begin(), ord.end());
    for (auto& p : ord) {
      out.keys[L].push_back(p.first);
      out.kinds[L].push_back(hk[p.second]);
      if (L == Lmax && (hk[p.second] & MT_KIND_DIS)) {
        out.dis_keys.push_back(p.first);
        out.dis_vals.insert(out.dis_vals.end(), hd.begin() + (Q - 1) * p.second, hd.begin() + (Q - 1) * (p.second + 1));
      }
    }
  }
  return MT_OK;
}
}  // namespace

extern "C" {

int mt_levelset_build(const mt_levelset_desc* d, mt_levelset** out) {
  if (!d || !out) { set_error("null argument"); return MT_ERR_ARG; }
  *out = nullptr;
  if (d->Q != 19 && d->Q != 27) { set_error("Q must be 19 or 27"); return MT_ERR_ARG; }
  if (d->Lmax < 0 || d->Lmax > 6) { set_error("Lmax out of range"); return MT_ERR_ARG; }
  for (int a = 0; a < 3; ++a)
    if (d->dims0[a] <= 0 || ((int64_t)d->dims0[a] << d->Lmax) > (1 << 20)) { set_error("bad domain size"); return MT_ERR_ARG; }
  if (d->shape != MT_SHAPE_SPHERE && d->shape != MT_SHAPE_MESH) { set_error("bad shape"); return MT_ERR_ARG; }
  if (d->shape == MT_SHAPE_MESH && (d->n_tri <= 0 || !d->tri)) { set_error("a mesh needs triangles"); return MT_ERR_ARG; }
  for (int f = 0; f < 6; ++f)
    if (d->face_kind[f] < MT_BC_NONE || d->face_kind[f] > MT_BC_NEE) { set_error("bad face kind"); return MT_ERR_ARG; }
  for (int a = 0; a < 3; ++a)
    if ((d->face_kind[2 * a] == MT_BC_PERIODIC) != (d->face_kind[2 * a + 1] == MT_BC_PERIODIC)) {
      set_error("periodic faces must come in pairs");
      return MT_ERR_ARG;
    }
  mt_levelset* ls = new mt_levelset();
  int r = build(d, *ls);
  if (r != MT_OK) {
    delete ls;
    return r;
  }
  *out = ls;
  return MT_OK;
}

int mt_levelset_counts(mt_levelset* ls, int64_t* n_cells, int64_t* n_dis) {
  if (!ls) { set_error("null level set"); return MT_ERR_ARG; }
  if (n_cells)
    for (int L = 0; L <= ls->Lmax; ++L) n_cells[L] = (int64_t)ls->keys[L].size();
  if (n_dis) *n_dis = (int64_t)ls->dis_keys.size();
  return MT_OK;
}

int mt_levelset_copy(mt_levelset* ls, int level, uint64_t* keys, uint16_t* kinds) {
  if (!ls || level < 0 || level > ls->Lmax) { set_error("bad level set or level"); return MT_ERR_ARG; }
  const size_t n = ls->keys[level].size();
  if (keys) std::memcpy(keys, ls->keys[level].data(), 8 * n);
  if (kinds) std::memcpy(kinds, ls->kinds[level].data(), 2 * n);
  return MT_OK;
}

int mt_levelset_copy_dis(mt_levelset* ls, uint64_t* keys, double* fractions) {
  if (!ls) { set_error("null level set"); return MT_ERR_ARG; }
  if (keys) std::memcpy(keys, ls->dis_keys.data(), 8 * ls->dis_keys.size());
  if (fractions) std::memcpy(fractions, ls->dis_vals.data(), sizeof(double) * ls->dis_vals.size());
  return MT_OK;
}

int mt_levelset_destroy(mt_levelset* ls) {
  delete ls;
  return MT_OK;
}

}  // extern "C"
