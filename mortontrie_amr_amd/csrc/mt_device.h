// mt_device.h -- device-side lattice constants, brick addressing and the
// per-cell physics shared by every kernel of the AMR LBM update.
//
// Layout.  Each refinement level stores its cells in 4x4x4 Morton bricks: brick
// b holds slots [64b, 64b+64) and a cell's local index is the 6-bit Morton code
// of its coordinates inside the brick, so slot order is Morton order.  Bricks
// are sorted by the Morton code of their brick coordinates.  Populations are
// SoA, pop[q][slot].  Neighbour addressing inside a brick is arithmetic; across
// bricks it goes through a 27-entry neighbour-brick table per brick.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mt_amr.h"

namespace mt {

constexpr int kEX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1};
constexpr int kEY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, 1, -1};
constexpr int kEZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1};

__host__ __device__ constexpr int qinv(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
__host__ __device__ constexpr int qnorm1(int q) {
  return (kEX[q] != 0) + (kEY[q] != 0) + (kEZ[q] != 0);
}
template <int Q>
__host__ __device__ constexpr double wq(int q) {
  return Q == 19 ? (qnorm1(q) == 0 ? 1. / 3. : (qnorm1(q) == 1 ? 1. / 18. : 1. / 36.))
                 : (qnorm1(q) == 0 ? 8. / 27. : (qnorm1(q) == 1 ? 2. / 27. : (qnorm1(q) == 2 ? 1. / 54. : 1. / 216.)));
}
// D3Q19 table used by the solid-boundary interpolation for any Q (entries >= 19 are 0)
__host__ __device__ constexpr double wib(int q) { return q < 19 ? wq<19>(q) : 0.0; }

// velocity (i,j,k) in {-1,0,1}^3 -> q (D3Q27 ordering)
__host__ __device__ constexpr int qof(int i, int j, int k) {
  int r = -1;
  for (int q = 0; q < 27; ++q)
    if (kEX[q] == i && kEY[q] == j && kEZ[q] == k) r = q;
  return r;
}

// ---------------------------------------------------------------- bricks
__device__ __forceinline__ int brick_local(int lx, int ly, int lz) {
  return (lx & 1) | ((ly & 1) << 1) | ((lz & 1) << 2) | ((lx >> 1) << 3) | ((ly >> 1) << 4) | ((lz >> 1) << 5);
}
__device__ __forceinline__ void brick_coords(int l, int& lx, int& ly, int& lz) {
  lx = (l & 1) | (((l >> 3) & 1) << 1);
  ly = ((l >> 1) & 1) | (((l >> 4) & 1) << 1);
  lz = ((l >> 2) & 1) | (((l >> 5) & 1) << 1);
}
// slot of the cell at offset (dx,dy,dz), |d| <= 4, from local coords in brick b;
// nb27 holds the 27 neighbour-brick indices of b (-1 = none)
__device__ __forceinline__ int slot_at(const int* nb27, int lx, int ly, int lz, int dx, int dy, int dz) {
  int nx = lx + dx, ny = ly + dy, nz = lz + dz;
  int ox = nx >> 2, oy = ny >> 2, oz = nz >> 2;
  int nb = nb27[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
  if (nb < 0) return -1;
  return (nb << 6) | brick_local(nx & 3, ny & 3, nz & 3);
}

// ---------------------------------------------------------------- storage
// Populations are stored as S; arithmetic is in R.  For the shifted formats the
// stored value is f_q - w_q (the deviation from the rest state).
template <typename S>
struct Store;
template <>
struct Store<double> {
  static constexpr bool shifted = false;
  __device__ static __forceinline__ double ld(const double* p) { return *p; }
  __device__ static __forceinline__ void st(double* p, double v) { *p = v; }
};
template <>
struct Store<float> {
  static constexpr bool shifted = true;
  __device__ static __forceinline__ float ld(const float* p) { return *p; }
  __device__ static __forceinline__ void st(float* p, float v) { *p = v; }
};
template <>
struct Store<_Float16> {
  static constexpr bool shifted = true;
  __device__ static __forceinline__ float ld(const _Float16* p) { return (float)*p; }
  __device__ static __forceinline__ void st(_Float16* p, float v) { *p = (_Float16)v; }
};
template <typename S>
struct RealOf { typedef float type; };
template <>
struct RealOf<double> { typedef double type; };

template <int Q, typename S, typename R>
__device__ __forceinline__ R load_pop(const S* base, int64_t ns, int q, int slot) {
  R v = (R)Store<S>::ld(base + (int64_t)q * ns + slot);
  if (Store<S>::shifted) v += (R)wq<Q>(q);
  return v;
}
// The value a population takes once stored and reloaded (identity for FP64).
// Values computed in registers and used again before any store (boundary
// reconstruction, coalescence, a collision feeding an explosion) are snapped so
// that the fused schedule computes exactly what the per-method kernels compute
// from the stored state.
template <int Q, typename S, typename R>
__device__ __forceinline__ R snap_pop(int q, R v) {
  if (!Store<S>::shifted) return v;
  const R w = (R)wq<Q>(q);
  S t;
  Store<S>::st(&t, v - w);
  return (R)Store<S>::ld(&t) + w;
}
template <int Q, typename S, typename R>
__device__ __forceinline__ void store_pop(S* base, int64_t ns, int q, int slot, R v) {
  if (Store<S>::shifted) v -= (R)wq<Q>(q);
  Store<S>::st(base + (int64_t)q * ns + slot, v);
}

// ---------------------------------------------------------------- physics
// moments with the +-1e-6 density floor (LBM_Manager.cpp:983-1001)
template <int Q, typename R>
__device__ __forceinline__ void cell_macros(const R* f, R& rho, R& ux, R& uy, R& uz) {
  R r = f[0], vx = 0, vy = 0, vz = 0;
#pragma unroll
  for (int q = 1; q < Q; ++q) {
    r += f[q];
    if (kEX[q] == 1) vx += f[q]; else if (kEX[q] == -1) vx -= f[q];
    if (kEY[q] == 1) vy += f[q]; else if (kEY[q] == -1) vy -= f[q];
    if (kEZ[q] == 1) vz += f[q]; else if (kEZ[q] == -1) vz -= f[q];
  }
  R s = (r > (R)-1e-6 && r < (R)1e-6) ? ((r < (R)0) ? (R)-1e-6 : (R)1e-6) : r;
  rho = r;
  ux = vx / s;
  uy = vy / s;
  uz = vz / s;
}
// compressible feq, direction q (LBM_Manager.cpp:205-233)
template <int Q, typename R>
__device__ __forceinline__ R feq_c(int q, R rho, R ux, R uy, R uz, R uu) {
  R eu = (R)kEX[q] * ux + (R)kEY[q] * uy + (R)kEZ[q] * uz;
  return rho * (R)wq<Q>(q) * ((R)1 + (R)3 * eu + (R)4.5 * eu * eu - (R)1.5 * uu);
}
// He-Luo feq (LBM_Manager.cpp:235-263)
template <int Q, typename R>
__device__ __forceinline__ R feq_hl(int q, R rho, R ux, R uy, R uz, R uu, R rho0) {
  R eu = (R)kEX[q] * ux + (R)kEY[q] * uy + (R)kEZ[q] * uz;
  return (R)wq<Q>(q) * (rho + rho0 * ((R)3 * eu + (R)4.5 * eu * eu - (R)1.5 * uu));
}

// Smagorinsky eddy relaxation time (lbm_kernel/les.cpp:12-44)
template <int Q, typename R>
__device__ __forceinline__ R smag_tau(const R* f, R rho, R ux, R uy, R uz, R tau, R C) {
  R uu = ux * ux + uy * uy + uz * uz;
  R p[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    R d = f[q] - feq_c<Q, R>(q, rho, ux, uy, uz, uu);
    const int c[3] = {kEX[q], kEY[q], kEZ[q]};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        if (c[a] * c[b] != 0) p[3 * a + b] += (R)(c[a] * c[b]) * d;
  }
  R m = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) m += p[i] * p[i];
  R qn = sqrt(m);
  R c2 = C * C;
  R nu = (tau - (R)0.5) / (R)3.0;
  R s = (sqrt(nu * nu + (R)18 * c2 * qn) - nu) / ((R)6.0 * c2);
  return (R)3 * (nu + c2 * s) + (R)0.5;
}

// LBGK (lbm_kernel/Collision_Manager.cpp:19-38), in place
template <int Q, typename R>
__device__ __forceinline__ void collide_lbgk(R* f, R rho, R ux, R uy, R uz, R tau) {
  R uu = ux * ux + uy * uy + uz * uz;
  R it = (R)1 / tau;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    R fe = feq_c<Q, R>(q, rho, ux, uy, uz, uu);
    f[q] = f[q] - (f[q] - fe) * it;
  }
}

// ---- D3Q27 cumulant (lbm_kernel/Collision_Cumulant.cpp:22-652) -------------
// m[9a+3b+c] holds the (a,b,c) moment; velocity index (i,j,k) in {0,1,2} maps
// to lattice velocity (i-1, j-1, k-1).
template <typename R>
__device__ __forceinline__ void fwd1(R& a, R& b, R& c, R v) {  // f(-1),f(0),f(+1) -> m0,m1,m2
  R m0 = a + b + c;
  R d = c - a;
  R m1 = d - v * m0;
  R m2 = (c + a) - (R)2 * v * d + v * v * m0;
  a = m0;
  b = m1;
  c = m2;
}
template <typename R>
__device__ __forceinline__ void bwd1(R& a, R& b, R& c, R v) {  // m0,m1,m2 -> f(-1),f(0),f(+1)
  R v2 = v * v;
  R m0 = a, m1 = b, m2 = c;
  R f0 = m0 * ((R)1 - v2) - m1 * (R)2 * v - m2;
  R fm = (m0 * (v2 - v) + m1 * ((R)2 * v - (R)1) + m2) * (R)0.5;
  R fp = (m0 * (v2 + v) + m1 * ((R)2 * v + (R)1) + m2) * (R)0.5;
  a = fm;
  b = f0;
  c = fp;
}
template <typename R>
__device__ __forceinline__ R safe_irho(R r) {
  return (R)1 / ((r > (R)-1e-6 && r < (R)1e-6) ? ((r < (R)0) ? (R)-1e-6 : (R)1e-6) : r);
}
#define MI(a, b, c) ((a)*9 + (b)*3 + (c))
template <typename R>
__device__ __forceinline__ void collide_cumulant(R* f, R rho, R ux, R uy, R uz, R tau) {
  R m[27];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) m[MI(i, j, k)] = f[qof(i - 1, j - 1, k - 1)];
  // forward central moments: collapse z, then y, then x
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) fwd1(m[MI(i, j, 0)], m[MI(i, j, 1)], m[MI(i, j, 2)], uz);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) fwd1(m[MI(i, 0, c)], m[MI(i, 1, c)], m[MI(i, 2, c)], uy);
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int c = 0; c < 3; ++c) fwd1(m[MI(0, b, c)], m[MI(1, b, c)], m[MI(2, b, c)], ux);
  // m is now the central-moment set k[a][b][c].  Relaxation sets every cumulant
  // of order >= 3 to zero, and the backward transform rebuilds the central
  // moments from the relaxed second-order members.  As in the reference,
  // k000, the first-order moments and k210, k021, k102 are never reassigned by
  // the cumulant transforms and keep their forward values
  // (Collision_Cumulant.cpp:378-381, 451-454); k120, k012, k201, k111 become 0.
  const R ir = safe_irho(rho);
  const R w1 = (R)1 / tau, w2 = (R)1;
  const R k000 = m[MI(0, 0, 0)];
  const R k210 = m[MI(2, 1, 0)], k021 = m[MI(0, 2, 1)], k102 = m[MI(1, 0, 2)];
  const R c200 = m[MI(2, 0, 0)], c020 = m[MI(0, 2, 0)], c002 = m[MI(0, 0, 2)];
  const R Dxu = -(w1 * ((R)2 * c200 - c020 - c002) + w2 * (c200 + c020 + c002 - k000)) * ir * (R)0.5;
  const R Dyv = Dxu + (R)3 * w1 * (c200 - c020) * (R)0.5 * ir;
  const R Dzw = Dxu + (R)3 * w1 * (c200 - c002) * (R)0.5 * ir;
  const R ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
  const R X1 = ((R)1 - w1) * (c200 - c020) - (R)3 * rho * ((R)1 - w1 * (R)0.5) * (ux2 * Dxu - uy2 * Dyv);
  const R X2 = ((R)1 - w1) * (c200 - c002) - (R)3 * rho * ((R)1 - w1 * (R)0.5) * (ux2 * Dxu - uz2 * Dzw);
  const R X3 = k000 * w2 + ((R)1 - w2) * (c200 + c020 + c002) -
               (R)3 * rho * ((R)1 - w2 * (R)0.5) * (ux2 * Dxu + uy2 * Dyv + uz2 * Dzw);
  const R k200 = (X1 + X2 + X3) / (R)3;
  const R k020 = k200 - X1;
  const R k002 = k200 - X2;
  const R k110 = m[MI(1, 1, 0)] * ((R)1 - w1);
  const R k101 = m[MI(1, 0, 1)] * ((R)1 - w1);
  const R k011 = m[MI(0, 1, 1)] * ((R)1 - w1);
  const R k211 = (k200 * k011 + (R)2 * k110 * k101) * ir;
  const R k121 = (k020 * k101 + (R)2 * k011 * k110) * ir;
  const R k112 = (k002 * k110 + (R)2 * k101 * k011) * ir;
  const R k220 = (k200 * k020 + (R)2 * k110 * k110) * ir;
  const R k022 = (k020 * k002 + (R)2 * k011 * k011) * ir;
  const R k202 = (k002 * k200 + (R)2 * k101 * k101) * ir;
  const R k122 = (k020 * k102 + (R)2 * k101 * k021) * ir;
  const R k212 = (k002 * k210 + (R)2 * k110 * k102) * ir;
  const R k221 = (k200 * k021 + (R)2 * k011 * k210) * ir;
  const R ka = (k200 * k022 + k020 * k202 + k002 * k220 + (R)4 * (k011 * k211 + k101 * k121 + k110 * k112)) * ir;
  const R kb = ((R)16 * k110 * k101 * k011 +
                (R)4 * (k101 * k101 * k020 + k011 * k011 * k200 + k110 * k110 * k002) + (R)2 * k200 * k020 * k002) *
               ir * ir;
  m[MI(2, 0, 0)] = k200; m[MI(0, 2, 0)] = k020; m[MI(0, 0, 2)] = k002;
  m[MI(1, 1, 0)] = k110; m[MI(1, 0, 1)] = k101; m[MI(0, 1, 1)] = k011;
  m[MI(1, 2, 0)] = 0; m[MI(0, 1, 2)] = 0; m[MI(2, 0, 1)] = 0; m[MI(1, 1, 1)] = 0;
  m[MI(2, 1, 1)] = k211; m[MI(1, 2, 1)] = k121; m[MI(1, 1, 2)] = k112;
  m[MI(2, 2, 0)] = k220; m[MI(0, 2, 2)] = k022; m[MI(2, 0, 2)] = k202;
  m[MI(1, 2, 2)] = k122; m[MI(2, 1, 2)] = k212; m[MI(2, 2, 1)] = k221;
  m[MI(2, 2, 2)] = ka - kb;
  // backward: x, then y, then z (Collision_Cumulant.cpp:476-652)
  // x pass
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int c = 0; c < 3; ++c) bwd1(m[MI(0, b, c)], m[MI(1, b, c)], m[MI(2, b, c)], ux);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) bwd1(m[MI(i, 0, c)], m[MI(i, 1, c)], m[MI(i, 2, c)], uy);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) bwd1(m[MI(i, j, 0)], m[MI(i, j, 1)], m[MI(i, j, 2)], uz);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) f[qof(i - 1, j - 1, k - 1)] = m[MI(i, j, k)];
}

}  // namespace mt
