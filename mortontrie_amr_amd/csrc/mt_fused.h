// mt_fused.h -- the fused per-level update (one brick per 64-lane block) and
// the device helpers it shares with the seam kernels of mt_kernels.hip.
//
// The fused update, for every lat_f / C2F cell of a level: pull fcol from the
// neighbours, apply the level-0 wet-node boundary or the finest level's solid
// boundary, compute the moments and collide, writing the next sub-step's fcol
// into the other buffer (and, on the last sub-step of a pair, coalesce into the
// coarser level's F2C cells).  The reference runs stream -> boundary -> (next
// sub-step) macros -> collide on the same f with no other writer in between
// (LBM_Simulator.cpp:20-229), so the fused order computes the same values while
// reading and writing each population once.
//
// The kernels are templates over the lattice (Q), the storage format (S), the
// collision kind (CK) and the boundary / transfer work a level carries; each
// (Q, S) pair is instantiated in its own translation unit (mt_fused_*.hip)
// through launch_fused_qs / launch_fused_levels_qs.
#pragma once
#include <algorithm>
#include <type_traits>
#include <vector>

#include <hip/hip_ext.h>

#include "mt_case.h"
#include "mt_device.h"
#include "mt_morton.h"

namespace mt {

#define TRY_K(x)                  \
  do {                            \
    int r_ = (x);                 \
    if (r_ != MT_OK) return r_;   \
  } while (0)

constexpr uint16_t kHasPop = MT_KIND_FLUID | MT_KIND_F2C | MT_KIND_C2F | MT_KIND_GHOST_C2F | MT_KIND_SOLID | MT_KIND_BCRES;

struct LvArgs {
  int64_t ns;
  int L;
  const MemberParams* __restrict__ mp;   // per-member parameters
  const int32_t* __restrict__ bmember;   // [n_bricks] member of each brick
  const int32_t* __restrict__ bnbr;
  const uint16_t* __restrict__ kind;
  const uint32_t* __restrict__ pmask;
  void* f;
  const void* fcol_in;
  void* fcol_out;
  const void* rho_in;
  const void* u_in;
  void* rho_out;
  void* u_out;
};

// parameters of the member owning slot s
__device__ __forceinline__ const MemberParams& member_of(const LvArgs& v, int64_t s) { return v.mp[v.bmember[s >> 6]]; }

// ------------------------------------------------------------------ addressing
// slot of the neighbour at (DX,DY,DZ) in {-1,0,1}^3; N holds the 27
// neighbour-brick indices (wave-uniform).  Written so that, once the caller's
// q loop is unrolled, every N[] index is a constant.
__device__ __forceinline__ int nbr_slot(const int* N, int lx, int ly, int lz, int DX, int DY, int DZ) {
  const bool cx = DX > 0 ? (lx == 3) : (DX < 0 ? (lx == 0) : false);
  const bool cy = DY > 0 ? (ly == 3) : (DY < 0 ? (ly == 0) : false);
  const bool cz = DZ > 0 ? (lz == 3) : (DZ < 0 ? (lz == 0) : false);
  const int OX = DX + 1, OY = DY + 1, OZ = DZ + 1;
  int nb = cz ? (cy ? (cx ? N[OX + 3 * OY + 9 * OZ] : N[1 + 3 * OY + 9 * OZ])
                    : (cx ? N[OX + 3 + 9 * OZ] : N[1 + 3 + 9 * OZ]))
              : (cy ? (cx ? N[OX + 3 * OY + 9] : N[1 + 3 * OY + 9]) : (cx ? N[OX + 3 + 9] : N[13]));
  int nl = brick_local((lx + DX) & 3, (ly + DY) & 3, (lz + DZ) & 3);
  return nb < 0 ? -1 : ((nb << 6) | nl);
}
// compile-time direction form of nbr_slot
template <int DX, int DY, int DZ>
__device__ __forceinline__ int nbr_slot_c(const int* N, int lx, int ly, int lz) {
  const bool cx = DX > 0 ? (lx == 3) : (DX < 0 ? (lx == 0) : false);
  const bool cy = DY > 0 ? (ly == 3) : (DY < 0 ? (ly == 0) : false);
  const bool cz = DZ > 0 ? (lz == 3) : (DZ < 0 ? (lz == 0) : false);
  constexpr int OX = DX + 1, OY = DY + 1, OZ = DZ + 1;
  const int n000 = N[13];
  const int n100 = N[OX + 3 + 9], n010 = N[1 + 3 * OY + 9], n001 = N[1 + 3 + 9 * OZ];
  const int n110 = N[OX + 3 * OY + 9], n101 = N[OX + 3 + 9 * OZ], n011 = N[1 + 3 * OY + 9 * OZ];
  const int n111 = N[OX + 3 * OY + 9 * OZ];
  const int nb = cz ? (cy ? (cx ? n111 : n011) : (cx ? n101 : n001)) : (cy ? (cx ? n110 : n010) : (cx ? n100 : n000));
  const int nl = brick_local((lx + DX) & 3, (ly + DY) & 3, (lz + DZ) & 3);
  return (nb << 6) | nl;  // callers only use it when the neighbour is present
}
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& fn) {
  if constexpr (B < E) {
    fn(std::integral_constant<int, B>{});
    static_for<B + 1, E>(fn);
  }
}

// general offset (|d| <= 4) through the table in memory (rare paths)
__device__ __forceinline__ int slot_off(const int32_t* bnbr, int s, int dx, int dy, int dz) {
  int lx, ly, lz;
  brick_coords(s & 63, lx, ly, lz);
  return slot_at(bnbr + (int64_t)(s >> 6) * 27, lx, ly, lz, dx, dy, dz);
}

template <typename T>
__device__ __forceinline__ T ld_real(const void* p, int64_t i) { return ((const T*)p)[i]; }
template <typename T>
__device__ __forceinline__ void st_real(void* p, int64_t i, T v) { ((T*)p)[i] = v; }

// ------------------------------------------------------------------ boundaries
// Wet-node level-0 boundary on one cell (definitions restated in the oracle;
// the uncompiled reference text is BC/ZouHe_Velocity.cpp:170-532 and
// BC/Outflow.cpp:12-136).  Under the pull f_q(x) <- fcol_q(x+e_q) the unknowns
// on a face with outward normal n are the e_q.n = +1 populations.
// Non-equilibrium extrapolation from one neighbour n: f_q = feq_q(rho_n, w) +
// fn_q - feq_q(rho_n, u_n) for every q, with w the wall velocity (own_u: w =
// u_n).  fn(q) is the neighbour's post-stream f_q; it may return f_q itself (a
// missing neighbour), so every fn_q is read before f is replaced.
template <int Q, typename R, typename FN>
__device__ __forceinline__ void nee_from(R* f, R wx, R wy, R wz, FN fn, bool own_u = false) {
  R g[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) g[q] = fn(q);
  R rn, ux, uy, uz;
  cell_macros<Q, R>(g, rn, ux, uy, uz);
  if (own_u) {
    wx = ux;
    wy = uy;
    wz = uz;
  }
  const R uu = ux * ux + uy * uy + uz * uz, ww = wx * wx + wy * wy + wz * wz;
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q] = feq_c<Q, R>(q, rn, wx, wy, wz, ww) + g[q] - feq_c<Q, R>(q, rn, ux, uy, uz, uu);
}

// Non-equilibrium extrapolation from the cell t (a slot, -1 when missing: the
// cell's own f stands in for it), val(t, q) its post-stream f_q.
template <int Q, typename R, typename VAL>
__device__ __forceinline__ void nee_at(R* f, int t, R wx, R wy, R wz, bool own_u, VAL val) {
  if (t < 0)
    nee_from<Q, R>(f, wx, wy, wz, [&](int q) { return f[q]; }, own_u);
  else
    nee_from<Q, R>(f, wx, wy, wz, [&](int q) { return val(t, q); }, own_u);
}

// nb(dx,dy,dz) returns the slot of the cell at offset d (-1: none, the cell's
// own value stands in for it); val(t, q) the post-stream f_q of the cell at
// slot t.
// NEE = false: a kernel for cases with no NEE face and no edge or corner of wet
// faces (the host checks), which leaves the extrapolation's code out.
template <int Q, typename R, bool NEE = true, typename NB, typename VAL>
__device__ __forceinline__ void wet_bc(R* f, uint16_t k, const BcCfg& bc, NB nb, VAL val) {
  const int act = ((k >> 8) & 63) & bc.wet_mask;
  if (!act) return;
  if (__popc(act) == 1) {
    const int face = __ffs(act) - 1;
    const int ax = face >> 1;
    const int sg = (face & 1) ? 1 : -1;
    if constexpr (NEE) {
      if (bc.kind[face] == MT_BC_NEE) {
        // f_q = feq_q(rho_n, u_w) + f_q(x - n) - feq_q(rho_n, u_n), all q
        const int nx = ax == 0 ? sg : 0, ny = ax == 1 ? sg : 0, nz = ax == 2 ? sg : 0;
        nee_at<Q, R>(f, nb(-nx, -ny, -nz), (R)bc.u[face][0], (R)bc.u[face][1], (R)bc.u[face][2], false, val);
        return;
      }
    }
    if (bc.kind[face] == MT_BC_OUTFLOW) {
      const int nx = ax == 0 ? sg : 0, ny = ax == 1 ? sg : 0, nz = ax == 2 ? sg : 0;
      const int t = nb(-nx, -ny, -nz);
      if (t >= 0) {
#pragma unroll
        for (int q = 1; q < Q; ++q) {
          const int ea = ax == 0 ? kEX[q] : (ax == 1 ? kEY[q] : kEZ[q]);
          if (ea * sg == 1) f[q] = val(t, q);
        }
      }
      return;
    }
    const R uw[3] = {(R)bc.u[face][0], (R)bc.u[face][1], (R)bc.u[face][2]};
    R S0 = 0, Sin = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int en = (ax == 0 ? kEX[q] : (ax == 1 ? kEY[q] : kEZ[q])) * sg;
      if (en == 0) S0 += f[q];
      else if (en == -1) Sin += f[q];
    }
    const R un = uw[ax] * (R)sg;
    const R rho = (S0 + (R)2 * Sin) / ((R)1 - un);
    const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
    R N1 = 0, N2 = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int en = (ax == 0 ? kEX[q] : (ax == 1 ? kEY[q] : kEZ[q])) * sg;
      const int e1 = t1 == 0 ? kEX[q] : kEY[q];
      if (en == 0 && e1) N1 += (R)e1 * f[q];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int en = (ax == 0 ? kEX[q] : (ax == 1 ? kEY[q] : kEZ[q])) * sg;
      const int e2 = t2 == 1 ? kEY[q] : kEZ[q];
      if (en == 0 && e2) N2 += (R)e2 * f[q];
    }
    N1 = (R)0.5 * N1 - rho * uw[t1] / (R)3.0;
    N2 = (R)0.5 * N2 - rho * uw[t2] / (R)3.0;
    // (in place: an unknown's opposite has e.n = -1 and is never overwritten)
#pragma unroll
    for (int q = 1; q < Q; ++q) {
      const int en = (ax == 0 ? kEX[q] : (ax == 1 ? kEY[q] : kEZ[q])) * sg;
      if (en == 1) {
        const int e1 = t1 == 0 ? kEX[q] : kEY[q];
        const int e2 = t2 == 1 ? kEY[q] : kEZ[q];
        const R w = (R)wq<Q>(q);
        const R eu = (R)kEX[q] * uw[0] + (R)kEY[q] * uw[1] + (R)kEZ[q] * uw[2];
        const R tr = (R)e1 * N1 + (R)e2 * N2;
        f[q] = f[qinv(q)] + (R)6 * w * rho * eu - (R)36 * w * tr;
      }
    }
    return;
  }
  if constexpr (NEE) {
    // edge / corner: non-equilibrium extrapolation from the inward diagonal cell
    int dx = 0, dy = 0, dz = 0, uface = -1;
    for (int ff = 0; ff < 6; ++ff)
      if (act & (1 << ff)) {
        const int sg = (ff & 1) ? 1 : -1, ax = ff >> 1;
        if (ax == 0) dx -= sg; else if (ax == 1) dy -= sg; else dz -= sg;
        if (bc.kind[ff] == MT_BC_ZOUHE || bc.kind[ff] == MT_BC_NEE) uface = ff;
      }
    const int t = nb(dx, dy, dz);
    if (uface >= 0)
      nee_at<Q, R>(f, t, (R)bc.u[uface][0], (R)bc.u[uface][1], (R)bc.u[uface][2], false, val);
    else
      nee_at<Q, R>(f, t, (R)0, (R)0, (R)0, true, val);
  }
}

// Solid boundary on one dis cell (SinglePointInterpolation::treat,
// ComplexBoundary_Manager.cpp:63-225): link interpolation (FH or distance
// weighted) on wall links, pull or bounce-back elsewhere, density clamp,
// eight multi-direct-forcing sweeps.  Only fcol and the density map are read,
// so the result is independent of the post-stream f it replaces.
// fc(q, dir) returns fcol_q at the neighbour along e_dir (dir 0 = the cell itself)
// or NaN-free fallback through 'present'.
template <int Q, typename R, typename FC>
__device__ __forceinline__ void ib_cell(R* f, const double* drow, R rho_b, uint32_t pm, int method, FC fc,
                                        double& fx, double& fy, double& fz) {
  f[0] = fc(0, 0);
  double sx = 0, sy = 0, sz = 0;
#pragma unroll
  for (int q = 1; q < Q; ++q) {
    const int qi = qinv(q);
    const double dist = drow[qi - 1];
    if (dist != -1.0) {
      const R finc = fc(qi, 0);
      const R fs = ((pm >> qi) & 1) ? fc(q, qi) : finc;  // fcol_q(x - e_q)
      R fnew;
      const R d = (R)dist;
      if (method == MT_IB_FH) {
        if (d < (R)0.5) {
          fnew = ((R)1 - (R)2 * d) * finc + (R)2 * d * fs;
        } else {
          const R fe = (R)wib(qi) * rho_b;
          const R i2d = (R)1 / ((R)2 * d);
          fnew = i2d * fs + ((R)1 - i2d) * fe;
        }
      } else {
        fnew = ((R)1 - d) * finc + d * fs;
      }
      f[q] = fnew;
      const double m = (double)finc + (double)fnew;
      if (kEX[q]) sx -= m * kEX[q];
      if (kEY[q]) sy -= m * kEY[q];
      if (kEZ[q]) sz -= m * kEZ[q];
    } else {
      f[q] = ((pm >> q) & 1) ? fc(q, q) : fc(qi, 0);
    }
  }
  fx += sx;
  fy += sy;
  fz += sz;
  R ra = 0;
#pragma unroll
  for (int q = 0; q < Q; ++q) ra += f[q];
  if (ra < (R)0.3) f[0] += (R)1.0 * ((R)0.3 - ra);
  else if (ra < (R)0.5) f[0] += (R)0.5 * ((R)0.5 - ra);
  else if (ra > (R)3.0) f[0] += (R)1.0 * ((R)3.0 - ra);
  else if (ra > (R)2.0) f[0] += (R)0.5 * ((R)2.0 - ra);
  for (int it = 0; it < 8; ++it) {
    R rho = 0, ux = 0, uy = 0, uz = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      rho += f[q];
      if (kEX[q]) ux += (R)kEX[q] * f[q];
      if (kEY[q]) uy += (R)kEY[q] * f[q];
      if (kEZ[q]) uz += (R)kEZ[q] * f[q];
    }
    if (rho <= (R)0) continue;
    ux /= rho;
    uy /= rho;
    uz /= rho;
    const R um = sqrt(ux * ux + uy * uy + uz * uz);
    if (um < (R)1e-10) continue;
    const R mx = (R)-0.5 * rho * ux, my = (R)-0.5 * rho * uy, mz = (R)-0.5 * rho * uz;
    R adj = 0;
#pragma unroll
    for (int q = 1; q < Q; ++q) {
      if (drow[qinv(q) - 1] == -1.0) continue;
      const R df = (R)wib(q) * (R)3.0 * ((R)kEX[q] * mx + (R)kEY[q] * my + (R)kEZ[q] * mz);
      f[q] += df;
      adj -= df;
    }
    f[0] += adj;
  }
}

// collision of one cell in place (LBGK or cumulant, optional Smagorinsky)
template <int Q, typename R>
__device__ __forceinline__ void collide_cell(R* f, R rho, R ux, R uy, R uz, R tau, R lesC, int model) {
  if (lesC > (R)0) tau = smag_tau<Q, R>(f, rho, ux, uy, uz, tau, lesC);
  if (Q == 27 && model == MT_COLLIDE_CUMULANT)
    collide_cumulant<R>(f, rho, ux, uy, uz, tau);
  else
    collide_lbgk<Q, R>(f, rho, ux, uy, uz, tau);
}

// Collision kinds the fused kernels are compiled for: the model and whether any
// member runs the Smagorinsky model.  A kernel built for one kind carries only
// that collision's code, so its registers are sized for it alone.
enum CollideKind { CK_LBGK = 0, CK_LBGK_LES = 1, CK_CUMULANT = 2, CK_CUMULANT_LES = 3 };

// the collision of collide_cell for a kind fixed at compile time (the same
// operations in the same order)
template <int Q, typename R, int CK>
__device__ __forceinline__ void collide_ck(R* f, R rho, R ux, R uy, R uz, R tau, R lesC) {
  if constexpr ((CK & 1) != 0) {
    if (lesC > (R)0) tau = smag_tau<Q, R>(f, rho, ux, uy, uz, tau, lesC);
  }
  if constexpr (Q == 27 && (CK & 2) != 0)
    collide_cumulant<R>(f, rho, ux, uy, uz, tau);
  else
    collide_lbgk<Q, R>(f, rho, ux, uy, uz, tau);
}

__device__ __forceinline__ void block_reduce3(double& x, double& y, double& z, double* lds) {
  // wave reduction then cross-wave in LDS (256 threads = 4 waves)
  for (int off = 32; off > 0; off >>= 1) {
    x += __shfl_down(x, off, 64);
    y += __shfl_down(y, off, 64);
    z += __shfl_down(z, off, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    lds[3 * w] = x;
    lds[3 * w + 1] = y;
    lds[3 * w + 2] = z;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    x = lds[0] + lds[3] + lds[6] + lds[9];
    y = lds[1] + lds[4] + lds[7] + lds[10];
    z = lds[2] + lds[5] + lds[8] + lds[11];
  }
}

// ------------------------------------------------------------------ fused step
struct FusedArgs {
  LvArgs v;
  const int32_t* __restrict__ active;
  const int32_t* __restrict__ slot_of;  // force slot of each list position (null: the position)
  int64_t n_active;
  int xmap;  // block -> brick map: 0 one contiguous range per XCD, 1 in order
  int model;
  int write_f;
  int do_bc;  // level 0 with a wet-node face in some member
  int do_ib;  // the finest level's solid-boundary links (SinglePointInterpolation::treat)
  int ib_method;
  const int32_t* __restrict__ dis_row;
  const double* __restrict__ dis;
  const uint32_t* __restrict__ wmask;  // wall links of the dis cells
  double* force_slots;  // [n_active][3] per-brick running momentum-exchange sums
  double* force_acc;
  int write_macros;
  int c2f_fcol;  // store the C2F cells' fcol (0: an explosion overwrites it before any read)
  int c2f_f;     // store the C2F cells' f (read by a separate coalescence kernel)
  // coalescence into the next coarser level (k_fused<.., true>)
  const int32_t* __restrict__ cpar;  // [n_slots / 8] parent slot of an aligned group of 8, -1 none
  int64_t ns_c;
  void* cf;         // coarse f
  void* cfcol_out;  // coarse fcol after the F2C cell's collision
  void* crho_out;
  void* cu_out;
  int c_write_macros;
  // explosions inside the update (Case::xin)
  const uint32_t* __restrict__ tmask;  // bit q: the pull source of q (x - e_q on a wall link) is an explosion target
                                       // (bit 0: the cell itself)
  const void* xbuf;                    // this pair's explosion targets (xg[pf of level L-1]),
                                       // one entry per aligned group of 8 slots
  const int32_t* __restrict__ gxrow;   // ghost_F2C parent -> row of gx_child, -1 none
  const int32_t* __restrict__ gx_child;
  void* gx_dst;                        // level L+1 targets of the next pair
  int64_t ns_f;
  void* cx_dst;                        // with COAL: this level's targets of the next pair
  // markers (finest level): the band cells' density and velocity of the new
  // fcol, [4][band capacity], as k_mk_moments computes them
  const int32_t* __restrict__ band_row;
  double* band_mom;
  int64_t n_band;  // the band's capacity (the row stride of band_mom)
};

// A value from another lane of the wave through a DPP lane permutation (CTRL:
// quad_perm [1,0,3,2] = 0xB1 reads lane ^ 1, [2,3,0,1] = 0x4E lane ^ 2,
// row_half_mirror = 0x141 lane 7 - i within each aligned group of 8).  Every
// lane must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_get(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_get(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
// AMR_transDDF_explosion (LBM_Manager.cpp:1059-1129) of one parent held in a
// lane: its post-collision fcol rescaled around the He-Luo equilibrium of its
// own moments, written to its (up to 8, -1 = none) targets.
template <int Q, typename S, typename R>
__device__ __forceinline__ void explode_lane(const R* fc, const int32_t* child, S* dst, int64_t ns, R sc, R rho0) {
  // the children are one aligned group of 8 slots of the finer level (the 2x2x2
  // block of a parent), and all receive the same values: the group's target
  // entry, [Q][ns / 8], is written once
  int t = -1;
#pragma unroll
  for (int j = 0; j < 8; ++j) t = t < 0 ? child[j] : t;
  if (t < 0) return;
  R rho, ux, uy, uz;
  cell_macros<Q, R>(fc, rho, ux, uy, uz);
  const R uu = ux * ux + uy * uy + uz * uz;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const R fe = feq_hl<Q, R>(q, rho, ux, uy, uz, uu, rho0);
    store_pop<Q, S, R>(dst, ns >> 3, q, t >> 3, fe + sc * (fc[q] - fe));
  }
}

// SinglePointInterpolation::treat (ComplexBoundary_Manager.cpp:63-225) on one
// treated cell, in the lane that streams it.  On entry f holds the pulled
// populations (fluid links: fcol_q(x + e_q), or the bounce-back fcol_inv(q)(x)
// when the source is missing).  Wall links (dis[inv(q) - 1] != -1) interpolate
// between fcol_inv(q)(x) and fcol_q(x - e_q) (Filippova-Haenel or distance
// weighted) and add their momentum exchange to the force; then the density
// clamp on f_0 and eight multi-direct-forcing sweeps.  A sweep adds
// df_q = 3 w_q e_q.dM (dM = -rho u / 2, D3Q19 weights, zero for q > 18) to the
// wall links and -sum df_q to f_0, so rho is invariant and the momentum obeys
// m <- m + 3 W dM with W = sum_wall w_q e_q e_q^T: the sweeps reduce to that
// 3-vector recurrence, and each wall link finally receives 3 w_q e_q.D with D
// the sum of the sweeps' dM (f_0: -3 V.D, V = sum_wall w_q e_q).
template <int Q, typename S, typename R, int CH, typename Args>
__device__ __forceinline__ void ib_links_cell(R* f, const Args& a, const int* N, int lx, int ly, int lz, int s,
                                              int row, uint32_t pm, double& fx, double& fy, double& fz) {
  const S* fin = (const S*)a.v.fcol_in;
  const int64_t ns = a.v.ns;
  const double* drow = a.dis + (int64_t)row * (Q - 1);
  const R rho_b = ld_real<R>(a.v.rho_in, s);
  const uint32_t wm = a.wmask[s];
  R wxx = 0, wyy = 0, wzz = 0, wxy = 0, wxz = 0, wyz = 0, vx = 0, vy = 0, vz = 0;
  // Wall links in chunks of CH directions: the chunk's loads (fraction,
  // fcol_inv(q)(x), fcol_q(x - e_q)) are issued together, then used, so a cell
  // waits for memory once per chunk instead of once per wall link (CH is set by
  // the registers the kernel can give them).
  static_for<0, (Q - 1 + CH - 1) / CH>([&](auto cc) {
    constexpr int c0 = 1 + decltype(cc)::value * CH;
    double dist_c[CH];
    R finc_c[CH], fs_c[CH];
    static_for<0, CH>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      constexpr int q = c0 + j;
      if constexpr (q < Q) {
        constexpr int qi = qinv(q);
        if ((wm >> q) & 1) {
          dist_c[j] = drow[qi - 1];
          finc_c[j] = load_pop<Q, S, R>(fin, ns, qi, s);
          fs_c[j] = f[q];  // the pull loaded fcol_q(x - e_q) (or fcol_inv(q)(x)) for a wall link
        }
      }
    });
    static_for<0, CH>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      constexpr int q = c0 + j;
      if constexpr (q < Q) {
        constexpr int qi = qinv(q);
        if ((wm >> q) & 1) {
          const R finc = finc_c[j];
          const R fs = fs_c[j];
          const R d = (R)dist_c[j];
          R fn;
          if (a.ib_method == MT_IB_FH) {
            if (d < (R)0.5) {
              fn = ((R)1 - (R)2 * d) * finc + (R)2 * d * fs;
            } else {
              // D3Q19 weight of inv(q) (|e_inv(q)| = |e_q|), zero beyond q = 18
              constexpr double w = qi < 19 ? (qnorm1(q) == 1 ? 1. / 18. : 1. / 36.) : 0.;
              const R i2d = (R)1 / ((R)2 * d);
              fn = i2d * fs + ((R)1 - i2d) * ((R)w * rho_b);
            }
          } else {
            fn = ((R)1 - d) * finc + d * fs;
          }
          f[q] = fn;
          const double m = (double)finc + (double)fn;
          fx -= m * kEX[q];
          fy -= m * kEY[q];
          fz -= m * kEZ[q];
        }
      }
    });
  });
  // W = sum_wall w_q e_q e_q^T and V = sum_wall w_q e_q over the wall links in
  // ascending q (after the links, so the sums are not held across their loads)
  static_for<1, (Q < 19 ? Q : 19)>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if ((wm >> q) & 1) {
      constexpr double w = qnorm1(q) == 1 ? 1. / 18. : 1. / 36.;
      wxx += (R)(w * kEX[q] * kEX[q]);
      wyy += (R)(w * kEY[q] * kEY[q]);
      wzz += (R)(w * kEZ[q] * kEZ[q]);
      wxy += (R)(w * kEX[q] * kEY[q]);
      wxz += (R)(w * kEX[q] * kEZ[q]);
      wyz += (R)(w * kEY[q] * kEZ[q]);
      vx += (R)(w * kEX[q]);
      vy += (R)(w * kEY[q]);
      vz += (R)(w * kEZ[q]);
    }
  });
  // density clamp through f_0 (ComplexBoundary_Manager.cpp:150-158)
  R ra = 0, m0 = 0, m1 = 0, m2 = 0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    ra += f[q];
    m0 += (R)kEX[q] * f[q];
    m1 += (R)kEY[q] * f[q];
    m2 += (R)kEZ[q] * f[q];
  }
  R dclamp = 0;
  if (ra < (R)0.3) dclamp = (R)1.0 * ((R)0.3 - ra);
  else if (ra < (R)0.5) dclamp = (R)0.5 * ((R)0.5 - ra);
  else if (ra > (R)3.0) dclamp = (R)1.0 * ((R)3.0 - ra);
  else if (ra > (R)2.0) dclamp = (R)0.5 * ((R)2.0 - ra);
  f[0] += dclamp;
  const R rho = ra + dclamp;
  R D0 = 0, D1 = 0, D2 = 0;
  if (rho > (R)0) {
    const R ir = (R)1 / rho;
    for (int it = 0; it < 8; ++it) {
      const R ux = m0 * ir, uy = m1 * ir, uz = m2 * ir;
      if (ux * ux + uy * uy + uz * uz < (R)1e-20) break;  // |u| < 1e-10: m no longer changes
      const R d0 = (R)-0.5 * rho * ux, d1 = (R)-0.5 * rho * uy, d2 = (R)-0.5 * rho * uz;
      D0 += d0;
      D1 += d1;
      D2 += d2;
      m0 += (R)3.0 * (wxx * d0 + wxy * d1 + wxz * d2);
      m1 += (R)3.0 * (wxy * d0 + wyy * d1 + wyz * d2);
      m2 += (R)3.0 * (wxz * d0 + wyz * d1 + wzz * d2);
    }
  }
  static_for<1, (Q < 19 ? Q : 19)>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr double w = qnorm1(q) == 1 ? 1. / 18. : 1. / 36.;
    if ((wm >> q) & 1) f[q] += (R)w * (R)3.0 * ((R)kEX[q] * D0 + (R)kEY[q] * D1 + (R)kEZ[q] * D2);
  });
  f[0] -= (R)3.0 * (vx * D0 + vy * D1 + vz * D2);
}

// One 64-lane block per brick.  Workgroups are dealt round-robin over the 8
// XCDs, each with its own L2; the block -> brick map below gives every XCD one
// contiguous Morton range of bricks, so a brick's neighbour bricks (its halo)
// are mostly served by the same L2.
//
// With COAL the block also coalesces (AMR_transDDF_coalescence,
// LBM_Manager.cpp:1131-1169) into the F2C parents whose children it holds:
// children of a parent are the 8 lanes of an aligned group, in the reference's
// child order; their post-stream f are summed as a DPP butterfly over the group,
// ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)), not in the reference's sequential order
// c0+c1+...+c7 (LBM_Manager.cpp:1151-1155, kept by the seam kernel k_coalesce),
// so the average can differ from the reference's in the last bit; the group's
// first lane rescales the average around its He-Luo
// equilibrium and runs the F2C cell's calculateMacros + collide, which the
// reference does at the start of the coarse level's next sub-step.
//
// fused_block is the update of block `blk` of a launch of `nb` blocks (a
// multiple of 8) for one level; k_fused launches one level, k_fused_levels
// several levels side by side.
// ROOM: the kernel already holds a variant whose registers leave no fourth wave
// per SIMD (the domain boundary's NEE extrapolation), so the solid boundary
// loads its wall links in fewer, larger chunks (fewer memory round trips).
template <int Q, typename S, int CK, bool COAL, bool WET, bool IB, bool ROOM = false, bool NEE = true>
__device__ __forceinline__ void fused_block(const FusedArgs& a, const uint32_t blk, const int64_t nb) {
  typedef typename RealOf<S>::type R;
  const int lane = threadIdx.x;
  const int64_t per_xcd = (nb + 7) >> 3;
  const int64_t bi = a.xmap ? (int64_t)blk : (int64_t)(blk & 7) * per_xcd + (blk >> 3);
  const S* fin = (const S*)a.v.fcol_in;
  S* fout = (S*)a.v.fcol_out;
  S* fst = (S*)a.v.f;
  const int64_t ns = a.v.ns;
  if (bi >= a.n_active) return;  // uniform over the block
  const int b = __builtin_amdgcn_readfirstlane(a.active[bi]);
  const int L = a.v.L;
  const MemberParams& P = a.v.mp[__builtin_amdgcn_readfirstlane(a.v.bmember[b])];
  const double p_tau = P.tau[L], p_les = P.les_C;
  R f[Q];
  if (COAL) {
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = 0;
  }
  double fx = 0, fy = 0, fz = 0;  // momentum exchange of this lane's treated cell
  const int s = (b << 6) | lane;
  // FP64: the wave issues its tables and its pull at a raised priority, so the
  // waves of a SIMD get their loads in flight before any of them takes the
  // VALU for a collision (measured 2% per coarse step on the sphere); FP32 and
  // FP16 storage, with half and a quarter of the bytes, lose by it.  (A wave
  // with no cell to stream keeps the priority to its end: it has little left.)
  if constexpr (sizeof(S) == 8) __builtin_amdgcn_s_setprio(2);
  // the cell's tables, loaded together up front (one memory round trip before
  // the pull instead of one per table)
  const uint16_t k = a.v.kind[s];
  const uint32_t pmx = a.v.pmask[s];
  const uint32_t tm = a.tmask ? a.tmask[s] : 0u;
  const int row0 = IB ? a.dis_row[s] : -1;
  const uint32_t wm0 = IB ? a.wmask[s] : 0u;
  // the coalescence parent of this lane's group and the ghost explosion row:
  // loaded with the tables, so the work after the collision waits for no load
  const int par = COAL ? a.cpar[(b << 3) | (lane >> 3)] : -1;
  const int gr = a.gxrow ? a.gxrow[s] : -1;
  {
    int N[27];
#pragma unroll
    for (int o = 0; o < 27; ++o) N[o] = __builtin_amdgcn_readfirstlane(a.v.bnbr[(int64_t)b * 27 + o]);
    // keep the cell's loads together with the brick's: the compiler would
    // otherwise sink each into the branch that first uses it, one memory round
    // trip after another
    asm volatile("" ::"v"(k), "v"(pmx), "v"(tm), "v"(row0), "v"(wm0), "v"(par), "v"(gr), "s"(p_tau), "s"(p_les));
    const int row = (IB && (k & MT_KIND_DIS) && (k & MT_KIND_FLUID)) ? row0 : -1;
    const bool treated = row >= 0;
    // A C2F cell's stream feeds the coalescence (COAL), its own stored f or
    // fcol (c2f_f, write_f, c2f_fcol) or, in dis, its fcol; in any other
    // sub-step nothing reads it before an explosion replaces it, so it is not
    // done.
    const bool c2f_live = COAL || a.c2f_f || a.write_f || a.c2f_fcol;
    if ((k & MT_KIND_FLUID) || ((k & MT_KIND_C2F) && (c2f_live || (k & MT_KIND_DIS)))) {
      int lx, ly, lz;
      brick_coords(lane, lx, ly, lz);
      // the pair's explosion targets: one entry per aligned group of 8 slots
      const S* xg = (const S*)a.xbuf;
      const int64_t ng = ns >> 3;
      f[0] = (tm & 1u) ? load_pop<Q, S, R>(xg, ng, 0, s >> 3) : load_pop<Q, S, R>(fin, ns, 0, s);
      // pull f_q(x) <- fcol_q(x + e_q); a missing source keeps the cell's own f_q
      // (Stream_Manager.cpp:53-62), except on a treated solid-boundary cell,
      // which bounces back its own fcol_inv(q) (ComplexBoundary_Manager.cpp:
      // 140-146).  The source is selected, not branched on, so all Q loads are
      // issued back to back.
      static_for<1, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int t = nbr_slot_c<kEX[q], kEY[q], kEZ[q]>(N, lx, ly, lz);
        const bool p = (pmx >> q) & 1;
        const S* pb = ((tm >> q) & 1) ? xg + (int64_t)q * ng + (t >> 3) : fin + (int64_t)q * ns + t;
        const S* src = p ? pb : (treated ? fin + (int64_t)qinv(q) * ns + s : fst + (int64_t)q * ns + s);
        if (IB && treated && ((wm0 >> q) & 1)) {
          // a wall link's population is interpolated from fcol_q(x - e_q) (or
          // fcol_inv(q)(x) when x - e_q is missing): load that here, with the
          // pull, instead of the source inside the wall; tmask bit q marks an
          // x - e_q that is an explosion target
          constexpr int qi = qinv(q);
          const int tb = nbr_slot_c<-kEX[q], -kEY[q], -kEZ[q]>(N, lx, ly, lz);
          src = ((pmx >> qi) & 1) ? (((tm >> q) & 1) ? xg + (int64_t)q * ng + (tb >> 3) : fin + (int64_t)q * ns + tb)
                                  : fin + (int64_t)qi * ns + s;
        }
        R v = (R)Store<S>::ld(src);
        if (Store<S>::shifted) v += (R)wq<Q>(q);
        f[q] = v;
      });
      if constexpr (sizeof(S) == 8) __builtin_amdgcn_s_setprio(0);
      if (treated) {
        ib_links_cell<Q, S, R, (ROOM ? (sizeof(R) == 8 ? 13 : 26) : 9)>(f, a, N, lx, ly, lz, s, row, pmx, fx, fy, fz);
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = snap_pop<Q, S, R>(q, f[q]);
      }
      if (WET && (k & 0x3F00)) {
        const int32_t* bn = a.v.bnbr;
        const uint32_t* pmk = a.v.pmask;
        auto nb = [&](int dx, int dy, int dz) { return slot_off(bn, s, dx, dy, dz); };
        // the post-stream f_q of cell t, as its own pull computes it
        auto val = [&](int t, int q) -> R {
          if (q == 0) return load_pop<Q, S, R>(fin, ns, 0, t);
          if ((pmk[t] >> q) & 1) {
            const int u = slot_off(bn, t, kEX[q], kEY[q], kEZ[q]);
            return load_pop<Q, S, R>(fin, ns, q, u);
          }
          return load_pop<Q, S, R>(fst, ns, q, t);
        };
        wet_bc<Q, R, NEE>(f, k, P.bc, nb, val);
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = snap_pop<Q, S, R>(q, f[q]);
      }
      // A C2F cell's stored f is read by a separate coalescence kernel and, for
      // the directions with no source, by its own next pull; those entries never
      // change, so with the coalescence done here only the final f is stored.
      if (((k & MT_KIND_C2F) && a.c2f_f) || a.write_f) {
#pragma unroll
        for (int q = 0; q < Q; ++q) store_pop<Q, S, R>(fst, ns, q, s, f[q]);
      }
      if (k & MT_KIND_FLUID) {
        // (in place: lat_f cells are never children in a coalescence group)
        R rho, ux, uy, uz;
        cell_macros<Q, R>(f, rho, ux, uy, uz);
        // the density/velocity maps are read only by the solid boundary (its own
        // cells, next sub-step), the reservoir NEE and the caller at the end
        if (a.write_macros || treated) {
          st_real<R>(a.v.rho_out, s, rho);
          st_real<R>(a.v.u_out, s, ux);
          st_real<R>(a.v.u_out, ns + s, uy);
          st_real<R>(a.v.u_out, 2 * ns + s, uz);
        }
        collide_ck<Q, R, CK>(f, rho, ux, uy, uz, (R)p_tau, (R)p_les);
#pragma unroll
        for (int q = 0; q < Q; ++q) store_pop<Q, S, R>(fout, ns, q, s, f[q]);
        // (loaded here, not with the cell's tables: a value live across the
        // pull costs the step more than the load's latency on marker cases)
        const int brow = a.band_row ? a.band_row[s] : -1;
        if (brow >= 0) {
          // a marker band cell: the moments of the fcol just stored (rounded
          // through the storage format, as a load would return them)
#pragma unroll
          for (int q = 0; q < Q; ++q) f[q] = snap_pop<Q, S, R>(q, f[q]);
          R br, bx, by, bz;
          cell_macros<Q, R>(f, br, bx, by, bz);
          a.band_mom[brow] = (double)br;
          a.band_mom[a.n_band + brow] = (double)bx;
          a.band_mom[2 * a.n_band + brow] = (double)by;
          a.band_mom[3 * a.n_band + brow] = (double)bz;
        }
        if (a.gxrow) {
          // a ghost_F2C parent explodes its next collision into the finer
          // level's ghost_C2F targets of the next pair
          if (gr >= 0) {
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = snap_pop<Q, S, R>(q, f[q]);
            explode_lane<Q, S, R>(f, a.gx_child + (int64_t)gr * 8, (S*)a.gx_dst, a.ns_f, (R)P.xscale[L + 1],
                                  (R)P.rho0);
          }
        }
      } else if (a.c2f_fcol || (k & MT_KIND_DIS)) {
        // (a C2F cell in dis is never exploded into: it keeps the stream-copy fcol)
#pragma unroll
        for (int q = 0; q < Q; ++q) store_pop<Q, S, R>(fout, ns, q, s, f[q]);
      }
    }
  }
  if (IB) {
    // this call's force: one running slot per brick, fixed block -> brick map
    for (int off = 32; off > 0; off >>= 1) {
      fx += __shfl_down(fx, off, 64);
      fy += __shfl_down(fy, off, 64);
      fz += __shfl_down(fz, off, 64);
    }
    if (lane == 0) {
      if (fx != 0.0 || fy != 0.0 || fz != 0.0) {
        const int64_t fs = a.slot_of ? (int64_t)a.slot_of[bi] : bi;
        a.force_slots[3 * fs] += fx;
        a.force_slots[3 * fs + 1] += fy;
        a.force_slots[3 * fs + 2] += fz;
      }
      if (bi == 0) a.force_acc[7] += 1.0;  // treat() calls folded into the slots
    }
  }
  if (!COAL) return;
  if (__ballot(par >= 0) == 0) return;  // uniform over the wave
  // every lane of a group with a parent is a streamed C2F cell holding its
  // post-stream f; the sums run over all 64 lanes (all active here), as a
  // butterfly within each aligned group of 8 (((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)),
  // the same bits in every lane of the group since the additions commute
  // exactly; the last step pairs lane i with lane 7 - i, whose partial sum is
  // that of lane i ^ 4)
  // (in place: a lane's own f is not read again; every lane of the group
  // holds the group's average)
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    R sum = f[q];
    sum += dpp_get<0xB1>(sum);
    sum += dpp_get<0x4E>(sum);
    sum += dpp_get<0x141>(sum);
    f[q] = sum / (R)8.0;
  }
  if (!(lane & 7) && par >= 0) {
    R* avg = f;
    R rho, ux, uy, uz;
    cell_macros<Q, R>(avg, rho, ux, uy, uz);
    R uu = ux * ux + uy * uy + uz * uz;
    const R sc = (R)P.cscale[L - 1];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const R fe = feq_hl<Q, R>(q, rho, ux, uy, uz, uu, (R)P.rho0);
      avg[q] = snap_pop<Q, S, R>(q, fe + sc * (avg[q] - fe));
      store_pop<Q, S, R>((S*)a.cf, a.ns_c, q, par, avg[q]);
    }
    cell_macros<Q, R>(avg, rho, ux, uy, uz);
    if (a.c_write_macros) {
      st_real<R>(a.crho_out, par, rho);
      st_real<R>(a.cu_out, par, ux);
      st_real<R>(a.cu_out, a.ns_c + par, uy);
      st_real<R>(a.cu_out, 2 * a.ns_c + par, uz);
    }
    collide_ck<Q, R, CK>(avg, rho, ux, uy, uz, (R)P.tau[L - 1], (R)P.les_C);
#pragma unroll
    for (int q = 0; q < Q; ++q) store_pop<Q, S, R>((S*)a.cfcol_out, a.ns_c, q, par, avg[q]);
    if (a.cx_dst) {
      // the parent's explosion for the next pair: its 8 children (this group)
      // all receive these values, kept once as the group's target entry
      // (children in dis, which are not exploded into (LBM_Manager.cpp:1095),
      // never read it)
#pragma unroll
      for (int q = 0; q < Q; ++q) avg[q] = snap_pop<Q, S, R>(q, avg[q]);
      cell_macros<Q, R>(avg, rho, ux, uy, uz);
      uu = ux * ux + uy * uy + uz * uz;
      const R sx = (R)P.xscale[L];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const R fe = feq_hl<Q, R>(q, rho, ux, uy, uz, uu, (R)P.rho0);
        store_pop<Q, S, R>((S*)a.cx_dst, ns >> 3, q, s >> 3, fe + sx * (avg[q] - fe));
      }
    }
  }
}

template <int Q, typename S, int CK, bool COAL, bool WET, bool IB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_fused(FusedArgs a) {
  fused_block<Q, S, CK, COAL, WET, IB>(a, blockIdx.x, gridDim.x);
}

// Several levels' fused updates in one launch: segment i owns the blocks
// [end[i-1], end[i]) (multiples of 8, so each segment keeps its XCD map) and
// runs the instantiation variant[i] = COAL | WET << 1 | IB << 2.  A coarse
// step issued this way is one chain of launches on one stream, with no
// cross-stream waits.
struct LevelsArgs {
  int n;
  int variant[kMaxLevels];
  uint32_t end[kMaxLevels];
  FusedArgs seg[kMaxLevels];
};

// The variants a launch of several levels may hold (bit 0: a level with the
// domain boundary, bit 1: a level with the solid boundary, bit 2: the domain
// boundary's NEE extrapolation): the kernel carries only those, so a launch
// without them is not sized for their registers.
enum { VM_WET = 1, VM_IB = 2, VM_NEE = 4 };

template <int Q, typename S, int CK, int VM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_fused_levels(LevelsArgs m) {
  const uint32_t blk = blockIdx.x;
  int i = 0;
  while (i + 1 < m.n && blk >= m.end[i]) ++i;
  const uint32_t b0 = i ? m.end[i - 1] : 0u;
  const FusedArgs& a = m.seg[i];
  const int64_t nb = (int64_t)(m.end[i] - b0);
  switch (m.variant[i]) {
    case 0: fused_block<Q, S, CK, false, false, false>(a, blk - b0, nb); break;
    case 1: fused_block<Q, S, CK, true, false, false>(a, blk - b0, nb); break;
    case 2:
      if constexpr ((VM & VM_WET) != 0) fused_block<Q, S, CK, false, true, false, false, (VM & VM_NEE) != 0>(a, blk - b0, nb);
      break;
    case 4:
      if constexpr ((VM & VM_IB) != 0) fused_block<Q, S, CK, false, false, true, (VM & VM_NEE) != 0>(a, blk - b0, nb);
      break;
    case 5:
      if constexpr ((VM & VM_IB) != 0) fused_block<Q, S, CK, true, false, true, (VM & VM_NEE) != 0>(a, blk - b0, nb);
      break;
    default: break;  // (a level with both boundaries is a single-level case)
  }
}

// Launch with start/stop events attached to the dispatch itself (timing mode):
// the events bracket the kernel's execution, not the queue around it.
template <typename Kern, typename Args>
static int launch_events(Kern kern, int64_t nb, int bs, hipStream_t st, Args& a, hipEvent_t e0, hipEvent_t e1) {
  if (e0) {
    void* args[] = {&a};
    MT_HIP(hipExtLaunchKernel((const void*)kern, dim3((uint32_t)nb), dim3(bs), args, 0, st, e0, e1, 0));
  } else {
    kern<<<nb, bs, 0, st>>>(a);
  }
  return MT_OK;
}



// Launchers of one (Q, S) pair, instantiated in mt_fused_*.hip: the fused
// update of one level (variant = COAL | WET << 1 | IB << 2) and of several
// levels in one launch (vm: the variants the launch holds), for collision kind
// ck.
template <int Q, typename S, int CK>
static int launch_fused_ck(int variant, FusedArgs& a, int64_t nb, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  switch (variant) {
    case 0: return launch_events(k_fused<Q, S, CK, false, false, false>, nb, 64, st, a, e0, e1);
    case 1: return launch_events(k_fused<Q, S, CK, true, false, false>, nb, 64, st, a, e0, e1);
    case 2: return launch_events(k_fused<Q, S, CK, false, true, false>, nb, 64, st, a, e0, e1);
    case 4: return launch_events(k_fused<Q, S, CK, false, false, true>, nb, 64, st, a, e0, e1);
    case 5: return launch_events(k_fused<Q, S, CK, true, false, true>, nb, 64, st, a, e0, e1);
    default: return launch_events(k_fused<Q, S, CK, false, true, true>, nb, 64, st, a, e0, e1);
  }
}

template <int Q, typename S, int CK>
static int launch_fused_levels_ck(int vm, LevelsArgs& m, int64_t nb, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  switch (vm & 7) {
    case 0: case 4: return launch_events(k_fused_levels<Q, S, CK, 0>, nb, 64, st, m, e0, e1);
    case 1: return launch_events(k_fused_levels<Q, S, CK, VM_WET>, nb, 64, st, m, e0, e1);
    case 5: return launch_events(k_fused_levels<Q, S, CK, VM_WET | VM_NEE>, nb, 64, st, m, e0, e1);
    case 2: case 6: return launch_events(k_fused_levels<Q, S, CK, VM_IB>, nb, 64, st, m, e0, e1);
    case 3: return launch_events(k_fused_levels<Q, S, CK, VM_WET | VM_IB>, nb, 64, st, m, e0, e1);
    default: return launch_events(k_fused_levels<Q, S, CK, VM_WET | VM_IB | VM_NEE>, nb, 64, st, m, e0, e1);
  }
}

template <int Q, typename S>
int launch_fused_qs(int ck, int variant, FusedArgs& a, int64_t nb, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
  if constexpr (Q == 27) {
    switch (ck) {
      case CK_LBGK: return launch_fused_ck<Q, S, CK_LBGK>(variant, a, nb, st, e0, e1);
      case CK_LBGK_LES: return launch_fused_ck<Q, S, CK_LBGK_LES>(variant, a, nb, st, e0, e1);
      case CK_CUMULANT: return launch_fused_ck<Q, S, CK_CUMULANT>(variant, a, nb, st, e0, e1);
      default: return launch_fused_ck<Q, S, CK_CUMULANT_LES>(variant, a, nb, st, e0, e1);
    }
  } else {
    if (ck & 2) {
      set_error("the cumulant collision needs Q = 27");
      return MT_ERR_STATE;
    }
    return (ck & 1) ? launch_fused_ck<Q, S, CK_LBGK_LES>(variant, a, nb, st, e0, e1)
                    : launch_fused_ck<Q, S, CK_LBGK>(variant, a, nb, st, e0, e1);
  }
}

template <int Q, typename S>
int launch_fused_levels_qs(int ck, int vm, LevelsArgs& m, int64_t nb, hipStream_t st, hipEvent_t e0,
                           hipEvent_t e1) {
  if constexpr (Q == 27) {
    switch (ck) {
      case CK_LBGK: return launch_fused_levels_ck<Q, S, CK_LBGK>(vm, m, nb, st, e0, e1);
      case CK_LBGK_LES: return launch_fused_levels_ck<Q, S, CK_LBGK_LES>(vm, m, nb, st, e0, e1);
      case CK_CUMULANT: return launch_fused_levels_ck<Q, S, CK_CUMULANT>(vm, m, nb, st, e0, e1);
      default: return launch_fused_levels_ck<Q, S, CK_CUMULANT_LES>(vm, m, nb, st, e0, e1);
    }
  } else {
    if (ck & 2) {
      set_error("the cumulant collision needs Q = 27");
      return MT_ERR_STATE;
    }
    return (ck & 1) ? launch_fused_levels_ck<Q, S, CK_LBGK_LES>(vm, m, nb, st, e0, e1)
                    : launch_fused_levels_ck<Q, S, CK_LBGK>(vm, m, nb, st, e0, e1);
  }
}

#define MT_FUSED_INSTANTIATE(Q, S)                                                                           \
  template int launch_fused_qs<Q, S>(int, int, FusedArgs&, int64_t, hipStream_t, hipEvent_t, hipEvent_t); \
  template int launch_fused_levels_qs<Q, S>(int, int, LevelsArgs&, int64_t, hipStream_t, hipEvent_t, hipEvent_t);
#define MT_FUSED_EXTERN(Q, S)                                                                                       \
  extern template int launch_fused_qs<Q, S>(int, int, FusedArgs&, int64_t, hipStream_t, hipEvent_t, hipEvent_t); \
  extern template int launch_fused_levels_qs<Q, S>(int, int, LevelsArgs&, int64_t, hipStream_t, hipEvent_t, hipEvent_t);

}  // namespace mt
