// mt_kernels.hip -- the seam kernels and the host side of the per-level update.
//
//  * seam kernels, one per reference method (calculateMacros, collide, stream,
//    explosion, coalescence, boundary strategies, treat), each reading and
//    writing the single logical f / fcol / rho / u state exactly where the
//    reference's method does;
//  * the launches of the fused update (mt_fused.h) that mt_step issues, the
//    markers' kernels, reductions, diagnostics and the remap.
#include "mt_fused.h"

namespace mt {

// the fused kernels are instantiated per (Q, S) in mt_fused_*.hip
MT_FUSED_EXTERN(19, double)
MT_FUSED_EXTERN(19, float)
MT_FUSED_EXTERN(19, _Float16)
MT_FUSED_EXTERN(27, double)
MT_FUSED_EXTERN(27, float)
MT_FUSED_EXTERN(27, _Float16)

// fixed-order sum of the per-block IB force partials into the accumulator
// (LBM_Manager::accumIBForce, LBM_Manager.cpp:846-852); last[3] gets this call
__global__ void k_force_reduce(const double* part, int64_t nb, double* acc) {
  __shared__ double lds[3 * 256];
  double x = 0, y = 0, z = 0;
  for (int64_t i = threadIdx.x; i < nb; i += 256) {
    x += part[3 * i];
    y += part[3 * i + 1];
    z += part[3 * i + 2];
  }
  lds[3 * threadIdx.x] = x;
  lds[3 * threadIdx.x + 1] = y;
  lds[3 * threadIdx.x + 2] = z;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sx = 0, sy = 0, sz = 0;
    for (int i = 0; i < 256; ++i) {
      sx += lds[3 * i];
      sy += lds[3 * i + 1];
      sz += lds[3 * i + 2];
    }
    acc[0] += sx;
    acc[1] += sy;
    acc[2] += sz;
    acc[3] += 1.0;
    acc[4] = sx;
    acc[5] = sy;
    acc[6] = sz;
  }
}

// ------------------------------------------------------------------ seams
struct SeamArgs {
  LvArgs v;
  int model;
};

// LBM_Manager::calculateMacros(level) (LBM_Manager.cpp:1003-1050)
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_macros(SeamArgs a) {
  typedef typename RealOf<S>::type R;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.v.ns) return;
  if (!(a.v.kind[s] & (MT_KIND_FLUID | MT_KIND_F2C))) return;
  R f[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q] = load_pop<Q, S, R>((const S*)a.v.f, a.v.ns, q, s);
  R rho, ux, uy, uz;
  cell_macros<Q, R>(f, rho, ux, uy, uz);
  st_real<R>(a.v.rho_out, s, rho);
  st_real<R>(a.v.u_out, s, ux);
  st_real<R>(a.v.u_out, a.v.ns + s, uy);
  st_real<R>(a.v.u_out, 2 * a.v.ns + s, uz);
}

// Collision_Model::collide(level): lat_f and F2C collide with the map moments,
// C2F copy f -> fcol (Collision_Manager.cpp:41-52)
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_collide(SeamArgs a) {
  typedef typename RealOf<S>::type R;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.v.ns) return;
  const uint16_t k = a.v.kind[s];
  if (!(k & (MT_KIND_FLUID | MT_KIND_F2C | MT_KIND_C2F))) return;
  R f[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q] = load_pop<Q, S, R>((const S*)a.v.f, a.v.ns, q, s);
  if (k & (MT_KIND_FLUID | MT_KIND_F2C)) {
    const R rho = ld_real<R>(a.v.rho_in, s);
    const R ux = ld_real<R>(a.v.u_in, s), uy = ld_real<R>(a.v.u_in, a.v.ns + s),
            uz = ld_real<R>(a.v.u_in, 2 * a.v.ns + s);
    const MemberParams& P = member_of(a.v, s);
    collide_cell<Q, R>(f, rho, ux, uy, uz, (R)P.tau[a.v.L], (R)P.les_C, a.model);
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) store_pop<Q, S, R>((S*)a.v.fcol_out, a.v.ns, q, s, f[q]);
}

// ABPatternPull::stream(level) (Stream_Manager.cpp:43-76)
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_stream(SeamArgs a) {
  typedef typename RealOf<S>::type R;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.v.ns) return;
  if (!(a.v.kind[s] & (MT_KIND_FLUID | MT_KIND_C2F))) return;
  const S* fin = (const S*)a.v.fcol_in;
  S* fst = (S*)a.v.f;
  const uint32_t pm = a.v.pmask[s];
  store_pop<Q, S, R>(fst, a.v.ns, 0, (int)s, load_pop<Q, S, R>(fin, a.v.ns, 0, (int)s));
#pragma unroll
  for (int q = 1; q < Q; ++q)
    if ((pm >> q) & 1) {
      const int t = slot_off(a.v.bnbr, (int)s, kEX[q], kEY[q], kEZ[q]);
      store_pop<Q, S, R>(fst, a.v.ns, q, (int)s, load_pop<Q, S, R>(fin, a.v.ns, q, t));
    }
}

// level-0 wet-node boundaries applied to f after stream
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_bc_wet(SeamArgs a) {
  typedef typename RealOf<S>::type R;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.v.ns) return;
  const uint16_t k = a.v.kind[s];
  const BcCfg& bc = member_of(a.v, s).bc;
  if (!(k & MT_KIND_FLUID) || !(((k >> 8) & 63) & bc.wet_mask)) return;
  S* fst = (S*)a.v.f;
  R f[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q] = load_pop<Q, S, R>(fst, a.v.ns, q, (int)s);
  const int32_t* bn = a.v.bnbr;
  const int64_t ns = a.v.ns;
  auto nb = [&](int dx, int dy, int dz) { return slot_off(bn, (int)s, dx, dy, dz); };
  auto val = [&](int t, int q) -> R { return load_pop<Q, S, R>(fst, ns, q, t); };
  wet_bc<Q, R>(f, k, bc, nb, val);
  // every cell this writes is read by no other boundary cell: in place is Jacobi
#pragma unroll
  for (int q = 0; q < Q; ++q) store_pop<Q, S, R>(fst, a.v.ns, q, (int)s, f[q]);
}

// Non-equilibrium extrapolation on the reservoir cells
// (BC/NonEquilibrumExtrapolation.cpp:50-66), Jacobi form: every read of f is
// of the pre-pass state; results staged in tmp[Q][n] then scattered.
template <int Q, typename S>
__global__ void k_nee(SeamArgs a, const int32_t* list, int64_t n, double* tmp) {
  typedef typename RealOf<S>::type R;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = list[i];
  const uint16_t k = a.v.kind[s];
  const BcCfg& bc = member_of(a.v, s).bc;
  const int fm = (k >> 8) & 63;
  const S* fst = (const S*)a.v.f;
  const R rb = ld_real<R>(a.v.rho_in, s);
  // the face strategy that owns this cell: West/East own their edges and corners,
  // South/North the remaining edges, Bot/Top their face
  // (header/BC/NonEquilibrumExtrapolation.hpp:24-109)
  int owner = (fm & 1) ? 0 : (fm & 2) ? 1 : (fm & 4) ? 2 : (fm & 8) ? 3 : (fm & 16) ? 4 : 5;
  const R bx = (R)bc.u[owner][0], by = (R)bc.u[owner][1], bz = (R)bc.u[owner][2];
  const R bb = bx * bx + by * by + bz * bz;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    bool ok = true;
    for (int ff = 0; ff < 6; ++ff)
      if (fm & (1 << ff)) {
        const int ax = ff >> 1, sg = (ff & 1) ? 1 : -1;
        const int ea = ax == 0 ? kEX[q] : (ax == 1 ? kEY[q] : kEZ[q]);
        if (ea * sg > 0) ok = false;
      }
    double out = (double)load_pop<Q, S, R>(fst, a.v.ns, q, s);
    const int t = ok ? slot_off(a.v.bnbr, s, kEX[q], kEY[q], kEZ[q]) : -1;
    if (t >= 0) {
      const R rn = ld_real<R>(a.v.rho_in, t);
      const R ux = ld_real<R>(a.v.u_in, t), uy = ld_real<R>(a.v.u_in, a.v.ns + t),
              uz = ld_real<R>(a.v.u_in, 2 * a.v.ns + t);
      const R fw = feq_c<Q, R>(q, rb, bx, by, bz, bb);
      const R fn = feq_c<Q, R>(q, rn, ux, uy, uz, ux * ux + uy * uy + uz * uz);
      out = (double)(fw + load_pop<Q, S, R>(fst, a.v.ns, q, t) - fn);
    }
    tmp[(int64_t)q * n + i] = out;
  }
}
template <int Q, typename S>
__global__ void k_nee_scatter(SeamArgs a, const int32_t* list, int64_t n, const double* tmp) {
  typedef typename RealOf<S>::type R;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int q = 0; q < Q; ++q) store_pop<Q, S, R>((S*)a.v.f, a.v.ns, q, list[i], (R)tmp[(int64_t)q * n + i]);
}

// SinglePointInterpolation::treat as a seam: f of dis cells from fcol + rho map
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_ib(SeamArgs a, const int32_t* dis_row, const double* dis, int method,
                                            double* force_part) {
  typedef typename RealOf<S>::type R;
  __shared__ double red[12];
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double fx = 0, fy = 0, fz = 0;
  if (s < a.v.ns && method == MT_IB_UNIFORM) {
    // UniformBoundary::treat (ComplexBoundary_Manager.cpp:15-40): every dis cell,
    // directions in order; a wall link q writes f_inv(q) =
    // (d (fcol_inv(q) + fcol_q) + (1 - d) fcol_q(x - e_q)) / (1 + d), any other
    // direction pulls f_q = fcol_q(x + e_q); later directions overwrite earlier
    // writes as the loop does; f_0 and the force are untouched
    const uint16_t k = a.v.kind[s];
    if ((k & MT_KIND_DIS) && dis_row[s] >= 0) {
      const S* fin = (const S*)a.v.fcol_in;
      const int64_t ns = a.v.ns;
      const int32_t* bn = a.v.bnbr;
      const uint32_t pm = a.v.pmask[s];
      const double* drow = dis + (int64_t)dis_row[s] * (Q - 1);
      R f[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) f[q] = load_pop<Q, S, R>((const S*)a.v.f, ns, q, (int)s);
#pragma unroll
      for (int q = 1; q < Q; ++q) {
        const int qi = qinv(q);
        const R d = (R)drow[qi - 1];
        if (drow[qi - 1] != -1.0) {
          if ((pm >> qi) & 1) {
            const int t = slot_off(bn, (int)s, kEX[qi], kEY[qi], kEZ[qi]);
            const R fo = load_pop<Q, S, R>(fin, ns, q, t);
            f[qi] = (d * (load_pop<Q, S, R>(fin, ns, qi, (int)s) + load_pop<Q, S, R>(fin, ns, q, (int)s)) +
                     ((R)1 - d) * fo) / ((R)1 + d);
          }
        } else if ((pm >> q) & 1) {
          f[q] = load_pop<Q, S, R>(fin, ns, q, slot_off(bn, (int)s, kEX[q], kEY[q], kEZ[q]));
        }
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) store_pop<Q, S, R>((S*)a.v.f, ns, q, (int)s, f[q]);
    }
  } else if (s < a.v.ns) {
    const uint16_t k = a.v.kind[s];
    if ((k & MT_KIND_DIS) && (k & MT_KIND_FLUID) && dis_row[s] >= 0) {
      const S* fin = (const S*)a.v.fcol_in;
      const int64_t ns = a.v.ns;
      const int32_t* bn = a.v.bnbr;
      const uint32_t pm = a.v.pmask[s];
      const double* drow = dis + (int64_t)dis_row[s] * (Q - 1);
      const R rho_b = ld_real<R>(a.v.rho_in, s);
      auto fc = [&](int q, int dir) -> R {
        int t = (int)s;
        if (dir) t = slot_off(bn, (int)s, kEX[dir], kEY[dir], kEZ[dir]);
        return load_pop<Q, S, R>(fin, ns, q, t);
      };
      R f[Q];
      ib_cell<Q, R>(f, drow, rho_b, pm, method, fc, fx, fy, fz);
#pragma unroll
      for (int q = 0; q < Q; ++q) store_pop<Q, S, R>((S*)a.v.f, ns, q, (int)s, f[q]);
    }
  }
  block_reduce3(fx, fy, fz, red);
  if (threadIdx.x == 0) {
    force_part[3 * blockIdx.x] = fx;
    force_part[3 * blockIdx.x + 1] = fy;
    force_part[3 * blockIdx.x + 2] = fz;
  }
}

// ------------------------------------------------------------------ transfers
struct XferArgs {
  int64_t ns_c, ns_f;
  const void* src;     // coarse fcol (explosion) / fine f (coalescence)
  void* dst;           // fine fcol (explosion) / coarse f (coalescence)
  void* dst2;          // explosion: second fine fcol buffer for ghost_C2F targets
  int group;           // explosion: dst holds one value per aligned group of 8 children
  const int32_t* list; // parent slots
  int64_t n;
  const int32_t* child;  // [n][8]
  // factors, relaxation and reference density from the parent's member
  const MemberParams* mp;
  const int32_t* bmember_c;  // coarse level's brick -> member
  int Lc;                    // coarse level
  // coalescence + collision of the F2C cell
  int collide;
  void* fcol_out;
  void* rho_out;
  void* u_out;
  int model;
  int write_macros;
};

// Explosion: one block of 256 threads serves 8 parents.  Per-parent arithmetic
// runs with lane q of a 32-lane group owning population q (the parent's moments
// come from an LDS copy, every lane reading the same words); global memory is
// touched (parent, child)-wise: a wave's 64 lanes are 8 parents x 8 children,
// so for one q they access 8 runs of 8 consecutive slots.

// AMR_transDDF_explosion (LBM_Manager.cpp:1059-1129): the parent's fcol is
// rescaled around its He-Luo equilibrium and written to its (up to 8) children;
// parents at index >= n_plain are ghost_F2C parents whose ghost_C2F targets are
// mirrored into the second fcol buffer (the layer is read by both fine sub-steps).
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_explode(XferArgs a, int64_t n_plain) {
  typedef typename RealOf<S>::type R;
  __shared__ R lf[8][32];
  __shared__ R lv[8][32];
  const int pl = threadIdx.x >> 5, q = threadIdx.x & 31;
  const int64_t i = (int64_t)blockIdx.x * 8 + pl;
  if (i < a.n && q < Q) lf[pl][q] = load_pop<Q, S, R>((const S*)a.src, a.ns_c, q, a.list[i]);
  __syncthreads();
  if (i < a.n && q < Q) {
    R fc[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) fc[r] = lf[pl][r];
    R rho, ux, uy, uz;
    cell_macros<Q, R>(fc, rho, ux, uy, uz);
    const R uu = ux * ux + uy * uy + uz * uz;
    const MemberParams& P = a.mp[a.bmember_c[a.list[i] >> 6]];
    const R fe = feq_hl<Q, R>(q, rho, ux, uy, uz, uu, (R)P.rho0);
    lv[pl][q] = fe + (R)P.xscale[a.Lc + 1] * (lf[pl][q] - fe);
  }
  __syncthreads();
  const int pair = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int pp = pair >> 3, j = pair & 7;
  const int64_t ip = (int64_t)blockIdx.x * 8 + pp;
  if (ip >= a.n) return;
  const int dst = a.child[ip * 8 + j];
  if (dst < 0) return;
  if (a.group) {
    // the children are one aligned group of 8 slots and all receive these
    // values: the group's entry is written once, by its first present child
    for (int jj = 0; jj < j; ++jj)
      if (a.child[ip * 8 + jj] >= 0) return;
    for (int r = w; r < Q; r += 4) store_pop<Q, S, R>((S*)a.dst, a.ns_f >> 3, r, dst >> 3, lv[pp][r]);
    return;
  }
  const bool ghost = ip >= n_plain;
  for (int r = w; r < Q; r += 4) {
    store_pop<Q, S, R>((S*)a.dst, a.ns_f, r, dst, lv[pp][r]);
    if (ghost) store_pop<Q, S, R>((S*)a.dst2, a.ns_f, r, dst, lv[pp][r]);
  }
}

// AMR_transDDF_coalescence (LBM_Manager.cpp:1131-1169), 16 parents per 64-lane
// block.  For a group of 8 parents the lanes are (parent, child): one load per
// population reads all 64 children (8 runs of 8 consecutive slots), and the 8
// children of a parent are summed in child order with lane shuffles and divided
// by 8.  Lane p < 16 then rescales its parent around the He-Luo equilibrium of
// the average and, with collide set, runs the F2C cell's next calculateMacros +
// collide, so the collision uses all 64 lanes' worth of parents in one wave.
template <int Q, typename S>
__global__ __launch_bounds__(64) void k_coalesce(XferArgs a) {
  typedef typename RealOf<S>::type R;
  constexpr int P = 16;
  __shared__ R ls[P][Q];
  const int lane = threadIdx.x, pp = lane >> 3, j = lane & 7, b0 = lane & ~7;
  const int64_t base = (int64_t)blockIdx.x * P;
#pragma unroll
  for (int g = 0; g < P / 8; ++g) {
    const int64_t ip = base + g * 8 + pp;
    const bool ok = ip < a.n;
    const int ch = ok ? a.child[ip * 8 + j] : 0;
    R v[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) v[r] = ok ? load_pop<Q, S, R>((const S*)a.src, a.ns_f, r, ch) : (R)0;
#pragma unroll
    for (int r = 0; r < Q; ++r) {
      R sum = __shfl(v[r], b0, 64);
#pragma unroll
      for (int t = 1; t < 8; ++t) sum += __shfl(v[r], b0 + t, 64);
      if (j == 0) ls[g * 8 + pp][r] = sum / (R)8.0;
    }
  }
  __syncthreads();
  const int64_t i = base + lane;
  if (lane >= P || i >= a.n) return;
  const int p = a.list[i];
  const MemberParams& M = a.mp[a.bmember_c[p >> 6]];
  R f[Q];
#pragma unroll
  for (int r = 0; r < Q; ++r) f[r] = ls[lane][r];
  R rho, ux, uy, uz;
  cell_macros<Q, R>(f, rho, ux, uy, uz);
  const R uu = ux * ux + uy * uy + uz * uz;
  const R sc = (R)M.cscale[a.Lc];
#pragma unroll
  for (int r = 0; r < Q; ++r) {
    const R fe = feq_hl<Q, R>(r, rho, ux, uy, uz, uu, (R)M.rho0);
    f[r] = snap_pop<Q, S, R>(r, fe + sc * (f[r] - fe));
    store_pop<Q, S, R>((S*)a.dst, a.ns_c, r, p, f[r]);
  }
  if (!a.collide) return;
  cell_macros<Q, R>(f, rho, ux, uy, uz);
  if (a.write_macros) {
    st_real<R>(a.rho_out, p, rho);
    st_real<R>(a.u_out, p, ux);
    st_real<R>(a.u_out, a.ns_c + p, uy);
    st_real<R>(a.u_out, 2 * a.ns_c + p, uz);
  }
  collide_cell<Q, R>(f, rho, ux, uy, uz, (R)M.tau[a.Lc], (R)M.les_C, a.model);
#pragma unroll
  for (int r = 0; r < Q; ++r) store_pop<Q, S, R>((S*)a.fcol_out, a.ns_c, r, p, f[r]);
}

// ------------------------------------------------------------------ state I/O
// allocMacrosAndDDF (LBM_Manager.cpp:482-592): f = fcol = feq(rho0,u0); solid
// cells zero with rho = 1, u = 0
template <int Q, typename S>
__global__ void k_init(LvArgs v, void* fcol2, void* rho2, void* u2) {
  typedef typename RealOf<S>::type R;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= v.ns) return;
  const uint16_t k = v.kind[s];
  const bool sol = k & MT_KIND_SOLID;
  const MemberParams& P = member_of(v, s);
  const double rho0 = P.rho0, u0x = P.u0[0], u0y = P.u0[1], u0z = P.u0[2];
  const R uu = (R)(u0x * u0x + u0y * u0y + u0z * u0z);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    R fe = (k && !sol) ? feq_c<Q, R>(q, (R)rho0, (R)u0x, (R)u0y, (R)u0z, uu) : (R)0;
    store_pop<Q, S, R>((S*)v.f, v.ns, q, (int)s, fe);
    store_pop<Q, S, R>((S*)v.fcol_out, v.ns, q, (int)s, fe);
    store_pop<Q, S, R>((S*)fcol2, v.ns, q, (int)s, fe);
  }
  R r = 0, a = 0, b = 0, c = 0;
  if (sol) r = 1;
  else if (k & (MT_KIND_FLUID | MT_KIND_F2C | MT_KIND_C2F | MT_KIND_BCRES)) {
    r = (R)rho0; a = (R)u0x; b = (R)u0y; c = (R)u0z;
  }
  void* rr[2] = {v.rho_out, rho2};
  void* uuv[2] = {v.u_out, u2};
  for (int j = 0; j < 2; ++j) {
    st_real<R>(rr[j], s, r);
    st_real<R>(uuv[j], s, a);
    st_real<R>(uuv[j], v.ns + s, b);
    st_real<R>(uuv[j], 2 * v.ns + s, c);
  }
}

template <int Q, typename S>
__global__ void k_gather(LvArgs v, int field, const int32_t* slots, int64_t n, double* out, const void* src) {
  typedef typename RealOf<S>::type R;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = slots[i];
  if (field <= MT_FIELD_FCOL) {
    for (int q = 0; q < Q; ++q) out[(int64_t)q * n + i] = (double)load_pop<Q, S, R>((const S*)src, v.ns, q, s);
  } else if (field == MT_FIELD_RHO) {
    out[i] = (double)ld_real<R>(src, s);
  } else {
    for (int c = 0; c < 3; ++c) out[(int64_t)c * n + i] = (double)ld_real<R>(src, (int64_t)c * v.ns + s);
  }
}
template <int Q, typename S>
__global__ void k_scatter(LvArgs v, int field, const int32_t* slots, int64_t n, const double* in, void* dst) {
  typedef typename RealOf<S>::type R;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = slots[i];
  if (field <= MT_FIELD_FCOL) {
    for (int q = 0; q < Q; ++q) store_pop<Q, S, R>((S*)dst, v.ns, q, s, (R)in[(int64_t)q * n + i]);
  } else if (field == MT_FIELD_RHO) {
    st_real<R>(dst, s, (R)in[i]);
  } else {
    for (int c = 0; c < 3; ++c) st_real<R>(dst, (int64_t)c * v.ns + s, (R)in[(int64_t)c * n + i]);
  }
}

// mass, momentum and NaN count over lat_f and F2C from f (FP64, fixed order)
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_reduce(LvArgs v, double* part) {
  typedef typename RealOf<S>::type R;
  __shared__ double lds[5 * 256];
  double m = 0, px = 0, py = 0, pz = 0, nn = 0;
  for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < v.ns; s += (int64_t)gridDim.x * 256) {
    if (!(v.kind[s] & (MT_KIND_FLUID | MT_KIND_F2C))) continue;
    for (int q = 0; q < Q; ++q) {
      const double fq = (double)load_pop<Q, S, R>((const S*)v.f, v.ns, q, (int)s);
      if (fq != fq) nn += 1;
      m += fq;
      px += kEX[q] * fq;
      py += kEY[q] * fq;
      pz += kEZ[q] * fq;
    }
  }
  double vals[5] = {m, px, py, pz, nn};
  for (int j = 0; j < 5; ++j) lds[j * 256 + threadIdx.x] = vals[j];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int j = 0; j < 5; ++j) {
      double t = 0;
      for (int i = 0; i < 256; ++i) t += lds[j * 256 + i];
      part[(int64_t)blockIdx.x * 5 + j] = t;
    }
}

// The NaN scans of LBMkernel (LBM_Simulator.cpp:33-41, 52-60, 78-86, 102-111,
// 128-137), per member of a packed case: counts the fluid and F2C cells holding a
// non-finite population in either fcol buffer.  One brick per wave: the wave's
// ballot is added to its member's counter by one lane.
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_member_scan(int64_t ns, const uint16_t* __restrict__ kind,
                                                     const int32_t* __restrict__ bmember, const S* __restrict__ fa,
                                                     const S* __restrict__ fb, unsigned long long* counts) {
  typedef typename RealOf<S>::type R;
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  bool bad = false;
  if (s < ns && (kind[s] & (MT_KIND_FLUID | MT_KIND_F2C))) {
    for (int q = 0; q < Q; ++q) {
      const R a = load_pop<Q, S, R>(fa, ns, q, (int)s);
      const R b = load_pop<Q, S, R>(fb, ns, q, (int)s);
      bad |= !isfinite(a) || !isfinite(b);
    }
  }
  const unsigned long long m = __ballot(bad);
  if ((threadIdx.x & 63) == 0 && s < ns && m) atomicAdd(&counts[bmember[s >> 6]], (unsigned long long)__popcll(m));
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

// infoPrint's velocity diagnostics (LBM_Manager.cpp:914-962) and, on the finest
// level, diagL2MaxF / diagL2RhoMin (:854-912), per block of 256 slots (grid
// stride).  Partials, 16 per block:
//   0 sum |v - v_old|^2 over lat_f, 1 sum |v|^2 over lat_f, 2 max |v| over
//   lat_f and F2C, 3 cells with |v| > 1; finest level only: 4 max |f| (5 its
//   slot, 6 its q), 7 max |fcol| (8, 9), 10 min sum f (11 its slot), 12 min
//   sum fcol, 13 cells with sum f <= 0, 14 with sum fcol <= 0, 15 with sum f < 0.5.
// Ties keep the lowest slot and q, so the result is independent of the order.
constexpr int kDiagVals = 16;
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_diag(LvArgs v, int finest, const void* u_prev, double* part) {
  typedef typename RealOf<S>::type R;
  __shared__ double lds[kDiagVals][256];
  double d[kDiagVals] = {0, 0, 0, 0, 0, -1, 0, 0, -1, 0, 1e300, -1, 1e300, 0, 0, 0};
  const int64_t ns = v.ns;
  for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < ns; s += (int64_t)gridDim.x * 256) {
    const uint16_t k = v.kind[s];
    if (!(k & (MT_KIND_FLUID | MT_KIND_F2C))) continue;
    const double ux = (double)ld_real<R>(v.u_in, s), uy = (double)ld_real<R>(v.u_in, ns + s),
                 uz = (double)ld_real<R>(v.u_in, 2 * ns + s);
    const double vmag = sqrt(ux * ux + uy * uy + uz * uz);
    if (vmag > d[2]) d[2] = vmag;
    if (vmag > 1.0) d[3] += 1;
    if (!(k & MT_KIND_FLUID)) continue;
    const double ox = (double)ld_real<R>(u_prev, s), oy = (double)ld_real<R>(u_prev, ns + s),
                 oz = (double)ld_real<R>(u_prev, 2 * ns + s);
    d[0] += (ux - ox) * (ux - ox) + (uy - oy) * (uy - oy) + (uz - oz) * (uz - oz);
    d[1] += ux * ux + uy * uy + uz * uz;
    if (!finest) continue;
    double rf = 0, rfc = 0;
    for (int q = 0; q < Q; ++q) {
      const double fv = (double)load_pop<Q, S, R>((const S*)v.f, ns, q, (int)s);
      const double fc = (double)load_pop<Q, S, R>((const S*)v.fcol_in, ns, q, (int)s);
      rf += fv;
      rfc += fc;
      if (fabs(fv) > d[4]) { d[4] = fabs(fv); d[5] = (double)s; d[6] = q; }
      if (fabs(fc) > d[7]) { d[7] = fabs(fc); d[8] = (double)s; d[9] = q; }
    }
    if (rf < d[10]) { d[10] = rf; d[11] = (double)s; }
    if (rfc < d[12]) d[12] = rfc;
    if (rf <= 0.0) d[13] += 1;
    if (rfc <= 0.0) d[14] += 1;
    if (rf < 0.5) d[15] += 1;
  }
  for (int j = 0; j < kDiagVals; ++j) lds[j][threadIdx.x] = d[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    // thread order: sums in a fixed order, extrema with the first (lowest
    // slot) holder, each thread's slots ascending
    for (int t = 1; t < 256; ++t) {
      d[0] += lds[0][t];
      d[1] += lds[1][t];
      d[2] = fmax(d[2], lds[2][t]);
      d[3] += lds[3][t];
      auto better_max = [&](int vi, int si) {
        const double x = lds[vi][t], xs = lds[si][t];
        return x > d[vi] || (x == d[vi] && xs >= 0 && (d[si] < 0 || xs < d[si]));
      };
      if (better_max(4, 5)) { d[4] = lds[4][t]; d[5] = lds[5][t]; d[6] = lds[6][t]; }
      if (better_max(7, 8)) { d[7] = lds[7][t]; d[8] = lds[8][t]; d[9] = lds[9][t]; }
      const double x = lds[10][t], xs = lds[11][t];
      if (x < d[10] || (x == d[10] && xs >= 0 && (d[11] < 0 || xs < d[11]))) { d[10] = x; d[11] = xs; }
      d[12] = fmin(d[12], lds[12][t]);
      d[13] += lds[13][t];
      d[14] += lds[14][t];
      d[15] += lds[15][t];
    }
    for (int j = 0; j < kDiagVals; ++j) part[(int64_t)blockIdx.x * kDiagVals + j] = d[j];
  }
}

// ================================================================== launchers
static inline int64_t nblk(int64_t n, int b = 256) { return n <= 0 ? 1 : (n + b - 1) / b; }

static LvArgs lvargs(Case& c, int L, int fbuf_in, int fbuf_out, int rbuf_in, int rbuf_out) {
  DevLevel& l = c.lv[L];
  LvArgs v;
  v.ns = l.n_slots;
  v.L = L;
  v.mp = c.d_mem;
  v.bmember = l.bmember;
  v.bnbr = l.bnbr;
  v.kind = l.kind;
  v.pmask = l.pmask;
  v.f = l.f;
  v.fcol_in = l.fcol[fbuf_in];
  v.fcol_out = l.fcol[fbuf_out];
  v.rho_in = l.rho[rbuf_in];
  v.u_in = l.u[rbuf_in];
  v.rho_out = l.rho[rbuf_out];
  v.u_out = l.u[rbuf_out];
  return v;
}

// dispatch on (Q, storage)
#define MT_DISPATCH(c, KERNEL_CALL)                                                         \
  do {                                                                                      \
    if ((c).Q == 19) {                                                                      \
      if ((c).prec == MT_FP64) { constexpr int Q_ = 19; typedef double S_; KERNEL_CALL; }   \
      else if ((c).prec == MT_FP32_SHIFTED) { constexpr int Q_ = 19; typedef float S_; KERNEL_CALL; } \
      else { constexpr int Q_ = 19; typedef _Float16 S_; KERNEL_CALL; }                     \
    } else {                                                                                \
      if ((c).prec == MT_FP64) { constexpr int Q_ = 27; typedef double S_; KERNEL_CALL; }   \
      else if ((c).prec == MT_FP32_SHIFTED) { constexpr int Q_ = 27; typedef float S_; KERNEL_CALL; } \
      else { constexpr int Q_ = 27; typedef _Float16 S_; KERNEL_CALL; }                     \
    }                                                                                       \
  } while (0)

// ------------------------------------------------------------------ remap
// The state of one case carried onto another case's cells (mt_case_remap): the
// same key holding populations in the source is copied; else the nearest
// coarser source ancestor is exploded, f = feq_HL + s (f_a - feq_HL) with s the
// product of the transfer factors 2 tau_l / tau_{l-1} down to this level (the
// explosion of LBM_Manager.cpp:1059-1129, composed); else the source children
// one level finer are averaged and coalesced, f = feq_HL + tau_L / (2 tau_{L+1})
// (mean - feq_HL) (LBM_Manager.cpp:1131-1169); else the cell is new to the
// flow (uncovered by a moving solid) and starts at feq(rho0, u_fill).  The
// same rule gives fcol (from the source's fcol) and f; density and velocity are
// the moments of the new f.  Solid and reservoir cells of the destination keep
// their own initial state.
struct RemapSrc {
  LevelView v;
  const void* f;
  const void* fcol;
  const void* rho;
  const void* u;
  int64_t ns;
};
struct RemapArgs {
  RemapSrc src[kMaxLevels];
  int L, Lmax;
  const uint64_t* key;
  const uint16_t* kind;
  int64_t ns;
  void* f;
  void* fcol[2];
  void* rho[2];
  void* u[2];
  double xscale[kMaxLevels], cscale[kMaxLevels], rho0, fill[3];
  int* count;  // [4] cells copied, exploded, coalesced, filled
};

template <int Q, typename R>
__device__ __forceinline__ void remap_transfer(R* f, R scale, R rho0) {
  R rho, ux, uy, uz;
  cell_macros<Q, R>(f, rho, ux, uy, uz);
  const R uu = ux * ux + uy * uy + uz * uz;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const R fe = feq_hl<Q, R>(q, rho, ux, uy, uz, uu, rho0);
    f[q] = fe + scale * (f[q] - fe);
  }
}

template <int Q, typename S>
__global__ __launch_bounds__(256) void k_remap(RemapArgs a) {
  typedef typename RealOf<S>::type R;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.ns) return;
  const uint16_t k = a.kind[s];
  if (!k || (k & (MT_KIND_SOLID | MT_KIND_BCRES))) return;
  const uint64_t key = a.key[s];
  R f[Q], g[Q];
  int how = 3;
  const RemapSrc& same = a.src[a.L];
  int32_t sl = find_slot(same.v, key);
  if (sl >= 0 && !(same.v.kind[sl] & MT_KIND_SOLID)) {
    how = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      f[q] = load_pop<Q, S, R>((const S*)same.f, same.ns, q, sl);
      g[q] = load_pop<Q, S, R>((const S*)same.fcol, same.ns, q, sl);
    }
  }
  double scale = 1.0;
  for (int A = a.L - 1; how == 3 && A >= 0; --A) {
    scale *= a.xscale[A + 1];
    const RemapSrc& up = a.src[A];
    const uint64_t pk = key & ~((1ull << (3 * (a.Lmax - A))) - 1);
    sl = find_slot(up.v, pk);
    if (sl >= 0 && !(up.v.kind[sl] & MT_KIND_SOLID)) {
      how = 1;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        f[q] = load_pop<Q, S, R>((const S*)up.f, up.ns, q, sl);
        g[q] = load_pop<Q, S, R>((const S*)up.fcol, up.ns, q, sl);
      }
      remap_transfer<Q, R>(f, (R)scale, (R)a.rho0);
      remap_transfer<Q, R>(g, (R)scale, (R)a.rho0);
    }
  }
  if (how == 3 && a.L < a.Lmax) {
    const RemapSrc& dn = a.src[a.L + 1];
    const int sh = 3 * (a.Lmax - a.L - 1);
    int n = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = g[q] = (R)0;
    for (int c = 0; c < 8; ++c) {
      sl = find_slot(dn.v, key | ((uint64_t)c << sh));
      if (sl < 0 || (dn.v.kind[sl] & MT_KIND_SOLID)) continue;
      ++n;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        f[q] += load_pop<Q, S, R>((const S*)dn.f, dn.ns, q, sl);
        g[q] += load_pop<Q, S, R>((const S*)dn.fcol, dn.ns, q, sl);
      }
    }
    if (n) {
      how = 2;
      const R inv = (R)1 / (R)n;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        f[q] *= inv;
        g[q] *= inv;
      }
      remap_transfer<Q, R>(f, (R)a.cscale[a.L], (R)a.rho0);
      remap_transfer<Q, R>(g, (R)a.cscale[a.L], (R)a.rho0);
    }
  }
  if (how == 3) {
    const R ux = (R)a.fill[0], uy = (R)a.fill[1], uz = (R)a.fill[2], uu = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = g[q] = feq_c<Q, R>(q, (R)a.rho0, ux, uy, uz, uu);
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    store_pop<Q, S, R>((S*)a.f, a.ns, q, (int)s, f[q]);
    store_pop<Q, S, R>((S*)a.fcol[0], a.ns, q, (int)s, g[q]);
    store_pop<Q, S, R>((S*)a.fcol[1], a.ns, q, (int)s, g[q]);
    f[q] = load_pop<Q, S, R>((const S*)a.f, a.ns, q, (int)s);  // as stored
  }
  R rho, ux, uy, uz;
  cell_macros<Q, R>(f, rho, ux, uy, uz);
  for (int j = 0; j < 2; ++j) {
    st_real<R>(a.rho[j], s, rho);
    st_real<R>(a.u[j], s, ux);
    st_real<R>(a.u[j], a.ns + s, uy);
    st_real<R>(a.u[j], 2 * a.ns + s, uz);
  }
  atomicAdd(&a.count[how], 1);
}

int launch_remap(Case& dst, const Case& src, const double* fill, int64_t* counts) {
  int* d_cnt = nullptr;
  MT_HIP(dev_malloc((void**)&d_cnt, 4 * sizeof(int), dst.stream));
  const int zero[4] = {0, 0, 0, 0};
  MT_HIP(hipMemcpyAsync(d_cnt, zero, sizeof zero, hipMemcpyHostToDevice, dst.stream));
  RemapArgs a = {};
  for (int L = 0; L <= src.Lmax; ++L) {
    const DevLevel& l = src.lv[L];
    a.src[L].v = LevelView{l.bcode, l.kind, l.n_bricks, 3 * (src.Lmax - L)};
    a.src[L].f = l.f;
    a.src[L].fcol = l.fcol[l.pf];
    a.src[L].rho = l.rho[l.pr];
    a.src[L].u = l.u[l.pr];
    a.src[L].ns = l.n_slots;
  }
  const MemberParams& P = dst.h_mem[0];
  for (int L = 0; L <= dst.Lmax; ++L) {
    a.xscale[L] = P.xscale[L];
    a.cscale[L] = P.cscale[L];
  }
  a.rho0 = P.rho0;
  for (int j = 0; j < 3; ++j) a.fill[j] = fill[j];
  a.Lmax = dst.Lmax;
  a.count = d_cnt;
  for (int L = 0; L <= dst.Lmax; ++L) {
    DevLevel& l = dst.lv[L];
    a.L = L;
    a.key = l.key;
    a.kind = l.kind;
    a.ns = l.n_slots;
    a.f = l.f;
    a.fcol[0] = l.fcol[0];
    a.fcol[1] = l.fcol[1];
    a.rho[0] = l.rho[0];
    a.rho[1] = l.rho[1];
    a.u[0] = l.u[0];
    a.u[1] = l.u[1];
    MT_DISPATCH(dst, (k_remap<Q_, S_><<<nblk(l.n_slots), 256, 0, dst.stream>>>(a)));
  }
  MT_HIP(hipGetLastError());
  int h[4];
  MT_HIP(hipMemcpyAsync(h, d_cnt, sizeof h, hipMemcpyDeviceToHost, dst.stream));
  MT_HIP(hipStreamSynchronize(dst.stream));
  dev_free(d_cnt, dst.stream);
  for (int j = 0; j < 4; ++j) counts[j] = h[j];
  return MT_OK;
}

int init_state(Case& c) {
  for (int L = 0; L <= c.Lmax; ++L) {
    DevLevel& l = c.lv[L];
    LvArgs v = lvargs(c, L, 0, 0, 0, 0);
    MT_DISPATCH(c, (k_init<Q_, S_><<<nblk(l.n_slots), 256, 0, c.stream>>>(v, l.fcol[1], l.rho[1], l.u[1])));
    l.pf = 0;
    l.pr = 0;
  }
  MT_HIP(hipGetLastError());
  return MT_OK;
}

static SeamArgs seam(Case& c, int L) {
  DevLevel& l = c.lv[L];
  SeamArgs a;
  a.v = lvargs(c, L, l.pf, l.pf, l.pr, l.pr);
  a.model = c.d.collision;
  return a;
}

int launch_macros(Case& c, int L) {
  SeamArgs a = seam(c, L);
  MT_DISPATCH(c, (k_macros<Q_, S_><<<nblk(a.v.ns), 256, 0, c.st[L]>>>(a)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}
int launch_collide(Case& c, int L) {
  SeamArgs a = seam(c, L);
  MT_DISPATCH(c, (k_collide<Q_, S_><<<nblk(a.v.ns), 256, 0, c.st[L]>>>(a)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}
int launch_stream(Case& c, int L) {
  SeamArgs a = seam(c, L);
  MT_DISPATCH(c, (k_stream<Q_, S_><<<nblk(a.v.ns), 256, 0, c.stream>>>(a)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}
static int launch_nee(Case& c) {
  DevLevel& l = c.lv[0];
  if (!l.n_bcres) return MT_OK;
  SeamArgs a = seam(c, 0);
  MT_DISPATCH(c, (k_nee<Q_, S_><<<nblk(l.n_bcres), 256, 0, c.stream>>>(a, l.bcres, l.n_bcres, c.nee_tmp)));
  MT_DISPATCH(c, (k_nee_scatter<Q_, S_><<<nblk(l.n_bcres), 256, 0, c.stream>>>(a, l.bcres, l.n_bcres, c.nee_tmp)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}
int launch_bc(Case& c) {
  if (c.res_any) {
    int r = launch_nee(c);
    if (r) return r;
  }
  if (c.wet_any) {
    SeamArgs a = seam(c, 0);
    MT_DISPATCH(c, (k_bc_wet<Q_, S_><<<nblk(a.v.ns), 256, 0, c.stream>>>(a)));
  }
  MT_HIP(hipGetLastError());
  return MT_OK;
}
int launch_ib(Case& c, bool acc) {
  int L = c.Lmax;
  DevLevel& l = c.lv[L];
  SeamArgs a = seam(c, L);
  int64_t nb = nblk(a.v.ns);
  if (nb > c.force_part_cap) {
    set_error("force partial buffer too small");
    return MT_ERR_STATE;
  }
  MT_DISPATCH(c, (k_ib<Q_, S_><<<nb, 256, 0, c.stream>>>(a, l.dis_row, l.dis, c.d.ib, c.force_part)));
  // UniformBoundary reports no force (it never calls accumIBForce)
  if (c.d.ib != MT_IB_UNIFORM) k_force_reduce<<<1, 256, 0, c.stream>>>(c.force_part, nb, c.force_acc);
  MT_HIP(hipGetLastError());
  return MT_OK;
}

int launch_explode(Case& c, int L, bool mirror_all) {
  if (L <= 0) return MT_OK;
  DevLevel& lc = c.lv[L - 1];
  DevLevel& lf = c.lv[L];
  if (!lc.n_xp) return MT_OK;
  XferArgs a;
  a.ns_c = lc.n_slots;
  a.ns_f = lf.n_slots;
  a.src = lc.fcol[lc.pf];
  a.dst = lf.fcol[lf.pf];
  a.dst2 = lf.fcol[1 - lf.pf];
  a.group = 0;
  a.mp = c.d_mem;
  a.bmember_c = lc.bmember;
  a.Lc = L - 1;
  a.collide = 0;
  a.list = lc.xp_list;
  a.n = lc.n_xp;
  a.child = lc.xp_child;
  // mirror_all: every target also goes to the second buffer, which then holds
  // what the next sub-step's explosion would write
  MT_DISPATCH(c, (k_explode<Q_, S_><<<(a.n + 7) / 8, 256, 0, c.st[L]>>>(a, mirror_all ? 0 : lc.n_f2c)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}

// All targets of level L's explosion, once, into its target groups xg[dst_buf]
// (explosions inside the fused updates: priming the first pair).
int launch_explode_to(Case& c, int L, int dst_buf) {
  if (L <= 0) return MT_OK;
  DevLevel& lc = c.lv[L - 1];
  DevLevel& lf = c.lv[L];
  if (!lc.n_xp) return MT_OK;
  XferArgs a;
  a.ns_c = lc.n_slots;
  a.ns_f = lf.n_slots;
  a.src = lc.fcol[lc.pf];
  a.dst = lf.xg[dst_buf];
  a.dst2 = nullptr;
  a.group = 1;
  a.mp = c.d_mem;
  a.bmember_c = lc.bmember;
  a.Lc = L - 1;
  a.collide = 0;
  a.list = lc.xp_list;
  a.n = lc.n_xp;
  a.child = lc.xp_child;
  MT_DISPATCH(c, (k_explode<Q_, S_><<<(a.n + 7) / 8, 256, 0, c.stream>>>(a, a.n)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}

template <int Q, typename S>
__global__ void k_copy_targets(const int32_t* list, int64_t n, int64_t ns, const S* from, S* to) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = list[i];
#pragma unroll
  for (int q = 0; q < Q; ++q) to[q * ns + s] = from[q * (ns >> 3) + (s >> 3)];
}

// Write level L's explosion targets of the target groups xg[from_buf] into the
// target slots of fcol[to_buf]: every target, or with the reference's
// explosion (sub-step A only) the ghost_C2F targets, since there the C2F
// cells' fcol is sub-step A's copy.
int launch_copy_targets(Case& c, int L, int from_buf, int to_buf) {
  DevLevel& l = c.lv[L];
  const int32_t* list = l.tmask_g ? l.xtg_list : l.xt_list;
  const int64_t n = l.tmask_g ? l.n_xtg : l.n_xt;
  if (!n || !l.xg[from_buf]) return MT_OK;
  MT_DISPATCH(c, (k_copy_targets<Q_, S_><<<nblk(n), 256, 0, c.stream>>>(
                     list, n, l.n_slots, (const S_*)l.xg[from_buf], (S_*)l.fcol[to_buf])));
  MT_HIP(hipGetLastError());
  return MT_OK;
}

int launch_coalesce(Case& c, int L, bool collide_after, bool write_macros) {
  if (L >= c.Lmax) return MT_OK;
  DevLevel& lc = c.lv[L];
  DevLevel& lf = c.lv[L + 1];
  if (!lc.n_f2c) return MT_OK;
  XferArgs a;
  a.ns_c = lc.n_slots;
  a.ns_f = lf.n_slots;
  a.src = lf.f;
  a.dst = lc.f;
  a.dst2 = nullptr;
  a.list = lc.f2c;
  a.n = lc.n_f2c;
  a.child = lc.f2c_child;
  a.mp = c.d_mem;
  a.bmember_c = lc.bmember;
  a.Lc = L;
  a.collide = collide_after ? 1 : 0;
  a.write_macros = (write_macros || c.diag) ? 1 : 0;
  a.fcol_out = lc.fcol[1 - lc.pf];
  a.rho_out = lc.rho[1 - lc.pr];
  a.u_out = lc.u[1 - lc.pr];
  a.model = c.d.collision;
  MT_DISPATCH(c, (k_coalesce<Q_, S_><<<(a.n + 15) / 16, 64, 0, c.st[L]>>>(a)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}

bool level_has_ib(const Case& c, int L) {
  return c.d.ib != MT_IB_NONE && L == c.Lmax && c.Lmax > 0 && c.lv[L].n_dis > 0 && c.lv[L].n_ibl > 0;
}

// The arguments of level L's fused update in its current parities, its
// instantiation (COAL | WET << 1 | IB << 2) and its grid.
static int fused_args(Case& c, int L, bool write_f, bool write_macros, bool c2f_fcol, const CoalReq* coal,
                      FusedArgs& a, int& variant, int64_t& nb) {
  DevLevel& l = c.lv[L];
  a.v = lvargs(c, L, l.pf, 1 - l.pf, l.pr, 1 - l.pr);
  a.active = (coal && l.active_b) ? l.active_b : l.active;
  a.slot_of = (coal && l.active_b) ? l.slot_b : nullptr;
  a.n_active = l.n_active;
  // a sub-step whose C2F cells' stream is dropped (see fused_block) runs only
  // the bricks that hold other work
  const bool c2f_live = coal || write_f || c2f_fcol || (L > 0 && !c.lv[L - 1].coal_in_fused);
  if (!c2f_live && L > 0 && l.active_a) {
    a.active = l.active_a;
    a.slot_of = l.slot_a;
    a.n_active = l.n_active_a;
  }
  a.xmap = c.xmap;
  a.model = c.d.collision;
  a.write_f = write_f ? 1 : 0;
  a.do_bc = (L == 0 && c.wet_any) ? 1 : 0;
  a.do_ib = level_has_ib(c, L) ? 1 : 0;
  a.ib_method = c.d.ib;
  a.dis_row = l.dis_row;
  a.dis = l.dis;
  a.wmask = l.wmask;
  a.force_slots = c.force_slots;
  a.force_acc = c.force_acc;
  if (a.do_ib && l.n_active > c.force_part_cap) {
    set_error("force slot buffer too small");
    return MT_ERR_STATE;
  }
  // with diagnostics on, every update leaves its macros, so the previous
  // update's stay in the other buffer (v_old of calculateMacros)
  a.write_macros = (write_macros || c.diag) ? 1 : 0;
  a.c2f_fcol = c2f_fcol ? 1 : 0;
  a.c2f_f = (L > 0 && !c.lv[L - 1].coal_in_fused) ? 1 : 0;
  a.cpar = nullptr;
  a.ns_c = 0;
  a.cf = a.cfcol_out = a.crho_out = a.cu_out = nullptr;
  a.c_write_macros = 0;
  a.tmask = nullptr;
  a.xbuf = nullptr;
  a.gxrow = nullptr;
  a.gx_child = nullptr;
  a.gx_dst = nullptr;
  a.ns_f = 0;
  a.cx_dst = nullptr;
  a.band_row = nullptr;
  a.band_mom = nullptr;
  a.n_band = 0;
  if (L == c.Lmax && c.n_mk && c.band_row) {
    a.band_row = c.band_row;
    a.band_mom = c.band_mom;
    a.n_band = c.band_cap;
  }
  if (c.xin) {
    if (L >= 1) {
      // the reference explodes before sub-step A only: sub-step B (the one that
      // coalesces) reads the C2F cells' fcol, which sub-step A wrote (C2F
      // copied), and only the ghost_C2F targets from the target groups
      a.tmask = (coal && l.tmask_g) ? l.tmask_g : l.tmask;
      a.xbuf = l.xg[c.lv[L - 1].pf];
    }
    if (L < c.Lmax && l.gxrow) {
      DevLevel& lf = c.lv[L + 1];
      a.gxrow = l.gxrow;
      a.gx_child = l.gf2c_child;
      a.gx_dst = lf.xg[1 - l.pf];
      a.ns_f = lf.n_slots;
    }
    if (coal) {
      a.cx_dst = l.xg[1 - c.lv[L - 1].pf];
    }
  }
  if (coal) {
    if (L == 0 || !c.lv[L - 1].coal_in_fused || !l.cpar) {
      set_error("coalescence in the fused update requested where it does not apply");
      return MT_ERR_STATE;
    }
    DevLevel& lc = c.lv[L - 1];
    a.cpar = l.cpar;
    a.ns_c = lc.n_slots;
    a.cf = lc.f;
    a.cfcol_out = lc.fcol[1 - lc.pf];
    a.crho_out = lc.rho[1 - lc.pr];
    a.cu_out = lc.u[1 - lc.pr];
    a.c_write_macros = (coal->write_macros || c.diag) ? 1 : 0;
  }
  // grid: a multiple of 8 blocks so the XCD remap covers every brick index
  nb = ((a.n_active + 7) / 8) * 8;
  if (nb < 8) nb = 8;
  // the domain boundary (level 0 only), the coalescence (finer levels only) and
  // the solid boundary (finest level only) are compiled into separate
  // instantiations, so no level's kernel carries the registers of another's
  if (coal && a.do_bc) {
    set_error("coalescence and domain boundary in one fused update");
    return MT_ERR_STATE;
  }
  variant = (coal ? 1 : 0) | (a.do_bc ? 2 : 0) | (a.do_ib ? 4 : 0);
  return MT_OK;
}

// The collision kind the fused kernels are launched for: the case's model and
// whether any member has a Smagorinsky constant.
static int collide_kind(const Case& c) {
  bool les = false;
  for (const MemberParams& p : c.h_mem) les |= p.les_C > 0.0;
  return (c.d.collision == MT_COLLIDE_CUMULANT ? CK_CUMULANT : CK_LBGK) | (les ? 1 : 0);
}

// a start/stop event pair for timing mode (null when not timing)
static int timing_events(Case& c, hipEvent_t& e0, hipEvent_t& e1) {
  e0 = e1 = nullptr;
  if (!c.timing) return MT_OK;
  if (c.ev_pool.size() < 2) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      MT_HIP(hipEventCreate(&e));
      c.ev_pool.push_back(e);
    }
  }
  e0 = c.ev_pool.back();
  c.ev_pool.pop_back();
  e1 = c.ev_pool.back();
  c.ev_pool.pop_back();
  return MT_OK;
}

int launch_fused(Case& c, int L, bool write_f, bool nee_after, bool write_macros, bool c2f_fcol,
                 const CoalReq* coal) {
  FusedArgs a;
  int variant = 0;
  int64_t nb = 0;
  TRY_K(fused_args(c, L, write_f, write_macros, c2f_fcol, coal, a, variant, nb));
  hipEvent_t e0, e1;
  TRY_K(timing_events(c, e0, e1));
  int r = MT_OK;
  const int ck = collide_kind(c);
  MT_DISPATCH(c, (r = launch_fused_qs<Q_, S_>(ck, variant, a, nb, c.st[L], e0, e1)));
  if (r) return r;
  MT_HIP(hipGetLastError());
  if (c.timing) c.ev_pending.push_back({L, {e0, e1}});
  if (nee_after) {
    // reservoir NEE reads the pre-sub-step maps (rho/u[pr]) and the f just written
    int r = launch_nee(c);
    if (r) return r;
  }
  return MT_OK;
}

int launch_fused_levels(Case& c, int n, const LevelUpdate* u, bool nee_after) {
  if (n < 1 || n > kMaxLevels) {
    set_error("fused update of an invalid number of levels");
    return MT_ERR_STATE;
  }
  LevelsArgs m;
  m.n = n;
  int64_t total = 0;
  for (int i = 0; i < n; ++i) {
    int64_t nb = 0;
    TRY_K(fused_args(c, u[i].L, u[i].write_f, u[i].write_macros, u[i].c2f_fcol, u[i].coal, m.seg[i], m.variant[i],
                     nb));
    if (m.variant[i] == 6) {
      set_error("a level with both the domain and the solid boundary in a launch of several levels");
      return MT_ERR_STATE;
    }
    total += nb;
    if (total >= (int64_t)1 << 31) {
      set_error("fused update grid too large");
      return MT_ERR_STATE;
    }
    m.end[i] = (uint32_t)total;
  }
  for (int i = n; i < kMaxLevels; ++i) {
    m.variant[i] = 0;
    m.end[i] = (uint32_t)total;
  }
  hipEvent_t e0, e1;
  TRY_K(timing_events(c, e0, e1));
  // every level's stream is c.stream here; the launch is timed under the finest
  // level taking part
  int Lf = 0;
  for (int i = 0; i < n; ++i) Lf = std::max(Lf, u[i].L);
  int vm = 0;
  for (int i = 0; i < n; ++i) vm |= (m.variant[i] & 2) ? VM_WET : 0, vm |= (m.variant[i] & 4) ? VM_IB : 0;
  if ((vm & VM_WET) && c.nee_code) vm |= VM_NEE;
  int r = MT_OK;
  const int ck = collide_kind(c);
  MT_DISPATCH(c, (r = launch_fused_levels_qs<Q_, S_>(ck, vm, m, total, c.st[Lf], e0, e1)));
  if (r) return r;
  MT_HIP(hipGetLastError());
  if (c.timing) c.ev_pending.push_back({Lf, {e0, e1}});
  if (nee_after) {
    int r2 = launch_nee(c);
    if (r2) return r2;
  }
  return MT_OK;
}

// ------------------------------------------------------------------ markers
// Explicit direct forcing of the finest level by Lagrangian markers (the
// definition in mt_amr.h), three kernels per finest sub-step:
//   k_mk_moments  one thread per band cell (a cell in some marker's support):
//                 its density and velocity from fcol (after a fused update of
//                 the finest level, that update has left them already);
//   k_mk_interp   one 32-lane half-wave per marker, lane j < 27 its j-th kernel
//                 cell: the weighted moments summed by a shuffle tree, then
//                 F_k = dV_k rho_k (ub - u_k);
//   k_mk_spread   one thread per band cell: the momentum of its markers
//                 (ascending marker order) turned into the exact-difference
//                 population shift; an extra block 0 sums the body force
//                 sum_k F_k in a fixed order beside them (mk_force_sum; F is
//                 final when the kernel starts).
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_mk_moments(const S* fcol, int64_t ns, const int32_t* band,
                                                    const int32_t* n_band, int64_t nb, double* mom) {
  typedef typename RealOf<S>::type R;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *n_band) return;
  const int sl = band[i];
  R f[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q] = load_pop<Q, S, R>(fcol, ns, q, sl);
  R rho, ux, uy, uz;
  cell_macros<Q, R>(f, rho, ux, uy, uz);
  mom[i] = (double)rho;
  mom[nb + i] = (double)ux;
  mom[2 * nb + i] = (double)uy;
  mom[3 * nb + i] = (double)uz;
}

__global__ __launch_bounds__(256) void k_mk_interp(const double* mom, int64_t nb, const int32_t* mk_band,
                                                   const double* w, const double* dv, int64_t n, const double* ub,
                                                   double* F) {
  const int64_t m = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
  const int j = threadIdx.x & 31;
  double r = 0, a = 0, b = 0, c = 0;
  if (m < n && j < 27) {
    // unused entries carry weight 0 and a valid band index
    const double wj = w[m * 27 + j];
    const int32_t i = mk_band[m * 27 + j];
    r = wj * mom[i];
    a = wj * mom[nb + i];
    b = wj * mom[2 * nb + i];
    c = wj * mom[3 * nb + i];
  }
  for (int off = 16; off > 0; off >>= 1) {
    r += __shfl_down(r, off, 32);
    a += __shfl_down(a, off, 32);
    b += __shfl_down(b, off, 32);
    c += __shfl_down(c, off, 32);
  }
  if (m < n && j == 0) {
    F[3 * m] = dv[m] * r * (ub[3 * m] - a);
    F[3 * m + 1] = dv[m] * r * (ub[3 * m + 1] - b);
    F[3 * m + 2] = dv[m] * r * (ub[3 * m + 2] - c);
  }
}

// the body force of this call, sum_k F_k (the frame of the link-wise force,
// whose velocities are the reported ones), into acc[0..2] (count in acc[3]) by
// one block of 256 threads: 1024 strided partial sums, then a fixed pairwise
// tree (the first two levels pair partial sums that one thread holds)
__device__ void mk_force_sum(const double* F, int64_t n, double* acc) {
  __shared__ double red[3][256];
  const int t = threadIdx.x;
  double s[4][3];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    s[p][0] = s[p][1] = s[p][2] = 0;
    for (int64_t m = t + 256 * p; m < n; m += 1024) {
      s[p][0] += F[3 * m];
      s[p][1] += F[3 * m + 1];
      s[p][2] += F[3 * m + 2];
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    s[0][a] += s[2][a];  // partial t + 512
    s[1][a] += s[3][a];  // partial t + 768 into t + 256
    red[a][t] = s[0][a] + s[1][a];
  }
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h)
      for (int a = 0; a < 3; ++a) red[a][t] += red[a][t + h];
    __syncthreads();
  }
  if (t == 0) {
    for (int a = 0; a < 3; ++a) {
      acc[a] += red[a][0];
      acc[4 + a] = red[a][0];
    }
    acc[3] += 1.0;
  }
}

template <int Q, typename S>
__global__ __launch_bounds__(256) void k_mk_spread(S* fcol, int64_t ns, const int32_t* band, const int32_t* ptr,
                                                   const int32_t* mk, const double* bw, const int32_t* n_band,
                                                   int64_t nb, const double* mom, const double* F, int64_t n_mk,
                                                   double* acc) {
  typedef typename RealOf<S>::type R;
  // block 0 sums the body force beside the band's blocks 1, 2, ...
  if (blockIdx.x == 0) {
    mk_force_sum(F, n_mk, acc);
    return;
  }
  const int64_t i = (int64_t)(blockIdx.x - 1) * blockDim.x + threadIdx.x;
  if (i >= *n_band) return;
  // the cell's moments and populations do not depend on the markers' forces:
  // their loads go out first and overlap the walk over the cell's markers
  const int sl = band[i];
  const int32_t e0 = ptr[i], e1 = ptr[i + 1];
  // the moments were computed in R from the same fcol: the round trip through
  // double is exact
  const R rho = (R)mom[i], ux = (R)mom[nb + i], uy = (R)mom[2 * nb + i], uz = (R)mom[3 * nb + i];
  R fq[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) fq[q] = load_pop<Q, S, R>(fcol, ns, q, sl);
  double mx = 0, my = 0, mz = 0;
  for (int32_t e = e0; e < e1; ++e) {
    const double we = bw[e];
    const int32_t k = mk[e];
    mx += we * F[3 * k];
    my += we * F[3 * k + 1];
    mz += we * F[3 * k + 2];
  }
  const R vx = ux + (R)(mx / (double)rho), vy = uy + (R)(my / (double)rho), vz = uz + (R)(mz / (double)rho);
  const R uu = ux * ux + uy * uy + uz * uz, vv = vx * vx + vy * vy + vz * vz;
#pragma unroll
  for (int q = 0; q < Q; ++q)
    store_pop<Q, S, R>(fcol, ns, q, sl,
                       fq[q] + (feq_c<Q, R>(q, rho, vx, vy, vz, vv) - feq_c<Q, R>(q, rho, ux, uy, uz, uu)));
}

int launch_markers(Case& c, int buf, hipStream_t st, bool moments_done) {
  if (!c.n_mk) return MT_OK;
  DevLevel& l = c.lv[c.Lmax];
  // grids cover the band's capacity (its size changes as the markers move,
  // mt_markers_move); the kernels read the current size from the device
  if (!moments_done)
    MT_DISPATCH(c, (k_mk_moments<Q_, S_><<<nblk(c.band_cap), 256, 0, st>>>((const S_*)l.fcol[buf], l.n_slots,
                                                                           c.band_slot, c.d_nband, c.band_cap,
                                                                           c.band_mom)));
  k_mk_interp<<<nblk(32 * c.n_mk), 256, 0, st>>>(c.band_mom, c.band_cap, c.mk_band, c.mk_w, c.mk_dv, c.n_mk, c.mk_v,
                                                 c.mk_F);
  MT_DISPATCH(c, (k_mk_spread<Q_, S_><<<nblk(c.band_cap) + 1, 256, 0, st>>>(
                     (S_*)l.fcol[buf], l.n_slots, c.band_slot, c.band_ptr, c.band_mk, c.band_w, c.d_nband, c.band_cap,
                     c.band_mom, c.mk_F, c.n_mk, c.force_acc)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}

__global__ void k_mk_unforce(double* acc) {
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a) acc[a] -= acc[4 + a];
    acc[3] -= 1.0;
  }
}

int launch_mk_unforce(Case& c) {
  k_mk_unforce<<<1, 64, 0, c.stream>>>(c.force_acc);
  MT_HIP(hipGetLastError());
  return MT_OK;
}

// Device tables of the markers: the slots of their 27 cells (each must be a
// streamed lat_f cell of the finest level) and the band of cells in their
// support with each cell's markers in ascending order (CSR).  The band arrays
// are sized for 27 cells per marker, so that moving markers (mt_markers_move)
// rebuilds them in place.
int build_markers(Case& c) {
  if (!c.n_mk) return MT_OK;
  const int64_t n = c.n_mk, ne = 27 * n;
  DevLevel& l = c.lv[c.Lmax];
  uint64_t* dk = nullptr;
  MT_HIP(dev_malloc((void**)&dk, 8 * ne, c.stream));
  c.band_cap = ne;
  c.mk_slot = c.alloc<int32_t>(ne);
  c.mk_w = c.alloc<double>(ne);
  c.mk_dv = c.alloc<double>(n);
  c.mk_F = c.alloc<double>(3 * n);
  c.mk_v = c.alloc<double>(3 * n);
  c.mk_ent = c.alloc<int32_t>(ne);
  c.d_nband = c.alloc<int32_t>(1);
  c.mk_x = c.alloc<double>(3 * n);
  c.mv_x = c.alloc<double>(3 * n);
  c.mv_v = c.alloc<double>(3 * n);
  c.mv_slot = c.alloc<int32_t>(ne);
  c.mv_w = c.alloc<double>(ne);
  c.mv_flags = c.alloc<int32_t>(4);
  if (!c.mk_slot || !c.mk_w || !c.mk_dv || !c.mk_F || !c.mk_v || !c.mk_ent || !c.d_nband || !c.mk_x || !c.mv_x ||
      !c.mv_v || !c.mv_slot || !c.mv_w || !c.mv_flags) {
    set_error("out of device memory");
    return MT_ERR_HIP;
  }
  MT_HIP(hipMemcpyAsync(dk, c.h_mk_keys.data(), 8 * ne, hipMemcpyHostToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.mk_w, c.h_mk_w.data(), 8 * ne, hipMemcpyHostToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.mk_dv, c.h_mk_dv.data(), 8 * n, hipMemcpyHostToDevice, c.stream));
  std::vector<double> hv(3 * n);
  for (int64_t m = 0; m < n; ++m)
    for (int a = 0; a < 3; ++a) hv[3 * m + a] = c.mk_ub[a];
  MT_HIP(hipMemcpyAsync(c.mk_v, hv.data(), 8 * 3 * n, hipMemcpyHostToDevice, c.stream));
  int r = find_slots(c, c.Lmax, dk, ne, c.mk_slot);
  if (r) return r;
  std::vector<int32_t> hs(ne);
  std::vector<uint16_t> kind(l.n_slots);
  MT_HIP(hipMemcpyAsync(hs.data(), c.mk_slot, 4 * ne, hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipMemcpyAsync(kind.data(), l.kind, 2 * l.n_slots, hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipStreamSynchronize(c.stream));
  dev_free(dk, c.stream);
  struct E { int32_t slot, e; double w; };
  std::vector<E> es;
  for (int64_t m = 0; m < n; ++m)
    for (int j = 0; j < 27; ++j) {
      const double w = c.h_mk_w[m * 27 + j];
      if (w == 0.0) continue;
      const int32_t sl = hs[m * 27 + j];
      if (sl < 0 || !(kind[sl] & MT_KIND_FLUID) || (kind[sl] & (MT_KIND_C2F | MT_KIND_DIS | MT_KIND_SOLID))) {
        set_error("marker " + std::to_string(m) + ": a kernel cell is not a streamed fluid cell of the finest level");
        return MT_ERR_ARG;
      }
      es.push_back({sl, (int32_t)(m * 27 + j), w});
    }
  // (slot, entry) ascending: within a cell, markers in ascending order
  std::sort(es.begin(), es.end(), [](const E& a, const E& b) { return a.slot != b.slot ? a.slot < b.slot : a.e < b.e; });
  std::vector<int32_t> band, ptr, mk, ent(ne, -1), mb(ne, 0);
  std::vector<double> bw;
  for (size_t i = 0; i < es.size(); ++i) {
    if (i == 0 || es[i].slot != es[i - 1].slot) {
      band.push_back(es[i].slot);
      ptr.push_back((int32_t)i);
    }
    mk.push_back(es[i].e / 27);
    bw.push_back(es[i].w);
    ent[es[i].e] = (int32_t)i;
    mb[es[i].e] = (int32_t)band.size() - 1;
  }
  ptr.push_back((int32_t)es.size());
  // unused entries point at the band cell of the marker's first used entry
  for (int64_t m = 0; m < n; ++m) {
    int32_t first = -1;
    for (int j = 0; j < 27 && first < 0; ++j)
      if (c.h_mk_w[m * 27 + j] != 0.0) first = mb[m * 27 + j];
    if (first < 0) {
      set_error("marker " + std::to_string(m) + " has no kernel cell with a nonzero weight");
      return MT_ERR_ARG;
    }
    for (int j = 0; j < 27; ++j)
      if (c.h_mk_w[m * 27 + j] == 0.0) mb[m * 27 + j] = first;
  }
  c.n_band = (int64_t)band.size();
  const int32_t nb32 = (int32_t)c.n_band;
  c.band_slot = c.alloc<int32_t>(ne);
  c.band_ptr = c.alloc<int32_t>(ne + 1);
  c.band_mk = c.alloc<int32_t>(ne);
  c.band_w = c.alloc<double>(ne);
  c.band_mom = c.alloc<double>(4 * ne);
  c.mk_band = c.alloc<int32_t>(ne);
  const int64_t ns_f = c.lv[c.Lmax].n_slots;
  c.band_row = c.alloc<int32_t>(ns_f);
  if (!c.band_slot || !c.band_ptr || !c.band_mk || !c.band_w || !c.band_mom || !c.mk_band || !c.band_row) {
    set_error("out of device memory");
    return MT_ERR_HIP;
  }
  MT_HIP(hipMemcpyAsync(c.band_slot, band.data(), 4 * band.size(), hipMemcpyHostToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.band_ptr, ptr.data(), 4 * ptr.size(), hipMemcpyHostToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.band_mk, mk.data(), 4 * mk.size(), hipMemcpyHostToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.band_w, bw.data(), 8 * bw.size(), hipMemcpyHostToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.mk_band, mb.data(), 4 * ne, hipMemcpyHostToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.mk_ent, ent.data(), 4 * ne, hipMemcpyHostToDevice, c.stream));
  MT_HIP(hipMemcpyAsync(c.d_nband, &nb32, 4, hipMemcpyHostToDevice, c.stream));
  std::vector<int32_t> row(ns_f, -1);
  for (int64_t i = 0; i < c.n_band; ++i) row[band[i]] = (int32_t)i;
  MT_HIP(hipMemcpyAsync(c.band_row, row.data(), 4 * ns_f, hipMemcpyHostToDevice, c.stream));
  MT_HIP(hipStreamSynchronize(c.stream));
  return MT_OK;
}

int launch_gather(Case& c, int L, int field, const int32_t* d_slots, int64_t n, double* d_out, int buf) {
  DevLevel& l = c.lv[L];
  LvArgs v = lvargs(c, L, 0, 0, 0, 0);
  const void* src = field == MT_FIELD_F ? l.f
                    : field == MT_FIELD_FCOL ? l.fcol[buf]
                    : field == MT_FIELD_RHO ? l.rho[buf] : l.u[buf];
  MT_DISPATCH(c, (k_gather<Q_, S_><<<nblk(n), 256, 0, c.stream>>>(v, field, d_slots, n, d_out, src)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}
int launch_scatter(Case& c, int L, int field, const int32_t* d_slots, int64_t n, const double* d_in, int buf) {
  DevLevel& l = c.lv[L];
  LvArgs v = lvargs(c, L, 0, 0, 0, 0);
  void* dst = field == MT_FIELD_F ? l.f
              : field == MT_FIELD_FCOL ? l.fcol[buf]
              : field == MT_FIELD_RHO ? l.rho[buf] : l.u[buf];
  MT_DISPATCH(c, (k_scatter<Q_, S_><<<nblk(n), 256, 0, c.stream>>>(v, field, d_slots, n, d_in, dst)));
  MT_HIP(hipGetLastError());
  return MT_OK;
}
int launch_member_scan(Case& c, int64_t* h_counts) {
  const int64_t nm = (int64_t)c.h_mem.size();
  k_fill_u64<<<nblk(nm), 256, 0, c.stream>>>(c.member_bad, nm);
  MT_HIP(hipGetLastError());
  for (int L = 0; L <= c.Lmax; ++L) {
    DevLevel& l = c.lv[L];
    MT_DISPATCH(c, (k_member_scan<Q_, S_><<<nblk(l.n_slots), 256, 0, c.stream>>>(
                       l.n_slots, l.kind, l.bmember, (const S_*)l.fcol[0], (const S_*)l.fcol[1], c.member_bad)));
    MT_HIP(hipGetLastError());
  }
  std::vector<unsigned long long> h(nm);
  MT_HIP(hipMemcpyAsync(h.data(), c.member_bad, sizeof(unsigned long long) * nm, hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipStreamSynchronize(c.stream));
  for (int64_t i = 0; i < nm; ++i) h_counts[i] = (int64_t)h[i];
  return MT_OK;
}

int launch_diag(Case& c, int L, double* h_out) {
  DevLevel& l = c.lv[L];
  LvArgs v = lvargs(c, L, l.pf, l.pf, l.pr, l.pr);
  const int64_t nb = std::min<int64_t>(nblk(l.n_slots), c.red_cap / kDiagVals);
  MT_DISPATCH(c, (k_diag<Q_, S_><<<nb, 256, 0, c.stream>>>(v, L == c.Lmax ? 1 : 0, l.u[1 - l.pr], c.red_tmp)));
  MT_HIP(hipGetLastError());
  std::vector<double> part(nb * kDiagVals);
  MT_HIP(hipMemcpyAsync(part.data(), c.red_tmp, sizeof(double) * nb * kDiagVals, hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipStreamSynchronize(c.stream));
  double* d = h_out;
  for (int j = 0; j < kDiagVals; ++j) d[j] = part[j];
  for (int64_t b = 1; b < nb; ++b) {
    const double* p = part.data() + b * kDiagVals;
    d[0] += p[0];
    d[1] += p[1];
    d[2] = std::max(d[2], p[2]);
    d[3] += p[3];
    // blocks hold ascending slot ranges per grid-stride pass; ties keep the lowest slot
    if (p[4] > d[4] || (p[4] == d[4] && p[5] >= 0 && (d[5] < 0 || p[5] < d[5]))) { d[4] = p[4]; d[5] = p[5]; d[6] = p[6]; }
    if (p[7] > d[7] || (p[7] == d[7] && p[8] >= 0 && (d[8] < 0 || p[8] < d[8]))) { d[7] = p[7]; d[8] = p[8]; d[9] = p[9]; }
    if (p[10] < d[10] || (p[10] == d[10] && p[11] >= 0 && (d[11] < 0 || p[11] < d[11]))) { d[10] = p[10]; d[11] = p[11]; }
    d[12] = std::min(d[12], p[12]);
    d[13] += p[13];
    d[14] += p[14];
    d[15] += p[15];
  }
  return MT_OK;
}

int launch_reduce(Case& c, int L, double* h_out) {
  DevLevel& l = c.lv[L];
  LvArgs v = lvargs(c, L, 0, 0, 0, 0);
  int64_t nb = std::min<int64_t>(nblk(l.n_slots), 1024);
  if (nb * 5 > c.red_cap) {
    set_error("reduction buffer too small");
    return MT_ERR_STATE;
  }
  MT_DISPATCH(c, (k_reduce<Q_, S_><<<nb, 256, 0, c.stream>>>(v, c.red_tmp)));
  std::vector<double> part(nb * 5);
  MT_HIP(hipMemcpyAsync(part.data(), c.red_tmp, sizeof(double) * nb * 5, hipMemcpyDeviceToHost, c.stream));
  MT_HIP(hipStreamSynchronize(c.stream));
  for (int j = 0; j < 5; ++j) h_out[j] = 0;
  for (int64_t i = 0; i < nb; ++i)
    for (int j = 0; j < 5; ++j) h_out[j] += part[i * 5 + j];
  return MT_OK;
}

}  // namespace mt
