// mt_oracle.cpp -- CPU restatement of the reference's per-time-step AMR LBM update.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP path is checked
// against, and the timed CPU baseline.  Nothing in the product links or calls it.
//
// Storage is deliberately map-faithful: one std::unordered_map<key,double> per
// population per level (the C_MAP_TYPE=1 layout of header/user.h:35-46 and
// header/Grid_Class.h:116-119), and every loop walks the same cell sets in the
// same order the reference's loops do (ascending key order stands in for map
// iteration order).  Each function cites the reference lines it restates.
//
// Where the reference has no compiled code for a feature the workloads need
// (Zou-He / outflow / periodic, LES), the oracle states the definition the
// device path implements; those are marked "definition" below and are pinned by
// closed forms in tests/ rather than by reference output.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <tuple>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

constexpr int EX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1};
constexpr int EY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, 1, -1};
constexpr int EZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1};
// e_inv swaps (2k-1, 2k) (header/Constants.h:520,538)
inline int inv(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }

// lattice weights (LBM_Manager.cpp:141-158)
inline double weight(int Q, int q) {
  int n = std::abs(EX[q]) + std::abs(EY[q]) + std::abs(EZ[q]);
  if (Q == 19) return n == 0 ? 1. / 3. : (n == 1 ? 1. / 18. : 1. / 36.);
  return n == 0 ? 8. / 27. : (n == 1 ? 2. / 27. : (n == 2 ? 1. / 54. : 1. / 216.));
}
// The IB uses a 19-entry D3Q19 table even for Q=27; entries >= 19 are zero
// (ComplexBoundary_Manager.cpp:102-107,184-189).
inline double ib_weight(int q) { return q < 19 ? weight(19, q) : 0.0; }

enum Kind : uint16_t {
  K_FLUID = 0x1, K_F2C = 0x2, K_C2F = 0x4, K_GHOST_C2F = 0x8, K_GHOST_F2C = 0x10,
  K_SOLID = 0x20, K_BCRES = 0x40, K_DIS = 0x80
};
enum FaceKind { BC_NONE = 0, BC_RESERVOIR = 1, BC_ZOUHE = 2, BC_OUTFLOW = 3, BC_PERIODIC = 4, BC_NEE = 5 };

// ---------------------------------------------------------------- Morton keys
// Dilated-integer neighbour arithmetic (header/Morton_Assist.h:229-289,
// Morton_Assist.cpp:20-106): x bits at 3i, y at 3i+1, z at 3i+2, i <= 20; a
// level-L step is 1 << 3(Lmax-L); finer bits are cleared by the masks.
struct Morton {
  int Lmax = 0;
  uint64_t xm[8], ym[8], zm[8], one[8];
  void init(int lmax) {
    Lmax = lmax;
    for (int L = 0; L <= Lmax; ++L) {
      xm[L] = ym[L] = zm[L] = 0;
      for (int i = Lmax - L; i <= 20; ++i) {
        xm[L] |= 1ull << (3 * i);
        ym[L] |= 1ull << (3 * i + 1);
        zm[L] |= 1ull << (3 * i + 2);
      }
      one[L] = 1ull << (3 * (Lmax - L));
    }
  }
  static uint64_t low_clear(uint64_t m, const uint64_t one) { return m & ~(one - 1); }
  uint64_t up(uint64_t key, int L, uint64_t m, uint64_t o1, uint64_t o2) const {
    uint64_t fill = low_clear(~m, one[L]);
    uint64_t t = (key | fill) + one[L];
    return (t & m) | (key & o1) | (key & o2);
  }
  uint64_t down(uint64_t key, int L, uint64_t m, uint64_t o1, uint64_t o2) const {
    uint64_t t = (key & m) - one[L];
    return (t & m) | (key & o1) | (key & o2);
  }
  uint64_t x1(uint64_t k, int L) const { return up(k, L, xm[L], ym[L], zm[L]); }
  uint64_t x0(uint64_t k, int L) const { return down(k, L, xm[L], ym[L], zm[L]); }
  uint64_t y1(uint64_t k, int L) const { return up(k, L, ym[L], xm[L], zm[L]); }
  uint64_t y0(uint64_t k, int L) const { return down(k, L, ym[L], xm[L], zm[L]); }
  uint64_t z1(uint64_t k, int L) const { return up(k, L, zm[L], xm[L], ym[L]); }
  uint64_t z0(uint64_t k, int L) const { return down(k, L, zm[L], xm[L], ym[L]); }
  // find_neighbor: x shift on the key, then y, then z (Morton_Assist.h:261-289)
  uint64_t neighbor(uint64_t k, int L, int dx, int dy, int dz) const {
    uint64_t r = k;
    if (dx == 1) r = x1(k, L); else if (dx == -1) r = x0(k, L);
    if (dy == 1) r = y1(r, L); else if (dy == -1) r = y0(r, L);
    if (dz == 1) r = z1(r, L); else if (dz == -1) r = z0(r, L);
    return r;
  }
};

inline uint64_t compact(uint64_t v) {
  v &= 0x1249249249249249ull;
  v = (v ^ (v >> 2)) & 0x10C30C30C30C30C3ull;
  v = (v ^ (v >> 4)) & 0x100F00F00F00F00Full;
  v = (v ^ (v >> 8)) & 0x1F0000FF0000FFull;
  v = (v ^ (v >> 16)) & 0x1F00000000FFFFull;
  v = (v ^ (v >> 32)) & 0x1FFFFFull;
  return v;
}
inline uint64_t spread(uint64_t v) {
  v &= 0x1FFFFF;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// ---------------------------------------------------------------- local physics
// feq, compressible (LBM_Manager.cpp:137-169)
void feq_c(int Q, double rho, const double* u, double* out) {
  double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  for (int q = 0; q < Q; ++q) {
    double eu = EX[q] * u[0] + EY[q] * u[1] + EZ[q] * u[2];
    out[q] = rho * weight(Q, q) * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * uu);
  }
}
// feq, He-Luo incompressible (LBM_Manager.cpp:171-203)
void feq_hl(int Q, double rho, const double* u, double rho0, double* out) {
  double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  for (int q = 0; q < Q; ++q) {
    double eu = EX[q] * u[0] + EY[q] * u[1] + EZ[q] * u[2];
    out[q] = weight(Q, q) * (rho + rho0 * (3. * eu + 4.5 * eu * eu - 1.5 * uu));
  }
}
// moments with the +-1e-6 density floor (LBM_Manager.cpp:983-1001)
void macros(int Q, const double* f, double& rho, double* u) {
  rho = f[0];
  double vx = 0, vy = 0, vz = 0;
  for (int q = 1; q < Q; ++q) {
    rho += f[q];
    vx += f[q] * EX[q];
    vy += f[q] * EY[q];
    vz += f[q] * EZ[q];
  }
  double s = (rho > -1e-6 && rho < 1e-6) ? ((rho < 0.0) ? -1e-6 : 1e-6) : rho;
  u[0] = vx / s;
  u[1] = vy / s;
  u[2] = vz / s;
}
// Smagorinsky: |Pi_neq|_F and the eddy tau (lbm_kernel/les.cpp:12-44)
double stress_norm(int Q, const double* f, const double* feq) {
  double e[9] = {0};
  for (int q = 0; q < Q; ++q) {
    double d = f[q] - feq[q];
    int c[3] = {EX[q], EY[q], EZ[q]};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) e[3 * a + b] += c[a] * c[b] * d;
  }
  double m = 0;
  for (int i = 0; i < 9; ++i) m += e[i] * e[i];
  return std::sqrt(m);
}
double smag_tau(double tau, double C, double qn) {
  double c2 = C * C;
  double nu = (tau - 0.5) / 3.0;
  double s = (std::sqrt(nu * nu + 18 * c2 * qn) - nu) / (6.0 * c2);
  return 3 * (nu + c2 * s) + 0.5;
}

// ---- D3Q27 cumulant (lbm_kernel/Collision_Cumulant.cpp:22-652) ------------
// Moments are indexed m[a][b][c] with a,b,c in {0,1,2}.
typedef double M3[3][3][3];
inline int qidx(int i, int j, int k) {  // velocity (i,j,k) in {-1,0,1}^3 -> q
  for (int q = 0; q < 27; ++q)
    if (EX[q] == i && EY[q] == j && EZ[q] == k) return q;
  return -1;
}
// literal forward central moments, accumulated over q in order (:320-367)
void central_fwd(const double* f, const double* u, M3 k) {
  std::memset(k, 0, sizeof(M3));
  for (int q = 0; q < 27; ++q) {
    double cx[3] = {1, EX[q] - u[0], (EX[q] - u[0]) * (EX[q] - u[0])};
    double cy[3] = {1, EY[q] - u[1], (EY[q] - u[1]) * (EY[q] - u[1])};
    double cz[3] = {1, EZ[q] - u[2], (EZ[q] - u[2]) * (EZ[q] - u[2])};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) k[a][b][c] += cx[a] * cy[b] * cz[c] * f[q];
  }
}
// backward: three 1-D passes x, y, z (:476-652).  For one axis with central
// moments (m0,m1,m2) about velocity v the populations at -1,0,+1 are
//   f0 = m0(1-v^2) - 2v m1 - m2,  f(+-1) = (m0(v^2 +- v) + m1(2v +- 1) + m2)/2.
inline void inv1(double m0, double m1, double m2, double v, double& fm, double& f0, double& fp) {
  double v2 = v * v;
  f0 = m0 * (1 - v2) - m1 * 2 * v - m2;
  fm = (m0 * (v2 - v) + m1 * (2 * v - 1) + m2) * 0.5;
  fp = (m0 * (v2 + v) + m1 * (2 * v + 1) + m2) * 0.5;
}
void central_bwd(const M3 k, const double* u, double* f) {
  double g[3][3][3], h[3][3][3];  // index 0,1,2 <-> velocity -1,0,+1 on resolved axes
  for (int b = 0; b < 3; ++b)
    for (int c = 0; c < 3; ++c) inv1(k[0][b][c], k[1][b][c], k[2][b][c], u[0], g[0][b][c], g[1][b][c], g[2][b][c]);
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) inv1(g[i][0][c], g[i][1][c], g[i][2][c], u[1], h[i][0][c], h[i][1][c], h[i][2][c]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double fm, f0, fp;
      inv1(h[i][j][0], h[i][j][1], h[i][j][2], u[2], fm, f0, fp);
      f[qidx(i - 1, j - 1, -1)] = fm;
      f[qidx(i - 1, j - 1, 0)] = f0;
      f[qidx(i - 1, j - 1, 1)] = fp;
    }
}
inline double safe_irho(double r) {
  return 1.0 / ((r > -1e-6 && r < 1e-6) ? ((r < 0.0) ? -1e-6 : 1e-6) : r);
}
// The correction brackets shared by the forward (C = k - B) and backward
// (k = C + B) cumulant transforms (Collision_Cumulant.cpp:383-397, 456-470).
struct Brackets {
  double b211, b121, b112, b220, b022, b202, b122, b212, b221, a, b;
};
Brackets brackets(const M3 k, double ir) {
  auto K = [&](int a, int b, int c) { return k[a][b][c]; };
  Brackets r;
  r.b211 = (K(2, 0, 0) * K(0, 1, 1) + 2 * K(1, 1, 0) * K(1, 0, 1)) * ir;
  r.b121 = (K(0, 2, 0) * K(1, 0, 1) + 2 * K(0, 1, 1) * K(1, 1, 0)) * ir;
  r.b112 = (K(0, 0, 2) * K(1, 1, 0) + 2 * K(1, 0, 1) * K(0, 1, 1)) * ir;
  r.b220 = (K(2, 0, 0) * K(0, 2, 0) + 2 * K(1, 1, 0) * K(1, 1, 0)) * ir;
  r.b022 = (K(0, 2, 0) * K(0, 0, 2) + 2 * K(0, 1, 1) * K(0, 1, 1)) * ir;
  r.b202 = (K(0, 0, 2) * K(2, 0, 0) + 2 * K(1, 0, 1) * K(1, 0, 1)) * ir;
  r.b122 = (K(0, 0, 2) * K(1, 2, 0) + K(0, 2, 0) * K(1, 0, 2) + 4 * K(0, 1, 1) * K(1, 1, 1) +
            2 * (K(1, 0, 1) * K(0, 2, 1) + K(1, 1, 0) * K(0, 1, 2))) * ir;
  r.b212 = (K(2, 0, 0) * K(0, 1, 2) + K(0, 0, 2) * K(2, 1, 0) + 4 * K(1, 0, 1) * K(1, 1, 1) +
            2 * (K(1, 1, 0) * K(1, 0, 2) + K(0, 1, 1) * K(2, 0, 1))) * ir;
  r.b221 = (K(0, 2, 0) * K(2, 0, 1) + K(2, 0, 0) * K(0, 2, 1) + 4 * K(1, 1, 0) * K(1, 1, 1) +
            2 * (K(0, 1, 1) * K(2, 1, 0) + K(1, 0, 1) * K(1, 2, 0))) * ir;
  r.a = (4 * K(1, 1, 1) * K(1, 1, 1) + K(2, 0, 0) * K(0, 2, 2) + K(0, 2, 0) * K(2, 0, 2) + K(0, 0, 2) * K(2, 2, 0) +
         4 * (K(0, 1, 1) * K(2, 1, 1) + K(1, 0, 1) * K(1, 2, 1) + K(1, 1, 0) * K(1, 1, 2)) +
         2 * (K(1, 2, 0) * K(1, 0, 2) + K(2, 1, 0) * K(0, 1, 2) + K(2, 0, 1) * K(0, 2, 1))) * ir;
  r.b = (16 * K(1, 1, 0) * K(1, 0, 1) * K(0, 1, 1) +
         4 * (K(1, 0, 1) * K(1, 0, 1) * K(0, 2, 0) + K(0, 1, 1) * K(0, 1, 1) * K(2, 0, 0) + K(1, 1, 0) * K(1, 1, 0) * K(0, 0, 2)) +
         2 * K(2, 0, 0) * K(0, 2, 0) * K(0, 0, 2)) * ir * ir;
  return r;
}
// Second- and third-order members copied between k and C by both transforms.
// k210, k021, k102 and the orders 0 and 1 are NOT among them: the forward
// transform leaves those cumulants zero and the backward transform leaves those
// central moments at their forward values (Collision_Cumulant.cpp:378-381, 451-454).
const int kCopied[10][3] = {{1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
                            {1, 2, 0}, {0, 1, 2}, {2, 0, 1}, {1, 1, 1}};
// central moments -> cumulants (fwd_cum_trans, :374-398); C starts zero
void to_cumulants(const M3 k, double rho, M3 C) {
  std::memset(C, 0, sizeof(M3));
  for (auto& i : kCopied) C[i[0]][i[1]][i[2]] = k[i[0]][i[1]][i[2]];
  Brackets B = brackets(k, safe_irho(rho));
  C[2][1][1] = k[2][1][1] - B.b211;
  C[1][2][1] = k[1][2][1] - B.b121;
  C[1][1][2] = k[1][1][2] - B.b112;
  C[2][2][0] = k[2][2][0] - B.b220;
  C[0][2][2] = k[0][2][2] - B.b022;
  C[2][0][2] = k[2][0][2] - B.b202;
  C[1][2][2] = k[1][2][2] - B.b122;
  C[2][1][2] = k[2][1][2] - B.b212;
  C[2][2][1] = k[2][2][1] - B.b221;
  C[2][2][2] = k[2][2][2] - B.a + B.b;
}
// cumulants -> central moments (bwd_cum_trans, :446-471), in place on k
void from_cumulants(const M3 C, double rho, M3 k) {
  for (auto& i : kCopied) k[i[0]][i[1]][i[2]] = C[i[0]][i[1]][i[2]];
  double ir = safe_irho(rho);
  Brackets B = brackets(k, ir);  // order-4 brackets read orders 2-3 only
  k[2][1][1] = C[2][1][1] + B.b211;
  k[1][2][1] = C[1][2][1] + B.b121;
  k[1][1][2] = C[1][1][2] + B.b112;
  k[2][2][0] = C[2][2][0] + B.b220;
  k[0][2][2] = C[0][2][2] + B.b022;
  k[2][0][2] = C[2][0][2] + B.b202;
  k[1][2][2] = C[1][2][2] + B.b122;
  k[2][1][2] = C[2][1][2] + B.b212;
  k[2][2][1] = C[2][2][1] + B.b221;
  B = brackets(k, ir);  // the order-6 bracket reads the restored order-4 members
  k[2][2][2] = C[2][2][2] + B.a - B.b;
}
// relaxation (cum_collide_kernel, :405-439): shear omega1 = 1/tau, bulk omega2 = 1,
// every cumulant of order 3..6 set to zero
void cum_relax(M3 C, const M3 k, double tau, const double* u, double rho) {
  double w1 = 1. / tau, w2 = 1.;
  double ir = safe_irho(rho);
  double c200 = C[2][0][0], c020 = C[0][2][0], c002 = C[0][0][2];
  double Dxu = -(w1 * (2 * c200 - c020 - c002) + w2 * (c200 + c020 + c002 - k[0][0][0])) * ir * 0.5;
  double Dyv = Dxu + 3 * w1 * (c200 - c020) * 0.5 * ir;
  double Dzw = Dxu + 3 * w1 * (c200 - c002) * 0.5 * ir;
  C[1][1][0] *= (1 - w1);
  C[1][0][1] *= (1 - w1);
  C[0][1][1] *= (1 - w1);
  double ux2 = u[0] * u[0], uy2 = u[1] * u[1], uz2 = u[2] * u[2];
  double X1 = (1 - w1) * (c200 - c020) - 3 * rho * (1 - w1 * 0.5) * (ux2 * Dxu - uy2 * Dyv);
  double X2 = (1 - w1) * (c200 - c002) - 3 * rho * (1 - w1 * 0.5) * (ux2 * Dxu - uz2 * Dzw);
  double X3 = k[0][0][0] * w2 + (1 - w2) * (c200 + c020 + c002) -
              3 * rho * (1 - w2 * 0.5) * (ux2 * Dxu + uy2 * Dyv + uz2 * Dzw);
  C[2][0][0] = (X1 + X2 + X3) / 3;
  C[0][2][0] = C[2][0][0] - X1;
  C[0][0][2] = C[2][0][0] - X2;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c)
        if (a + b + c >= 3) C[a][b][c] = 0;
}
void cumulant_collide(const double* f, double rho, const double* u, double tau, double* out) {
  M3 k, C;
  central_fwd(f, u, k);
  to_cumulants(k, rho, C);
  cum_relax(C, k, tau, u, rho);
  from_cumulants(C, rho, k);
  central_bwd(k, u, out);
}

// ---------------------------------------------------------------- the case
struct Level {
  std::vector<uint64_t> lat_f, f2c, c2f, ghost_c2f, ghost_f2c, solid, bcres;
  std::unordered_set<uint64_t> lat_f_set, ghost_c2f_set;
  std::unordered_map<uint64_t, uint16_t> kind;
  std::vector<std::unordered_map<uint64_t, double>> f, fcol;  // one map per population
  std::unordered_map<uint64_t, double> rho;
  std::unordered_map<uint64_t, std::array<double, 3>> u, v_old;
};

struct Case {
  int Q = 19, Lmax = 0, collision = 0, ib = 0, nee_seq = 1, options = 0;
  int mdf_sweeps = 8;  // the reference's MDF_ITERS; tests set 0 to read the pre-forcing state
  double tau0 = 0.6, les_C = 0, rho0 = 1;
  double u0[3] = {0, 0, 0};
  int dims0[3] = {0, 0, 0};
  int face_kind[6] = {0};
  double face_u[6][3] = {{0}};
  std::vector<double> tau;
  Morton mt;
  std::vector<Level> lv;
  std::map<uint64_t, std::array<double, 26>> dis;  // ascending-key iteration
  std::vector<std::array<double, 3>> force_log;
  // Lagrangian markers on the finest level (definition, no reference
  // counterpart): per marker up to 27 (cell key, weight) pairs of a 3-point
  // kernel, its surface element dV and the body velocity ub.
  int64_t n_mk = 0;
  std::vector<uint64_t> mk_keys;  // [n][27]
  std::vector<double> mk_w;       // [n][27]
  std::vector<double> mk_dv;      // [n]
  double mk_ub[3] = {0, 0, 0};

  // Explicit direct forcing on the post-collision populations of the finest
  // level: the marker velocity and density are kernel-weighted sums of the
  // cells' moments (taken from fcol), the marker force density is
  // F_k = dV_k rho_k (ub - u_k), each cell receives the momentum
  // dm(x) = sum_k w_k(x) F_k and its populations shift by the exact-difference
  // term feq(rho, u + dm/rho) - feq(rho, u) (Kupershtokh's method).  The force
  // logged like a treat() force is sum_k F_k: the body's reaction -sum_k F_k
  // expressed, like the link-wise force, with the reported velocities' sign
  // flipped back (the populations move along -e_q).
  void marker_correct() {
    if (!n_mk) return;
    Level& l = lv[Lmax];
    std::vector<std::array<double, 3>> F(n_mk);
    auto moments = [&](uint64_t k, double& rho, double* u) {
      double f[27];
      load(Lmax, k, f, true);
      macros(Q, f, rho, u);
    };
    for (int64_t m = 0; m < n_mk; ++m) {
      double rk = 0, uk[3] = {0, 0, 0};
      for (int j = 0; j < 27; ++j) {
        const double w = mk_w[m * 27 + j];
        if (w == 0.0) continue;
        double rho, u[3];
        moments(mk_keys[m * 27 + j], rho, u);
        rk += w * rho;
        for (int a = 0; a < 3; ++a) uk[a] += w * u[a];
      }
      for (int a = 0; a < 3; ++a) F[m][a] = mk_dv[m] * rk * (mk_ub[a] - uk[a]);
    }
    // spread in ascending (cell key, marker) order
    std::map<uint64_t, std::array<double, 3>> dm;
    std::vector<std::tuple<uint64_t, int64_t, double>> pairs;
    for (int64_t m = 0; m < n_mk; ++m)
      for (int j = 0; j < 27; ++j)
        if (mk_w[m * 27 + j] != 0.0) pairs.emplace_back(mk_keys[m * 27 + j], m, mk_w[m * 27 + j]);
    std::sort(pairs.begin(), pairs.end(), [](auto& a, auto& b) {
      return std::get<0>(a) != std::get<0>(b) ? std::get<0>(a) < std::get<0>(b) : std::get<1>(a) < std::get<1>(b);
    });
    for (auto& p : pairs) {
      auto& d = dm[std::get<0>(p)];
      for (int a = 0; a < 3; ++a) d[a] += std::get<2>(p) * F[std::get<1>(p)][a];
    }
    for (auto& kv : dm) {
      double f[27], rho, u[3], u2[3], fe1[27], fe2[27];
      load(Lmax, kv.first, f, true);
      macros(Q, f, rho, u);
      for (int a = 0; a < 3; ++a) u2[a] = u[a] + kv.second[a] / rho;
      feq_c(Q, rho, u2, fe2);
      feq_c(Q, rho, u, fe1);
      for (int q = 0; q < Q; ++q) l.fcol[q].at(kv.first) = f[q] + (fe2[q] - fe1[q]);
    }
    double fx = 0, fy = 0, fz = 0;
    for (int64_t m = 0; m < n_mk; ++m) {
      fx += F[m][0];
      fy += F[m][1];
      fz += F[m][2];
    }
    force_log.push_back({fx, fy, fz});
  }

  // neighbour with periodic wrap on periodic axes (definition: periodic = wrapped
  // neighbour lookup, the limit of the ghost-copy text in BC/Periodic.cpp:8-445)
  uint64_t nbr(uint64_t key, int L, int dx, int dy, int dz) const {
    bool per[3] = {face_kind[0] == BC_PERIODIC, face_kind[2] == BC_PERIODIC, face_kind[4] == BC_PERIODIC};
    if (!(per[0] || per[1] || per[2])) return mt.neighbor(key, L, dx, dy, dz);
    int s = Lmax - L;
    int64_t c[3] = {(int64_t)(compact(key) >> s), (int64_t)(compact(key >> 1) >> s), (int64_t)(compact(key >> 2) >> s)};
    int d[3] = {dx, dy, dz};
    bool wrapped = false;
    for (int a = 0; a < 3; ++a) {
      int64_t n = (int64_t)dims0[a] << L;
      int64_t v = c[a] + d[a];
      if (per[a] && (v < 0 || v >= n)) {
        c[a] = (v + n) % n;
        wrapped = true;
      } else {
        c[a] = v;
      }
    }
    if (!wrapped) return mt.neighbor(key, L, dx, dy, dz);
    return spread((uint64_t)c[0] << s) | (spread((uint64_t)c[1] << s) << 1) | (spread((uint64_t)c[2] << s) << 2);
  }

  void load(int L, uint64_t k, double* out, bool col) const {
    auto& m = col ? lv[L].fcol : lv[L].f;
    for (int q = 0; q < Q; ++q) out[q] = m[q].at(k);
  }
  void children(uint64_t k, int fine, uint64_t* c) const {
    // child order of LBM_Manager.cpp:1074-1082
    c[0] = k;
    c[1] = mt.x1(k, fine);
    c[2] = mt.y1(k, fine);
    c[3] = mt.y1(mt.x1(k, fine), fine);
    c[4] = mt.z1(k, fine);
    c[5] = mt.z1(mt.x1(k, fine), fine);
    c[6] = mt.z1(mt.y1(k, fine), fine);
    c[7] = mt.z1(mt.y1(mt.x1(k, fine), fine), fine);
  }

  // allocMacrosAndDDF (LBM_Manager.cpp:482-592) + initialBC (:296-304)
  void init() {
    double fe[27];
    feq_c(Q, rho0, u0, fe);
    for (int L = 0; L <= Lmax; ++L) {
      Level& l = lv[L];
      l.f.assign(Q, {});
      l.fcol.assign(Q, {});
      for (auto& kv : l.kind) {
        uint64_t k = kv.first;
        uint16_t kd = kv.second;
        bool sol = kd & K_SOLID;
        for (int q = 0; q < Q; ++q) {
          l.f[q][k] = sol ? 0.0 : fe[q];
          l.fcol[q][k] = sol ? 0.0 : fe[q];
        }
        if (kd & (K_FLUID | K_F2C | K_C2F | K_BCRES)) {
          l.rho[k] = rho0;
          l.u[k] = {u0[0], u0[1], u0[2]};
          l.v_old[k] = {0, 0, 0};
        } else if (sol) {
          l.rho[k] = 1.0;
          l.u[k] = {0, 0, 0};
          l.v_old[k] = {0, 0, 0};
        }
      }
    }
    init_tau();
  }
  void init_tau() {
    tau.resize(Lmax + 1);
    for (int L = 0; L <= Lmax; ++L) tau[L] = double(1 << L) * (tau0 - 0.5) + 0.5;
  }

  // calculateMacros(level) (LBM_Manager.cpp:1003-1050)
  void calc_macros(int L) {
    Level& l = lv[L];
    double f[27], u[3];
    for (uint64_t k : l.lat_f) {
      l.v_old[k] = l.u.at(k);
      load(L, k, f, false);
      macros(Q, f, l.rho.at(k), u);
      l.u.at(k) = {u[0], u[1], u[2]};
    }
    for (uint64_t k : l.f2c) {
      load(L, k, f, false);
      macros(Q, f, l.rho.at(k), u);
      l.u.at(k) = {u[0], u[1], u[2]};
    }
  }

  // Collision_LBGK::collide / Collision_Cumulant::collide
  // (lbm_kernel/Collision_Manager.cpp:13-53, lbm_kernel/Collision_Cumulant.cpp:22-159)
  void collide(int L) {
    Level& l = lv[L];
    auto one = [&](uint64_t k) {
      double f[27], fe[27], out[27];
      load(L, k, f, false);
      double rho = l.rho.at(k);
      auto uu = l.u.at(k);
      double u[3] = {uu[0], uu[1], uu[2]};
      double t = tau[L];
      if (les_C > 0) {  // definition: the #ifdef LES path with a compiled stress
        feq_c(Q, rho, u, fe);
        t = smag_tau(t, les_C, stress_norm(Q, f, fe));
      }
      if (collision == 0) {
        feq_c(Q, rho, u, fe);
        for (int q = 0; q < Q; ++q) out[q] = f[q] - (f[q] - fe[q]) / t;
      } else {
        cumulant_collide(f, rho, u, t, out);
      }
      for (int q = 0; q < Q; ++q) l.fcol[q].at(k) = out[q];
    };
    for (uint64_t k : l.lat_f) one(k);
    for (uint64_t k : l.f2c) one(k);
    for (uint64_t k : l.c2f)
      for (int q = 0; q < Q; ++q) l.fcol[q].at(k) = l.f[q].at(k);
  }

  // ABPatternPull::stream (lbm_kernel/Stream_Manager.cpp:43-76): f_q(x) <- fcol_q(x+e_q)
  void stream(int L) {
    Level& l = lv[L];
    auto one = [&](uint64_t k) {
      l.f[0].at(k) = l.fcol[0].at(k);
      for (int q = 1; q < Q; ++q) {
        uint64_t n = nbr(k, L, EX[q], EY[q], EZ[q]);
        auto it = l.fcol[q].find(n);
        if (it == l.fcol[q].end()) continue;
        l.f[q].at(k) = it->second;
      }
    };
    for (uint64_t k : l.lat_f) one(k);
    for (uint64_t k : l.c2f) one(k);
  }

  // AMR_transDDF_explosion (LBM_Manager.cpp:1059-1129)
  void explode(int fine) {
    int coarse = fine - 1;
    double s = (options & 1) ? tau[fine] / (2.0 * tau[coarse]) : (2.0 * tau[fine]) / tau[coarse];
    Level& lc = lv[coarse];
    Level& lf = lv[fine];
    double dc[27], fe[27], u[3], rho;
    uint64_t ch[8];
    for (uint64_t k : lc.f2c) {
      children(k, fine, ch);
      load(coarse, k, dc, true);
      macros(Q, dc, rho, u);
      feq_hl(Q, rho, u, rho0, fe);
      for (int q = 0; q < Q; ++q) {
        double v = fe[q] + s * (dc[q] - fe[q]);
        for (int c = 0; c < 8; ++c) {
          if (dis.count(ch[c])) continue;
          lf.fcol[q].at(ch[c]) = v;
        }
      }
    }
    for (uint64_t k : lc.ghost_f2c) {
      children(k, fine, ch);
      load(coarse, k, dc, true);
      macros(Q, dc, rho, u);
      feq_hl(Q, rho, u, rho0, fe);
      for (int c = 0; c < 8; ++c) {
        if (dis.count(ch[c])) continue;
        if (lf.ghost_c2f_set.count(ch[c]))
          for (int q = 0; q < Q; ++q) lf.fcol[q].at(ch[c]) = fe[q] + s * (dc[q] - fe[q]);
      }
    }
  }

  // AMR_transDDF_coalescence (LBM_Manager.cpp:1131-1169)
  void coalesce(int coarse) {
    int fine = coarse + 1;
    double s = (options & 1) ? (2.0 * tau[coarse]) / tau[fine] : tau[coarse] / (2.0 * tau[fine]);
    Level& lc = lv[coarse];
    Level& lf = lv[fine];
    double avg[27], fe[27], u[3], rho;
    uint64_t ch[8];
    for (uint64_t k : lc.f2c) {
      children(k, fine, ch);
      for (int q = 0; q < Q; ++q) {
        avg[q] = 0;
        for (int c = 0; c < 8; ++c) avg[q] += lf.f[q].at(ch[c]);
        avg[q] /= 8.0;
      }
      macros(Q, avg, rho, u);
      feq_hl(Q, rho, u, rho0, fe);
      for (int q = 0; q < Q; ++q) lc.f[q].at(k) = fe[q] + s * (avg[q] - fe[q]);
    }
  }

  // ---- level-0 boundaries, applied to f after stream --------------------
  // which faces (W,E,S,N,B,T) a key lies on, from its face bits
  static int faces_of(uint16_t kd) { return (kd >> 8) & 0x3F; }
  static void normal(int face, int* n) {
    n[0] = n[1] = n[2] = 0;
    n[face / 2] = (face & 1) ? 1 : -1;
  }

  // NEE, compiled eflow_stable_bc form (BC/NonEquilibrumExtrapolation.cpp:50-66):
  // f_q(x_b) = feq_q(rho_b, u_bc) + f_q(x_b+e_q) - feq_q(rho_n, u_n) for the
  // directions whose neighbour exists (header/BC/D3Q19_BC_Manager.hpp:60-122).
  // Sequential mode reproduces the in-place chain along a face in ascending key
  // order over the face's set list; Jacobi mode reads pre-pass f everywhere.
  void nee_face(int face, const std::vector<uint64_t>& cells,
                const std::unordered_map<uint64_t, std::vector<double>>* snapshot) {
    Level& l = lv[0];
    const double* ub = face_u[face];
    for (uint64_t k : cells) {
      int fm = faces_of(l.kind.at(k));
      for (int q = 0; q < Q; ++q) {
        // the direction exists iff it does not leave the box through a face the cell is on
        bool ok = true;
        int e[3] = {EX[q], EY[q], EZ[q]};
        for (int ff = 0; ff < 6; ++ff)
          if (fm & (1 << ff)) {
            int n[3];
            normal(ff, n);
            if (e[0] * n[0] + e[1] * n[1] + e[2] * n[2] > 0) ok = false;
          }
        if (!ok) continue;
        uint64_t nb = nbr(k, 0, EX[q], EY[q], EZ[q]);
        double rb = l.rho.at(k);
        double un3[3] = {l.u.at(nb)[0], l.u.at(nb)[1], l.u.at(nb)[2]};
        double fw[27], fn[27];
        feq_c(Q, rb, ub, fw);
        feq_c(Q, l.rho.at(nb), un3, fn);
        double fnb = snapshot ? snapshot->at(nb)[q] : l.f[q].at(nb);
        l.f[q].at(k) = fw[q] + fnb - fn[q];
      }
    }
  }
  // set lists per face (header/BC/NonEquilibrumExtrapolation.hpp:24-109): West
  // and East take their face plus the 4 edges and 4 corners they touch, South
  // and North their face plus the two top/bottom edges, Top/Bot the face only.
  std::vector<uint64_t> nee_cells(int face) {
    // enumerate in set order, each set in ascending key order
    Level& l = lv[0];
    std::vector<std::vector<uint64_t>> sets;
    auto on = [&](uint64_t k, int m) { return faces_of(l.kind.at(k)) == m; };
    const int W = 1, E = 2, S = 4, N = 8, B = 16, T = 32;
    std::vector<int> masks;
    if (face == 0) masks = {W, W | N, W | S, W | T, W | B, W | N | T, W | S | T, W | N | B, W | S | B};
    if (face == 1) masks = {E, E | N, E | S, E | T, E | B, E | N | T, E | S | T, E | N | B, E | S | B};
    if (face == 2) masks = {S, S | T, S | B};
    if (face == 3) masks = {N, N | T, N | B};
    if (face == 4) masks = {B};
    if (face == 5) masks = {T};
    std::vector<uint64_t> out;
    for (int m : masks)
      for (uint64_t k : l.bcres)
        if (on(k, m)) out.push_back(k);
    return out;
  }

  // Wet-node boundary definitions (the reference has these as uncompiled text:
  // BC/ZouHe_Velocity.cpp:170-532, BC/Outflow.cpp:12-136).  Under the reference's
  // pull convention f_q(x) <- fcol_q(x+e_q) the unknown populations on a face with
  // outward normal n are those with e_q.n = +1.
  //  * Zou-He (definition): rho = (S0 + 2 S_in)/(1 - u_w.n) with S0 = sum over
  //    e.n = 0, S_in = sum over e.n = -1; for each unknown q
  //    f_q = f_inv(q) + 6 w_q rho (e_q.u_w) - 36 w_q sum_t (e_q.t) N_t,
  //    N_t = 1/2 sum_{e.n=0} f_j (e_j.t) - rho (u_w.t)/3 over the two tangents t.
  //    For D3Q19 this is Hecht-Harting; it reproduces rho and u_w exactly.
  //  * Outflow (definition): f_q(x_b) = f_q(x_b - n) for the unknown q.
  //  * Wet-node NEE (definition): f_q = feq_q(rho_n, u_w) + f_q(x_n) - feq_q(rho_n, u_n)
  //    for all q, x_n = x_b - n.
  //  * Edges/corners (>= 2 wet faces, definition): f_q = feq_q(rho_n, u_w) +
  //    f_q(x_n) - feq_q(rho_n, u_n) for all q, x_n = x_b - sum n, u_w the
  //    velocity of the last Zou-He face in W,E,S,N,B,T order (u_n if none).
  // All reads are of post-stream f of cells no boundary rule modifies.
  void wet_bc() {
    Level& l = lv[0];
    std::vector<std::pair<uint64_t, std::vector<double>>> updates;
    for (uint64_t k : l.lat_f) {
      int fm = faces_of(l.kind.at(k));
      int act = 0, nact = 0;
      for (int ff = 0; ff < 6; ++ff)
        if ((fm & (1 << ff)) && (face_kind[ff] == BC_ZOUHE || face_kind[ff] == BC_OUTFLOW || face_kind[ff] == BC_NEE)) {
          act |= 1 << ff;
          ++nact;
        }
      if (!nact) continue;
      std::vector<double> f(Q);
      for (int q = 0; q < Q; ++q) f[q] = l.f[q].at(k);
      if (nact == 1) {
        int face = __builtin_ctz(act);
        int n[3];
        normal(face, n);
        auto en = [&](int q) { return EX[q] * n[0] + EY[q] * n[1] + EZ[q] * n[2]; };
        if (face_kind[face] == BC_NEE) {
          // wet-node NEE (definition): the NonEquilibrumExtrapolation rule of
          // BC/NonEquilibrumExtrapolation.cpp:50-66 applied on the wall cell, from
          // its inward neighbour x - n for every q
          uint64_t kin = nbr(k, 0, -n[0], -n[1], -n[2]);
          double fn[27], un[3], rn, fw[27], fq[27];
          load(0, kin, fn, false);
          macros(Q, fn, rn, un);
          feq_c(Q, rn, face_u[face], fw);
          feq_c(Q, rn, un, fq);
          for (int q = 0; q < Q; ++q) f[q] = fw[q] + fn[q] - fq[q];
        } else if (face_kind[face] == BC_OUTFLOW) {
          uint64_t kin = nbr(k, 0, -n[0], -n[1], -n[2]);
          for (int q = 0; q < Q; ++q)
            if (en(q) == 1) f[q] = l.f[q].at(kin);
        } else {
          const double* uw = face_u[face];
          double S0 = 0, Sin = 0;
          for (int q = 0; q < Q; ++q) {
            if (en(q) == 0) S0 += f[q];
            else if (en(q) == -1) Sin += f[q];
          }
          double un = uw[0] * n[0] + uw[1] * n[1] + uw[2] * n[2];
          double rho = (S0 + 2 * Sin) / (1 - un);
          int ax = face / 2;
          int t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
          if (t1 > t2) std::swap(t1, t2);
          double Nt[3] = {0, 0, 0};
          for (int t : {t1, t2}) {
            double s = 0;
            for (int q = 0; q < Q; ++q)
              if (en(q) == 0) {
                int e[3] = {EX[q], EY[q], EZ[q]};
                s += f[q] * e[t];
              }
            Nt[t] = 0.5 * s - rho * uw[t] / 3.0;
          }
          for (int q = 0; q < Q; ++q) {
            if (en(q) != 1) continue;
            int e[3] = {EX[q], EY[q], EZ[q]};
            double w = weight(Q, q);
            double eu = e[0] * uw[0] + e[1] * uw[1] + e[2] * uw[2];
            double tr = e[t1] * Nt[t1] + e[t2] * Nt[t2];
            f[q] = f[inv(q)] + 6 * w * rho * eu - 36 * w * tr;
          }
        }
      } else {
        int d[3] = {0, 0, 0};
        const double* uw = nullptr;
        for (int ff = 0; ff < 6; ++ff)
          if (act & (1 << ff)) {
            int n[3];
            normal(ff, n);
            for (int a = 0; a < 3; ++a) d[a] -= n[a];
            if (face_kind[ff] == BC_ZOUHE || face_kind[ff] == BC_NEE) uw = face_u[ff];
          }
        uint64_t kn = nbr(k, 0, d[0], d[1], d[2]);
        double fn[27], un[3], rn;
        load(0, kn, fn, false);
        macros(Q, fn, rn, un);
        double fw[27], fq[27];
        feq_c(Q, rn, uw ? uw : un, fw);
        feq_c(Q, rn, un, fq);
        for (int q = 0; q < Q; ++q) f[q] = fw[q] + fn[q] - fq[q];
      }
      updates.emplace_back(k, std::move(f));
    }
    for (auto& up : updates)
      for (int q = 0; q < Q; ++q) l.f[q].at(up.first) = up.second[q];
  }

  // applyBoundaryCondition(face, "AfterStream") for W,E,S,N,Bot,Top in turn
  // (LBM_Simulator.cpp:89-96, header/BC/D3Q19_BC_Manager.hpp:220-244)
  void bc_after_stream() {
    std::unordered_map<uint64_t, std::vector<double>> snap;
    bool any_res = false;
    for (int face = 0; face < 6; ++face) any_res |= face_kind[face] == BC_RESERVOIR;
    if (any_res && !nee_seq) {
      Level& l = lv[0];
      for (auto& kv : l.kind) {
        std::vector<double> v(Q);
        for (int q = 0; q < Q; ++q) v[q] = l.f[q].at(kv.first);
        snap[kv.first] = v;
      }
    }
    for (int face = 0; face < 6; ++face)
      if (face_kind[face] == BC_RESERVOIR) nee_face(face, nee_cells(face), nee_seq ? nullptr : &snap);
    wet_bc();
  }

  // UniformBoundary::treat (ComplexBoundary_Manager.cpp:15-40): every dis cell,
  // directions in order; a missing neighbour throws std::out_of_range as .at() does
  void uniform_treat() {
    Level& l = lv[Lmax];
    for (auto& d : dis) {
      uint64_t k = d.first;
      for (int q = 1; q < Q; ++q) {
        double dist = d.second[inv(q) - 1];
        if (dist != -1) {
          uint64_t ki = nbr(k, Lmax, EX[inv(q)], EY[inv(q)], EZ[inv(q)]);
          double fc = l.fcol[q].at(k);
          double fci = l.fcol[inv(q)].at(k);
          double fo = l.fcol[q].at(ki);
          l.f[inv(q)].at(k) = (dist * (fci + fc) + (1. - dist) * fo) / (1. + dist);
        } else {
          uint64_t kn = nbr(k, Lmax, EX[q], EY[q], EZ[q]);
          l.f[q].at(k) = l.fcol[q].at(kn);
        }
      }
    }
  }

  // SinglePointInterpolation::treat (ComplexBoundary_Manager.cpp:43-225)
  void ib_treat(double* F) {
    if (ib == 3) {
      uniform_treat();
      if (F) F[0] = F[1] = F[2] = 0.0;
      return;
    }
    int L = Lmax;
    Level& l = lv[L];
    double fx = 0, fy = 0, fz = 0;
    for (auto& d : dis) {
      uint64_t k = d.first;
      if (!l.lat_f_set.count(k)) continue;
      auto ub = l.u.at(k);
      double rb = l.rho.at(k);
      (void)ub;
      l.f[0].at(k) = l.fcol[0].at(k);
      for (int q = 1; q < Q; ++q) {
        double dist = d.second[inv(q) - 1];
        if (dist != -1.) {
          double finc = l.fcol[inv(q)].at(k);
          uint64_t kn = nbr(k, L, EX[inv(q)], EY[inv(q)], EZ[inv(q)]);
          auto it = l.fcol[q].find(kn);
          double fs = (it != l.fcol[q].end()) ? it->second : finc;
          double fnew;
          if (ib == 1) {
            if (dist < 0.5) {
              fnew = (1.0 - 2.0 * dist) * finc + 2.0 * dist * fs;
            } else {
              double fe = ib_weight(inv(q)) * rb;
              double i2d = 1.0 / (2.0 * dist);
              fnew = i2d * fs + (1.0 - i2d) * fe;
            }
          } else {
            fnew = (1.0 - dist) * finc + dist * fs;
          }
          l.f[q].at(k) = fnew;
          double m = finc + fnew;
          fx += -m * EX[q];
          fy += -m * EY[q];
          fz += -m * EZ[q];
        } else {
          uint64_t kn = nbr(k, L, EX[q], EY[q], EZ[q]);
          auto it = l.fcol[q].find(kn);
          if (it == l.fcol[q].end())
            l.f[q].at(k) = l.fcol[inv(q)].at(k);
          else
            l.f[q].at(k) = it->second;
        }
      }
      double ra = 0;
      for (int q = 0; q < Q; ++q) ra += l.f[q].at(k);
      double& f0 = l.f[0].at(k);
      if (ra < 0.3) f0 += 1.0 * (0.3 - ra);
      else if (ra < 0.5) f0 += 0.5 * (0.5 - ra);
      else if (ra > 3.0) f0 += 1.0 * (3.0 - ra);
      else if (ra > 2.0) f0 += 0.5 * (2.0 - ra);
    }
    force_log.push_back({fx, fy, fz});
    if (F) { F[0] = fx; F[1] = fy; F[2] = fz; }
    // multi-direct forcing, 8 sweeps of relaxation 0.5 (:180-225)
    for (int it = 0; it < mdf_sweeps; ++it) {
      for (auto& d : dis) {
        uint64_t k = d.first;
        if (!l.lat_f_set.count(k)) continue;
        double rho = 0, u[3] = {0, 0, 0};
        for (int q = 0; q < Q; ++q) {
          double fq = l.f[q].at(k);
          rho += fq;
          u[0] += fq * EX[q];
          u[1] += fq * EY[q];
          u[2] += fq * EZ[q];
        }
        if (rho <= 0.) continue;
        u[0] /= rho; u[1] /= rho; u[2] /= rho;
        double um = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        if (um < 1e-10) continue;
        double dM[3] = {-0.5 * rho * u[0], -0.5 * rho * u[1], -0.5 * rho * u[2]};
        double adj = 0;
        for (int q = 1; q < Q; ++q) {
          if (d.second[inv(q) - 1] == -1.) continue;
          double df = ib_weight(q) * 3.0 * (EX[q] * dM[0] + EY[q] * dM[1] + EZ[q] * dM[2]);
          l.f[q].at(k) += df;
          adj -= df;
        }
        l.f[0].at(k) += adj;
      }
    }
  }

  // LBMkernel (LBM_Simulator.cpp:20-229)
  void kernel(int L) {
    calc_macros(L);
    collide(L);
    if (L == Lmax) marker_correct();
    if (L < Lmax) kernel(L + 1);
    if (L > 0) explode(L);
    stream(L);
    if (L == 0) bc_after_stream();
    else if (L == Lmax && ib) ib_treat(nullptr);
    if (L < Lmax) coalesce(L);
    if (L == 0) return;
    calc_macros(L);
    collide(L);
    if (L == Lmax) marker_correct();
    if (L < Lmax) kernel(L + 1);
    if (options & 2) explode(L);
    stream(L);
    if (L == 0) bc_after_stream();
    else if (L == Lmax && ib) ib_treat(nullptr);
    if (L < Lmax) coalesce(L);
  }
  // fluidSimulate body (LBM_Simulator.cpp:11-17)
  void step() {
    kernel(0);
    for (int L = 0; L <= Lmax; ++L) calc_macros(L);
  }
};


// ---------------------------------------------------------------- dense mode
// The same algorithm on sorted arrays: every cell set of a level becomes a list
// of indices into the level's keys in ascending key order (the order the map mode
// walks), every map lookup becomes an index resolved once from the same Morton
// arithmetic, and the per-cell arithmetic is the map mode's own functions.  Loops
// whose cells do not interact (all but the NEE chain and the fixed-order force
// sums) run in parallel; those two run in the map mode's order.  The result is
// the map mode's bits (tests compare them) at a speed that reaches full-size runs.
struct Dense {
  struct Lev {
    std::vector<uint64_t> keys;
    std::unordered_map<uint64_t, int64_t> idx;
    std::vector<uint16_t> kind;
    std::vector<double> f, fcol, rho, u, v_old;  // f/fcol [Q][n], u/v_old [3][n]
    std::vector<uint8_t> has_macro;
    std::vector<int64_t> lat_f, f2c, c2f, bcres;
    std::vector<int64_t> strm_cells;  // lat_f then c2f
    std::vector<int64_t> strm_src;    // [cell][Q]
    std::vector<int64_t> ex_parent, ex_child;  // coarse F2C parent -> 8 fine children (-1: skipped)
    std::vector<int64_t> gx_parent, gx_child;  // coarse ghost-F2C parent -> 8 ghost children (-1: none)
    std::vector<int64_t> co_parent, co_child;  // this level's F2C -> 8 children on the finer level
    int64_t n = 0;
  };
  const Case* src = nullptr;
  int Q = 19, Lmax = 0;
  std::vector<Lev> lv;
  // finest-level treated cells (ascending key order, lat_f members only)
  std::vector<int64_t> tr_cell, tr_inc_src, tr_pull_src;  // tr_*_src [cell][Q]
  std::vector<std::array<double, 26>> tr_dis;
  // level-0 wet boundary cells and reservoir chains
  std::vector<int64_t> wet_cell, wet_in;  // wet_in [cell][2]: x - n (one face) or x - sum n
  std::vector<int> wet_act;
  std::vector<std::vector<int64_t>> nee_cells, nee_nb;  // per face: cells, [cell][Q] neighbour or -1 (not a direction)
  // markers
  std::vector<int64_t> mk_idx;  // [n][27] index or -1
  std::vector<int64_t> band_cell, band_ptr, band_mk;
  std::vector<double> band_w;
  std::vector<std::array<double, 3>> force_log;

  int64_t at(int L, uint64_t k) const {
    auto it = lv[L].idx.find(k);
    return it == lv[L].idx.end() ? -1 : it->second;
  }

  // fresh: the initial state of Case::init (LBM_Manager.cpp:482-592) is set here and
  // the case's maps are not read (they need not be allocated)
  int build(const Case& c, bool fresh) {
    src = &c;
    if (c.ib == 3) return 4;  // the uniform boundary runs in the map mode only
    Q = c.Q;
    Lmax = c.Lmax;
    lv.assign(Lmax + 1, {});
    for (int L = 0; L <= Lmax; ++L) {
      const Level& m = c.lv[L];
      Lev& d = lv[L];
      for (auto& kv : m.kind) d.keys.push_back(kv.first);
      std::sort(d.keys.begin(), d.keys.end());
      d.n = (int64_t)d.keys.size();
      d.idx.reserve(d.n * 2);
      for (int64_t i = 0; i < d.n; ++i) d.idx[d.keys[i]] = i;
      d.kind.resize(d.n);
      d.f.resize(Q * d.n);
      d.fcol.resize(Q * d.n);
      d.rho.assign(d.n, NAN);
      d.u.assign(3 * d.n, NAN);
      d.v_old.assign(3 * d.n, NAN);
      d.has_macro.assign(d.n, 0);
      double fe[27];
      feq_c(Q, c.rho0, c.u0, fe);
      for (int64_t i = 0; i < d.n && fresh; ++i) {
        uint16_t kd = m.kind.at(d.keys[i]);
        d.kind[i] = kd;
        bool sol = kd & K_SOLID;
        for (int q = 0; q < Q; ++q) d.f[q * d.n + i] = d.fcol[q * d.n + i] = sol ? 0.0 : fe[q];
        if (kd & (K_FLUID | K_F2C | K_C2F | K_BCRES)) {
          d.has_macro[i] = 1;
          d.rho[i] = c.rho0;
          for (int a = 0; a < 3; ++a) {
            d.u[a * d.n + i] = c.u0[a];
            d.v_old[a * d.n + i] = 0.0;
          }
        } else if (sol) {
          d.has_macro[i] = 1;
          d.rho[i] = 1.0;
          for (int a = 0; a < 3; ++a) d.u[a * d.n + i] = d.v_old[a * d.n + i] = 0.0;
        }
      }
      for (int64_t i = 0; i < d.n && !fresh; ++i) {
        uint64_t k = d.keys[i];
        d.kind[i] = m.kind.at(k);
        for (int q = 0; q < Q; ++q) {
          d.f[q * d.n + i] = m.f[q].at(k);
          d.fcol[q * d.n + i] = m.fcol[q].at(k);
        }
        auto r = m.rho.find(k);
        if (r != m.rho.end()) {
          d.has_macro[i] = 1;
          d.rho[i] = r->second;
          auto& uu = m.u.at(k);
          auto& vo = m.v_old.at(k);
          for (int a = 0; a < 3; ++a) {
            d.u[a * d.n + i] = uu[a];
            d.v_old[a * d.n + i] = vo[a];
          }
        }
      }
      auto to_idx = [&](const std::vector<uint64_t>& ks, std::vector<int64_t>& out) {
        out.clear();
        for (uint64_t k : ks) out.push_back(d.idx.at(k));
      };
      to_idx(m.lat_f, d.lat_f);
      to_idx(m.f2c, d.f2c);
      to_idx(m.c2f, d.c2f);
      to_idx(m.bcres, d.bcres);
      d.strm_cells = d.lat_f;
      d.strm_cells.insert(d.strm_cells.end(), d.c2f.begin(), d.c2f.end());
      d.strm_src.assign(d.strm_cells.size() * Q, -1);
      for (size_t j = 0; j < d.strm_cells.size(); ++j) {
        uint64_t k = d.keys[d.strm_cells[j]];
        for (int q = 1; q < Q; ++q) d.strm_src[j * Q + q] = at(L, c.nbr(k, L, EX[q], EY[q], EZ[q]));
      }
    }
    // transfers
    for (int L = 1; L <= Lmax; ++L) {
      const Level& mc = c.lv[L - 1];
      const Level& mf = c.lv[L];
      Lev& dc = lv[L - 1];
      Lev& df = lv[L];
      uint64_t ch[8];
      for (uint64_t k : mc.f2c) {
        c.children(k, L, ch);
        df.ex_parent.push_back(dc.idx.at(k));
        for (int j = 0; j < 8; ++j) {
          int64_t i = at(L, ch[j]);
          if (i < 0) return 1;  // the map mode's .at() would throw
          df.ex_child.push_back(c.dis.count(ch[j]) ? -1 : i);
        }
      }
      for (uint64_t k : mc.ghost_f2c) {
        c.children(k, L, ch);
        df.gx_parent.push_back(dc.idx.at(k));
        for (int j = 0; j < 8; ++j) {
          bool ok = !c.dis.count(ch[j]) && mf.ghost_c2f_set.count(ch[j]);
          df.gx_child.push_back(ok ? df.idx.at(ch[j]) : -1);
        }
      }
      for (uint64_t k : mc.f2c) {
        c.children(k, L, ch);
        dc.co_parent.push_back(dc.idx.at(k));
        for (int j = 0; j < 8; ++j) {
          int64_t i = at(L, ch[j]);
          if (i < 0) return 1;
          dc.co_child.push_back(i);
        }
      }
    }
    // treated cells
    const Level& top = c.lv[Lmax];
    for (auto& dd : c.dis) {
      uint64_t k = dd.first;
      if (!top.lat_f_set.count(k)) continue;
      tr_cell.push_back(lv[Lmax].idx.at(k));
      tr_dis.push_back(dd.second);
      for (int q = 0; q < Q; ++q) {
        tr_inc_src.push_back(q ? at(Lmax, c.nbr(k, Lmax, EX[inv(q)], EY[inv(q)], EZ[inv(q)])) : -1);
        tr_pull_src.push_back(q ? at(Lmax, c.nbr(k, Lmax, EX[q], EY[q], EZ[q])) : -1);
      }
    }
    // level-0 wet boundaries (the cells and neighbours Case::wet_bc reads)
    {
      const Level& m0 = c.lv[0];
      for (uint64_t k : m0.lat_f) {
        int fm = Case::faces_of(m0.kind.at(k));
        int act = 0, nact = 0;
        for (int ff = 0; ff < 6; ++ff)
          if ((fm & (1 << ff)) && (c.face_kind[ff] == BC_ZOUHE || c.face_kind[ff] == BC_OUTFLOW || c.face_kind[ff] == BC_NEE)) {
            act |= 1 << ff;
            ++nact;
          }
        if (!nact) continue;
        wet_cell.push_back(lv[0].idx.at(k));
        wet_act.push_back(act);
        int d[3] = {0, 0, 0};
        for (int ff = 0; ff < 6; ++ff)
          if (act & (1 << ff)) {
            int nn[3];
            Case::normal(ff, nn);
            for (int a = 0; a < 3; ++a) d[a] -= nn[a];
          }
        int64_t kin = at(0, c.nbr(k, 0, d[0], d[1], d[2]));
        if (kin < 0) return 2;
        wet_in.push_back(kin);
      }
      nee_cells.assign(6, {});
      nee_nb.assign(6, {});
      Case& cc = const_cast<Case&>(c);
      for (int face = 0; face < 6; ++face) {
        if (c.face_kind[face] != BC_RESERVOIR) continue;
        for (uint64_t k : cc.nee_cells(face)) {
          nee_cells[face].push_back(lv[0].idx.at(k));
          int fm = Case::faces_of(m0.kind.at(k));
          for (int q = 0; q < Q; ++q) {
            bool ok = true;
            int e[3] = {EX[q], EY[q], EZ[q]};
            for (int ff = 0; ff < 6; ++ff)
              if (fm & (1 << ff)) {
                int nn[3];
                Case::normal(ff, nn);
                if (e[0] * nn[0] + e[1] * nn[1] + e[2] * nn[2] > 0) ok = false;
              }
            int64_t nb = ok ? at(0, c.nbr(k, 0, EX[q], EY[q], EZ[q])) : -2;
            if (ok && nb < 0) return 3;  // .at() would throw
            nee_nb[face].push_back(nb);
          }
        }
      }
    }
    // markers: cell indices and the band's (marker, weight) lists in ascending (cell key, marker) order
    if (c.n_mk) {
      mk_idx.resize(27 * c.n_mk);
      std::vector<std::tuple<uint64_t, int64_t, double>> pairs;
      for (int64_t m = 0; m < c.n_mk; ++m)
        for (int j = 0; j < 27; ++j) {
          double w = c.mk_w[m * 27 + j];
          mk_idx[m * 27 + j] = w == 0.0 ? -1 : lv[Lmax].idx.at(c.mk_keys[m * 27 + j]);
          if (w != 0.0) pairs.emplace_back(c.mk_keys[m * 27 + j], m, w);
        }
      std::sort(pairs.begin(), pairs.end(), [](auto& a, auto& b) {
        return std::get<0>(a) != std::get<0>(b) ? std::get<0>(a) < std::get<0>(b) : std::get<1>(a) < std::get<1>(b);
      });
      for (size_t i = 0; i < pairs.size(); ++i) {
        if (i == 0 || std::get<0>(pairs[i]) != std::get<0>(pairs[i - 1])) {
          band_cell.push_back(lv[Lmax].idx.at(std::get<0>(pairs[i])));
          band_ptr.push_back((int64_t)i);
        }
        band_mk.push_back(std::get<1>(pairs[i]));
        band_w.push_back(std::get<2>(pairs[i]));
      }
      band_ptr.push_back((int64_t)pairs.size());
    }
    return 0;
  }

  void load(const Lev& d, int64_t i, double* out, bool col) const {
    const double* a = col ? d.fcol.data() : d.f.data();
    for (int q = 0; q < Q; ++q) out[q] = a[q * d.n + i];
  }

  void calc_macros(int L) {
    Lev& d = lv[L];
    const int64_t n = d.n;
    auto one = [&](int64_t i, bool vold) {
      double f[27], u[3];
      if (vold)
        for (int a = 0; a < 3; ++a) d.v_old[a * n + i] = d.u[a * n + i];
      load(d, i, f, false);
      macros(Q, f, d.rho[i], u);
      for (int a = 0; a < 3; ++a) d.u[a * n + i] = u[a];
    };
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < d.lat_f.size(); ++j) one(d.lat_f[j], true);
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < d.f2c.size(); ++j) one(d.f2c[j], false);
  }

  void collide(int L) {
    Lev& d = lv[L];
    const int64_t n = d.n;
    const double tL = src->tau[L];
    const int coll = src->collision;
    const double lesC = src->les_C;
    auto one = [&](int64_t i) {
      double f[27], fe[27], out[27];
      load(d, i, f, false);
      double rho = d.rho[i];
      double u[3] = {d.u[i], d.u[n + i], d.u[2 * n + i]};
      double t = tL;
      if (lesC > 0) {
        feq_c(Q, rho, u, fe);
        t = smag_tau(t, lesC, stress_norm(Q, f, fe));
      }
      if (coll == 0) {
        feq_c(Q, rho, u, fe);
        for (int q = 0; q < Q; ++q) out[q] = f[q] - (f[q] - fe[q]) / t;
      } else {
        cumulant_collide(f, rho, u, t, out);
      }
      for (int q = 0; q < Q; ++q) d.fcol[q * n + i] = out[q];
    };
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < d.lat_f.size(); ++j) one(d.lat_f[j]);
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < d.f2c.size(); ++j) one(d.f2c[j]);
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < d.c2f.size(); ++j)
      for (int q = 0; q < Q; ++q) d.fcol[q * n + d.c2f[j]] = d.f[q * n + d.c2f[j]];
  }

  void stream(int L) {
    Lev& d = lv[L];
    const int64_t n = d.n;
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < d.strm_cells.size(); ++j) {
      int64_t i = d.strm_cells[j];
      d.f[i] = d.fcol[i];
      for (int q = 1; q < Q; ++q) {
        int64_t s = d.strm_src[j * Q + q];
        if (s >= 0) d.f[q * n + i] = d.fcol[q * n + s];
      }
    }
  }

  void explode(int fine) {
    const int coarse = fine - 1;
    const int opt = src->options;
    const double* tau = src->tau.data();
    double s = (opt & 1) ? tau[fine] / (2.0 * tau[coarse]) : (2.0 * tau[fine]) / tau[coarse];
    Lev& dc = lv[coarse];
    Lev& df = lv[fine];
    const double rho0 = src->rho0;
    auto run = [&](const std::vector<int64_t>& par, const std::vector<int64_t>& chl) {
#pragma omp parallel for schedule(static)
      for (size_t p = 0; p < par.size(); ++p) {
        double cf[27], fe[27], u[3], rho;
        load(dc, par[p], cf, true);
        macros(Q, cf, rho, u);
        feq_hl(Q, rho, u, rho0, fe);
        for (int j = 0; j < 8; ++j) {
          int64_t c = chl[8 * p + j];
          if (c < 0) continue;
          for (int q = 0; q < Q; ++q) df.fcol[q * df.n + c] = fe[q] + s * (cf[q] - fe[q]);
        }
      }
    };
    run(df.ex_parent, df.ex_child);
    run(df.gx_parent, df.gx_child);
  }

  void coalesce(int coarse) {
    const int fine = coarse + 1;
    const double* tau = src->tau.data();
    double s = (src->options & 1) ? (2.0 * tau[coarse]) / tau[fine] : tau[coarse] / (2.0 * tau[fine]);
    Lev& dc = lv[coarse];
    Lev& df = lv[fine];
    const double rho0 = src->rho0;
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < dc.co_parent.size(); ++p) {
      double avg[27], fe[27], u[3], rho;
      for (int q = 0; q < Q; ++q) {
        avg[q] = 0;
        for (int j = 0; j < 8; ++j) avg[q] += df.f[q * df.n + dc.co_child[8 * p + j]];
        avg[q] /= 8.0;
      }
      macros(Q, avg, rho, u);
      feq_hl(Q, rho, u, rho0, fe);
      for (int q = 0; q < Q; ++q) dc.f[q * dc.n + dc.co_parent[p]] = fe[q] + s * (avg[q] - fe[q]);
    }
  }

  void bc_after_stream() {
    Lev& d = lv[0];
    const int64_t n = d.n;
    bool any_res = false;
    for (int face = 0; face < 6; ++face) any_res |= src->face_kind[face] == BC_RESERVOIR;
    std::vector<double> snap;
    if (any_res && !src->nee_seq) snap = d.f;
    for (int face = 0; face < 6; ++face) {
      if (src->face_kind[face] != BC_RESERVOIR) continue;
      const double* ub = src->face_u[face];
      const auto& cells = nee_cells[face];
      const auto& nbs = nee_nb[face];
      for (size_t j = 0; j < cells.size(); ++j) {
        int64_t i = cells[j];
        for (int q = 0; q < Q; ++q) {
          int64_t nb = nbs[j * Q + q];
          if (nb == -2) continue;
          double un3[3] = {d.u[nb], d.u[n + nb], d.u[2 * n + nb]};
          double fw[27], fn[27];
          feq_c(Q, d.rho[i], ub, fw);
          feq_c(Q, d.rho[nb], un3, fn);
          double fnb = src->nee_seq ? d.f[q * n + nb] : snap[q * n + nb];
          d.f[q * n + i] = fw[q] + fnb - fn[q];
        }
      }
    }
    // wet-node faces: every update reads post-stream f of cells no rule modifies
    std::vector<double> upd(wet_cell.size() * Q);
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < wet_cell.size(); ++j) {
      int64_t i = wet_cell[j];
      int act = wet_act[j];
      double f[27];
      load(d, i, f, false);
      int nact = __builtin_popcount(act);
      if (nact == 1) {
        int face = __builtin_ctz(act);
        int nn[3];
        Case::normal(face, nn);
        auto en = [&](int q) { return EX[q] * nn[0] + EY[q] * nn[1] + EZ[q] * nn[2]; };
        int64_t kin = wet_in[j];
        if (src->face_kind[face] == BC_NEE) {
          double fn[27], un[3], rn, fw[27], fq[27];
          load(d, kin, fn, false);
          macros(Q, fn, rn, un);
          feq_c(Q, rn, src->face_u[face], fw);
          feq_c(Q, rn, un, fq);
          for (int q = 0; q < Q; ++q) f[q] = fw[q] + fn[q] - fq[q];
        } else if (src->face_kind[face] == BC_OUTFLOW) {
          for (int q = 0; q < Q; ++q)
            if (en(q) == 1) f[q] = d.f[q * n + kin];
        } else {
          const double* uw = src->face_u[face];
          double S0 = 0, Sin = 0;
          for (int q = 0; q < Q; ++q) {
            if (en(q) == 0) S0 += f[q];
            else if (en(q) == -1) Sin += f[q];
          }
          double un = uw[0] * nn[0] + uw[1] * nn[1] + uw[2] * nn[2];
          double rho = (S0 + 2 * Sin) / (1 - un);
          int ax = face / 2;
          int t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
          if (t1 > t2) std::swap(t1, t2);
          double Nt[3] = {0, 0, 0};
          for (int t : {t1, t2}) {
            double sacc = 0;
            for (int q = 0; q < Q; ++q)
              if (en(q) == 0) {
                int e[3] = {EX[q], EY[q], EZ[q]};
                sacc += f[q] * e[t];
              }
            Nt[t] = 0.5 * sacc - rho * uw[t] / 3.0;
          }
          double g[27];
          for (int q = 0; q < Q; ++q) g[q] = f[q];
          for (int q = 0; q < Q; ++q) {
            if (en(q) != 1) continue;
            int e[3] = {EX[q], EY[q], EZ[q]};
            double w = weight(Q, q);
            double eu = e[0] * uw[0] + e[1] * uw[1] + e[2] * uw[2];
            double tr = e[t1] * Nt[t1] + e[t2] * Nt[t2];
            g[q] = f[inv(q)] + 6 * w * rho * eu - 36 * w * tr;
          }
          for (int q = 0; q < Q; ++q) f[q] = g[q];
        }
      } else {
        const double* uw = nullptr;
        for (int ff = 0; ff < 6; ++ff)
          if ((act & (1 << ff)) && (src->face_kind[ff] == BC_ZOUHE || src->face_kind[ff] == BC_NEE)) uw = src->face_u[ff];
        double fn[27], un[3], rn, fw[27], fq[27];
        load(d, wet_in[j], fn, false);
        macros(Q, fn, rn, un);
        feq_c(Q, rn, uw ? uw : un, fw);
        feq_c(Q, rn, un, fq);
        for (int q = 0; q < Q; ++q) f[q] = fw[q] + fn[q] - fq[q];
      }
      for (int q = 0; q < Q; ++q) upd[j * Q + q] = f[q];
    }
    for (size_t j = 0; j < wet_cell.size(); ++j)
      for (int q = 0; q < Q; ++q) d.f[q * n + wet_cell[j]] = upd[j * Q + q];
  }

  void ib_treat(double* F) {
    Lev& d = lv[Lmax];
    const int64_t n = d.n;
    const int ib = src->ib;
    const size_t nt = tr_cell.size();
    std::vector<double> mterm(nt * Q, 0.0);
    std::vector<uint8_t> link(nt * Q, 0);
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < nt; ++j) {
      int64_t i = tr_cell[j];
      const auto& dist_of = tr_dis[j];
      double rb = d.rho[i];
      d.f[i] = d.fcol[i];
      for (int q = 1; q < Q; ++q) {
        double dist = dist_of[inv(q) - 1];
        if (dist != -1.) {
          double finc = d.fcol[inv(q) * n + i];
          int64_t sidx = tr_inc_src[j * Q + q];
          double fs = sidx >= 0 ? d.fcol[q * n + sidx] : finc;
          double fnew;
          if (ib == 1) {
            if (dist < 0.5) {
              fnew = (1.0 - 2.0 * dist) * finc + 2.0 * dist * fs;
            } else {
              double fe = ib_weight(inv(q)) * rb;
              double i2d = 1.0 / (2.0 * dist);
              fnew = i2d * fs + (1.0 - i2d) * fe;
            }
          } else {
            fnew = (1.0 - dist) * finc + dist * fs;
          }
          d.f[q * n + i] = fnew;
          mterm[j * Q + q] = finc + fnew;
          link[j * Q + q] = 1;
        } else {
          int64_t sidx = tr_pull_src[j * Q + q];
          d.f[q * n + i] = sidx < 0 ? d.fcol[inv(q) * n + i] : d.fcol[q * n + sidx];
        }
      }
      double ra = 0;
      for (int q = 0; q < Q; ++q) ra += d.f[q * n + i];
      double& f0 = d.f[i];
      if (ra < 0.3) f0 += 1.0 * (0.3 - ra);
      else if (ra < 0.5) f0 += 0.5 * (0.5 - ra);
      else if (ra > 3.0) f0 += 1.0 * (3.0 - ra);
      else if (ra > 2.0) f0 += 0.5 * (2.0 - ra);
    }
    double fx = 0, fy = 0, fz = 0;
    for (size_t j = 0; j < nt; ++j)
      for (int q = 1; q < Q; ++q)
        if (link[j * Q + q]) {
          double m = mterm[j * Q + q];
          fx += -m * EX[q];
          fy += -m * EY[q];
          fz += -m * EZ[q];
        }
    force_log.push_back({fx, fy, fz});
    if (F) { F[0] = fx; F[1] = fy; F[2] = fz; }
    const int sweeps = src->mdf_sweeps;
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < nt; ++j) {
      int64_t i = tr_cell[j];
      const auto& dist_of = tr_dis[j];
      for (int it = 0; it < sweeps; ++it) {
        double rho = 0, u[3] = {0, 0, 0};
        for (int q = 0; q < Q; ++q) {
          double fq = d.f[q * n + i];
          rho += fq;
          u[0] += fq * EX[q];
          u[1] += fq * EY[q];
          u[2] += fq * EZ[q];
        }
        if (rho <= 0.) continue;
        u[0] /= rho; u[1] /= rho; u[2] /= rho;
        double um = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        if (um < 1e-10) continue;
        double dM[3] = {-0.5 * rho * u[0], -0.5 * rho * u[1], -0.5 * rho * u[2]};
        double adj = 0;
        for (int q = 1; q < Q; ++q) {
          if (dist_of[inv(q) - 1] == -1.) continue;
          double df = ib_weight(q) * 3.0 * (EX[q] * dM[0] + EY[q] * dM[1] + EZ[q] * dM[2]);
          d.f[q * n + i] += df;
          adj -= df;
        }
        d.f[i] += adj;
      }
    }
  }

  void marker_correct() {
    const Case& c = *src;
    if (!c.n_mk) return;
    Lev& d = lv[Lmax];
    const int64_t n = d.n;
    std::vector<std::array<double, 3>> F(c.n_mk);
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < c.n_mk; ++m) {
      double rk = 0, uk[3] = {0, 0, 0};
      for (int j = 0; j < 27; ++j) {
        const double w = c.mk_w[m * 27 + j];
        if (w == 0.0) continue;
        double f[27], rho, u[3];
        load(d, mk_idx[m * 27 + j], f, true);
        macros(Q, f, rho, u);
        rk += w * rho;
        for (int a = 0; a < 3; ++a) uk[a] += w * u[a];
      }
      for (int a = 0; a < 3; ++a) F[m][a] = c.mk_dv[m] * rk * (c.mk_ub[a] - uk[a]);
    }
    const size_t nb = band_cell.size();
    std::vector<double> newf(nb * Q);
#pragma omp parallel for schedule(static)
    for (size_t b = 0; b < nb; ++b) {
      double dm[3] = {0, 0, 0};
      for (int64_t p = band_ptr[b]; p < band_ptr[b + 1]; ++p)
        for (int a = 0; a < 3; ++a) dm[a] += band_w[p] * F[band_mk[p]][a];
      double f[27], rho, u[3], u2[3], fe1[27], fe2[27];
      load(d, band_cell[b], f, true);
      macros(Q, f, rho, u);
      for (int a = 0; a < 3; ++a) u2[a] = u[a] + dm[a] / rho;
      feq_c(Q, rho, u2, fe2);
      feq_c(Q, rho, u, fe1);
      for (int q = 0; q < Q; ++q) newf[b * Q + q] = f[q] + (fe2[q] - fe1[q]);
    }
    for (size_t b = 0; b < nb; ++b)
      for (int q = 0; q < Q; ++q) d.fcol[q * n + band_cell[b]] = newf[b * Q + q];
    double fx = 0, fy = 0, fz = 0;
    for (int64_t m = 0; m < c.n_mk; ++m) {
      fx += F[m][0];
      fy += F[m][1];
      fz += F[m][2];
    }
    force_log.push_back({fx, fy, fz});
  }

  void kernel(int L) {
    const int opt = src->options;
    const bool ib = src->ib != 0;
    calc_macros(L);
    collide(L);
    if (L == Lmax) marker_correct();
    if (L < Lmax) kernel(L + 1);
    if (L > 0) explode(L);
    stream(L);
    if (L == 0) bc_after_stream();
    else if (L == Lmax && ib) ib_treat(nullptr);
    if (L < Lmax) coalesce(L);
    if (L == 0) return;
    calc_macros(L);
    collide(L);
    if (L == Lmax) marker_correct();
    if (L < Lmax) kernel(L + 1);
    if (opt & 2) explode(L);
    stream(L);
    if (L == 0) bc_after_stream();
    else if (L == Lmax && ib) ib_treat(nullptr);
    if (L < Lmax) coalesce(L);
  }
  void step() {
    kernel(0);
    for (int L = 0; L <= Lmax; ++L) calc_macros(L);
  }
};
}  // namespace

extern "C" {

void* mto_create(int Q, int Lmax, double tau0, int collision, double les_C, double rho0, const double* u0,
                 const int* dims0, const int* face_kind, const double* face_u, int ib, int nee_sequential) {
  Case* c = new Case;
  c->Q = Q; c->Lmax = Lmax; c->tau0 = tau0; c->collision = collision; c->les_C = les_C; c->rho0 = rho0;
  for (int a = 0; a < 3; ++a) { c->u0[a] = u0[a]; c->dims0[a] = dims0[a]; }
  for (int f = 0; f < 6; ++f) {
    c->face_kind[f] = face_kind[f];
    for (int a = 0; a < 3; ++a) c->face_u[f][a] = face_u[3 * f + a];
  }
  c->ib = ib; c->nee_seq = nee_sequential;
  c->mt.init(Lmax);
  c->lv.resize(Lmax + 1);
  return c;
}
void mto_destroy(void* h) { delete (Case*)h; }
void mto_set_options(void* h, int options) { ((Case*)h)->options = options; }
int mto_set_mdf_sweeps(void* h, int n) {
  ((Case*)h)->mdf_sweeps = n;
  return 0;
}
// overrides tau_L after mto_init (tests that run the reference's compiled path
// use the relaxation times its setUnit computed, LBM_Manager.cpp:763-844)
int mto_set_tau(void* h, int L, double tau) {
  Case* c = (Case*)h;
  if (L < 0 || L > c->Lmax || (int)c->tau.size() <= L) return 1;
  c->tau[L] = tau;
  return 0;
}

int mto_add_level(void* h, int L, const uint64_t* keys, const uint16_t* kinds, int64_t n) {
  Case* c = (Case*)h;
  Level& l = c->lv[L];
  std::vector<std::pair<uint64_t, uint16_t>> v(n);
  for (int64_t i = 0; i < n; ++i) v[i] = {keys[i], kinds[i]};
  std::sort(v.begin(), v.end());
  for (auto& p : v) {
    l.kind[p.first] = p.second;
    if (p.second & K_FLUID) { l.lat_f.push_back(p.first); l.lat_f_set.insert(p.first); }
    if (p.second & K_F2C) l.f2c.push_back(p.first);
    if (p.second & K_C2F) l.c2f.push_back(p.first);
    if (p.second & K_GHOST_C2F) { l.ghost_c2f.push_back(p.first); l.ghost_c2f_set.insert(p.first); }
    if (p.second & K_GHOST_F2C) l.ghost_f2c.push_back(p.first);
    if (p.second & K_SOLID) l.solid.push_back(p.first);
    if (p.second & K_BCRES) l.bcres.push_back(p.first);
  }
  return 0;
}
int mto_set_markers(void* h, int64_t n, const uint64_t* keys, const double* w, const double* dv, const double* ub) {
  Case* c = (Case*)h;
  c->n_mk = n;
  c->mk_keys.assign(keys, keys + 27 * n);
  c->mk_w.assign(w, w + 27 * n);
  c->mk_dv.assign(dv, dv + n);
  for (int a = 0; a < 3; ++a) c->mk_ub[a] = ub[a];
  return 0;
}

int mto_set_dis(void* h, const uint64_t* keys, const double* vals, int64_t n) {
  Case* c = (Case*)h;
  int m = c->Q - 1;
  for (int64_t i = 0; i < n; ++i) {
    std::array<double, 26> a;
    a.fill(-1.0);
    for (int j = 0; j < m; ++j) a[j] = vals[i * m + j];
    c->dis[keys[i]] = a;
  }
  return 0;
}
int mto_init(void* h) { ((Case*)h)->init(); return 0; }
int mto_step(void* h, int n) {
  Case* c = (Case*)h;
  for (int i = 0; i < n; ++i) c->step();
  return 0;
}
// what: 0 f [Q][n], 1 fcol [Q][n], 2 rho [n], 3 u [3][n], 4 v_old [3][n]
int mto_get(void* h, int L, int what, const uint64_t* keys, int64_t n, double* out) {
  Case* c = (Case*)h;
  Level& l = c->lv[L];
  for (int64_t i = 0; i < n; ++i) {
    uint64_t k = keys[i];
    if (what <= 1) {
      auto& m = what == 0 ? l.f : l.fcol;
      for (int q = 0; q < c->Q; ++q) out[q * n + i] = m[q].at(k);
    } else if (what == 2) {
      auto it = l.rho.find(k);
      out[i] = it == l.rho.end() ? NAN : it->second;
    } else {
      auto& m = what == 3 ? l.u : l.v_old;
      auto it = m.find(k);
      for (int a = 0; a < 3; ++a) out[a * n + i] = it == m.end() ? NAN : it->second[a];
    }
  }
  return 0;
}
int mto_set(void* h, int L, int what, const uint64_t* keys, int64_t n, const double* in) {
  Case* c = (Case*)h;
  Level& l = c->lv[L];
  for (int64_t i = 0; i < n; ++i) {
    uint64_t k = keys[i];
    if (what <= 1) {
      auto& m = what == 0 ? l.f : l.fcol;
      for (int q = 0; q < c->Q; ++q) m[q].at(k) = in[q * n + i];
    } else if (what == 2) {
      l.rho[k] = in[i];
    } else {
      auto& m = what == 3 ? l.u : l.v_old;
      m[k] = {in[i], in[n + i], in[2 * n + i]};
    }
  }
  return 0;
}
int mto_macros(void* h, int L) { ((Case*)h)->calc_macros(L); return 0; }
int mto_collide(void* h, int L) { ((Case*)h)->collide(L); return 0; }
int mto_stream(void* h, int L) { ((Case*)h)->stream(L); return 0; }
int mto_explode(void* h, int L) { ((Case*)h)->explode(L); return 0; }
int mto_coalesce(void* h, int L) { ((Case*)h)->coalesce(L); return 0; }
int mto_bc_after_stream(void* h) { ((Case*)h)->bc_after_stream(); return 0; }
int mto_ib_treat(void* h, double* F) {
  try {
    ((Case*)h)->ib_treat(F);
  } catch (const std::out_of_range&) {
    return 1;  // a neighbour the reference's .at() would not find
  }
  return 0;
}
int mto_marker_correct(void* h) { ((Case*)h)->marker_correct(); return 0; }
int64_t mto_force_log(void* h, double* out, int64_t max) {
  Case* c = (Case*)h;
  int64_t n = (int64_t)c->force_log.size();
  for (int64_t i = 0; i < n && i < max; ++i)
    for (int a = 0; a < 3; ++a) out[3 * i + a] = c->force_log[i][a];
  return n;
}


// ---- dense mode (same bits as the map mode, sorted arrays, threads)
// fresh = 1: from the case's level sets alone (mto_init need not have run; the
// case must outlive the dense copy); fresh = 0: from the map mode's current state
void* mto_dense_create(void* h, int fresh) {
  Case* c = (Case*)h;
  if (fresh) c->init_tau();
  Dense* d = new Dense;
  if (d->build(*c, fresh != 0)) {
    delete d;
    return nullptr;
  }
  return d;
}
void mto_dense_destroy(void* d) { delete (Dense*)d; }
int mto_dense_step(void* h, int n) {
  Dense* d = (Dense*)h;
  for (int i = 0; i < n; ++i) d->step();
  return 0;
}
int mto_dense_get(void* h, int L, int what, const uint64_t* keys, int64_t n, double* out) {
  Dense* d = (Dense*)h;
  Dense::Lev& l = d->lv[L];
  for (int64_t j = 0; j < n; ++j) {
    int64_t i = d->at(L, keys[j]);
    if (i < 0) return 1;
    if (what <= 1) {
      const auto& a = what == 0 ? l.f : l.fcol;
      for (int q = 0; q < d->Q; ++q) out[q * n + j] = a[q * l.n + i];
    } else if (what == 2) {
      out[j] = l.has_macro[i] ? l.rho[i] : NAN;
    } else {
      const auto& a = what == 3 ? l.u : l.v_old;
      for (int c = 0; c < 3; ++c) out[c * n + j] = l.has_macro[i] ? a[c * l.n + i] : NAN;
    }
  }
  return 0;
}
int mto_dense_set(void* h, int L, int what, const uint64_t* keys, int64_t n, const double* in) {
  Dense* d = (Dense*)h;
  Dense::Lev& l = d->lv[L];
  for (int64_t j = 0; j < n; ++j) {
    int64_t i = d->at(L, keys[j]);
    if (i < 0) return 1;
    if (what <= 1) {
      auto& a = what == 0 ? l.f : l.fcol;
      for (int q = 0; q < d->Q; ++q) a[q * l.n + i] = in[q * n + j];
    } else if (what == 2) {
      l.rho[i] = in[j];
      l.has_macro[i] = 1;
    } else {
      auto& a = what == 3 ? l.u : l.v_old;
      for (int c = 0; c < 3; ++c) a[c * l.n + i] = in[c * n + j];
    }
  }
  return 0;
}
int64_t mto_dense_force_log(void* h, double* out, int64_t max) {
  Dense* d = (Dense*)h;
  int64_t n = (int64_t)d->force_log.size();
  for (int64_t i = 0; i < n && i < max; ++i)
    for (int a = 0; a < 3; ++a) out[3 * i + a] = d->force_log[i][a];
  return n;
}
// infoPrint's velocity diagnostics (LBM_Manager.cpp:914-962) and diagL2MaxF /
// diagL2RhoMin (:854-912) of the current state, cells in ascending key order
// (ties keep the lowest key).  vals: 0 velocity error, 1..8 max |v| per level,
// 9..16 cells with |v| > 1 per level, 17 max |f|, 18 max |fcol|, 19 min sum f,
// 20 min sum fcol, 21 n sum f <= 0, 22 n sum fcol <= 0, 23 n sum f < 0.5,
// 24 the worst cell is a dis cell; keys: max |f|, max |fcol|, worst; qs: their q.
int mto_diagnostics(void* h, double* vals, uint64_t* keys, int* qs) {
  Case* c = (Case*)h;
  for (int i = 0; i < 25; ++i) vals[i] = 0;
  double gu = 0, gr = 0;
  for (int L = 0; L <= c->Lmax && L < 8; ++L) {
    Level& l = c->lv[L];
    double mx = 0, nb = 0;
    auto vel = [&](uint64_t k) {
      auto v = l.u.at(k);
      double m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      if (m > mx) mx = m;
      if (m > 1.0) nb += 1;
    };
    for (uint64_t k : l.lat_f) {
      vel(k);
      auto v = l.u.at(k);
      auto o = l.v_old.at(k);
      gu += (v[0] - o[0]) * (v[0] - o[0]) + (v[1] - o[1]) * (v[1] - o[1]) + (v[2] - o[2]) * (v[2] - o[2]);
      gr += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    }
    for (uint64_t k : l.f2c) vel(k);
    vals[1 + L] = mx;
    vals[9 + L] = nb;
  }
  vals[0] = std::sqrt(gu) / std::sqrt(gr + 1e-16);
  Level& l = c->lv[c->Lmax];
  double maxf = 0, maxfc = 0, minf = 1e300, minfc = 1e300, nnf = 0, nnfc = 0, nlow = 0;
  keys[0] = keys[1] = keys[2] = 0;
  qs[0] = qs[1] = 0;
  for (uint64_t k : l.lat_f) {
    double rf = 0, rfc = 0;
    for (int q = 0; q < c->Q; ++q) {
      double fv = l.f[q].at(k), fc = l.fcol[q].at(k);
      rf += fv;
      rfc += fc;
      if (std::fabs(fv) > maxf) { maxf = std::fabs(fv); keys[0] = k; qs[0] = q; }
      if (std::fabs(fc) > maxfc) { maxfc = std::fabs(fc); keys[1] = k; qs[1] = q; }
    }
    if (rf < minf) { minf = rf; keys[2] = k; }
    if (rfc < minfc) minfc = rfc;
    if (rf <= 0.0) nnf += 1;
    if (rfc <= 0.0) nnfc += 1;
    if (rf < 0.5) nlow += 1;
  }
  vals[17] = maxf;
  vals[18] = maxfc;
  vals[19] = minf;
  vals[20] = minfc;
  vals[21] = nnf;
  vals[22] = nnfc;
  vals[23] = nlow;
  vals[24] = c->dis.count(keys[2]) ? 1 : 0;
  return 0;
}

// ---- primitives for closed-form tests
void mto_feq(int Q, double rho, const double* u, double* out) { feq_c(Q, rho, u, out); }
void mto_feq_heluo(int Q, double rho, const double* u, double rho0, double* out) { feq_hl(Q, rho, u, rho0, out); }
void mto_macros_cell(int Q, const double* f, double* rho, double* u) { macros(Q, f, *rho, u); }
void mto_cumulant(const double* f, double rho, const double* u, double tau, double* out) {
  cumulant_collide(f, rho, u, tau, out);
}
void mto_central_fwd(const double* f, const double* u, double* k) { central_fwd(f, u, *(M3*)k); }
void mto_central_bwd(const double* k, const double* u, double* f) { central_bwd(*(const M3*)k, u, f); }
void mto_to_cumulants(const double* k, double rho, double* C) { to_cumulants(*(const M3*)k, rho, *(M3*)C); }
void mto_from_cumulants(const double* C, double rho, double* k) { from_cumulants(*(const M3*)C, rho, *(M3*)k); }  // k in/out
double mto_stress_norm(int Q, const double* f, const double* feq) { return stress_norm(Q, f, feq); }
double mto_smag_tau(double tau, double C, double qn) { return smag_tau(tau, C, qn); }
uint64_t mto_find_neighbor(uint64_t key, int L, int Lmax, int dx, int dy, int dz) {
  Morton m;
  m.init(Lmax);
  return m.neighbor(key, L, dx, dy, dz);
}
}
