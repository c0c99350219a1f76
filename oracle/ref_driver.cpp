// ref_driver.cpp -- runs the reference's own compiled hot path on a level-set
// layout handed in from outside.
//
// TEST INFRASTRUCTURE ONLY.  This file is linked, by oracle/Makefile.ref, with
// objects compiled from the reference's sources where they lie (every
// translation unit of its Makefile except HDF5.cpp, whose header is absent, and
// Solid_Manager.cpp, which does not compile: an `#endif` without `#if` at
// Solid_Manager.cpp:18).  The only symbol left unresolved is
// Solid_Manager::initial, which nothing here calls (the library is loaded with
// lazy binding).  The result, oracle/_ref/libref_amr.so, is the checker the
// restated oracle (oracle/mt_oracle.cpp) is compared with in tests/test_ref.py.
//
// The reference is compiled as it ships: D3Q19 (C_Q 19), C_MAP_TYPE 1, the
// `sphere` configuration with C_max_level = 2, LBGK collision, ABPatternPull
// streaming, NEE reservoir faces and SinglePointInterpolation (FH) on the finest
// level (LBM_Manager.cpp:32-95).  This driver only:
//   * fills the level sets of the global Lat_Manager (header/Lat_Manager.hpp:245-316),
//   * initialises the Morton helpers the way Grid_Manager::initial does
//     (Grid_Manager.cpp:43-48),
//   * constructs one LBM_Manager (its constructor builds the collision, stream,
//     solid-boundary and BC models and allocates the per-level maps),
//   * reads and writes the per-level maps, and calls LBMkernel(0),
//     calculateMacros(), and the single seams.
// Private members are reached by compiling this file (only this one) with
// `private`/`protected` defined to `public`; the reference objects are built
// unmodified.
#include <algorithm>
#include <array>
#include <bitset>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#define private public
#define protected public
#include "General.h"
#include "Constants.h"
#include "user.h"
#include "Morton_Assist.h"
#include "Lat_Manager.hpp"
#include "LBM_Manager.hpp"
#include "lbm_kernel/Collision_Manager.hpp"
#include "lbm_kernel/Stream_Manager.hpp"
#include "ComplexBoundary_Manager.hpp"
#include "BC/D3Q19_BC_Manager.hpp"
#include "BC/NonEquilibrumExtrapolation.hpp"
#include "tecplot.h"
#undef private
#undef protected

extern Lat_Manager lat_manager;  // Obj_Manager.cpp:26

namespace {
Morton_Assist g_morton;
LBM_Manager* g_lbm = nullptr;

D_morton key_of(uint64_t k) { return D_morton(k); }

D_mapLat* lat_set(int set, int L) {
  Lat_Manager* lm = &lat_manager;
  switch (set) {
    case 0: return &lm->lat_f.at(L);
    case 1: return &lm->lat_overlap_F2C.at(L);
    case 2: return &lm->lat_overlap_C2F.at(L);
    case 3: return &lm->ghost_overlap_C2F.at(L);
    case 4: return &lm->ghost_overlap_F2C.at(L);
    case 5: return &lm->lat_sf;
    default: return nullptr;
  }
}
}  // namespace

extern "C" {

int ref_q() { return C_Q; }
int ref_lmax() { return (int)C_max_level; }

// set: 0 lat_f, 1 lat_overlap_F2C, 2 lat_overlap_C2F, 3 ghost_overlap_C2F,
// 4 ghost_overlap_F2C, 5 lat_sf (level ignored), 6 + b: lat_bc[b] (level 0)
int ref_add(int set, int L, const uint64_t* keys, int64_t n) {
  if (L < 0 || L > (int)C_max_level) return 1;
  if (set >= 6) {
    int b = set - 6;
    if (b >= 26) return 1;
    for (int64_t i = 0; i < n; ++i) lat_manager.lat_bc.at(b).insert(key_of(keys[i]));
    return 0;
  }
  D_mapLat* m = lat_set(set, L);
  if (!m) return 1;
  for (int64_t i = 0; i < n; ++i) m->insert(std::make_pair(key_of(keys[i]), Cell(0)));
  return 0;
}

// dis: per finest cell C_Q-1 link fractions (or -1), index q-1 for direction q
int ref_add_dis(const uint64_t* keys, const double* vals, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    std::array<D_real, C_Q - 1> a;
    for (int j = 0; j < C_Q - 1; ++j) a[j] = vals[i * (C_Q - 1) + j];
    lat_manager.dis[key_of(keys[i])] = a;
  }
  return 0;
}

int ref_set_dx(const double* dx) {
  for (int L = 0; L <= (int)C_max_level; ++L) lat_manager.dx.at(L) = dx[L];
  return 0;
}

// ib: 1 FH, 2 distance-weighted, 3 UniformBoundary.  May be called once per process: the
// collision model is a function-local singleton that captures the first
// manager's maps and relaxation times (Collision_Manager.hpp:33-34, 57-61).
int ref_construct(int ib) {
  if (g_lbm) return 2;
  Lat_Manager::pointer_me = &lat_manager;
  Morton_Assist::pointer_me = &g_morton;
  g_morton.morton_initial();
  for (unsigned L = 0; L <= C_max_level; ++L) Morton_Assist::compute_ref(L);
  g_lbm = new LBM_Manager();
  if (ib == 2) {
    delete g_lbm->solidBCModel;
    g_lbm->solidBCModel = new SinglePointInterpolation(g_lbm->user, IB_Method::DISTANCE_WEIGHTED);
  } else if (ib == 3) {
    // the alternative the reference keeps commented out (LBM_Manager.cpp:46)
    delete g_lbm->solidBCModel;
    g_lbm->solidBCModel = new UniformBoundary(g_lbm->user);
  }
  return 0;
}

int ref_tau(double* out) {
  if (!g_lbm) return 1;
  for (int L = 0; L <= (int)C_max_level; ++L) out[L] = g_lbm->tau[L];
  return 0;
}

// what: 0 f [Q][n], 1 fcol [Q][n], 2 rho [n], 3 u [3][n]; absent keys read NaN
int ref_get(int L, int what, const uint64_t* keys, int64_t n, double* out) {
  if (!g_lbm) return 1;
  User& u = g_lbm->user[L];
  for (int64_t i = 0; i < n; ++i) {
    D_morton k = key_of(keys[i]);
    if (what <= 1) {
      auto* m = what == 0 ? u.df.f : u.df.fcol;
      for (int q = 0; q < C_Q; ++q) {
        auto it = m[q].find(k);
        out[q * n + i] = it == m[q].end() ? NAN : it->second;
      }
    } else if (what == 2) {
      auto it = u.density.find(k);
      out[i] = it == u.density.end() ? NAN : it->second;
    } else {
      auto it = u.velocity.find(k);
      out[i] = it == u.velocity.end() ? NAN : it->second.x;
      out[n + i] = it == u.velocity.end() ? NAN : it->second.y;
      out[2 * n + i] = it == u.velocity.end() ? NAN : it->second.z;
    }
  }
  return 0;
}

// writes only keys the reference's maps already hold
int ref_set(int L, int what, const uint64_t* keys, int64_t n, const double* in) {
  if (!g_lbm) return 1;
  User& u = g_lbm->user[L];
  for (int64_t i = 0; i < n; ++i) {
    D_morton k = key_of(keys[i]);
    if (what <= 1) {
      auto* m = what == 0 ? u.df.f : u.df.fcol;
      for (int q = 0; q < C_Q; ++q) {
        auto it = m[q].find(k);
        if (it != m[q].end()) it->second = in[q * n + i];
      }
    } else if (what == 2) {
      auto it = u.density.find(k);
      if (it != u.density.end()) it->second = in[i];
    } else {
      auto it = u.velocity.find(k);
      if (it != u.velocity.end()) it->second = D_uvw(in[i], in[n + i], in[2 * n + i]);
    }
  }
  return 0;
}

// n coarse steps of fluidSimulate's loop body without its prints and output
// (LBM_Simulator.cpp:11-17): LBMkernel(0); calculateMacros()
int ref_step(int n) {
  if (!g_lbm) return 1;
  for (int i = 0; i < n; ++i) {
    g_lbm->LBMkernel(0);
    g_lbm->calculateMacros();
  }
  return 0;
}

// single seams: 0 calculateMacros(L), 1 collide(L), 2 stream(L),
// 3 AMR_transDDF_explosion(L), 4 AMR_transDDF_coalescence(L),
// 5 the six AfterStream face strategies (W,E,S,N,Bot,Top), 6 treat()
int ref_call(int op, int L) {
  if (!g_lbm) return 1;
  switch (op) {
    case 0: g_lbm->calculateMacros(L); break;
    case 1: g_lbm->collideModel->collide(L); break;
    case 2: g_lbm->streamModel->stream(L); break;
    case 3: g_lbm->AMR_transDDF_explosion(L); break;
    case 4: g_lbm->AMR_transDDF_coalescence(L); break;
    case 5:
      for (const char* s : {"West", "East", "South", "North", "Bot", "Top"})
        g_lbm->bc_manager_->applyBoundaryCondition(s, "AfterStream");
      break;
    case 6: g_lbm->solidBCModel->treat(); break;
    default: return 2;
  }
  return 0;
}

// the wall velocity each face strategy holds, lattice units, W,E,S,N,Bot,Top
// (its constructor divides the given velocity by CF_U,
// header/BC/NonEquilibrumExtrapolation.hpp:36)
int ref_bc_velocity(double* out) {
  if (!g_lbm) return 1;
  const char* names[6] = {"West", "East", "South", "North", "Bot", "Top"};
  for (int i = 0; i < 6; ++i) {
    auto it = g_lbm->bc_manager_->boundaryConditionStrategies_.find(names[i]);
    if (it == g_lbm->bc_manager_->boundaryConditionStrategies_.end()) return 2;
    D3Q19_BC_Strategy* s = it->second.get();
    D_uvw v(NAN, NAN, NAN);
    if (auto* p = dynamic_cast<nonEquilibrumExtrapolation_West*>(s)) v = p->bc_velocity;
    if (auto* p = dynamic_cast<nonEquilibrumExtrapolation_East*>(s)) v = p->bc_velocity;
    if (auto* p = dynamic_cast<nonEquilibrumExtrapolation_South*>(s)) v = p->bc_velocity;
    if (auto* p = dynamic_cast<nonEquilibrumExtrapolation_North*>(s)) v = p->bc_velocity;
    if (auto* p = dynamic_cast<nonEquilibrumExtrapolation_Bot*>(s)) v = p->bc_velocity;
    if (auto* p = dynamic_cast<nonEquilibrumExtrapolation_Top*>(s)) v = p->bc_velocity;
    out[3 * i] = v.x;
    out[3 * i + 1] = v.y;
    out[3 * i + 2] = v.z;
  }
  return 0;
}

// the reference's printed diagnostics of the current state: 0 infoPrint(0)
// (LBM_Manager.cpp:914-981), 1 diagL2MaxF, 2 diagL2RhoMin (:854-912)
int ref_info(int which) {
  if (!g_lbm) return 1;
  if (which == 0) g_lbm->infoPrint(0);
  else if (which == 1) g_lbm->diagL2MaxF("ref");
  else if (which == 2) g_lbm->diagL2RhoMin("ref");
  else return 2;
  std::cout.flush();
  return 0;
}

// IO_TECPLOTH::write(0) (tecplot.cpp:16-177): output/sphereFlow_0.000000.dat
// under the current directory; *cf_u receives the velocity factor it applies
int ref_write_tecplot(double* cf_u) {
  if (!g_lbm) return 1;
  ensureOutputDir();
  IO_TECPLOTH t;
  t.write(0);
  *cf_u = g_lbm->CF_U;
  return 0;
}

// accumIBForce's running sums and count (LBM_Manager.cpp:846-852); reset=1 clears them
int ref_force(double* out, int reset) {
  if (!g_lbm) return 1;
  out[0] = g_lbm->m_ibFx_sum;
  out[1] = g_lbm->m_ibFy_sum;
  out[2] = g_lbm->m_ibFz_sum;
  out[3] = (double)g_lbm->m_ibForce_count;
  if (reset) {
    g_lbm->m_ibFx_sum = g_lbm->m_ibFy_sum = g_lbm->m_ibFz_sum = 0;
    g_lbm->m_ibForce_count = 0;
  }
  return 0;
}

// cell primitives (LBM_Manager.cpp:137-203, 983-1001)
void ref_feq(double rho, const double* u, double* out) {
  LBM_Manager::calculateFeq(rho, D_uvw(u[0], u[1], u[2]), out);
}
void ref_feq_heluo(double rho, const double* u, double* out) {
  LBM_Manager::calculateFeq_HeLuo(rho, D_uvw(u[0], u[1], u[2]), out);
}
void ref_macros_cell(const double* f, double* rho, double* u) {
  D_uvw v;
  LBM_Manager::calculateMacros(f, *rho, v);
  u[0] = v.x;
  u[1] = v.y;
  u[2] = v.z;
}
uint64_t ref_find_neighbor(uint64_t key, int L, int dx, int dy, int dz) {
  return Morton_Assist::find_neighbor(key_of(key), (uint)L, dx, dy, dz).to_ullong();
}
int ref_init_morton() {
  Morton_Assist::pointer_me = &g_morton;
  g_morton.morton_initial();
  for (unsigned L = 0; L <= C_max_level; ++L) Morton_Assist::compute_ref(L);
  return 0;
}
}
