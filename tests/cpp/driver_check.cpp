// driver_check.cpp -- drives a lid-driven cavity through the C++ host driver
// twice, once with the seam-by-seam LBMkernel recursion and once with the fused
// fluidSimulate schedule, and checks that both give the same populations.
// Usage: driver_check [n] [steps]; exit status 0 on agreement.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mt_lbm_manager.hpp"

static uint64_t spread(uint64_t v) {
  uint64_t r = 0;
  for (int i = 0; i < 21; ++i) r |= ((v >> i) & 1ull) << (3 * i);
  return r;
}

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 12;
  const int steps = argc > 2 ? std::atoi(argv[2]) : 20;
  mt_case_desc d = {};
  d.Q = 19;
  d.Lmax = 0;
  d.collision = MT_COLLIDE_LBGK;
  d.ib = MT_IB_NONE;
  d.precision = MT_FP64;
  d.dims0[0] = d.dims0[1] = d.dims0[2] = n;
  for (int f = 0; f < 6; ++f) d.face_kind[f] = MT_BC_ZOUHE;
  d.face_u[5][0] = -0.05;  // lid, reported frame
  d.tau0 = 0.6;
  d.rho0 = 1.0;
  std::vector<mt::LevelCells> levels(1);
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        uint16_t k = MT_KIND_FLUID;
        if (x == 0) k |= MT_FACE_W;
        if (x == n - 1) k |= MT_FACE_E;
        if (y == 0) k |= MT_FACE_S;
        if (y == n - 1) k |= MT_FACE_N;
        if (z == 0) k |= MT_FACE_B;
        if (z == n - 1) k |= MT_FACE_T;
        levels[0].keys.push_back(spread(x) | (spread(y) << 1) | (spread(z) << 2));
        levels[0].kinds.push_back(k);
      }
  try {
    mt::LBMManager a(d, levels, {}, {});
    mt::LBMManager b(d, levels, {}, {});
    for (int s = 0; s < steps; ++s) {
      a.LBMkernel(0, false);
      a.calculateMacros();
    }
    b.fluidSimulate(steps);
    auto fa = a.get(0, MT_FIELD_F, levels[0].keys);
    auto fb = b.get(0, MT_FIELD_F, levels[0].keys);
    double num = 0, den = 0;
    for (size_t i = 0; i < fa.size(); ++i) {
      num += (fa[i] - fb[i]) * (fa[i] - fb[i]);
      den += fb[i] * fb[i];
    }
    double rel = std::sqrt(num / den);
    std::printf("driver_check n=%d steps=%d rel_l2=%.3e\n", n, steps, rel);
    // an absent key raises MT_ERR_LOOKUP through mt::Error
    try {
      a.get(0, MT_FIELD_RHO, {~0ull >> 1});
      std::printf("missing-key lookup did not fail\n");
      return 2;
    } catch (const mt::Error& e) {
      if (e.code != MT_ERR_LOOKUP) return 3;
    }
    return rel < 1e-13 ? 0 : 1;
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 4;
  }
}
