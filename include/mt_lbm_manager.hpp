// mt_lbm_manager.hpp -- C++ host driver over the C-ABI of mt_amr.h.
//
// The class keeps the shape of the reference's LBM_Manager
// (header/LBM_Manager.hpp:58-88, LBM_Simulator.cpp:8-229): the same method
// names, the same recursive sub-cycling schedule, the same per-level seams, so
// a caller written against LBM_Manager drives the device path unchanged.
// Errors: every failing C call throws mt::Error carrying mt_last_error().
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "mt_amr.h"

namespace mt {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc) {
  if (rc != MT_OK) throw Error(rc, std::string("mt_amr: ") + mt_last_error());
}

// One level's cell set in the reference's terms (Lat_Manager::lat_f, lat_overlap_*,
// ghost_overlap_*, lat_sf, lat_bc): keys plus MT_KIND_* flags.
struct LevelCells {
  std::vector<uint64_t> keys;
  std::vector<uint16_t> kinds;
};

class LBMManager {
 public:
  LBMManager(const mt_case_desc& desc, const std::vector<LevelCells>& levels, const std::vector<uint64_t>& dis_keys,
             const std::vector<double>& dis_fractions)
      : Lmax_(desc.Lmax), Q_(desc.Q) {
    check(mt_case_create(&desc, &c_));
    try {
      for (int L = 0; L <= Lmax_; ++L)
        check(mt_case_upload_level(c_, L, levels.at(L).keys.data(), levels.at(L).kinds.data(),
                                   (int64_t)levels.at(L).keys.size()));
      if (!dis_keys.empty())
        check(mt_case_upload_dis(c_, dis_keys.data(), dis_fractions.data(), (int64_t)dis_keys.size()));
      check(mt_case_finalize(c_));  // allocMacrosAndDDF + initialBC
    } catch (...) {
      mt_case_destroy(c_);
      throw;
    }
  }
  ~LBMManager() { mt_case_destroy(c_); }
  LBMManager(const LBMManager&) = delete;
  LBMManager& operator=(const LBMManager&) = delete;

  int maxLevel() const { return Lmax_; }
  mt_case* handle() { return c_; }

  // ---- seams (LBM_Manager.cpp / strategy methods) ----
  void calculateMacros(int level) { check(mt_macros(c_, level)); }
  void calculateMacros() {
    for (int L = 0; L <= Lmax_; ++L) calculateMacros(L);
  }
  void collide(int level) { check(mt_collide(c_, level)); }
  void stream(int level) { check(mt_stream(c_, level)); }
  void AMR_transDDF_explosion(int level) { check(mt_explode(c_, level)); }
  void AMR_transDDF_coalescence(int level) { check(mt_coalesce(c_, level)); }
  void applyBoundaryCondition() { check(mt_apply_bc(c_)); }
  void treat(double* force3 = nullptr) { check(mt_ib_treat(c_, force3)); }

  // LBMkernel(level) through the seams, in the reference's order
  // (LBM_Simulator.cpp:20-229); sub-step B has no explosion.
  void LBMkernel(int L, bool has_ib) {
    calculateMacros(L);
    collide(L);
    if (L < Lmax_) LBMkernel(L + 1, has_ib);
    if (L > 0) AMR_transDDF_explosion(L);
    stream(L);
    if (L == 0) applyBoundaryCondition();
    else if (L == Lmax_ && has_ib) treat();
    if (L < Lmax_) AMR_transDDF_coalescence(L);
    if (L == 0) return;
    calculateMacros(L);
    collide(L);
    if (L < Lmax_) LBMkernel(L + 1, has_ib);
    stream(L);
    if (L == Lmax_ && has_ib) treat();
    if (L < Lmax_) AMR_transDDF_coalescence(L);
  }

  // fluidSimulate's loop body n times (LBM_Simulator.cpp:11-17) on the fused,
  // graph-captured device schedule.
  void fluidSimulate(int n_steps) { check(mt_step(c_, n_steps)); }

  // ---- state access ----
  std::vector<double> get(int level, int field, const std::vector<uint64_t>& keys) {
    size_t rows = field <= MT_FIELD_FCOL ? (size_t)Q_ : (field == MT_FIELD_RHO ? 1 : 3);
    std::vector<double> out(rows * keys.size());
    check(mt_get_field(c_, level, field, keys.data(), (int64_t)keys.size(), out.data()));
    return out;
  }
  void set(int level, int field, const std::vector<uint64_t>& keys, const std::vector<double>& v) {
    check(mt_set_field(c_, level, field, keys.data(), (int64_t)keys.size(), v.data()));
  }
  // infoPrint's drag coefficient (LBM_Manager.cpp:964-979): mean IB force since
  // the last call, Cd = 2 Fx / (U^2 pi d^2 / 4) with rho_lb = 1
  double dragCoefficient(double U_lb, double d_lb) {
    double s[3];
    int64_t n = 0;
    check(mt_ib_force(c_, s, &n, 1));
    if (n == 0) return 0.0;
    const double A = 3.14159265358979323846 * d_lb * d_lb / 4.0;
    return 2.0 * (s[0] / (double)n) / (U_lb * U_lb * A);
  }

 private:
  mt_case* c_ = nullptr;
  int Lmax_, Q_;
};

}  // namespace mt
