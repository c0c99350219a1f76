// mt_case.h -- host-side representation of one device-resident AMR case.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/mt_amr.h"

namespace mt {

constexpr int kMaxLevels = 7;

// level-0 wet-node boundary configuration of one member
struct BcCfg {
  int kind[6];
  double u[6][3];
  int wet_mask;  // faces with MT_BC_ZOUHE or MT_BC_OUTFLOW
};

// Per-member parameters.  A case holds one or more independent members (packed
// batching); every brick belongs to one member, and every kernel takes its
// relaxation times, transfer factors, boundary configuration and initial state
// from its brick's member.  A single case is member 0.
struct MemberParams {
  double tau[kMaxLevels];     // tau_L = 2^L (tau0 - 1/2) + 1/2
  double xscale[kMaxLevels];  // explosion factor into level L (L >= 1)
  double cscale[kMaxLevels];  // coalescence factor into level L (from L + 1)
  double les_C, rho0, u0[3];
  BcCfg bc;
  int32_t lo[kMaxLevels][3];  // member box per level, cells [lo, hi)
  int32_t hi[kMaxLevels][3];
  int32_t periodic[3];
};

struct DevLevel {
  int64_t n_cells = 0, n_bricks = 0, n_slots = 0;
  int32_t dims[3] = {0, 0, 0};
  // the uploaded level set, staged on the device until mt_case_finalize sorts
  // it into the tables (n_up = -1: nothing uploaded)
  uint64_t* up_keys = nullptr;
  uint16_t* up_kinds = nullptr;
  int64_t n_up = -1;
  uint64_t up_max = 0;
  // tables
  uint64_t* bcode = nullptr;  // [n_bricks] sorted brick Morton codes
  int32_t* bnbr = nullptr;    // [n_bricks][27] neighbour bricks (-1 none)
  uint64_t* key = nullptr;    // [n_slots] reference key, ~0 when empty
  uint16_t* kind = nullptr;   // [n_slots] MT_KIND_* (0 = empty)
  uint32_t* pmask = nullptr;  // [n_slots] bit q: neighbour along e_q holds populations
  // state (storage type depends on precision)
  void* f = nullptr;          // [Q][n_slots]
  void* fcol[2] = {nullptr, nullptr};
  void* rho[2] = {nullptr, nullptr};  // [n_slots]
  void* u[2] = {nullptr, nullptr};    // [3][n_slots]
  int pf = 0;                 // current fcol buffer
  int pr = 0;                 // current rho/u buffer
  // work lists
  int32_t* active = nullptr;  // bricks holding lat_f or C2F cells
  int64_t n_active = 0;
  // the same bricks in the order of a sub-step that coalesces into the coarser
  // level (coalescing bricks first), with each one's position in `active`
  // (its force slot)
  int32_t* active_b = nullptr;
  int32_t* slot_b = nullptr;
  // the bricks a sub-step whose C2F cells neither feed a coalescence nor keep
  // their own stream needs (some lat_f cell, or a C2F cell in dis), in the
  // order of `active`, with each one's position there
  int32_t* active_a = nullptr;
  int32_t* slot_a = nullptr;
  int64_t n_active_a = 0;
  int32_t* f2c = nullptr;     // F2C slots of this level (coarse side)
  int64_t n_f2c = 0;
  int32_t* f2c_child = nullptr;        // [n_f2c][8] child slots on level+1
  int32_t* f2c_child_write = nullptr;  // [n_f2c][8] explosion target or -1 (child in dis)
  int32_t* gf2c = nullptr;             // ghost_F2C slots
  int64_t n_gf2c = 0;
  int32_t* gf2c_child = nullptr;       // [n_gf2c][8] ghost_C2F target or -1
  // explosion work list: F2C parents then ghost_F2C parents, with their targets
  int32_t* xp_list = nullptr;
  int32_t* xp_child = nullptr;         // [n_xp][8]
  int64_t n_xp = 0;
  int32_t* dis_row = nullptr;          // [n_slots] row in dis table or -1 (finest level)
  int32_t* ibl = nullptr;              // treated dis slots (dis and lat_f), finest level
  int64_t n_ibl = 0;
  double* dis = nullptr;               // [n_dis][Q-1]
  uint32_t* wmask = nullptr;           // [n_slots] bit q: link q of a dis cell is a wall link
  int64_t n_dis = 0;
  int32_t* bcres = nullptr;
  int64_t n_bcres = 0;
  // coalescence inside the finer level's fused update (coarse side: whether it
  // applies; fine side: parent slot per aligned group of 8 slots, -1 none)
  bool coal_in_fused = false;
  int32_t* cpar = nullptr;
  // explosions inside the fused updates (Case::xin).  Fine side: per slot, bit q
  // set when the pull source x + e_q (x - e_q on a treated cell's wall link) is
  // an explosion target (C2F / ghost_C2F, not in dis), read from the pair's
  // target groups; the target slots.  Coarse side:
  // per slot, the row of a ghost_F2C parent in gf2c_child, else -1.
  uint32_t* tmask = nullptr;
  // with the reference's explosion (sub-step A only), sub-step B reads the C2F
  // cells' own fcol (sub-step A's copy) and only the ghost_C2F targets from the
  // target groups: the same mask and slot list restricted to ghost_C2F targets
  uint32_t* tmask_g = nullptr;
  int32_t* xtg_list = nullptr;
  int64_t n_xtg = 0;
  // the explosion targets of the current and the next pair: one value per
  // population per aligned group of 8 slots (the 8 children of one parent all
  // receive the same values), [Q][n_slots / 8] of the storage format
  void* xg[2] = {nullptr, nullptr};
  int32_t* xt_list = nullptr;
  int64_t n_xt = 0;
  int32_t* gxrow = nullptr;
  int32_t* bmember = nullptr;          // [n_bricks] member of the brick
  std::vector<int32_t> h_members;      // upload staging: member of each cell (empty: all 0)
};

// a coalescence into level L-1 requested from level L's last sub-step
struct CoalReq {
  int write_macros;
};

struct Case {
  mt_case_desc d;
  int Q = 19, Lmax = 0, prec = MT_FP64;
  double tau[kMaxLevels];
  DevLevel lv[kMaxLevels];
  std::vector<mt_member_desc> h_mdesc;  // members as given (one by default)
  std::vector<MemberParams> h_mem;
  MemberParams* d_mem = nullptr;
  // Lagrangian markers on the finest level (explicit direct forcing)
  int64_t n_mk = 0;
  std::vector<uint64_t> h_mk_keys;  // [n][27] upload staging
  std::vector<double> h_mk_w;
  std::vector<double> h_mk_dv;
  double mk_ub[3] = {0, 0, 0};
  int32_t* mk_slot = nullptr;       // [n][27]
  int32_t* mk_band = nullptr;       // [n][27] band index of each kernel cell
  double* band_mom = nullptr;       // [4][n_band] rho, u of the band cells (this call)
  double* mk_w = nullptr;           // [n][27]
  double* mk_dv = nullptr;          // [n]
  double* mk_F = nullptr;           // [n][3] this call's marker forces
  int32_t* band_slot = nullptr;     // cells in some marker's support
  int32_t* band_row = nullptr;      // [finest n_slots] band index of the slot, -1 none
  int64_t n_band = 0;
  int32_t* band_ptr = nullptr;      // [n_band + 1]
  int32_t* band_mk = nullptr;       // markers of each cell, ascending
  double* band_w = nullptr;
  int64_t band_cap = 0;             // 27 n: capacity of the band arrays
  int32_t* d_nband = nullptr;       // the band's size on the device
  int32_t* mk_ent = nullptr;        // [n][27] position of the entry in the band CSR, -1 unused
  double* mk_v = nullptr;           // [n][3] marker velocities (reported frame)
  double* mk_x = nullptr;           // [n][3] positions, finest cells
  bool mk_x_set = false;            // mk_x holds positions (after a move)
  // the accumulated force holds the marker call of the collision computed ahead
  // (mt_step's state), which is the last marker call (force_acc[4..6])
  bool mk_acc_ahead = false;
  // staging of a move: positions, velocities, the new kernel cells and weights,
  // and flags: bad markers, first bad marker, support changed, used entries
  double* mv_x = nullptr;
  double* mv_v = nullptr;
  int32_t* mv_slot = nullptr;
  double* mv_w = nullptr;
  int32_t* mv_flags = nullptr;
  // the uploaded dis keys and fractions, staged on the device like the level
  // sets
  uint64_t* up_dis_keys = nullptr;
  double* up_dis_vals = nullptr;
  int64_t n_up_dis = 0;
  uint64_t up_dis_max = 0;
  bool finalized = false;
  bool ahead = false;  // fcol[pf] already holds the next collision (mt_step state)
  bool graphs = true;
  bool timing = false;
  bool diag = false;  // every update leaves its macros (mt_set_diagnostics)
  int64_t diag_steps = 0;
  // launch configuration (results do not depend on it; mt_autotune picks it)
  int xmap = 0;               // fused kernel block -> brick map (FusedArgs::xmap)
  bool level_streams = true;  // levels on concurrent streams in mt_step
  // with explosions inside the updates (xin): mt_step issues one launch per
  // finest sub-step on one stream, carrying every coarser level's update whose
  // span ends with it (takes precedence over level_streams)
  bool merge_levels = true;
  bool wet_any = false;  // some member has a Zou-He / outflow face
  bool nee_code = false;  // some wet-node cell may need the NEE extrapolation (fused kernels carry it)
  bool res_any = false;  // some member has a reservoir face
  // Explosions inside the fused updates (with MT_OPT_EXPLODE_BOTH).  Both
  // explosions of a fine pair write the same values, so the pair's targets are
  // written once, before the pair, into fcol[pf of the coarser level] of the
  // fine level (never otherwise written there), and the fine pulls read them
  // from that buffer: F2C parents are exploded by the fine level's coalescing
  // update, ghost_F2C parents by the coarse level's update.  xprimed: the
  // targets for the next pair are in place (an explosion kernel primes them
  // at the start of mt_step).
  bool xin = false;
  bool xprimed = false;
  hipStream_t stream = nullptr;
  // per-level streams used by mt_step (level 0 uses `stream`); st[L] is the
  // stream a level's kernels go to: `stream` for the seam entry points
  hipStream_t lstream[kMaxLevels] = {nullptr};
  hipStream_t st[kMaxLevels] = {nullptr};
  hipEvent_t ev_ready[kMaxLevels] = {nullptr};  // level L's current fcol final for its children
  hipEvent_t ev_done[kMaxLevels] = {nullptr};   // level L's sub-step tree finished
  bool join_pending[kMaxLevels] = {false};       // level L has not yet waited for ev_done[L + 1]
  std::vector<void*> allocs;
  // IB force
  double* force_part = nullptr;  // per-block partials of the finest fused kernel
  int64_t force_part_cap = 0;
  double* force_acc = nullptr;   // [8]: seam-path sum x,y,z and count, last call x,y,z,
                                 // [7] treat() calls folded into force_slots
  double* force_slots = nullptr; // [force_part_cap][3] per-block running sums of the fused path
  double* nee_tmp = nullptr;     // NEE Jacobi staging
  double* red_tmp = nullptr;     // reduction partials
  int64_t red_cap = 0;
  unsigned long long* member_bad = nullptr;  // [members] cells with a non-finite population
  // graphs: key = (start parity of level 0) | (final << 1)
  std::map<int, hipGraphExec_t> graph_exec;
  // timing
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  double t_ms[kMaxLevels] = {0};
  int64_t t_calls[kMaxLevels] = {0};
  size_t pop_bytes() const { return prec == MT_FP64 ? 8 : (prec == MT_FP32_SHIFTED ? 4 : 2); }
  size_t real_bytes() const { return prec == MT_FP64 ? 8 : 4; }

  template <typename T>
  T* alloc(int64_t n);
};

// Process-wide caches of device memory and streams (mt_api.hip).  A serving
// loop builds and destroys cases continually, and hipMalloc / hipFree /
// hipStreamCreate cost from tens of microseconds to milliseconds per call.
// A block freed with a stream is reused only by allocations on that stream
// (stream order) until pool_quiesce(stream), called once the stream has
// drained; a block freed with no stream is reusable at once.
hipError_t dev_malloc(void** p, size_t bytes, hipStream_t s);
void dev_free(void* p, hipStream_t s);
void pool_quiesce(hipStream_t s);
hipError_t stream_acquire(hipStream_t* s);
void stream_release(hipStream_t s);

// Frees the device staging of an uploaded level set (of the uploaded dis),
// in the order of stream s.
inline void drop_level_upload(DevLevel& l, hipStream_t s) {
  if (l.up_keys) dev_free(l.up_keys, s);
  if (l.up_kinds) dev_free(l.up_kinds, s);
  l.up_keys = nullptr;
  l.up_kinds = nullptr;
  l.n_up = -1;
  l.up_max = 0;
}
inline void drop_dis_upload(Case& c, hipStream_t s) {
  if (c.up_dis_keys) dev_free(c.up_dis_keys, s);
  if (c.up_dis_vals) dev_free(c.up_dis_vals, s);
  c.up_dis_keys = nullptr;
  c.up_dis_vals = nullptr;
  c.n_up_dis = 0;
  c.up_dis_max = 0;
}

template <typename T>
T* Case::alloc(int64_t n) {
  void* p = nullptr;
  if (n <= 0) n = 1;
  if (dev_malloc(&p, sizeof(T) * n, nullptr) != hipSuccess) return nullptr;
  allocs.push_back(p);
  return (T*)p;
}

void set_error(const std::string& s);
#define MT_HIP(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      mt::set_error(std::string(#x) + ": " + hipGetErrorString(e_));               \
      return MT_ERR_HIP;                                                            \
    }                                                                               \
  } while (0)

// tables (mt_tables.hip)
int build_members(Case& c);
int copy_member(Case& dst, int32_t dm, Case& src, int32_t sm);
int build_tables(Case& c);
int find_slots(Case& c, int level, const uint64_t* d_keys, int64_t n, int32_t* d_out);
// kernels (mt_kernels.hip)
int init_state(Case& c);
int launch_macros(Case& c, int L);
int launch_collide(Case& c, int L);
int launch_stream(Case& c, int L);
int launch_explode(Case& c, int L, bool mirror_all = false);
int launch_explode_to(Case& c, int L, int dst_buf);
int launch_copy_targets(Case& c, int L, int from_buf, int to_buf);
int launch_coalesce(Case& c, int L, bool collide_after, bool write_macros = true);
int launch_bc(Case& c);
int launch_ib(Case& c, bool reduce_into_acc);
int launch_fused(Case& c, int L, bool write_f, bool bc_nee_after, bool write_macros = true, bool c2f_fcol = true,
                 const CoalReq* coal = nullptr);
// one level's part of a launch of several levels' fused updates
struct LevelUpdate {
  int L;
  bool write_f;
  bool write_macros;
  bool c2f_fcol;
  const CoalReq* coal;
};
// The fused updates of n levels in one launch on the finest one's stream (every
// level's stream must be that one), then the reservoir NEE if nee_after.
int launch_fused_levels(Case& c, int n, const LevelUpdate* u, bool nee_after);
bool level_has_ib(const Case& c, int L);
int build_markers(Case& c);
// Moves the markers to mv_x with velocities mv_v (staged on the device; rigid:
// first move mk_x rigidly into mv_x, see mt_markers_rigid) and rebuilds their
// kernel cells and band (mt_tables.hip); on an invalid support returns
// MT_ERR_ARG and changes nothing.
int move_markers(Case& c, bool rigid, double dt, const double* center, const double* vel, const double* omega);
// moments_done: the finest fused update has already left the band's moments
// (FusedArgs::band_row)
int launch_markers(Case& c, int buf, hipStream_t st, bool moments_done = false);
// the state of src carried onto dst's cells (mt_case_remap); counts[4]: cells
// copied, exploded, coalesced, filled
int launch_remap(Case& dst, const Case& src, const double* fill, int64_t* counts);
// removes the last marker call from the accumulated force
int launch_mk_unforce(Case& c);
int launch_gather(Case& c, int L, int field, const int32_t* d_slots, int64_t n, double* d_out, int buf);
int launch_scatter(Case& c, int L, int field, const int32_t* d_slots, int64_t n, const double* d_in, int buf);
int launch_reduce(Case& c, int L, double* h_out);
// per member: fluid and F2C cells of any level holding a non-finite population
// in either fcol buffer
int launch_member_scan(Case& c, int64_t* h_counts);
void set_member_tau(const Case& c, MemberParams& P, const double* tau);
// infoPrint / diagL2* partial values of one level (k_diag's 16 values, combined)
int launch_diag(Case& c, int L, double* h_out);

}  // namespace mt

// the C-ABI handle is the case itself
struct mt_case : public mt::Case {};
