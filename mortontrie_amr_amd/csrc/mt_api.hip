// mt_api.hip -- the C-ABI (include/mt_amr.h) and the coarse-step schedule.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "mt_case.h"

namespace mt {

static thread_local std::string g_err;
void set_error(const std::string& s) { g_err = s; }

// ---------------------------------------------------------------- caches
namespace {
struct Pool {
  std::mutex mu;
  // cached blocks by size; a block carries the stream it was freed on
  // (nullptr: reusable by anyone)
  std::multimap<size_t, std::pair<void*, hipStream_t>> free_blocks;
  std::unordered_map<void*, size_t> size_of;
  std::vector<hipStream_t> streams;
};
Pool& pool() {
  static Pool* p = new Pool();  // lives until the process ends
  return *p;
}
size_t block_size(size_t n) {
  if (n == 0) n = 1;
  if (n <= (size_t(1) << 20)) return (n + 4095) & ~size_t(4095);
  return (n + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
}
void release_cached_locked(Pool& P) {
  for (auto& b : P.free_blocks) {
    P.size_of.erase(b.second.first);
    hipFree(b.second.first);
  }
  P.free_blocks.clear();
}
}  // namespace

hipError_t dev_malloc(void** p, size_t bytes, hipStream_t s) {
  Pool& P = pool();
  const size_t n = block_size(bytes);
  {
    std::lock_guard<std::mutex> g(P.mu);
    for (auto it = P.free_blocks.lower_bound(n); it != P.free_blocks.end() && it->first <= n + n / 8; ++it) {
      if (it->second.second == nullptr || it->second.second == s) {
        *p = it->second.first;
        P.free_blocks.erase(it);
        return hipSuccess;
      }
    }
  }
  hipError_t e = hipMalloc(p, n);
  if (e == hipErrorOutOfMemory) {
    // give the cached blocks back and try once more
    hipDeviceSynchronize();
    std::lock_guard<std::mutex> g(P.mu);
    release_cached_locked(P);
    (void)hipGetLastError();
    e = hipMalloc(p, n);
  }
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> g(P.mu);
    P.size_of[*p] = n;
  }
  return e;
}

void dev_free(void* p, hipStream_t s) {
  if (!p) return;
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  auto it = P.size_of.find(p);
  if (it == P.size_of.end()) {
    hipFree(p);
    return;
  }
  P.free_blocks.emplace(it->second, std::make_pair(p, s));
}

void pool_quiesce(hipStream_t s) {
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  for (auto& b : P.free_blocks)
    if (b.second.second == s) b.second.second = nullptr;
}

hipError_t stream_acquire(hipStream_t* s) {
  Pool& P = pool();
  {
    std::lock_guard<std::mutex> g(P.mu);
    if (!P.streams.empty()) {
      *s = P.streams.back();
      P.streams.pop_back();
      return hipSuccess;
    }
  }
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

// the stream must have drained
void stream_release(hipStream_t s) {
  if (!s) return;
  pool_quiesce(s);
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  P.streams.push_back(s);
}

#define CHECK_CASE(c)                                   \
  do {                                                  \
    if (!(c)) { set_error("null case"); return MT_ERR_ARG; } \
  } while (0)
#define CHECK_FINAL(c)                                                         \
  do {                                                                         \
    CHECK_CASE(c);                                                             \
    if (!(c)->finalized) { set_error("case not finalized"); return MT_ERR_STATE; } \
  } while (0)
#define CHECK_LEVEL(c, L)                                                      \
  do {                                                                         \
    if ((L) < 0 || (L) > (c)->Lmax) { set_error("level out of range"); return MT_ERR_ARG; } \
  } while (0)
#define TRY(x)                    \
  do {                            \
    int r_ = (x);                 \
    if (r_ != MT_OK) return r_;   \
  } while (0)

static int run_level(Case& c, int L, bool last, const CoalReq* coal);

static bool has_reservoir(const Case& c) { return c.res_any; }

// Return the case to reference-state buffers: after mt_step the current fcol
// buffer already holds the next collision and the reference's fcol is the other.
static int settle(Case& c) {
  if (!c.ahead) return MT_OK;
  if (c.mk_acc_ahead) {
    // the discarded collision's marker force leaves the accumulated force (its
    // collision is redone, and counted, by the next step)
    TRY(launch_mk_unforce(c));
    c.mk_acc_ahead = false;
  }
  if (c.xin && c.xprimed) {
    // the last pair's explosion targets sit in fcol[1 - pf of the coarser level]
    // (the other buffer already holds the next pair's): move them to the
    // buffer that becomes the level's reference fcol
    for (int L = 1; L <= c.Lmax; ++L) TRY(launch_copy_targets(c, L, 1 - c.lv[L - 1].pf, 1 - c.lv[L].pf));
  }
  c.xprimed = false;
  for (int L = 0; L <= c.Lmax; ++L) c.lv[L].pf ^= 1;
  c.ahead = false;
  return MT_OK;
}

// The prologue of a step: calculateMacros(L) + collide(L) on every level, the
// first two calls of each level's LBMkernel sub-step (LBM_Simulator.cpp:26-28).
static int prologue(Case& c) {
  for (int L = 0; L <= c.Lmax; ++L) {
    TRY(launch_macros(c, L));
    TRY(launch_collide(c, L));
  }
  // the markers act on every finest-level collision
  TRY(launch_markers(c, c.lv[c.Lmax].pf, c.st[c.Lmax]));
  c.ahead = true;
  c.mk_acc_ahead = c.n_mk > 0;
  return MT_OK;
}

// LBMkernel(L) (LBM_Simulator.cpp:20-229) over the fused kernels.  On entry
// fcol[pf] of every level holds the collision of the sub-step about to stream.
//
// Each level runs on its own stream.  Level L's explosion and stream+collide
// read level L-1's current fcol and level L's current fcol and write level L's
// C2F / ghost_C2F cells and the other fcol buffer; nothing in LBMkernel(L+1)
// writes level-L data, and what it reads of level L (the current fcol of the
// F2C and ghost_F2C parents) is neither of those cell sets (an interface cell is
// never both, mt_case_upload_level rejects such a level).  So the finer levels'
// sub-tree runs concurrently with level L's explosion and fused update; only
// the coalescence, which reads the finer level's final f, joins them.  Issue
// order is the reference's (finer levels, explosion, stream, coalescence).
// One sub-step of level L.  explode: 0 none, 1 the reference's explosion,
// 2 explosion whose targets also fill the second buffer.  coal: the coalescence
// into level L-1 that this sub-step's fused update performs (last sub-step of
// the pair), or null.
static int substep(Case& c, int L, bool child_last, bool write_f, bool nee, bool write_macros, int explode,
                   bool c2f_fcol, const CoalReq* coal) {
  DevLevel& l = c.lv[L];
  hipStream_t s = c.st[L];
  if (L < c.Lmax) {
    MT_HIP(hipEventRecord(c.ev_ready[L], s));
    MT_HIP(hipStreamWaitEvent(c.st[L + 1], c.ev_ready[L], 0));
    // The previous sub-step's finer pair coalesced into this level's F2C
    // cells, so it must finish before this sub-step's update of level L.
    // Waiting here, after the fork, lets the next finer pair follow the
    // previous one directly on its stream, without a round trip through this
    // level's stream.
    if (c.join_pending[L]) {
      MT_HIP(hipStreamWaitEvent(s, c.ev_done[L + 1], 0));
      c.join_pending[L] = false;
    }
    const CoalReq cr = {write_macros ? 1 : 0};
    TRY(run_level(c, L + 1, child_last, l.coal_in_fused ? &cr : nullptr));
  }
  if (explode) TRY(launch_explode(c, L, explode == 2));
  TRY(launch_fused(c, L, write_f, nee, write_macros, c2f_fcol, coal));
  if (L == c.Lmax) TRY(launch_markers(c, 1 - l.pf, s, true));
  if (L < c.Lmax) {
    MT_HIP(hipEventRecord(c.ev_done[L + 1], c.st[L + 1]));
    if (l.coal_in_fused) {
      c.join_pending[L] = true;
    } else {
      MT_HIP(hipStreamWaitEvent(s, c.ev_done[L + 1], 0));
      TRY(launch_coalesce(c, L, true, write_macros));
    }
  }
  l.pf ^= 1;
  l.pr ^= 1;
  return MT_OK;
}
// The joins the levels still owe their finer pairs, finest first, so that
// c.stream (level 0's) ends after everything issued.  A level's own next
// update waits for its finer pair (the coalescence into its F2C cells) inside
// `substep`, after the fork; nothing coarser needs more than the finer level's
// own updates (a finer level's work writes only its own data and the F2C cells
// of the level above it), so between coarse steps a join is carried over and
// only the end of the issued steps closes every stream.
static int join_all(Case& c) {
  for (int L = c.Lmax - 1; L >= 0; --L) {
    if (c.st[L] != c.st[L + 1]) {
      MT_HIP(hipEventRecord(c.ev_done[L + 1], c.st[L + 1]));
      MT_HIP(hipStreamWaitEvent(c.st[L], c.ev_done[L + 1], 0));
    }
    c.join_pending[L] = false;
  }
  return MT_OK;
}
static int run_level(Case& c, int L, bool last, const CoalReq* coal) {
  const bool res0 = (L == 0) && has_reservoir(c);
  if (L == 0) {
    // level 0 has one sub-step per coarse step
    return substep(c, 0, last, last || res0, res0, last || res0, 0, true, nullptr);
  }
  // Sub-step B: the reference does not explode (LBM_Simulator.cpp:152-217).
  // With MT_OPT_EXPLODE_BOTH it explodes again from the same parent fcol (the
  // parent level streams only after both sub-steps), i.e. it rewrites the values
  // of sub-step A's explosion: A's explosion fills both buffers instead, and the
  // fused updates skip the C2F cells' fcol (always overwritten before a read).
  const bool both = (c.d.options & MT_OPT_EXPLODE_BOTH) != 0;
  // (with c.xin the pair's targets are already in place: no explosion kernel)
  TRY(substep(c, L, false, false, false, false, c.xin ? 0 : (both ? 2 : 1), !both, nullptr));
  return substep(c, L, last, last, false, last, 0, !both, coal);
}

// One coarse step as 2^Lmax launches on one stream (Case::merge_levels): launch
// t carries the finest level's sub-step t and every coarser level L's sub-step
// whose span of 2^(Lmax-L) finest sub-steps ends with t, with the flags the
// recursion in run_level gives it.  Level L's update j then runs beside the
// last update of its finer pair j (the stream schedule runs it beside the
// whole pair), after its finer pair j-1 (an earlier span) and before its finer
// pair j+1 (the next span): the orders the stream schedule's fork and join
// keep, without any wait between streams.  Any launch of the span would keep
// them; the last one (the finest level's sub-step B, whose coalescing bricks
// leave the longest tail) measured fastest, with the coarser levels' blocks
// dispatched first.  A level's buffers flip when its span ends, as in
// run_level, where a level flips after its finer pair.  Needs the explosions
// and coalescences inside the updates (Case::xin).
static bool merged(const Case& c) { return c.merge_levels && c.xin && c.Lmax >= 1; }

static int merged_step(Case& c, bool last) {
  const int Lm = c.Lmax;
  const bool res0 = has_reservoir(c);
  const bool both = (c.d.options & MT_OPT_EXPLODE_BOTH) != 0;
  CoalReq cr[kMaxLevels];
  LevelUpdate u[kMaxLevels];
  for (int t = 0; t < (1 << Lm); ++t) {
    int n = 0;
    bool nee = false;
    for (int L = 0; L <= Lm; ++L) {
      const int span = 1 << (Lm - L);
      if (t % span != span - 1) continue;
      const int j = t / span;  // level L's sub-step within the coarse step
      LevelUpdate& x = u[n++];
      x.L = L;
      x.coal = nullptr;
      if (L == 0) {
        x.write_f = x.write_macros = last || res0;
        x.c2f_fcol = true;
        nee = res0;
      } else {
        // sub-step B (odd j) of the level's last pair in the last step writes f
        // and the maps; B coalesces into level L-1 with the maps flag of
        // level L-1's current sub-step
        const bool b = (j & 1) != 0;
        x.write_f = x.write_macros = b && last && (j >> 1) == (1 << (L - 1)) - 1;
        x.c2f_fcol = !both;
        if (b) {
          const int jp = j >> 1;
          const bool wm = (L == 1) ? (last || res0) : ((jp & 1) != 0 && last && (jp >> 1) == (1 << (L - 2)) - 1);
          cr[L].write_macros = wm ? 1 : 0;
          x.coal = &cr[L];
        }
      }
    }
    TRY(launch_fused_levels(c, n, u, nee));
    TRY(launch_markers(c, 1 - c.lv[Lm].pf, c.st[Lm], true));
    // the levels whose span ends here flip (their finer pair's updates in the
    // span address the buffers of the sub-step in progress)
    for (int i = 0; i < n; ++i) {
      c.lv[u[i].L].pf ^= 1;
      c.lv[u[i].L].pr ^= 1;
    }
  }
  return MT_OK;
}

static int one_step(Case& c, bool last) { return merged(c) ? merged_step(c, last) : run_level(c, 0, last, nullptr); }

static void use_level_streams(Case& c, bool on) {
  on = on && c.level_streams && !merged(c);
  for (int L = 0; L <= c.Lmax; ++L) c.st[L] = (on && L > 0) ? c.lstream[L] : c.stream;
}

static void drop_graphs(Case& c) {
  for (auto& g : c.graph_exec) hipGraphExecDestroy(g.second);
  c.graph_exec.clear();
}

static int parity_key(const Case& c, bool last) {
  int k = (last ? 1 : 0) | (c.diag ? (1 << 20) : 0);
  for (int L = 0; L <= c.Lmax; ++L) k |= (c.lv[L].pf << (1 + 2 * L)) | (c.lv[L].pr << (2 + 2 * L));
  return k;
}

// k consecutive coarse steps (the last of them `last`), replayed as one HIP
// graph: the step boundaries inside a graph cost no host launch, and a graph
// per (parities, k, last) is captured once.  Consecutive graphs leave a gap of
// about 8 us on the device, so a graph holds 16 coarse steps (0.5 us per step
// of the 24 us 64^3 cavity step, against 2 us with 4).
constexpr int kStepsPerGraph = 16;
static int coarse_steps(Case& c, int k, bool last) {
  if (!c.graphs || c.timing) {
    for (int j = 0; j < k; ++j) TRY(one_step(c, last && j == k - 1));
    return join_all(c);
  }
  const int key = parity_key(c, last) | (k << 24);
  auto it = c.graph_exec.find(key);
  int pf[kMaxLevels], pr[kMaxLevels];
  bool jp[kMaxLevels];
  for (int L = 0; L <= c.Lmax; ++L) { pf[L] = c.lv[L].pf; pr[L] = c.lv[L].pr; jp[L] = c.join_pending[L]; }
  const bool xprimed = c.xprimed, ahead = c.ahead;
  // capture records the schedule without running it: the host-side state the
  // recording advanced goes back to where the device still is, also on failure
  auto restore = [&]() {
    for (int L = 0; L <= c.Lmax; ++L) { c.lv[L].pf = pf[L]; c.lv[L].pr = pr[L]; c.join_pending[L] = jp[L]; }
    c.xprimed = xprimed;
    c.ahead = ahead;
  };
  if (it == c.graph_exec.end()) {
    hipGraph_t g = nullptr;
    MT_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    int r = MT_OK;
    for (int j = 0; j < k && r == MT_OK; ++j) r = one_step(c, last && j == k - 1);
    if (r == MT_OK) r = join_all(c);
    hipError_t e = hipStreamEndCapture(c.stream, &g);
    const bool xprimed_after = c.xprimed, ahead_after = c.ahead;
    bool jp_after[kMaxLevels];
    for (int L = 0; L <= c.Lmax; ++L) jp_after[L] = c.join_pending[L];
    restore();
    if (r != MT_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (r == MT_OK) { set_error(std::string("graph capture: ") + hipGetErrorString(e)); r = MT_ERR_HIP; }
      return r;
    }
    hipGraphExec_t ex;
    hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) { set_error(std::string("graph instantiate: ") + hipGetErrorString(ei)); return MT_ERR_HIP; }
    c.graph_exec[key] = ex;
    it = c.graph_exec.find(key);
    // the replay below leaves the flags as the recording did
    c.xprimed = xprimed_after;
    c.ahead = ahead_after;
    for (int L = 0; L <= c.Lmax; ++L) c.join_pending[L] = jp_after[L];
  }
  MT_HIP(hipGraphLaunch(it->second, c.stream));
  // parities advance as in the captured schedule: per coarse step level L
  // flips 2^L times, level 0 once
  for (int L = 0; L <= c.Lmax; ++L) {
    const int flips = k * (L == 0 ? 1 : (1 << L));
    c.lv[L].pf = pf[L] ^ (flips & 1);
    c.lv[L].pr = pr[L] ^ (flips & 1);
  }
  return MT_OK;
}

static int drain_timing(Case& c) {
  if (c.ev_pending.empty()) return MT_OK;
  MT_HIP(hipStreamSynchronize(c.stream));
  for (auto& p : c.ev_pending) {
    float ms = 0;
    MT_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    c.t_ms[p.first] += ms;
    c.t_calls[p.first] += 1;
    c.ev_pool.push_back(p.second.first);
    c.ev_pool.push_back(p.second.second);
  }
  c.ev_pending.clear();
  return MT_OK;
}

// ---------------------------------------------------------------- allocation-free lookup
// Device memory of mt_find_noalloc: the query keys and their slots for up to
// MT_FIND_NOALLOC_MAX_KEYS keys, keys first (8 B each), slots after them (4 B each).
// Each GPU holds its own copy until the process ends.  In host code the name is
// only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
static __device__ __align__(256) unsigned char scratch[MT_FIND_NOALLOC_MAX_KEYS * 12];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// the device address of `scratch` on the device `stream` runs on
int scratch_for_stream(hipStream_t stream, unsigned char** out) {
  int dev = -1;
  MT_HIP(hipStreamGetDevice(stream, &dev));
  if (dev < 0 || dev >= kMaxDevices) { set_error("device index out of range"); return MT_ERR_ARG; }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    MT_HIP(hipGetDevice(&cur));
    if (cur != dev) MT_HIP(hipSetDevice(dev));
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return MT_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return MT_OK;
}

}  // namespace mt

using namespace mt;


extern "C" {

const char* mt_last_error(void) { return g_err.c_str(); }

int mt_case_create(const mt_case_desc* d, mt_case** out) {
  if (!d || !out) { set_error("null argument"); return MT_ERR_ARG; }
  if (d->Q != 19 && d->Q != 27) { set_error("Q must be 19 or 27"); return MT_ERR_ARG; }
  if (d->Lmax < 0 || d->Lmax >= kMaxLevels) { set_error("Lmax out of range"); return MT_ERR_ARG; }
  if (d->collision == MT_COLLIDE_CUMULANT && d->Q != 27) {
    // Collision_Cumulant::collide requires D3Q27 (Collision_Cumulant.cpp:24-26)
    set_error("the cumulant collision needs Q = 27");
    return MT_ERR_ARG;
  }
  if (d->precision < MT_FP64 || d->precision > MT_FP16_STORAGE) { set_error("bad precision"); return MT_ERR_ARG; }
  if (d->tau0 <= 0.5) { set_error("tau0 must exceed 1/2"); return MT_ERR_ARG; }
  if (d->ib < MT_IB_NONE || d->ib > MT_IB_UNIFORM) { set_error("bad solid-boundary method"); return MT_ERR_ARG; }
  for (int f = 0; f < 6; ++f)
    if (d->face_kind[f] < MT_BC_NONE || d->face_kind[f] > MT_BC_NEE) { set_error("bad face kind"); return MT_ERR_ARG; }
  for (int a = 0; a < 3; ++a) {
    bool p0 = d->face_kind[2 * a] == MT_BC_PERIODIC, p1 = d->face_kind[2 * a + 1] == MT_BC_PERIODIC;
    if (p0 != p1) { set_error("periodic faces must come in pairs"); return MT_ERR_ARG; }
  }
  mt_case* c = new mt_case;
  c->d = *d;
  c->Q = d->Q;
  c->Lmax = d->Lmax;
  c->prec = d->precision;
  // the case is one member until mt_case_set_members says otherwise
  mt_member_desc m = {};
  for (int a = 0; a < 3; ++a) {
    m.origin0[a] = 0;
    m.dims0[a] = d->dims0[a];
    m.u0[a] = d->u0[a];
  }
  for (int f = 0; f < 6; ++f) {
    m.face_kind[f] = d->face_kind[f];
    for (int a = 0; a < 3; ++a) m.face_u[f][a] = d->face_u[f][a];
  }
  m.tau0 = d->tau0;
  m.les_C = d->les_C;
  m.rho0 = d->rho0;
  c->h_mdesc.assign(1, m);
  hipError_t e = stream_acquire(&c->stream);
  for (int L = 1; L <= c->Lmax && e == hipSuccess; ++L) e = stream_acquire(&c->lstream[L]);
  for (int L = 0; L <= c->Lmax && e == hipSuccess; ++L) {
    e = hipEventCreateWithFlags(&c->ev_ready[L], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_done[L], hipEventDisableTiming);
  }

  if (e != hipSuccess) {
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    mt_case_destroy(c);
    return MT_ERR_HIP;
  }
  use_level_streams(*c, false);
  *out = c;
  return MT_OK;
}

int mt_case_destroy(mt_case* c) {
  if (!c) return MT_OK;
  // every stream of the case drains before its memory and streams are reused
  if (c->stream) hipStreamSynchronize(c->stream);
  for (int L = 0; L < kMaxLevels; ++L)
    if (c->lstream[L]) hipStreamSynchronize(c->lstream[L]);
  for (auto& g : c->graph_exec) hipGraphExecDestroy(g.second);
  for (auto e : c->ev_pool) hipEventDestroy(e);
  for (auto& p : c->ev_pending) {
    hipEventDestroy(p.second.first);
    hipEventDestroy(p.second.second);
  }
  for (void* p : c->allocs) dev_free(p, nullptr);
  for (int L = 0; L < kMaxLevels; ++L) drop_level_upload(c->lv[L], nullptr);
  drop_dis_upload(*c, nullptr);
  for (int L = 0; L < kMaxLevels; ++L) {
    if (c->ev_ready[L]) hipEventDestroy(c->ev_ready[L]);
    if (c->ev_done[L]) hipEventDestroy(c->ev_done[L]);
    stream_release(c->lstream[L]);
  }
  stream_release(c->stream);
  delete c;
  return MT_OK;
}

int mt_case_upload_level(mt_case* c, int level, const uint64_t* keys, const uint16_t* kinds, int64_t n) {
  CHECK_CASE(c);
  CHECK_LEVEL(c, level);
  if (c->finalized) { set_error("case already finalized"); return MT_ERR_STATE; }
  if (n < 0 || (n && (!keys || !kinds))) { set_error("bad key array"); return MT_ERR_ARG; }
  if (n > (int64_t)1 << 30) { set_error("too many cells on one level"); return MT_ERR_ARG; }
  // an interface cell is either on the coarse side (F2C, ghost_F2C) or on the
  // fine side (C2F, ghost_C2F) of a level boundary, never both
  const uint16_t coarse_side = MT_KIND_F2C | MT_KIND_GHOST_F2C, fine_side = MT_KIND_C2F | MT_KIND_GHOST_C2F;
  for (int64_t i = 0; i < n; ++i) {
    if ((kinds[i] & coarse_side) && (kinds[i] & fine_side)) {
      set_error("level " + std::to_string(level) + ": cell " + std::to_string(i) +
                " is on both sides of a level interface (refinement shells too close)");
      return MT_ERR_ARG;
    }
    // F2C cells are taken out of lat_f (Lat_Voxelize.cpp:628-632) and C2F cells are
    // never in it: a fluid cell of either kind would be updated by two levels
    if ((kinds[i] & MT_KIND_FLUID) && (kinds[i] & (MT_KIND_F2C | MT_KIND_C2F))) {
      set_error("level " + std::to_string(level) + ": cell " + std::to_string(i) +
                " is both a fluid cell and an interface (F2C / C2F) cell");
      return MT_ERR_ARG;
    }
  }
  uint64_t mx = 0;
  for (int64_t i = 0; i < n; ++i) mx = std::max(mx, keys[i]);
  // staged on the device at once (the caller may free its arrays on return)
  DevLevel& l = c->lv[level];
  drop_level_upload(l, c->stream);
  if (n > 0) {
    MT_HIP(dev_malloc((void**)&l.up_keys, 8 * n, c->stream));
    MT_HIP(dev_malloc((void**)&l.up_kinds, 2 * n, c->stream));
    MT_HIP(hipMemcpyAsync(l.up_keys, keys, 8 * n, hipMemcpyHostToDevice, c->stream));
    MT_HIP(hipMemcpyAsync(l.up_kinds, kinds, 2 * n, hipMemcpyHostToDevice, c->stream));
    MT_HIP(hipStreamSynchronize(c->stream));
  }
  l.n_up = n;
  l.up_max = mx;
  return MT_OK;
}

int mt_case_upload_dis(mt_case* c, const uint64_t* keys, const double* vals, int64_t n) {
  CHECK_CASE(c);
  if (c->finalized) { set_error("case already finalized"); return MT_ERR_STATE; }
  if (n < 0 || (n && (!keys || !vals))) { set_error("bad dis array"); return MT_ERR_ARG; }
  uint64_t mx = 0;
  for (int64_t i = 0; i < n; ++i) mx = std::max(mx, keys[i]);
  drop_dis_upload(*c, c->stream);
  if (n > 0) {
    const int64_t m = c->Q - 1;
    MT_HIP(dev_malloc((void**)&c->up_dis_keys, 8 * n, c->stream));
    MT_HIP(dev_malloc((void**)&c->up_dis_vals, 8 * n * m, c->stream));
    MT_HIP(hipMemcpyAsync(c->up_dis_keys, keys, 8 * n, hipMemcpyHostToDevice, c->stream));
    MT_HIP(hipMemcpyAsync(c->up_dis_vals, vals, 8 * n * m, hipMemcpyHostToDevice, c->stream));
    MT_HIP(hipStreamSynchronize(c->stream));
  }
  c->n_up_dis = n;
  c->up_dis_max = mx;
  return MT_OK;
}

static int check_member(const mt_case_desc& d, const mt_member_desc& m, std::string& why) {
  if (m.tau0 <= 0.5) { why = "tau0 must exceed 1/2"; return MT_ERR_ARG; }
  for (int f = 0; f < 6; ++f)
    if (m.face_kind[f] < MT_BC_NONE || m.face_kind[f] > MT_BC_NEE) { why = "bad face kind"; return MT_ERR_ARG; }
  for (int a = 0; a < 3; ++a) {
    bool p0 = m.face_kind[2 * a] == MT_BC_PERIODIC, p1 = m.face_kind[2 * a + 1] == MT_BC_PERIODIC;
    if (p0 != p1) { why = "periodic faces must come in pairs"; return MT_ERR_ARG; }
    if (p0 && (m.dims0[a] % 4)) { why = "periodic axis extent must be a multiple of 4"; return MT_ERR_ARG; }
    if (m.origin0[a] % 4 || m.origin0[a] < 0) { why = "member origin must be a non-negative multiple of 4"; return MT_ERR_ARG; }
    if (m.dims0[a] <= 0 || m.origin0[a] + m.dims0[a] > d.dims0[a]) { why = "member box outside the case"; return MT_ERR_ARG; }
  }
  return MT_OK;
}

int mt_case_set_members(mt_case* c, const mt_member_desc* m, int32_t n) {
  CHECK_CASE(c);
  if (c->finalized) { set_error("case already finalized"); return MT_ERR_STATE; }
  if (n <= 0 || !m) { set_error("no members"); return MT_ERR_ARG; }
  for (int32_t i = 0; i < n; ++i) {
    std::string why;
    if (check_member(c->d, m[i], why) != MT_OK) {
      set_error("member " + std::to_string(i) + ": " + why);
      return MT_ERR_ARG;
    }
  }
  // boxes are disjoint in whole bricks: each brick belongs to one member
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = i + 1; j < n; ++j) {
      bool overlap = true;
      for (int a = 0; a < 3; ++a) {
        const int64_t ai = m[i].origin0[a], bi = ai + m[i].dims0[a];
        const int64_t aj = m[j].origin0[a], bj = aj + m[j].dims0[a];
        // brick-rounded extents
        if ((bi + 3) / 4 * 4 <= aj || (bj + 3) / 4 * 4 <= ai) overlap = false;
      }
      if (overlap) {
        set_error("member boxes " + std::to_string(i) + " and " + std::to_string(j) + " share a brick");
        return MT_ERR_ARG;
      }
    }
  c->h_mdesc.assign(m, m + n);
  return MT_OK;
}

int mt_case_upload_markers(mt_case* c, int64_t n, const uint64_t* keys27, const double* w27, const double* dv,
                           const double* ub3) {
  CHECK_CASE(c);
  if (c->finalized) { set_error("case already finalized"); return MT_ERR_STATE; }
  if (n < 0 || (n && (!keys27 || !w27 || !dv || !ub3))) { set_error("bad marker arrays"); return MT_ERR_ARG; }
  c->n_mk = n;
  c->h_mk_keys.assign(keys27, keys27 + 27 * n);
  c->h_mk_w.assign(w27, w27 + 27 * n);
  c->h_mk_dv.assign(dv, dv + n);
  for (int a = 0; a < 3 && n; ++a) c->mk_ub[a] = ub3[a];
  return MT_OK;
}

int mt_markers_move(mt_case* c, int64_t n, const double* xyz, const double* ub) {
  CHECK_FINAL(c);
  if (!c->n_mk) { set_error("the case has no markers"); return MT_ERR_STATE; }
  if (n != c->n_mk || !xyz) { set_error("positions of all n markers needed"); return MT_ERR_ARG; }
  TRY(settle(*c));
  MT_HIP(hipMemcpyAsync(c->mv_x, xyz, 8 * 3 * n, hipMemcpyHostToDevice, c->stream));
  if (ub)
    MT_HIP(hipMemcpyAsync(c->mv_v, ub, 8 * 3 * n, hipMemcpyHostToDevice, c->stream));
  else
    MT_HIP(hipMemcpyAsync(c->mv_v, c->mk_v, 8 * 3 * n, hipMemcpyDeviceToDevice, c->stream));
  return move_markers(*c, false, 0.0, nullptr, nullptr, nullptr);
}

int mt_markers_rigid(mt_case* c, double dt, const double* center3, const double* vel3, const double* omega3) {
  CHECK_FINAL(c);
  if (!c->n_mk) { set_error("the case has no markers"); return MT_ERR_STATE; }
  if (!c->mk_x_set) { set_error("marker positions unknown: set them with mt_markers_move first"); return MT_ERR_STATE; }
  if (!center3 || !vel3 || !omega3) { set_error("null motion argument"); return MT_ERR_ARG; }
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(center3[a]) || !std::isfinite(vel3[a]) || !std::isfinite(omega3[a]) || !std::isfinite(dt)) {
      set_error("non-finite motion argument");
      return MT_ERR_ARG;
    }
  TRY(settle(*c));
  return move_markers(*c, true, dt, center3, vel3, omega3);
}

int mt_case_remap(mt_case* dst, mt_case* src, const double* u_fill3, int64_t* counts4) {
  CHECK_FINAL(dst);
  CHECK_FINAL(src);
  if (dst == src) { set_error("remap needs two cases"); return MT_ERR_ARG; }
  if (dst->Q != src->Q || dst->prec != src->prec || dst->Lmax != src->Lmax) {
    set_error("the cases differ in Q, precision or number of levels");
    return MT_ERR_ARG;
  }
  if (dst->h_mem.size() != 1 || src->h_mem.size() != 1) { set_error("remap takes single-member cases"); return MT_ERR_ARG; }
  const double zero[3] = {0, 0, 0};
  const double* fill = u_fill3 ? u_fill3 : zero;
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(fill[a])) { set_error("non-finite fill velocity"); return MT_ERR_ARG; }
  TRY(settle(*src));
  TRY(settle(*dst));
  MT_HIP(hipStreamSynchronize(src->stream));
  MT_HIP(hipStreamSynchronize(dst->stream));
  int64_t cnt[4];
  TRY(launch_remap(*dst, *src, fill, cnt));
  if (counts4)
    for (int j = 0; j < 4; ++j) counts4[j] = cnt[j];
  return MT_OK;
}

int mt_markers_get(mt_case* c, int64_t n, double* xyz, double* ub) {
  CHECK_FINAL(c);
  if (!c->n_mk) { set_error("the case has no markers"); return MT_ERR_STATE; }
  if (n != c->n_mk) { set_error("n differs from the case's marker count"); return MT_ERR_ARG; }
  if (xyz && !c->mk_x_set) { set_error("marker positions unknown: set them with mt_markers_move first"); return MT_ERR_STATE; }
  if (xyz) MT_HIP(hipMemcpyAsync(xyz, c->mk_x, 8 * 3 * n, hipMemcpyDeviceToHost, c->stream));
  if (ub) MT_HIP(hipMemcpyAsync(ub, c->mk_v, 8 * 3 * n, hipMemcpyDeviceToHost, c->stream));
  MT_HIP(hipStreamSynchronize(c->stream));
  return MT_OK;
}

int mt_case_finalize(mt_case* c) {
  CHECK_CASE(c);
  if (c->finalized) { set_error("case already finalized"); return MT_ERR_STATE; }
  if (c->n_mk && c->h_mdesc.size() > 1) { set_error("markers need a single-member case"); return MT_ERR_ARG; }
  TRY(build_members(*c));
  TRY(build_tables(*c));
  TRY(build_markers(*c));
  int64_t max_fp = 1;
  for (int L = 0; L <= c->Lmax; ++L) {
    DevLevel& l = c->lv[L];
    int64_t ns = l.n_slots;
    l.f = c->alloc<char>(ns * c->Q * (int64_t)c->pop_bytes());
    for (int j = 0; j < 2; ++j) {
      l.fcol[j] = c->alloc<char>(ns * c->Q * (int64_t)c->pop_bytes());
      l.rho[j] = c->alloc<char>(ns * (int64_t)c->real_bytes());
      l.u[j] = c->alloc<char>(3 * ns * (int64_t)c->real_bytes());
      if (!l.fcol[j] || !l.rho[j] || !l.u[j]) { set_error("out of device memory"); return MT_ERR_HIP; }
    }
    if (!l.f) { set_error("out of device memory"); return MT_ERR_HIP; }
    max_fp = std::max<int64_t>(max_fp, (ns + 255) / 256);
    max_fp = std::max<int64_t>(max_fp, l.n_active);
    max_fp = std::max<int64_t>(max_fp, (l.n_ibl + 31) / 32);
  }
  c->force_part_cap = max_fp;
  c->force_part = c->alloc<double>(3 * max_fp);
  c->force_acc = c->alloc<double>(8);
  c->force_slots = c->alloc<double>(3 * max_fp);
  c->nee_tmp = c->alloc<double>((int64_t)c->Q * std::max<int64_t>(1, c->lv[0].n_bcres));
  c->red_cap = 5 * 1024;
  c->red_tmp = c->alloc<double>(c->red_cap);
  c->member_bad = c->alloc<unsigned long long>((int64_t)c->h_mem.size());
  if (!c->force_part || !c->force_acc || !c->force_slots || !c->nee_tmp || !c->red_tmp || !c->member_bad) { set_error("out of device memory"); return MT_ERR_HIP; }
  MT_HIP(hipMemsetAsync(c->force_acc, 0, 8 * sizeof(double), c->stream));
  MT_HIP(hipMemsetAsync(c->force_slots, 0, 3 * max_fp * sizeof(double), c->stream));
  TRY(init_state(*c));
  MT_HIP(hipStreamSynchronize(c->stream));
  c->finalized = true;
  return MT_OK;
}

int mt_case_level_info(mt_case* c, int level, int64_t* n_cells, int64_t* n_bricks) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  if (n_cells) *n_cells = c->lv[level].n_cells;
  if (n_bricks) *n_bricks = c->lv[level].n_bricks;
  return MT_OK;
}

int mt_find(mt_case* c, int level, const uint64_t* keys, int64_t n, int32_t* slot_out) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  if (n <= 0) return MT_OK;
  uint64_t* dk = nullptr;
  int32_t* ds = nullptr;
  MT_HIP(dev_malloc((void**)&dk, 8 * n, c->stream));
  MT_HIP(dev_malloc((void**)&ds, 4 * n, c->stream));
  MT_HIP(hipMemcpyAsync(dk, keys, 8 * n, hipMemcpyHostToDevice, c->stream));
  int r = find_slots(*c, level, dk, n, ds);
  if (r == MT_OK) {
    hipError_t e = hipMemcpyAsync(slot_out, ds, 4 * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); r = MT_ERR_HIP; }
  }
  hipStreamSynchronize(c->stream);
  dev_free(dk, c->stream);
  dev_free(ds, c->stream);
  return r;
}

int mt_find_noalloc(mt_case* c, int level, const uint64_t* keys, int64_t n, int32_t* slot_out) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  if (n < 0) { set_error("mt_find_noalloc: negative key count"); return MT_ERR_ARG; }
  if (n > MT_FIND_NOALLOC_MAX_KEYS) {
    set_error("mt_find_noalloc: " + std::to_string(n) + " keys exceed its scratch capacity of " +
              std::to_string(MT_FIND_NOALLOC_MAX_KEYS) + " keys (use mt_find)");
    return MT_ERR_ARG;
  }
  if (n == 0) return MT_OK;
  if (!keys || !slot_out) { set_error("mt_find_noalloc: null keys or output"); return MT_ERR_ARG; }
  unsigned char* base = nullptr;
  TRY(scratch_for_stream(c->stream, &base));
  uint64_t* dk = reinterpret_cast<uint64_t*>(base);
  int32_t* ds = reinterpret_cast<int32_t*>(base + 8 * (size_t)MT_FIND_NOALLOC_MAX_KEYS);
  // every slot the lookup returns is written by it (found or -1): nothing to clear
  MT_HIP(hipMemcpyAsync(dk, keys, 8 * n, hipMemcpyHostToDevice, c->stream));
  TRY(find_slots(*c, level, dk, n, ds));
  MT_HIP(hipMemcpyAsync(slot_out, ds, 4 * n, hipMemcpyDeviceToHost, c->stream));
  MT_HIP(hipStreamSynchronize(c->stream));
  return MT_OK;
}

static int field_io(mt_case* c, int level, int field, const uint64_t* keys, int64_t n, double* out,
                    const double* in) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  if (field < MT_FIELD_F || field > MT_FIELD_U) { set_error("bad field"); return MT_ERR_ARG; }
  if (n <= 0) return MT_OK;
  int rows = field <= MT_FIELD_FCOL ? c->Q : (field == MT_FIELD_RHO ? 1 : 3);
  uint64_t* dk = nullptr;
  int32_t* ds = nullptr;
  double* dv = nullptr;
  MT_HIP(dev_malloc((void**)&dk, 8 * n, c->stream));
  MT_HIP(dev_malloc((void**)&ds, 4 * n, c->stream));
  MT_HIP(dev_malloc((void**)&dv, 8 * n * rows, c->stream));
  MT_HIP(hipMemcpyAsync(dk, keys, 8 * n, hipMemcpyHostToDevice, c->stream));
  int r = find_slots(*c, level, dk, n, ds);
  std::vector<int32_t> hs(n);
  if (r == MT_OK) {
    hipError_t e = hipMemcpyAsync(hs.data(), ds, 4 * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); r = MT_ERR_HIP; }
  }
  if (r == MT_OK)
    for (int64_t i = 0; i < n; ++i)
      if (hs[i] < 0) {
        char buf[96];
        snprintf(buf, sizeof buf, "key %llu absent on level %d", (unsigned long long)keys[i], level);
        set_error(buf);
        r = MT_ERR_LOOKUP;
        break;
      }
  // with the explosions inside the fused updates, an interface cell's fcol is
  // the reference's only after settling
  if (r == MT_OK && out && field == MT_FIELD_FCOL && c->xin) r = settle(*c);
  if (r == MT_OK) {
    DevLevel& l = c->lv[level];
    int buf = field == MT_FIELD_FCOL ? (c->ahead ? 1 - l.pf : l.pf) : l.pr;
    if (out) {
      r = launch_gather(*c, level, field, ds, n, dv, buf);
      if (r == MT_OK) {
        hipError_t e = hipMemcpyAsync(out, dv, 8 * n * rows, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); r = MT_ERR_HIP; }
      }
    } else {
      r = settle(*c);
      buf = field == MT_FIELD_FCOL ? l.pf : l.pr;
      hipError_t e = hipMemcpyAsync(dv, in, 8 * n * rows, hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) { set_error(hipGetErrorString(e)); r = MT_ERR_HIP; }
      if (r == MT_OK) r = launch_scatter(*c, level, field, ds, n, dv, buf);
      // rho/u live in both buffers for cells the step does not rewrite
      if (r == MT_OK && field >= MT_FIELD_RHO) r = launch_scatter(*c, level, field, ds, n, dv, 1 - buf);
      if (r == MT_OK && field == MT_FIELD_FCOL) {
        // keep the second fcol buffer coherent for cells no kernel rewrites
        r = launch_scatter(*c, level, field, ds, n, dv, 1 - buf);
      }
    }
  }
  hipStreamSynchronize(c->stream);
  dev_free(dk, c->stream);
  dev_free(ds, c->stream);
  dev_free(dv, c->stream);
  return r;
}

int mt_get_field(mt_case* c, int level, int field, const uint64_t* keys, int64_t n, double* out) {
  if (!out && n > 0) { set_error("null output"); return MT_ERR_ARG; }
  return field_io(c, level, field, keys, n, out, nullptr);
}
int mt_set_field(mt_case* c, int level, int field, const uint64_t* keys, int64_t n, const double* in) {
  if (!in && n > 0) { set_error("null input"); return MT_ERR_ARG; }
  return field_io(c, level, field, keys, n, nullptr, in);
}

int mt_macros(mt_case* c, int level) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  TRY(settle(*c));
  return launch_macros(*c, level);
}
int mt_collide(mt_case* c, int level) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  TRY(settle(*c));
  TRY(launch_collide(*c, level));
  if (level == c->Lmax) TRY(launch_markers(*c, c->lv[level].pf, c->st[level]));
  return MT_OK;
}
int mt_stream(mt_case* c, int level) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  TRY(settle(*c));
  return launch_stream(*c, level);
}
int mt_explode(mt_case* c, int level) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  if (level == 0) { set_error("explosion needs a coarser level"); return MT_ERR_ARG; }
  TRY(settle(*c));
  // the seam writes the single logical fcol; the ghost layer is mirrored into
  // the second buffer as the schedule does
  return launch_explode(*c, level);
}
int mt_coalesce(mt_case* c, int level) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  if (level == c->Lmax) { set_error("coalescence needs a finer level"); return MT_ERR_ARG; }
  TRY(settle(*c));
  return launch_coalesce(*c, level, false);
}
int mt_apply_bc(mt_case* c) {
  CHECK_FINAL(c);
  TRY(settle(*c));
  return launch_bc(*c);
}
int mt_ib_treat(mt_case* c, double* force3) {
  CHECK_FINAL(c);
  TRY(settle(*c));
  if (c->d.ib == MT_IB_NONE || c->lv[c->Lmax].n_dis == 0) {
    if (force3) force3[0] = force3[1] = force3[2] = 0;
    return MT_OK;
  }
  TRY(launch_ib(*c, true));
  if (force3) {
    double acc[8];
    MT_HIP(hipMemcpyAsync(acc, c->force_acc, sizeof acc, hipMemcpyDeviceToHost, c->stream));
    MT_HIP(hipStreamSynchronize(c->stream));
    force3[0] = acc[4];
    force3[1] = acc[5];
    force3[2] = acc[6];
  }
  return MT_OK;
}

int mt_step(mt_case* c, int n) {
  CHECK_FINAL(c);
  if (n < 0) { set_error("negative step count"); return MT_ERR_ARG; }
  if (c->d.ib == MT_IB_UNIFORM) {
    set_error("the uniform solid boundary runs through the seam entry points (mt_ib_treat), not mt_step");
    return MT_ERR_ARG;
  }
  if (n == 0) return MT_OK;
  if (!c->ahead) TRY(prologue(*c));
  if (c->xin && !c->xprimed) {
    // the first pair's explosion targets, from the coarser levels' next collision
    for (int L = 1; L <= c->Lmax; ++L) TRY(launch_explode_to(*c, L, c->lv[L - 1].pf));
    c->xprimed = true;
  }
  use_level_streams(*c, true);
  int r = MT_OK;
  if (c->diag) c->diag_steps += n;
  for (int i = 0; i < n && r == MT_OK;) {
    const int k = std::min(kStepsPerGraph, n - i);
    r = coarse_steps(*c, k, i + k == n);
    i += k;
  }
  use_level_streams(*c, false);
  if (r != MT_OK) return r;
  if (c->timing) TRY(drain_timing(*c));
  return MT_OK;
}

// Autotuner of the launch configuration.  Each candidate (fused-kernel block ->
// brick map x level concurrency) runs `steps` coarse steps of the real
// simulation after one untimed step that captures its graphs; every candidate
// computes the same values, so the tuning steps are ordinary steps.  The
// fastest candidate is kept; ms_out[i] (may be NULL) receives candidate i's
// time per coarse step, in the order (map 0, streams), (map 1, streams),
// (map 0, merged), (map 1, merged).
int mt_autotune(mt_case* c, int steps, double* ms_out) {
  CHECK_FINAL(c);
  if (steps <= 0) { set_error("steps must be positive"); return MT_ERR_ARG; }
  const bool timing = c->timing;
  c->timing = false;
  hipEvent_t e0, e1;
  MT_HIP(hipEventCreate(&e0));
  MT_HIP(hipEventCreate(&e1));
  double best = 1e300;
  int best_i = 0;
  int r = MT_OK;
  // two rounds over the candidates, each candidate's figure the faster of its
  // two: the first round also brings the clocks up
  double per_best[4] = {1e300, 1e300, 1e300, 1e300};
  for (int k = 0; k < 8 && r == MT_OK; ++k) {
    const int i = k & 3;
    c->xmap = i & 1;
    c->level_streams = i < 2;
    c->merge_levels = i >= 2;
    drop_graphs(*c);
    r = mt_step(c, 1);
    if (r != MT_OK) break;
    if (hipEventRecord(e0, c->stream) != hipSuccess) { r = MT_ERR_HIP; break; }
    r = mt_step(c, steps);
    if (r != MT_OK) break;
    if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { r = MT_ERR_HIP; break; }
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { r = MT_ERR_HIP; break; }
    per_best[i] = std::min(per_best[i], ms / (double)steps);
  }
  for (int i = 0; i < 4; ++i) {
    if (ms_out) ms_out[i] = per_best[i];
    if (per_best[i] < best) {
      best = per_best[i];
      best_i = i;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  c->xmap = best_i & 1;
  c->level_streams = best_i < 2;
  c->merge_levels = best_i >= 2;
  drop_graphs(*c);
  c->timing = timing;
  if (r == MT_ERR_HIP) set_error("autotune: HIP event failure");
  return r;
}

int mt_set_graphs(mt_case* c, int on) {
  CHECK_CASE(c);
  c->graphs = on != 0;
  return MT_OK;
}
int mt_set_schedule(mt_case* c, int schedule) {
  CHECK_CASE(c);
  if (schedule < MT_SCHED_MERGED || schedule > MT_SCHED_SERIAL) {
    set_error("unknown schedule");
    return MT_ERR_ARG;
  }
  c->merge_levels = schedule == MT_SCHED_MERGED;
  c->level_streams = schedule == MT_SCHED_STREAMS;
  drop_graphs(*c);
  return MT_OK;
}
int mt_steps_per_graph(void) { return kStepsPerGraph; }
int mt_get_schedule(mt_case* c, int32_t* schedule) {
  CHECK_FINAL(c);
  if (!schedule) { set_error("null output"); return MT_ERR_ARG; }
  *schedule = merged(*c) ? MT_SCHED_MERGED : c->level_streams ? MT_SCHED_STREAMS : MT_SCHED_SERIAL;
  return MT_OK;
}
int mt_set_timing(mt_case* c, int on) {
  CHECK_CASE(c);
  c->timing = on != 0;
  for (int L = 0; L < kMaxLevels; ++L) {
    c->t_ms[L] = 0;
    c->t_calls[L] = 0;
  }
  return MT_OK;
}
int mt_get_timing(mt_case* c, double* ms, int64_t* calls, int max_levels) {
  CHECK_CASE(c);
  TRY(drain_timing(*c));
  for (int L = 0; L < max_levels && L < kMaxLevels; ++L) {
    if (ms) ms[L] = c->t_ms[L];
    if (calls) calls[L] = c->t_calls[L];
  }
  return MT_OK;
}
int mt_sync(mt_case* c) {
  CHECK_CASE(c);
  MT_HIP(hipStreamSynchronize(c->stream));
  return MT_OK;
}

int mt_reduce(mt_case* c, int level, double* mass, double* momentum3, int64_t* nan_count) {
  CHECK_FINAL(c);
  CHECK_LEVEL(c, level);
  double out[5];
  TRY(launch_reduce(*c, level, out));
  if (mass) *mass = out[0];
  if (momentum3) {
    momentum3[0] = out[1];
    momentum3[1] = out[2];
    momentum3[2] = out[3];
  }
  if (nan_count) *nan_count = (int64_t)out[4];
  return MT_OK;
}

int mt_set_tau(mt_case* c, int32_t member, const double* tau, int32_t n_levels) {
  CHECK_FINAL(c);
  if (member < 0 || member >= (int32_t)c->h_mem.size()) { set_error("member out of range"); return MT_ERR_ARG; }
  if (!tau || n_levels != c->Lmax + 1) { set_error("tau needs one value per level"); return MT_ERR_ARG; }
  for (int L = 0; L <= c->Lmax; ++L)
    if (!(tau[L] > 0.5) || !std::isfinite(tau[L])) { set_error("tau must exceed 1/2"); return MT_ERR_ARG; }
  MemberParams& P = c->h_mem[member];
  set_member_tau(*c, P, tau);
  MT_HIP(hipMemcpyAsync(c->d_mem + member, &P, sizeof P, hipMemcpyHostToDevice, c->stream));
  MT_HIP(hipStreamSynchronize(c->stream));
  return MT_OK;
}

int mt_set_diagnostics(mt_case* c, int on) {
  CHECK_FINAL(c);
  if (c->diag != (on != 0)) {
    c->diag = on != 0;
    c->diag_steps = 0;
    drop_graphs(*c);
  }
  return MT_OK;
}

int mt_diagnostics(mt_case* c, mt_diag* out) {
  CHECK_FINAL(c);
  if (!out) { set_error("null output"); return MT_ERR_ARG; }
  if (c->Lmax >= 8) { set_error("mt_diag holds 8 levels"); return MT_ERR_ARG; }
  TRY(settle(*c));
  std::memset(out, 0, sizeof *out);
  double gu = 0, gr = 0;
  for (int L = 0; L <= c->Lmax; ++L) {
    double d[16];
    TRY(launch_diag(*c, L, d));
    gu += d[0];
    gr += d[1];
    out->max_v[L] = d[2];
    out->n_big[L] = (int64_t)d[3];
    if (L == c->Lmax) {
      DevLevel& l = c->lv[L];
      auto key_at = [&](double slot, uint64_t* k) -> int {
        if (slot < 0) { *k = 0; return MT_OK; }
        MT_HIP(hipMemcpyAsync(k, l.key + (int64_t)slot, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
        return MT_OK;
      };
      out->max_f = d[4];
      out->max_f_q = (int32_t)d[6];
      out->max_fcol = d[7];
      out->max_fcol_q = (int32_t)d[9];
      out->min_rho_f = d[10];
      out->min_rho_fcol = d[12];
      out->n_neg_f = (int64_t)d[13];
      out->n_neg_fcol = (int64_t)d[14];
      out->n_low_f = (int64_t)d[15];
      TRY(key_at(d[5], &out->max_f_key));
      TRY(key_at(d[8], &out->max_fcol_key));
      TRY(key_at(d[11], &out->worst_key));
      uint16_t kd = 0;
      if (d[11] >= 0) MT_HIP(hipMemcpyAsync(&kd, l.kind + (int64_t)d[11], 2, hipMemcpyDeviceToHost, c->stream));
      MT_HIP(hipStreamSynchronize(c->stream));
      out->worst_is_dis = (kd & MT_KIND_DIS) ? 1 : 0;
    }
  }
  out->velocity_error = std::sqrt(gu) / std::sqrt(gr + 1e-16);
  out->valid_v_old = (c->diag && c->diag_steps > 0) ? 1 : 0;
  return MT_OK;
}

int mt_scan_members(mt_case* c, int32_t n_members, int64_t* bad_cells) {
  CHECK_FINAL(c);
  if (n_members != (int32_t)c->h_mem.size() || !bad_cells) {
    set_error("mt_scan_members: n_members must be the case's member count and the output non-null");
    return MT_ERR_ARG;
  }
  TRY(settle(*c));
  return launch_member_scan(*c, bad_cells);
}

int mt_ib_force(mt_case* c, double* sum3, int64_t* count, int reset) {
  CHECK_FINAL(c);
  double acc[8];
  MT_HIP(hipMemcpyAsync(acc, c->force_acc, sizeof acc, hipMemcpyDeviceToHost, c->stream));
  std::vector<double> slots(3 * c->force_part_cap);
  MT_HIP(hipMemcpyAsync(slots.data(), c->force_slots, sizeof(double) * slots.size(), hipMemcpyDeviceToHost,
                        c->stream));
  MT_HIP(hipStreamSynchronize(c->stream));
  double s[3] = {acc[0], acc[1], acc[2]};
  for (int64_t b = 0; b < c->force_part_cap; ++b)
    for (int a = 0; a < 3; ++a) s[a] += slots[3 * b + a];
  if (sum3) {
    sum3[0] = s[0];
    sum3[1] = s[1];
    sum3[2] = s[2];
  }
  if (count) *count = (int64_t)(acc[3] + acc[7]);
  if (reset) {
    c->mk_acc_ahead = false;
    MT_HIP(hipMemsetAsync(c->force_acc, 0, 4 * sizeof(double), c->stream));
    MT_HIP(hipMemsetAsync(c->force_acc + 7, 0, sizeof(double), c->stream));
    MT_HIP(hipMemsetAsync(c->force_slots, 0, sizeof(double) * slots.size(), c->stream));
    MT_HIP(hipStreamSynchronize(c->stream));
  }
  return MT_OK;
}

int mt_case_copy_member(mt_case* dst, int32_t dst_member, mt_case* src, int32_t src_member) {
  CHECK_FINAL(dst);
  CHECK_FINAL(src);
  if (dst_member < 0 || dst_member >= (int32_t)dst->h_mem.size() || src_member < 0 ||
      src_member >= (int32_t)src->h_mem.size()) {
    set_error("member index out of range");
    return MT_ERR_ARG;
  }
  TRY(settle(*src));
  TRY(settle(*dst));
  MT_HIP(hipStreamSynchronize(src->stream));
  return copy_member(*dst, dst_member, *src, src_member);
}

int mt_ib_force_members(mt_case* c, double* sum3, int32_t n_members, int64_t* count, int reset) {
  CHECK_FINAL(c);
  const int32_t nm = (int32_t)c->h_mem.size();
  if (n_members != nm || (!sum3 && nm)) { set_error("member count mismatch"); return MT_ERR_ARG; }
  double acc[8];
  MT_HIP(hipMemcpyAsync(acc, c->force_acc, sizeof acc, hipMemcpyDeviceToHost, c->stream));
  std::vector<double> slots(3 * c->force_part_cap);
  MT_HIP(hipMemcpyAsync(slots.data(), c->force_slots, sizeof(double) * slots.size(), hipMemcpyDeviceToHost,
                        c->stream));
  DevLevel& l = c->lv[c->Lmax];
  std::vector<int32_t> act(l.n_active), bm(l.n_bricks);
  if (l.n_active) {
    MT_HIP(hipMemcpyAsync(act.data(), l.active, 4 * l.n_active, hipMemcpyDeviceToHost, c->stream));
    MT_HIP(hipMemcpyAsync(bm.data(), l.bmember, 4 * l.n_bricks, hipMemcpyDeviceToHost, c->stream));
  }
  MT_HIP(hipStreamSynchronize(c->stream));
  if (acc[3] != 0 && nm > 1) {
    set_error("per-member forces cover mt_step only (mt_ib_treat was called on a packed case)");
    return MT_ERR_STATE;
  }
  for (int32_t m = 0; m < nm; ++m)
    for (int a = 0; a < 3; ++a) sum3[3 * m + a] = 0;
  for (int a = 0; a < 3; ++a) sum3[a] += acc[a];
  // the fused update's slot of block bi belongs to brick active[bi]
  for (int64_t b = 0; b < l.n_active && b < c->force_part_cap; ++b)
    for (int a = 0; a < 3; ++a) sum3[3 * bm[act[b]] + a] += slots[3 * b + a];
  if (count) *count = (int64_t)(acc[3] + acc[7]);
  if (reset) {
    c->mk_acc_ahead = false;
    MT_HIP(hipMemsetAsync(c->force_acc, 0, 4 * sizeof(double), c->stream));
    MT_HIP(hipMemsetAsync(c->force_acc + 7, 0, sizeof(double), c->stream));
    MT_HIP(hipMemsetAsync(c->force_slots, 0, sizeof(double) * slots.size(), c->stream));
    MT_HIP(hipStreamSynchronize(c->stream));
  }
  return MT_OK;
}

}  // extern "C"
