// mt_serve.hip -- the continuous-batching server behind the C-ABI (no
// counterpart in the reference, which runs one flow per process).
//
// Every running request is one member of a single packed case, so all of them
// share each level sub-step's launch.  Each request keeps a slot (a box of
// kSlot level-0 cells along x) for its life; its keys are translated into the
// packed coordinates once, at submission.  mt_server_step runs one chunk:
//   * a repack is due when requests wait and `admit_every` chunks have passed
//     since the last one, or at once when no running request is unfinished;
//     it admits waiting requests in order while their device footprint fits
//     the budget, builds the new packed case, and copies every continuing
//     member's state into it on the device (mt_case_copy_member), so a member
//     computes exactly what it computes alone;
//   * the packed case advances `chunk` coarse steps; each unfinished member's
//     drag over the chunk is recorded, and a member finishes at max_steps or
//     when its drag moved by less than the tolerance since the previous chunk;
//     finished members keep stepping with the batch until the next repack
//     (their results are final) and wait in the results queue for
//     mt_server_poll;
//   * after every chunk the device scans each member for non-finite
//     populations (mt_scan_members, the reference's NaN checks per member); a
//     member that has any, or whose drag is not finite, finishes at once with
//     the status MT_REQ_DIVERGED;
//   * a repack that evicts the finished and diverged members is due, whatever
//     the admission interval, when it pays: their share of the batch's
//     footprint times the chunks the batch may still run (the most any running
//     member has left) reaches kRepackChunks, the cost of a repack in chunks.
#include <algorithm>
#include <cmath>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "mt_case.h"
#include "mt_morton.h"

namespace mt {
namespace {

constexpr int kSlotX = 100, kSlotY = 48, kSlotZ = 48;  // level-0 cells per slot
// Evicting pays when the retired share x the chunks left reaches this many
// chunks.  A repack (build the packed case's tables, copy the running members'
// state, capture the new case's graphs on its first chunk) costs about 30 ms on
// one MI355X against about 38 ms per 100-step chunk of a full batch of served
// requests; thresholds of 0.75, 1.5 and 3 chunks served alike within the
// run-to-run spread.
constexpr double kRepackChunks = 1.5;

// a level-L key (keys of every level live on the finest grid, shifted by
// 3 (Lmax - L)) moved by (dx, dy, dz) level-L cells
uint64_t shift_key(uint64_t k, int L, int Lmax, uint64_t dx) {
  const int sh = Lmax - L;
  const uint64_t x = (compact3(k) >> sh) + dx, y = compact3(k >> 1) >> sh, z = compact3(k >> 2) >> sh;
  return spread3(x << sh) | (spread3(y << sh) << 1) | (spread3(z << sh) << 2);
}

struct Request {
  int64_t id = 0;
  mt_case_desc d{};
  mt_member_desc m{};
  std::vector<std::vector<uint64_t>> keys;  // per level, the request's own coordinates
  std::vector<std::vector<uint16_t>> kinds;
  std::vector<uint64_t> dis_keys;
  std::vector<double> dis_vals;
  int32_t max_steps = 0;
  double tol = 0, cd_U = 0, cd_d = 1;
  int64_t footprint = 0;
  int slot = -1;
  int32_t steps = 0;
  std::vector<double> cd;
  bool done = false;
  int32_t status = MT_REQ_FINISHED;
};

}  // namespace

struct Server {
  int32_t precision = MT_FP32_SHIFTED, chunk = 100, admit_every = 1;
  int64_t budget = 0;
  bool have_model = false;
  mt_case_desc model{};  // Q, Lmax, collision, IB, options shared by every request
  std::deque<std::unique_ptr<Request>> waiting;
  std::vector<std::unique_ptr<Request>> running;  // member order of `batch`
  std::vector<std::unique_ptr<Request>> results;
  mt_case* batch = nullptr;
  int since = 1 << 30;
  std::string broken;  // set when a repack or a step failed; every later call reports it
};

}  // namespace mt

using namespace mt;

struct mt_server : public mt::Server {};

namespace {

int64_t footprint_of(const Request& r, int precision) {
  const int64_t pb = precision == MT_FP64 ? 8 : (precision == MT_FP32_SHIFTED ? 4 : 2);
  const int64_t rb = precision == MT_FP64 ? 8 : 4;
  int64_t cells = 0;
  for (auto& k : r.keys) cells += (int64_t)k.size();
  return (int64_t)(cells * 1.3) * (3 * r.d.Q * pb + 8 * rb + 24);
}

#define TRYS(x)                  \
  do {                           \
    int r_ = (x);                \
    if (r_ != MT_OK) return r_;  \
  } while (0)

// the packed case of `members` (slots already assigned); on success *out owns it
int build_batch(mt_server& s, const std::vector<Request*>& members, mt_case** out) {
  int max_slot = 0;
  for (auto* r : members) max_slot = std::max(max_slot, r->slot);
  mt_case_desc d = s.model;
  d.precision = s.precision;
  d.dims0[0] = (max_slot + 1) * kSlotX;
  d.dims0[1] = kSlotY;
  d.dims0[2] = kSlotZ;
  for (int f = 0; f < 6; ++f) {
    d.face_kind[f] = MT_BC_NONE;
    for (int a = 0; a < 3; ++a) d.face_u[f][a] = 0;
  }
  mt_case* c = nullptr;
  TRYS(mt_case_create(&d, &c));
  std::vector<mt_member_desc> md;
  for (auto* r : members) {
    mt_member_desc m = r->m;
    m.origin0[0] = r->slot * kSlotX;
    m.origin0[1] = m.origin0[2] = 0;
    md.push_back(m);
  }
  int rc = mt_case_set_members(c, md.data(), (int32_t)md.size());
  const int Lmax = d.Lmax;
  for (int L = 0; L <= Lmax && rc == MT_OK; ++L) {
    std::vector<uint64_t> keys;
    std::vector<uint16_t> kinds;
    for (auto* r : members) {
      const uint64_t dx = (uint64_t)(r->slot * kSlotX) << L;
      for (uint64_t k : r->keys[L]) keys.push_back(shift_key(k, L, Lmax, dx));
      kinds.insert(kinds.end(), r->kinds[L].begin(), r->kinds[L].end());
    }
    rc = mt_case_upload_level(c, L, keys.data(), kinds.data(), (int64_t)keys.size());
  }
  if (rc == MT_OK) {
    std::vector<uint64_t> dk;
    std::vector<double> dv;
    for (auto* r : members) {
      const uint64_t dx = (uint64_t)(r->slot * kSlotX) << Lmax;
      for (uint64_t k : r->dis_keys) dk.push_back(shift_key(k, Lmax, Lmax, dx));
      dv.insert(dv.end(), r->dis_vals.begin(), r->dis_vals.end());
    }
    if (!dk.empty()) rc = mt_case_upload_dis(c, dk.data(), dv.data(), (int64_t)dk.size());
  }
  if (rc == MT_OK) rc = mt_case_finalize(c);
  if (rc == MT_OK) rc = mt_set_graphs(c, 1);
  if (rc != MT_OK) {
    std::string e = mt_last_error();
    mt_case_destroy(c);
    set_error(e);
    return rc;
  }
  *out = c;
  return MT_OK;
}

// `old` is the current batch's member order (raw pointers taken before the
// continuing requests moved into `next`)
int repack(mt_server& s, std::vector<std::unique_ptr<Request>> next, const std::vector<Request*>& old) {
  std::vector<Request*> members;
  for (auto& r : next) {
    if (r->slot < 0) r->slot = -2;  // placeholder until assigned below
    members.push_back(r.get());
  }
  // new requests take the lowest slots no continuing request holds
  for (auto* r : members)
    if (r->slot == -2) {
      for (int k = 0;; ++k) {
        bool used = false;
        for (auto* o : members)
          if (o->slot == k) used = true;
        if (!used) {
          r->slot = k;
          break;
        }
      }
    }
  mt_case* nb = nullptr;
  if (!members.empty()) {
    int rc = build_batch(s, members, &nb);
    if (rc != MT_OK) return rc;
    for (size_t j = 0; j < members.size(); ++j)
      for (size_t i = 0; i < old.size(); ++i)
        if (old[i] == members[j]) {
          rc = mt_case_copy_member(nb, (int32_t)j, s.batch, (int32_t)i);
          if (rc != MT_OK) {
            std::string e = mt_last_error();
            mt_case_destroy(nb);
            set_error(e);
            return rc;
          }
        }
  }
  if (s.batch) mt_case_destroy(s.batch);
  s.batch = nb;
  // finished members are dropped here (their results went out already)
  std::vector<std::unique_ptr<Request>> keep;
  for (auto& r : next) keep.push_back(std::move(r));
  s.running = std::move(keep);
  s.since = 0;
  return MT_OK;
}

}  // namespace

extern "C" {

int mt_server_create(int32_t precision, int32_t chunk_steps, int32_t admit_every, int64_t hbm_budget,
                     mt_server** out) {
  if (!out) { set_error("null output"); return MT_ERR_ARG; }
  *out = nullptr;
  if (precision < MT_FP64 || precision > MT_FP16_STORAGE) { set_error("bad precision"); return MT_ERR_ARG; }
  if (chunk_steps <= 0 || admit_every <= 0 || hbm_budget <= 0) { set_error("chunk, admission interval and budget must be positive"); return MT_ERR_ARG; }
  mt_server* s = new mt_server();
  s->precision = precision;
  s->chunk = chunk_steps;
  s->admit_every = admit_every;
  s->budget = hbm_budget;
  *out = s;
  return MT_OK;
}

int mt_server_submit(mt_server* s, mt_case* req, int64_t id, int32_t max_steps, double drag_tol, double cd_U,
                     double cd_d) {
  if (!s || !req) {
    if (req) mt_case_destroy(req);
    set_error("null server or request");
    return MT_ERR_ARG;
  }
  if (!s->broken.empty()) {
    mt_case_destroy(req);
    set_error(s->broken);
    return MT_ERR_STATE;
  }
  Case& c = *req;
  auto bad = [&](const char* why) {
    mt_case_destroy(req);
    set_error(why);
    return MT_ERR_ARG;
  };
  if (c.finalized) return bad("a request must not be finalized (the server packs it)");
  if (c.h_mdesc.size() != 1) return bad("a request is a single-member case");
  if (c.n_mk) return bad("requests with markers are not served");
  if (max_steps <= 0) return bad("max_steps must be positive");
  if (cd_U == 0.0 || cd_d <= 0.0) return bad("drag reference velocity and diameter must be nonzero");
  const mt_case_desc& d = c.d;
  if (d.dims0[0] > kSlotX - 4 || d.dims0[1] > kSlotY || d.dims0[2] > kSlotZ)
    return bad("request does not fit a serving slot (96 x 48 x 48 level-0 cells)");
  if (!s->have_model) {
    s->model = d;
    s->have_model = true;
  } else if (d.Q != s->model.Q || d.Lmax != s->model.Lmax || d.collision != s->model.collision ||
             d.ib != s->model.ib || d.options != s->model.options) {
    return bad("requests must share Q, Lmax, collision, IB method and options");
  }
  auto r = std::make_unique<Request>();
  r->id = id;
  r->d = d;
  r->m = c.h_mdesc[0];
  // the request's level sets and dis, staged on the device by its uploads,
  // come back to the host: every repack translates them into the packed case
  hipError_t e = hipSuccess;
  for (int L = 0; L <= d.Lmax; ++L) {
    const DevLevel& l = c.lv[L];
    const int64_t n = std::max<int64_t>(l.n_up, 0);
    std::vector<uint64_t> keys(n);
    std::vector<uint16_t> kinds(n);
    if (n && e == hipSuccess) e = hipMemcpyAsync(keys.data(), l.up_keys, 8 * n, hipMemcpyDeviceToHost, c.stream);
    if (n && e == hipSuccess) e = hipMemcpyAsync(kinds.data(), l.up_kinds, 2 * n, hipMemcpyDeviceToHost, c.stream);
    r->keys.push_back(std::move(keys));
    r->kinds.push_back(std::move(kinds));
  }
  const int64_t nd = c.n_up_dis;
  r->dis_keys.resize(nd);
  r->dis_vals.resize(nd * (d.Q - 1));
  if (nd && e == hipSuccess) e = hipMemcpyAsync(r->dis_keys.data(), c.up_dis_keys, 8 * nd, hipMemcpyDeviceToHost, c.stream);
  if (nd && e == hipSuccess)
    e = hipMemcpyAsync(r->dis_vals.data(), c.up_dis_vals, 8 * nd * (d.Q - 1), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (e != hipSuccess) {
    mt_case_destroy(req);
    set_error(std::string("request download: ") + hipGetErrorString(e));
    return MT_ERR_HIP;
  }
  r->max_steps = max_steps;
  r->tol = drag_tol;
  r->cd_U = cd_U;
  r->cd_d = cd_d;
  r->footprint = footprint_of(*r, s->precision);
  mt_case_destroy(req);
  s->waiting.push_back(std::move(r));
  return MT_OK;
}

int mt_server_step(mt_server* s, int32_t* n_running, int32_t* n_waiting) {
  if (!s) { set_error("null server"); return MT_ERR_ARG; }
  if (!s->broken.empty()) { set_error(s->broken); return MT_ERR_STATE; }
  int64_t active = 0, used = 0, retired = 0, left = 0;
  for (auto& r : s->running)
    if (!r->done) {
      ++active;
      used += r->footprint;
      left = std::max<int64_t>(left, (r->max_steps - r->steps + s->chunk - 1) / s->chunk);
    } else {
      retired += r->footprint;
    }
  const bool due = !s->waiting.empty() && (s->since >= s->admit_every || active == 0);
  const bool evict = active > 0 && retired > 0 && (double)retired / (double)(retired + used) * (double)left >= kRepackChunks;
  if (due || evict || (!s->running.empty() && active == 0)) {
    std::vector<Request*> old;
    for (auto& r : s->running) old.push_back(r.get());
    std::vector<std::unique_ptr<Request>> next;
    for (auto& r : s->running)
      if (!r->done) next.push_back(std::move(r));
    while (!s->waiting.empty()) {
      const int64_t need = s->waiting.front()->footprint;
      if (!next.empty() && used + need > s->budget) break;
      used += need;
      next.push_back(std::move(s->waiting.front()));
      s->waiting.pop_front();
    }
    int rc = repack(*s, std::move(next), old);
    if (rc != MT_OK) {
      // the running requests left with the failed repack: the server stops
      s->broken = std::string("server stopped: a repack failed: ") + mt_last_error();
      s->running.clear();
      if (s->batch) mt_case_destroy(s->batch);
      s->batch = nullptr;
      set_error(s->broken);
      return rc;
    }
  }
  if (s->batch) {
    const int32_t nm = (int32_t)s->running.size();
    std::vector<double> sums(3 * (size_t)nm);
    std::vector<int64_t> bad(nm);
    int64_t calls = 0;
    int rc = mt_step(s->batch, s->chunk);
    if (rc == MT_OK) rc = mt_ib_force_members(s->batch, sums.data(), nm, &calls, 1);
    if (rc == MT_OK) rc = mt_scan_members(s->batch, nm, bad.data());
    if (rc != MT_OK) {
      // the batch's state is unknown after a failed step: the server stops
      s->broken = std::string("server stopped: a step failed: ") + mt_last_error();
      s->running.clear();
      mt_case_destroy(s->batch);
      s->batch = nullptr;
      set_error(s->broken);
      return rc;
    }
    s->since += 1;
    for (int32_t i = 0; i < nm; ++i) {
      Request& r = *s->running[i];
      if (r.done) continue;
      r.steps += s->chunk;
      const double fx = sums[3 * i] / (double)std::max<int64_t>(calls, 1);
      r.cd.push_back(2.0 * fx / (r.cd_U * r.cd_U * M_PI * r.cd_d * r.cd_d / 4.0));
      const size_t k = r.cd.size();
      const bool diverged = bad[i] > 0 || !std::isfinite(r.cd[k - 1]);
      if (diverged || r.steps >= r.max_steps || (k >= 2 && std::fabs(r.cd[k - 1] - r.cd[k - 2]) < r.tol)) {
        r.done = true;
        r.status = diverged ? MT_REQ_DIVERGED : MT_REQ_FINISHED;
        auto res = std::make_unique<Request>();
        res->id = r.id;
        res->steps = r.steps;
        res->cd = r.cd;
        res->status = r.status;
        s->results.push_back(std::move(res));
      }
    }
  }
  int32_t nr = 0;
  for (auto& r : s->running)
    if (!r->done) ++nr;
  if (n_running) *n_running = nr;
  if (n_waiting) *n_waiting = (int32_t)s->waiting.size();
  return MT_OK;
}

int mt_server_poll(mt_server* s, int64_t* ids, int32_t* steps, double* cd_last, int32_t* status,
                   int32_t max_results, int32_t* n_results) {
  if (!s || !n_results) { set_error("null server or count"); return MT_ERR_ARG; }
  int32_t n = 0;
  while (n < max_results && !s->results.empty()) {
    Request& r = *s->results.front();
    if (ids) ids[n] = r.id;
    if (steps) steps[n] = r.steps;
    if (cd_last) cd_last[n] = r.cd.empty() ? NAN : r.cd.back();
    if (status) status[n] = r.status;
    s->results.erase(s->results.begin());
    ++n;
  }
  *n_results = n;
  return MT_OK;
}

int mt_server_destroy(mt_server* s) {
  if (!s) return MT_OK;
  if (s->batch) mt_case_destroy(s->batch);
  delete s;
  return MT_OK;
}

}  // extern "C"
