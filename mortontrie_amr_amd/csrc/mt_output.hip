// mt_output.hip -- field snapshots and the output files of a run.
//
// A snapshot is one level's cells of the chosen kinds with their density and
// velocity, compacted on the device in slot order (= ascending key order) and
// copied to the host in one transfer.  The writers build the reference's files
// from snapshots:
//   * IO_TECPLOTH::write (tecplot.cpp:16-177): one FEBRICK zone whose elements are
//     every level's lat_f and lat_overlap_F2C cells (level by level, lat_f first),
//     whose nodes are the elements' corners numbered in order of first use, and
//     whose cell-centred U, V, W are the velocities times a conversion factor;
//   * the AMR node dump of Obj_Manager::initial (Obj_Manager.cpp:483-489): the
//     finest level's keys, one 64-bit binary string per line.
#include <cstdio>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "mt_case.h"
#include "mt_device.h"
#include "mt_morton.h"

namespace mt {
namespace {

// per block of 256 slots: how many match
__global__ void k_snap_count(const uint16_t* kind, int64_t ns, uint32_t mask, int32_t* counts) {
  __shared__ int32_t wsum[4];
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool hit = s < ns && (kind[s] & mask);
  const unsigned long long b = __ballot(hit);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// the matching slots of each block at its offset, in slot order
template <typename R>
__global__ void k_snap_write(const uint16_t* kind, const uint64_t* key, const R* rho, const R* u, int64_t ns,
                             uint32_t mask, const int64_t* offs, int64_t n, uint64_t* okey, double* orho, double* ou) {
  __shared__ int32_t wsum[4];
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool hit = s < ns && (kind[s] & mask);
  const unsigned long long b = __ballot(hit);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) wsum[w] = __popcll(b);
  __syncthreads();
  if (!hit) return;
  int64_t at = offs[blockIdx.x] + __popcll(b & ((1ull << l) - 1));
  for (int j = 0; j < w; ++j) at += wsum[j];
  okey[at] = key[s];
  orho[at] = (double)rho[s];
  ou[at] = (double)u[s];
  ou[n + at] = (double)u[ns + s];
  ou[2 * n + at] = (double)u[2 * ns + s];
}

struct Snap {
  std::vector<uint64_t> key;
  std::vector<double> rho, u;  // u: [3][n]
};

int snapshot(Case& c, int L, uint32_t mask, Snap& out) {
  DevLevel& l = c.lv[L];
  const int64_t ns = l.n_slots, nb = (ns + 255) / 256;
  int32_t* counts = nullptr;
  int64_t* offs = nullptr;
  MT_HIP(dev_malloc((void**)&counts, 4 * nb, c.stream));
  k_snap_count<<<nb, 256, 0, c.stream>>>(l.kind, ns, mask, counts);
  std::vector<int32_t> hc(nb);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(hc.data(), counts, 4 * nb, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  dev_free(counts, c.stream);
  MT_HIP(e);
  std::vector<int64_t> ho(nb);
  int64_t n = 0;
  for (int64_t b = 0; b < nb; ++b) {
    ho[b] = n;
    n += hc[b];
  }
  out.key.assign(n, 0);
  out.rho.assign(n, 0);
  out.u.assign(3 * n, 0);
  if (n == 0) return MT_OK;
  uint64_t* dk = nullptr;
  double *dr = nullptr, *du = nullptr;
  MT_HIP(dev_malloc((void**)&offs, 8 * nb, c.stream));
  MT_HIP(dev_malloc((void**)&dk, 8 * n, c.stream));
  MT_HIP(dev_malloc((void**)&dr, 8 * n, c.stream));
  MT_HIP(dev_malloc((void**)&du, 24 * n, c.stream));
  MT_HIP(hipMemcpyAsync(offs, ho.data(), 8 * nb, hipMemcpyHostToDevice, c.stream));
  if (c.prec == MT_FP64)
    k_snap_write<double><<<nb, 256, 0, c.stream>>>(l.kind, l.key, (const double*)l.rho[l.pr],
                                                   (const double*)l.u[l.pr], ns, mask, offs, n, dk, dr, du);
  else
    k_snap_write<float><<<nb, 256, 0, c.stream>>>(l.kind, l.key, (const float*)l.rho[l.pr], (const float*)l.u[l.pr],
                                                  ns, mask, offs, n, dk, dr, du);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out.key.data(), dk, 8 * n, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out.rho.data(), dr, 8 * n, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out.u.data(), du, 24 * n, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  dev_free(offs, c.stream);
  dev_free(dk, c.stream);
  dev_free(dr, c.stream);
  dev_free(du, c.stream);
  MT_HIP(e);
  return MT_OK;
}

// the key moved by (dx,dy,dz) cells of level L (find_x1 / find_y1 / find_z1 chains,
// header/Morton_Assist.h:229-259), with keys in finest-level units
uint64_t corner(uint64_t k, int L, int Lmax, int dx, int dy, int dz) {
  const uint64_t st = 1ull << (Lmax - L);
  const uint64_t x = compact3(k) + dx * st, y = compact3(k >> 1) + dy * st, z = compact3(k >> 2) + dz * st;
  return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

}  // namespace
}  // namespace mt

using namespace mt;

extern "C" {

int mt_snapshot(mt_case* c, int level, uint32_t kind_mask, int64_t* n, uint64_t* keys, double* rho, double* u) {
  if (!c || !c->finalized) { set_error("case not finalized"); return MT_ERR_STATE; }
  if (level < 0 || level > c->Lmax) { set_error("level out of range"); return MT_ERR_ARG; }
  if (!n) { set_error("null count"); return MT_ERR_ARG; }
  if (!kind_mask) kind_mask = MT_KIND_FLUID | MT_KIND_F2C;
  Snap s;
  int r = snapshot(*c, level, kind_mask, s);
  if (r != MT_OK) return r;
  const int64_t m = (int64_t)s.key.size();
  if (keys || rho || u) {
    if (*n < m) { set_error("snapshot buffers hold fewer cells than the level has"); return MT_ERR_ARG; }
    for (int64_t i = 0; i < m; ++i) {
      if (keys) keys[i] = s.key[i];
      if (rho) rho[i] = s.rho[i];
      if (u)
        for (int a = 0; a < 3; ++a) u[a * m + i] = s.u[a * m + i];
    }
  }
  *n = m;
  return MT_OK;
}

int mt_write_tecplot(mt_case* c, const char* path, double dx0, double convert_u) {
  if (!c || !c->finalized) { set_error("case not finalized"); return MT_ERR_STATE; }
  if (!path) { set_error("null path"); return MT_ERR_ARG; }
  const int Lmax = c->Lmax;
  // elements: per level lat_f then lat_overlap_F2C, each in ascending key order
  struct Elem {
    uint64_t key;
    int L;
    double u[3];
  };
  std::vector<Elem> el;
  for (int L = 0; L <= Lmax; ++L)
    for (uint32_t m : {(uint32_t)MT_KIND_FLUID, (uint32_t)MT_KIND_F2C}) {
      Snap s;
      int r = snapshot(*c, L, m, s);
      if (r != MT_OK) return r;
      const int64_t n = (int64_t)s.key.size();
      for (int64_t i = 0; i < n; ++i) el.push_back({s.key[i], L, {s.u[i], s.u[n + i], s.u[2 * n + i]}});
    }
  // nodes in order of first use, corners taken as the reference does
  std::unordered_map<uint64_t, int64_t> node;
  std::vector<uint64_t> node_key;
  node.reserve(el.size() * 2);
  const int first[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  for (auto& e : el)
    for (auto& o : first) {
      const uint64_t k = corner(e.key, e.L, Lmax, o[0], o[1], o[2]);
      if (node.emplace(k, (int64_t)node_key.size() + 1).second) node_key.push_back(k);
    }
  std::ofstream f(path);
  if (!f.is_open()) {
    set_error(std::string("cannot open ") + path);
    return MT_ERR_ARG;
  }
  f << "TITLE = \"Flow Field Example\"\n";
  f << "VARIABLES = \"X\", \"Y\", \"Z\", \"U\", \"V\", \"W\"\n";
  f << "ZONE, NODES=" << node_key.size() << ", ELEMENTS=" << el.size()
    << ", DATAPACKING=BLOCK, VARLOCATION=([4,5,6]=CELLCENTERED), ZONETYPE=FEBRICK\n";
  const double h = dx0 / (double)(1ull << Lmax);
  for (int a = 0; a < 3; ++a)
    for (uint64_t k : node_key) f << (double)compact3(k >> a) * h << "\n";
  for (int a = 0; a < 3; ++a)
    for (auto& e : el) f << e.u[a] * convert_u << "\n";
  // connectivity in FEBRICK order
  const int conn[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (auto& e : el) {
    for (auto& o : conn) f << node.at(corner(e.key, e.L, Lmax, o[0], o[1], o[2])) << " ";
    f << "\n";
  }
  if (!f.good()) {
    set_error(std::string("write failed: ") + path);
    return MT_ERR_HIP;
  }
  return MT_OK;
}

int mt_write_nodes(mt_case* c, const char* path) {
  if (!c || !c->finalized) { set_error("case not finalized"); return MT_ERR_STATE; }
  if (!path) { set_error("null path"); return MT_ERR_ARG; }
  Snap s;
  int r = snapshot(*c, c->Lmax, 0xFFFFu, s);
  if (r != MT_OK) return r;
  std::ofstream f(path);
  if (!f.is_open()) {
    set_error(std::string("cannot open ") + path);
    return MT_ERR_ARG;
  }
  char line[66];
  for (uint64_t k : s.key) {
    for (int b = 0; b < 64; ++b) line[b] = ((k >> (63 - b)) & 1) ? '1' : '0';
    line[64] = '\n';
    f.write(line, 65);
  }
  return f.good() ? MT_OK : MT_ERR_HIP;
}

}  // extern "C"
