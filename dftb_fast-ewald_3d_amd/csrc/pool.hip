// pool.hip -- many resident structures evaluated together: the batched and serving forms.
//
// A pool holds the resident state of many structures (one ew_handle each, all on the pool's
// stream) and a device table of their descriptors.  One evaluation of any subset is one launch
// per kernel kind (gather charges, spread, five FFT passes, pair sum, gather, energy) whose
// workgroups cover every structure of the subset; a block prefix per kind maps a workgroup to
// its structure.  Repeated evaluations of the same subset into the same buffers are replayed
// from a captured hipGraph.
//
// The per-structure work is the single-structure engine's (same kernel bodies, same variants),
// so a structure gives the same shift and energy alone and inside any batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fftewald_amd.h"
#include "engine.hpp"
#include "handle.hpp"
#include "kernels.hpp"
#include "group.hpp"
#include "pool.hpp"

using namespace ewamd;

namespace {

struct Captured {
  std::vector<int> ids;
  std::vector<const void*> ptrs;  // dq and shift per entry, then energies
  hipGraphExec_t exec = nullptr;
  void* staging = nullptr;  // map and io arrays the graph's kernels read
};

}  // namespace

struct ew_pool {
  ew_options opt{};
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;  // pair sums beside the SPME chain
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // completion of the last grouped evaluation on whichever stream it ran (ew_pool_eval_device
  // takes a caller's stream): staging, the descriptor table and the captured graphs are reused
  // or destroyed only after it
  hipEvent_t last_ev = nullptr;
  hipStream_t last_stream = nullptr;
  bool last_pending = false;
  std::string err = "no error";
  bool fp32 = false;
  int order = 0;  // B-spline order shared by every member (0 until the first member)
  long long capacity = 0;
  long long used = 0;
  std::vector<ew_handle*> members;  // by id; null = free id
  std::vector<StructDesc> desc;
  StructDesc* desc_dev = nullptr;
  size_t desc_cap = 0;
  bool desc_dirty = true;
  // staging of eager launches (map + io) and of the host-array entry point
  void* staging = nullptr;
  size_t staging_cap = 0;
  double* io_dq = nullptr;
  double* io_shift = nullptr;
  double* io_energy = nullptr;
  size_t io_atoms_cap = 0, io_struct_cap = 0;
  std::vector<Captured> graphs;

  int fail(int code, const std::string& msg) {
    err = msg;
    global_error() = msg;
    return code;
  }
  hipError_t after_last(hipStream_t s) const {
    if (!last_pending || last_stream == s) return hipSuccess;
    return hipStreamWaitEvent(s, last_ev, 0);
  }
  hipError_t mark_last(hipStream_t s) {
    const hipError_t e = hipEventRecord(last_ev, s);
    if (e == hipSuccess) {
      last_stream = s;
      last_pending = true;
    }
    return e;
  }
  void quiesce() {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamSynchronize(aux);
    if (last_pending) (void)hipEventSynchronize(last_ev);
    last_pending = false;
  }
  // destroys the captured graphs once nothing that launched them can still be running
  void drop_graphs() {
    if (!graphs.empty()) quiesce();
    for (auto& g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      if (g.staging) (void)hipFree(g.staging);
    }
    graphs.clear();
  }
};

#define PCHK(p, expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return (p)->fail(EW_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) +   \
                                          " at " #expr);                                     \
  } while (0)

namespace {

int ensure(ew_pool* P, void** buf, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return EW_OK;
  const size_t want = std::max(bytes, 2 * *cap);
  if (*buf) PCHK(P, hipFree(*buf));
  *buf = nullptr;
  *cap = 0;
  if (hipMalloc(buf, want) != hipSuccess) {
    (void)hipGetLastError();
    *buf = nullptr;
    return P->fail(EW_ERR_MEMORY, "pool staging allocation failed");
  }
  *cap = want;
  return EW_OK;
}

int sync_desc(ew_pool* P) {
  if (!P->desc_dirty) return EW_OK;
  const size_t bytes = sizeof(StructDesc) * std::max<size_t>(1, P->desc.size());
  if (P->desc_cap < bytes) {
    // the captured graphs and the last evaluation read the old table
    P->quiesce();
    P->drop_graphs();
    void* b = P->desc_dev;
    int st = ensure(P, &b, &P->desc_cap, bytes);
    P->desc_dev = static_cast<StructDesc*>(b);
    if (st) return st;
  }
  if (!P->desc.empty())
    PCHK(P, hipMemcpyAsync(P->desc_dev, P->desc.data(), sizeof(StructDesc) * P->desc.size(),
                           hipMemcpyHostToDevice, P->stream));
  PCHK(P, hipStreamSynchronize(P->stream));
  P->desc_dirty = false;
  return EW_OK;
}

// Mesh points from which a pool member's register transforms run as a batched sequence.
constexpr long long kPoolWaveMinPoints = 96LL * 96 * 96;

int stage(ew_pool* P, int count, const int* ids, const double* const* dq, double* const* shift,
          double* energies, Staged* S) {
  std::vector<IoSlot> io((size_t)count);
  for (int i = 0; i < count; ++i) io[i] = IoSlot{dq[i], shift[i], energies + i, ids[i], 0};
  // members that take register transforms on a large mesh run their FFTs as one batched
  // sequence per (mode, mesh shape, lanes), whatever else is in the launch; members with small
  // meshes take the grouped Stockham kinds, one launch per kind for all of them (a sequence per
  // small mesh shape would multiply the launches of a mixed batch).  Either way the choice is
  // the member's own, so its kernels do not depend on the rest of the launch.
  std::vector<unsigned char> by_wave((size_t)count, 0);
  std::vector<std::vector<int>> groups;
  auto key_eq = [&](const ew_handle* a, const ew_handle* b) {
    bool eq = a->wave == b->wave;
    for (int d = 0; d < 3; ++d)
      eq = eq && a->prm.K[d] == b->prm.K[d] && a->lanes[d] == b->lanes[d];
    return eq;
  };
  for (int i = 0; i < count; ++i) {
    const ew_handle* h = P->members[ids[i]];
    if (!h->wave || (long long)h->prm.K[0] * h->prm.K[1] * h->prm.K[2] < kPoolWaveMinPoints)
      continue;
    int g = 0;
    while (g < (int)groups.size() && !key_eq(P->members[ids[groups[g][0]]], h)) ++g;
    if (g == (int)groups.size()) groups.emplace_back();
    groups[g].push_back(i);
    by_wave[i] = 1;
  }
  if (!stage_group(P->desc.data(), io.data(), count, P->fp32, false, S, by_wave.data()))
    return P->fail(EW_ERR_ARGUMENT, "structure exceeds a workgroup's LDS");
  for (const std::vector<int>& g : groups) {
    std::vector<WaveIo> w;
    for (int i : g) {
      const ew_handle* h = P->members[ids[i]];
      w.push_back(WaveIo{h->grid, h->cgrid, h->inf_tab});
    }
    stage_wave_io(w, ids[g[0]], S);
  }
  return EW_OK;
}

GroupLaunch launch_of(const ew_pool* P, const Staged& S, const void* dev) {
  GroupLaunch G = group_launch_of(S, P->desc_dev, dev, P->order, P->opt.group_grid == 1);
  for (size_t g = 0; g < G.wave.size(); ++g)
    G.wave[g].l = spme_launch_of(P->members[S.wave_sid[g]]);
  return G;
}

int check_ids(ew_pool* P, int count, const int* ids) {
  if (count <= 0 || !ids) return P->fail(EW_ERR_ARGUMENT, "empty structure list");
  std::vector<char> seen(P->members.size(), 0);
  for (int i = 0; i < count; ++i) {
    const int id = ids[i];
    if (id < 0 || id >= (int)P->members.size() || !P->members[id])
      return P->fail(EW_ERR_ARGUMENT, "unknown structure id " + std::to_string(id));
    if (seen[id]) return P->fail(EW_ERR_ARGUMENT, "structure id listed twice in one evaluation");
    seen[id] = 1;
  }
  return EW_OK;
}

int enqueue(ew_pool* P, GroupLaunch G, hipStream_t s, hipStream_t aux, hipEvent_t fork_ev,
            hipEvent_t join_ev) {
  G.aux = aux;
  G.fork_ev = fork_ev;
  G.join_ev = join_ev;
  PCHK(P, P->fp32 ? launch_group<float>(G, s) : launch_group<double>(G, s));
  return EW_OK;
}

int launch_run(ew_pool* P, int count, const int* ids, const double* const* dq,
               double* const* shift, double* energies, hipStream_t s);

// One grouped evaluation on stream s: graph replay when the same ids and buffers were seen.
// It is ordered after the pool's previous evaluation when that ran on another stream (the
// members' buffers serve one evaluation at a time).
int run(ew_pool* P, int count, const int* ids, const double* const* dq, double* const* shift,
        double* energies, hipStream_t s) {
  int st = sync_desc(P);
  if (st) return st;
  PCHK(P, P->after_last(s));
  st = launch_run(P, count, ids, dq, shift, energies, s);
  if (st) return st;
  PCHK(P, P->mark_last(s));
  return EW_OK;
}

int launch_run(ew_pool* P, int count, const int* ids, const double* const* dq,
               double* const* shift, double* energies, hipStream_t s) {
  int st = EW_OK;
  const bool eager = P->opt.eager || P->opt.profile;
  std::vector<const void*> ptrs;
  const std::vector<int> key(ids, ids + count);
  if (!eager) {
    ptrs.reserve(2 * (size_t)count + 1);
    for (int i = 0; i < count; ++i) {
      ptrs.push_back(dq[i]);
      ptrs.push_back(shift[i]);
    }
    ptrs.push_back(energies);
    for (auto& g : P->graphs)
      if (g.ids == key && g.ptrs == ptrs) {
        PCHK(P, hipGraphLaunch(g.exec, s));
        return EW_OK;
      }
  }
  Staged S;
  st = stage(P, count, ids, dq, shift, energies, &S);
  if (st) return st;
  const size_t bytes = S.words.size() * sizeof(int);
  if (eager) {
    if (P->staging_cap < bytes) P->quiesce();  // an earlier launch may still read the old one
    st = ensure(P, &P->staging, &P->staging_cap, bytes);
    if (st) return st;
    // a pageable source: the copy has read the host image when the call returns
    PCHK(P, hipMemcpyAsync(P->staging, S.words.data(), bytes, hipMemcpyHostToDevice, s));
    return enqueue(P, launch_of(P, S, P->staging), s, P->aux, P->fork_ev, P->join_ev);
  }
  if (P->graphs.size() >= 8) {
    // the oldest capture goes once its launches are done
    P->quiesce();
    Captured& old = P->graphs.front();
    (void)hipGraphExecDestroy(old.exec);
    (void)hipFree(old.staging);
    P->graphs.erase(P->graphs.begin());
  }
  Captured c;
  c.ids = key;
  c.ptrs = ptrs;
  if (hipMalloc(&c.staging, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return P->fail(EW_ERR_MEMORY, "pool staging allocation failed");
  }
  PCHK(P, hipMemcpy(c.staging, S.words.data(), bytes, hipMemcpyHostToDevice));
  const GroupLaunch G = launch_of(P, S, c.staging);
  PCHK(P, hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
  st = enqueue(P, G, P->stream, P->aux, P->fork_ev, P->join_ev);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(P->stream, &graph);
  if (st != EW_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipFree(c.staging);
    if (st != EW_OK) return st;
    return P->fail(EW_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(ec));
  }
  const hipError_t ei = hipGraphInstantiate(&c.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    (void)hipFree(c.staging);
    return P->fail(EW_ERR_DEVICE, std::string("graph instantiate failed: ") + hipGetErrorString(ei));
  }
  P->graphs.push_back(c);
  PCHK(P, hipGraphLaunch(c.exec, s));
  return EW_OK;
}

}  // namespace

int ewamd::pool_launch_staged(ew_pool* P, int count, const int* ids, const double* const* dq,
                              double* const* shift, double* energies, void** staging,
                              size_t* staging_cap, hipStream_t s, hipStream_t aux,
                              hipEvent_t fork_ev, hipEvent_t join_ev) {
  int st = check_ids(P, count, ids);
  if (st) return st;
  st = sync_desc(P);
  if (st) return st;
  Staged S;
  st = stage(P, count, ids, dq, shift, energies, &S);
  if (st) return st;
  const size_t bytes = S.words.size() * sizeof(int);
  st = ensure(P, staging, staging_cap, bytes);
  if (st) return st;
  PCHK(P, hipMemcpyAsync(*staging, S.words.data(), bytes, hipMemcpyHostToDevice, s));
  return enqueue(P, launch_of(P, S, *staging), s, aux, fork_ev, join_ev);
}

int ewamd::pool_atoms(const ew_pool* P, int id) {
  if (id < 0 || id >= (int)P->members.size() || !P->members[id]) return -1;
  return P->members[id]->n;
}

int ewamd::pool_slots(const ew_pool* P) { return (int)P->members.size(); }

const char* ewamd::pool_error(const ew_pool* P) { return P->err.c_str(); }

int ewamd::pool_device(const ew_pool* P) { return P->opt.device; }

bool ewamd::pool_charged(const ew_pool* P) { return P->opt.charged != 0; }

extern "C" int ew_pool_create(const ew_options* opts, long long capacity_bytes, ew_pool** out) {
  if (!out) return set_global_error(EW_ERR_ARGUMENT, "ew_pool_create: out is NULL");
  *out = nullptr;
  ew_options o{};
  if (opts) o = *opts;
  if (o.precision != EW_PRECISION_FP64 && o.precision != EW_PRECISION_MIXED)
    return set_global_error(EW_ERR_ARGUMENT, "ew_pool_create: unknown precision mode");
  if (capacity_bytes < 0)
    return set_global_error(EW_ERR_ARGUMENT, "ew_pool_create: negative capacity");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return set_global_error(EW_ERR_DEVICE, "ew_pool_create: no HIP device available");
  }
  if (o.device < 0 || o.device >= ndev)
    return set_global_error(EW_ERR_ARGUMENT, "ew_pool_create: device ordinal out of range");
  if (hipSetDevice(o.device) != hipSuccess)
    return set_global_error(EW_ERR_DEVICE, "ew_pool_create: hipSetDevice failed");
  ew_pool* P = new ew_pool();
  P->opt = o;
  P->opt.profile = 0;  // per-phase events are a single-structure feature
  P->fp32 = o.precision == EW_PRECISION_MIXED;
  P->capacity = capacity_bytes;
  if (hipStreamCreate(&P->stream) != hipSuccess || hipStreamCreate(&P->aux) != hipSuccess ||
      hipEventCreateWithFlags(&P->fork_ev, kSyncEventFlags) != hipSuccess ||
      hipEventCreateWithFlags(&P->join_ev, kSyncEventFlags) != hipSuccess ||
      hipEventCreateWithFlags(&P->last_ev, kSyncEventFlags) != hipSuccess) {
    (void)hipGetLastError();
    if (P->last_ev) (void)hipEventDestroy(P->last_ev);
    if (P->join_ev) (void)hipEventDestroy(P->join_ev);
    if (P->fork_ev) (void)hipEventDestroy(P->fork_ev);
    if (P->aux) (void)hipStreamDestroy(P->aux);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
    return set_global_error(EW_ERR_DEVICE, "ew_pool_create: hipStreamCreate failed");
  }
  *out = P;
  return EW_OK;
}

extern "C" int ew_pool_destroy(ew_pool* P) {
  if (!P) return EW_OK;
  (void)hipSetDevice(P->opt.device);
  P->quiesce();
  P->drop_graphs();
  for (auto* h : P->members)
    if (h) (void)ew_destroy(h);
  if (P->desc_dev) (void)hipFree(P->desc_dev);
  if (P->staging) (void)hipFree(P->staging);
  if (P->io_dq) (void)hipFree(P->io_dq);
  if (P->io_shift) (void)hipFree(P->io_shift);
  if (P->io_energy) (void)hipFree(P->io_energy);
  (void)hipEventDestroy(P->fork_ev);
  (void)hipEventDestroy(P->join_ev);
  (void)hipEventDestroy(P->last_ev);
  (void)hipStreamDestroy(P->aux);
  (void)hipStreamDestroy(P->stream);
  delete P;
  return EW_OK;
}

extern "C" const char* ew_pool_last_error(const ew_pool* P) {
  return P ? P->err.c_str() : global_error().c_str();
}

extern "C" int ew_pool_add(ew_pool* P, int n, const double* coords, const double* latvecs,
                           const int* species, const double* hubbard_u, int n_species,
                           const double* charges, const int* grid, int* id) {
  if (!P || !id) return set_global_error(EW_ERR_ARGUMENT, "ew_pool_add: NULL argument");
  *id = -1;
  PCHK(P, hipSetDevice(P->opt.device));
  ew_options o = P->opt;
  if (grid)
    for (int a = 0; a < 3; ++a) o.grid[a] = grid[a];
  ew_handle* h = nullptr;
  int st = handle_create(o, P->stream, &h);
  if (st) return P->fail(st, global_error());
  st = ew_set_geometry(h, n, coords, latvecs, species, hubbard_u, n_species, charges);
  if (st) {
    const std::string msg = h->err;
    (void)ew_destroy(h);
    return P->fail(st, msg);
  }
  if (P->order != 0 && h->prm.order != P->order) {
    (void)ew_destroy(h);
    return P->fail(EW_ERR_ARGUMENT, "every structure of a pool must use the same B-spline order");
  }
  if (P->capacity > 0 && P->used + (long long)h->device_bytes > P->capacity) {
    const long long need = (long long)h->device_bytes;
    (void)ew_destroy(h);
    return P->fail(EW_ERR_MEMORY, "pool capacity exceeded: structure needs " +
                                      std::to_string(need) + " bytes, " +
                                      std::to_string(P->capacity - P->used) + " free");
  }
  P->order = h->prm.order;
  int slot = -1;
  for (size_t k = 0; k < P->members.size(); ++k)
    if (!P->members[k]) {
      slot = (int)k;
      break;
    }
  if (slot < 0) {
    slot = (int)P->members.size();
    P->members.push_back(nullptr);
    P->desc.emplace_back();
  }
  P->members[slot] = h;
  fill_desc(h, &P->desc[slot]);
  P->used += (long long)h->device_bytes;
  P->desc_dirty = true;
  *id = slot;
  return EW_OK;
}

extern "C" int ew_pool_remove(ew_pool* P, int id) {
  if (!P) return set_global_error(EW_ERR_ARGUMENT, "ew_pool_remove: pool is NULL");
  if (id < 0 || id >= (int)P->members.size() || !P->members[id])
    return P->fail(EW_ERR_ARGUMENT, "unknown structure id " + std::to_string(id));
  PCHK(P, hipSetDevice(P->opt.device));
  // the last evaluation and the graphs may name the structure's buffers
  P->quiesce();
  P->drop_graphs();
  P->used -= (long long)P->members[id]->device_bytes;
  (void)ew_destroy(P->members[id]);
  P->members[id] = nullptr;
  P->desc[id] = StructDesc{};
  P->desc_dirty = true;
  if (P->used <= 0) P->order = 0;
  return EW_OK;
}

extern "C" int ew_pool_parameters(const ew_pool* P, int id, ew_parameters* out) {
  if (!P || !out) return set_global_error(EW_ERR_ARGUMENT, "ew_pool_parameters: NULL argument");
  if (id < 0 || id >= (int)P->members.size() || !P->members[id])
    return set_global_error(EW_ERR_ARGUMENT, "unknown structure id " + std::to_string(id));
  return ew_get_parameters(P->members[id], out);
}

extern "C" int ew_pool_usage(const ew_pool* P, long long* used, long long* capacity, int* count) {
  if (!P) return set_global_error(EW_ERR_ARGUMENT, "ew_pool_usage: pool is NULL");
  if (used) *used = P->used;
  if (capacity) *capacity = P->capacity;
  if (count) {
    int c = 0;
    for (auto* h : P->members) c += h != nullptr;
    *count = c;
  }
  return EW_OK;
}

extern "C" int ew_pool_shift_device(ew_pool* P, int count, const int* ids,
                                    const double* const* d_dq, double* const* d_shift,
                                    double* d_energies, void* stream) {
  if (!P) return set_global_error(EW_ERR_ARGUMENT, "ew_pool_shift_device: pool is NULL");
  int st = check_ids(P, count, ids);
  if (st) return st;
  if (!d_dq || !d_shift || !d_energies)
    return P->fail(EW_ERR_ARGUMENT, "ew_pool_shift_device: NULL array");
  for (int i = 0; i < count; ++i)
    if (!d_dq[i] || !d_shift[i]) return P->fail(EW_ERR_ARGUMENT, "ew_pool_shift_device: NULL buffer");
  PCHK(P, hipSetDevice(P->opt.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : P->stream;
  return run(P, count, ids, d_dq, d_shift, d_energies, s);
}

extern "C" int ew_pool_shift(ew_pool* P, int count, const int* ids, const double* dq,
                             double* shift, double* energies) {
  if (!P) return set_global_error(EW_ERR_ARGUMENT, "ew_pool_shift: pool is NULL");
  int st = check_ids(P, count, ids);
  if (st) return st;
  if (!dq || !shift || !energies) return P->fail(EW_ERR_ARGUMENT, "ew_pool_shift: NULL array");
  std::vector<size_t> off((size_t)count + 1, 0);
  for (int i = 0; i < count; ++i) {
    const int n = P->members[ids[i]]->n;
    off[i + 1] = off[i] + n;
    double net = 0.0, absq = 0.0;
    for (int k = 0; k < n; ++k) {
      const double v = dq[off[i] + k];
      if (!std::isfinite(v)) return P->fail(EW_ERR_NOT_FINITE, "non-finite charge");
      if (std::fabs(v) > kMaxCharge) return P->fail(EW_ERR_ARGUMENT, "|charge| above 1e4");
      net += v;
      absq += std::fabs(v);
    }
    if (!P->opt.charged && std::fabs(net) > kNeutralityTol * absq)
      return P->fail(EW_ERR_NOT_NEUTRAL, "structure " + std::to_string(ids[i]) +
                                             ": the cell must be charge neutral");
  }
  PCHK(P, hipSetDevice(P->opt.device));
  const size_t atoms = off[count];
  if (P->io_atoms_cap < atoms || P->io_struct_cap < (size_t)count) {
    // the last evaluation and the graphs may point at the old buffers
    P->quiesce();
    P->drop_graphs();
    for (double** b : {&P->io_dq, &P->io_shift, &P->io_energy})
      if (*b) {
        (void)hipFree(*b);
        *b = nullptr;
      }
    P->io_atoms_cap = std::max(atoms, 2 * P->io_atoms_cap);
    P->io_struct_cap = std::max((size_t)count, 2 * P->io_struct_cap);
    if (hipMalloc(&P->io_dq, sizeof(double) * P->io_atoms_cap) != hipSuccess ||
        hipMalloc(&P->io_shift, sizeof(double) * P->io_atoms_cap) != hipSuccess ||
        hipMalloc(&P->io_energy, sizeof(double) * P->io_struct_cap) != hipSuccess) {
      (void)hipGetLastError();
      P->io_atoms_cap = P->io_struct_cap = 0;
      return P->fail(EW_ERR_MEMORY, "pool i/o allocation failed");
    }
  }
  std::vector<const double*> pin((size_t)count);
  std::vector<double*> pout((size_t)count);
  for (int i = 0; i < count; ++i) {
    pin[i] = P->io_dq + off[i];
    pout[i] = P->io_shift + off[i];
  }
  hipStream_t s = P->stream;
  PCHK(P, hipMemcpyAsync(P->io_dq, dq, sizeof(double) * atoms, hipMemcpyHostToDevice, s));
  st = run(P, count, ids, pin.data(), pout.data(), P->io_energy, s);
  if (st) return st;
  PCHK(P, hipMemcpyAsync(shift, P->io_shift, sizeof(double) * atoms, hipMemcpyDeviceToHost, s));
  PCHK(P, hipMemcpyAsync(energies, P->io_energy, sizeof(double) * count, hipMemcpyDeviceToHost, s));
  PCHK(P, hipStreamSynchronize(s));
  return EW_OK;
}

extern "C" int ew_pool_sync(ew_pool* P) {
  if (!P) return set_global_error(EW_ERR_ARGUMENT, "ew_pool_sync: pool is NULL");
  PCHK(P, hipStreamSynchronize(P->stream));
  return EW_OK;
}
