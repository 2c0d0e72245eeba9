// handle.hpp -- the resident state of one structure (ew_handle) and host helpers shared by
// the single-structure engine (engine.hip) and the structure pool (pool.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/fftewald_amd.h"
#include "engine.hpp"
#include "kernels.hpp"

namespace ewamd {

// Events that order work between streams of one device (fork, join, the last evaluation): no
// timing, and a device-scope release when recorded -- their consumers are kernels on the same
// device or a host wait for completion, neither needs the default system-scope fence (a cache
// write-back at every record).
constexpr unsigned kSyncEventFlags = hipEventDisableTiming | hipEventReleaseToDevice;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Message slot of ew_last_error(NULL) (per host thread).
std::string& global_error();
int set_global_error(int code, const std::string& msg);

}  // namespace ewamd

struct ew_handle {
  ew_options opt{};
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: the stream belongs to the pool that holds this handle
  hipStream_t aux = nullptr;      // runs the pair sum beside the SPME chain (own-stream handles)
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  hipEvent_t spread_ev = nullptr;  // after the spreading (the pair sum's late fork)
  // Completion of the last evaluation, on whichever stream it ran (ew_shift_device and
  // ew_shift_multi_device take a caller's stream): the geometry's buffers, the multi-vector
  // views and the captured graphs are released, rewritten or destroyed only after it.
  hipEvent_t last_ev = nullptr;
  hipStream_t last_stream = nullptr;
  bool last_pending = false;
  std::string err = "no error";
  bool have_geom = false;

  int n = 0;
  ewamd::Cell cell{};
  ewamd::Params prm{};
  bool fp32 = false;
  int ns = 0;
  int pair_mode = 0;
  long long nnz = 0;
  int T[3] = {0, 0, 0}, nt[3] = {0, 0, 0}, ntmax = 0;
  int max_tile_atoms = 0;  // atoms of the fullest sort tile
  int spread_slots = 0;  // atoms staged per chunk by the interval spread (0: checked spread)
  ewamd::RegionArgs region{};   // privatised spreading (scratch == null: not used)
  ewamd::FftPlan px{}, py{}, pz{};
  bool plane = false;        // fused x-y plane transforms
  ew_tuning tuning{};        // launch shapes in use (what px/py/pz, plane and region hold)
  double self_coef = 0.0;
  double background = 0.0;   // -pi / (V alpha^2), applied when opt.charged
  double* qsum = nullptr;    // total charge of the last evaluation (opt.charged)

  // device state of the current geometry
  std::vector<ewamd::DevBuf> bufs;
  // buffers of the previous geometry, reused by alloc() while the next one is built (a geometry
  // step of a trajectory needs the same sizes again) and freed when it is done
  std::vector<ewamd::DevBuf> spare;
  // host copies of the FFT tables of the current geometry, kept until its uploads are done (the
  // next ew_set_geometry synchronises the stream before rebuilding them)
  std::vector<double> tab_bsq[3];
  std::vector<float> tab_tw32[3];
  std::vector<double> tab_tw64[3];
  std::vector<float> tab_twp32[3];
  std::vector<double> tab_twp64[3];
  int* perm = nullptr;
  int4* k0 = nullptr;
  double* frac = nullptr;
  void* theta = nullptr;
  int* tile_of = nullptr;
  int* tile_start = nullptr;
  int* src_tab = nullptr;
  int* species = nullptr;
  double* hub = nullptr;
  double* gcut2 = nullptr;
  double* onsite = nullptr;
  long long* row_ptr = nullptr;
  int* col = nullptr;
  short* col16 = nullptr;  // signed 16-bit partner offsets (when every row fits), else null
  bool half = false;       // symmetric pair storage (hp) in place of the row list
  ewamd::HalfPairs hp{};
  long long half_entries = 0;
  void* coef = nullptr;
  void* grid = nullptr;
  void* cgrid = nullptr;
  void* tw[3] = {nullptr, nullptr, nullptr};
  double* bsq[3] = {nullptr, nullptr, nullptr};
  // register line transforms: W_K^m per axis and the influence function on the half spectrum
  void* twp[3] = {nullptr, nullptr, nullptr};
  void* inf_tab = nullptr;
  int wave = 0;  // 0: Stockham passes, 1: register line passes, 2: register x-y planes + z lines
  int lanes[3] = {0, 0, 0};
  double* qs = nullptr;
  double* phi_real = nullptr;
  double* gmesh = nullptr;  // mesh potential per sorted atom (split gather)
  double* epart = nullptr;
  double* energy = nullptr;
  double* dq_io = nullptr;
  double* shift_io = nullptr;
  int* err_flag = nullptr;
  double* force = nullptr;  // [n][3], allocated on the first force request
  ewamd::PairArgs pa{};
  size_t device_bytes = 0;

  // captured evaluations, one per (dq, shift, energy) buffer triple
  struct Captured {
    const double* dq;
    double* shift;
    double* energy;
    hipGraphExec_t exec;
  };
  std::vector<Captured> graphs;

  // Multi-vector evaluation (several charge vectors through this geometry in one grouped
  // launch): per-vector charge, potential, mesh and partial buffers, allocated on first use.
  struct Views {
    int cap = 0;                 // vectors the buffers hold
    double* qs = nullptr;        // [cap][n] sorted charges (contiguous: the pair sum reads them so)
    double* phi = nullptr;       // [cap][n] real-space potentials
    double* qil = nullptr;       // [n][spmm_stride(cap)] charges interleaved for the pair sum
    double* epart = nullptr;     // [cap][nparts]
    double* qsum = nullptr;      // [cap] (opt.charged)
    void* grid = nullptr;        // [cap][K0 K1 K2]
    void* cgrid = nullptr;       // [cap][Kh K1 K2] complex
    void* region = nullptr;      // [cap][blocks][region_stride] privatised spreading
    double* dq = nullptr;        // [cap][n] host-array staging
    double* shift = nullptr;     // [cap][n]
    double* energy = nullptr;    // [cap]
    ewamd::StructDesc* desc = nullptr;  // [cap] device descriptors, one per vector
    ewamd::WaveIo* wio = nullptr;       // [cap] meshes and spectra for the batched FFTs
    void* staging = nullptr;     // maps and IoSlots of eager launches
  } views;
  std::vector<ewamd::StructDesc> views_host;  // host image of views.desc
  size_t views_staging_cap = 0;
  struct CapturedMulti {
    int nvec;
    const double* dq;
    double* shift;
    double* energy;
    hipGraphExec_t exec;
    void* staging;  // maps and IoSlots the graph's kernels read
  };
  std::vector<CapturedMulti> mgraphs;

  // Orders work about to be enqueued on s after the last evaluation when that ran on another
  // stream (one handle's buffers serve one evaluation at a time).
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
  // Waits for the handle's streams and for the last evaluation on any stream.
  void quiesce() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (aux) (void)hipStreamSynchronize(aux);
    if (last_pending) (void)hipEventSynchronize(last_ev);
    last_pending = false;
  }

  // destroys the captured graphs once nothing that launched them can still be running
  void record_trace(int i, hipStream_t s) {
    if (opt.profile == 2) (void)hipEventRecord(tev[i], s);
  }
  void drop_graphs() {
    if (!graphs.empty() || !mgraphs.empty()) quiesce();
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
    for (auto& g : mgraphs) {
      (void)hipGraphExecDestroy(g.exec);
      (void)hipFree(g.staging);
    }
    mgraphs.clear();
  }

  // per-phase timing (opt.profile = 1: every phase alone on the stream)
  hipEvent_t ev[EW_NUM_PHASES + 1] = {};
  double phase_ms[EW_NUM_PHASES] = {};
  long long phase_calls = 0;
  // timing in the evaluation's own launch shape (opt.profile = 2), see TraceEvent
  enum TraceEvent {
    kTrStart = 0,  // stream s: start of the evaluation
    kTrChg0,       // charges: before the gather (on the stream that runs them)
    kTrChg1,       //          after the gather and charge sum
    kTrPair0,      // pair sum: before (auxiliary stream when it runs beside the chain)
    kTrPair1,      //           after
    kTrChain0,     // stream s: before the spreading
    kTrRegion,     //           after the region kernel
    kTrSpread, kTrX, kTrY, kTrZ, kTrYi, kTrXi,  // after each chain phase
    kTrGathered,   //           after the mesh gather (split gather; else right after kTrXi)
    kTrJoined,     //           after the wait for the pair sum
    kTrEnd,        //           after the gather and energy
    kTrEvents
  };
  hipEvent_t tev[kTrEvents] = {};
  double trace_ms[EW_NUM_TRACE] = {};
  long long trace_calls = 0;

  int fail(int code, const std::string& msg) {
    err = msg;
    ewamd::global_error() = msg;
    return code;
  }
  void release() {
    // work queued on the handle's streams, or an evaluation on a caller's stream, may still
    // read the buffers
    quiesce();
    drop_graphs();
    for (auto& b : bufs)
      if (b.p) spare.push_back(b);
    bufs.clear();
    device_bytes = 0;
    // every pointer below pointed into bufs: the next geometry may hand that memory to another
    // buffer or free it, and the optional ones (species terms, forces) are tested for null
    perm = nullptr;
    k0 = nullptr;
    frac = nullptr;
    theta = nullptr;
    tile_of = tile_start = src_tab = nullptr;
    species = nullptr;
    hub = gcut2 = onsite = nullptr;
    row_ptr = nullptr;
    col = nullptr;
    col16 = nullptr;
    half = false;
    hp = ewamd::HalfPairs{};
    half_entries = 0;
    coef = grid = cgrid = nullptr;
    for (int a = 0; a < 3; ++a) {
      tw[a] = nullptr;
      bsq[a] = nullptr;
      twp[a] = nullptr;
      lanes[a] = 0;
    }
    inf_tab = nullptr;
    wave = 0;
    qs = qsum = phi_real = gmesh = epart = energy = dq_io = shift_io = nullptr;
    err_flag = nullptr;
    force = nullptr;
    region = ewamd::RegionArgs{};
    views = Views{};
    views_host.clear();
    views_staging_cap = 0;
    have_geom = false;
  }
  // frees the spare buffers (nothing queued uses them: release() synchronised the streams)
  void trim() {
    for (auto& b : spare) (void)hipFree(b.p);
    spare.clear();
  }
  void free_buf(void* p) {
    for (size_t k = 0; k < bufs.size(); ++k)
      if (bufs[k].p == p) {
        (void)hipFree(p);
        device_bytes -= bufs[k].bytes;
        bufs.erase(bufs.begin() + k);
        return;
      }
  }
  template <typename T>
  int alloc(T** out, size_t count) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    // the smallest spare buffer that fits and wastes at most half of itself
    size_t best = spare.size();
    for (size_t k = 0; k < spare.size(); ++k)
      if (spare[k].bytes >= bytes && spare[k].bytes <= 2 * bytes + (1u << 20) &&
          (best == spare.size() || spare[k].bytes < spare[best].bytes))
        best = k;
    if (best < spare.size()) {
      const ewamd::DevBuf b = spare[best];
      spare.erase(spare.begin() + best);
      bufs.push_back(b);
      device_bytes += b.bytes;
      *out = static_cast<T*>(b.p);
      return EW_OK;
    }
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      *out = nullptr;
      return fail(EW_ERR_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    bufs.push_back({p, bytes});
    device_bytes += bytes;
    *out = static_cast<T*>(p);
    return EW_OK;
  }
};

#define HIPCHK(h, expr)                                                                      \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return (h)->fail(EW_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) +   \
                                          " at " #expr);                                     \
  } while (0)

#define CHK(expr)             \
  do {                        \
    int s_ = (expr);          \
    if (s_ != EW_OK) return s_; \
  } while (0)

namespace ewamd {

// A handle on a borrowed stream (pool members); ew_destroy leaves the stream alone.
int handle_create(const ew_options& o, hipStream_t borrowed, ew_handle** out);
// The grouped kernels' view of a handle with a geometry.
void fill_desc(const ew_handle* h, StructDesc* d);
// The reciprocal chain's launch description of a handle with a geometry.
SpmeLaunch spme_launch_of(const ew_handle* h);

}  // namespace ewamd
