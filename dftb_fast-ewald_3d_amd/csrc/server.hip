// server.hip -- the serving form: a request loop over a structure pool with continuous batching.
//
// Callers submit SCC charge-update requests (structure id + charges) from any thread.  One worker
// thread forms a batch whenever a launch slot is free: every pending request whose structure
// has no evaluation in flight, in arrival order, up to max_batch.  Each batch is one grouped
// pool evaluation on its own stream (charges in through pinned staging, shifts and energies back
// the same way); up to max_inflight batches run at once, so requests that arrive while others are
// on the GPU are launched as soon as a slot frees, and their work overlaps on the device.  A
// host function enqueued after each batch's copies completes its requests: results land in the
// callers' buffers, waiters wake and the latency (submit to result) is recorded.
//
// The reference has no serving form: it evaluates one structure per call
// (fe/fft/ewald_fft_3d.f90:36-134); each request here is that call for one resident structure.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fftewald_amd.h"
#include "engine.hpp"
#include "handle.hpp"
#include "pool.hpp"

using namespace ewamd;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  long long ticket = 0;
  int id = 0;
  int n = 0;
  std::vector<double> dq;
  double* shift = nullptr;
  double* energy = nullptr;
  Clock::time_point t_submit;
};

struct Batch {
  ew_server* server = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;  // the batch's pair sums, beside its SPME chain
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  double* h_dq = nullptr;  // pinned
  double* h_shift = nullptr;
  double* h_energy = nullptr;
  double* d_dq = nullptr;
  double* d_shift = nullptr;
  double* d_energy = nullptr;
  void* staging = nullptr;
  size_t staging_cap = 0;
  std::vector<Request> reqs;
  std::vector<size_t> offs;
};

constexpr int kPending = -1;

}  // namespace

struct ew_server {
  ew_pool* pool = nullptr;
  int max_batch = 64;
  int max_inflight = 2;
  size_t max_atoms = 0;
  std::string err = "no error";

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request> pending;
  std::vector<char> busy;  // per structure id: an evaluation is in flight
  std::vector<Batch*> free_batches;
  std::vector<Batch> batches;
  std::unordered_map<long long, int> status;  // ticket -> kPending or an EW_* code
  long long next_ticket = 1;
  bool stop = false;
  int inflight = 0;
  std::thread worker;

  // statistics
  long long submitted = 0, completed = 0, failed = 0, nbatches = 0, batched_requests = 0;
  std::vector<double> latency_ms;

  int fail(int code, const std::string& msg) {
    err = msg;
    global_error() = msg;
    return code;
  }
};

namespace {

// Completes the requests of a batch (runs on the HIP runtime's callback thread; no HIP calls).
void batch_done(void* arg) {
  Batch* b = static_cast<Batch*>(arg);
  ew_server* S = b->server;
  const Clock::time_point t = Clock::now();
  for (size_t k = 0; k < b->reqs.size(); ++k) {
    const Request& r = b->reqs[k];
    std::memcpy(r.shift, b->h_shift + b->offs[k], sizeof(double) * r.n);
    if (r.energy) *r.energy = b->h_energy[k];
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (const Request& r : b->reqs) {
      S->status[r.ticket] = EW_OK;
      S->busy[r.id] = 0;
      S->latency_ms.push_back(std::chrono::duration<double, std::milli>(t - r.t_submit).count());
      ++S->completed;
    }
    b->reqs.clear();
    S->free_batches.push_back(b);
    --S->inflight;
  }
  S->cv_done.notify_all();
  S->cv_work.notify_one();
}

// Enqueues one batch (worker thread, server lock not held).  On failure the requests are
// failed and the slot returned.
void launch(ew_server* S, Batch* b) {
  const int count = (int)b->reqs.size();
  std::vector<int> ids(count);
  std::vector<const double*> dq(count);
  std::vector<double*> sh(count);
  b->offs.assign(count, 0);
  size_t off = 0;
  for (int k = 0; k < count; ++k) {
    const Request& r = b->reqs[k];
    ids[k] = r.id;
    b->offs[k] = off;
    std::memcpy(b->h_dq + off, r.dq.data(), sizeof(double) * r.n);
    dq[k] = b->d_dq + off;
    sh[k] = b->d_shift + off;
    off += r.n;
  }
  int st = EW_OK;
  std::string msg;
  if (hipMemcpyAsync(b->d_dq, b->h_dq, sizeof(double) * off, hipMemcpyHostToDevice, b->stream) !=
      hipSuccess) {
    st = EW_ERR_DEVICE;
    msg = "charge upload failed";
  }
  if (st == EW_OK) {
    st = pool_launch_staged(S->pool, count, ids.data(), dq.data(), sh.data(), b->d_energy,
                            &b->staging, &b->staging_cap, b->stream, b->aux, b->fork_ev,
                            b->join_ev);
    if (st) msg = pool_error(S->pool);
  }
  if (st == EW_OK &&
      (hipMemcpyAsync(b->h_shift, b->d_shift, sizeof(double) * off, hipMemcpyDeviceToHost,
                      b->stream) != hipSuccess ||
       hipMemcpyAsync(b->h_energy, b->d_energy, sizeof(double) * count, hipMemcpyDeviceToHost,
                      b->stream) != hipSuccess ||
       hipLaunchHostFunc(b->stream, batch_done, b) != hipSuccess)) {
    st = EW_ERR_DEVICE;
    msg = "result download failed";
  }
  if (st == EW_OK) return;
  // the batch never reached its host function: fail its requests here
  (void)hipStreamSynchronize(b->stream);
  (void)hipGetLastError();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->err = msg;
    for (const Request& r : b->reqs) {
      S->status[r.ticket] = st;
      S->busy[r.id] = 0;
      ++S->failed;
    }
    b->reqs.clear();
    S->free_batches.push_back(b);
    --S->inflight;
  }
  S->cv_done.notify_all();
}

bool launchable(const ew_server* S) {
  for (const Request& r : S->pending)
    if (!S->busy[r.id]) return true;
  return false;
}

void worker_loop(ew_server* S) {
  (void)hipSetDevice(pool_device(S->pool));
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv_work.wait(lk, [S] {
      return (S->stop && S->pending.empty()) || (!S->free_batches.empty() && launchable(S));
    });
    if (S->stop && S->pending.empty()) break;
    Batch* b = S->free_batches.back();
    S->free_batches.pop_back();
    // arrival order; one request per structure per batch; structures in flight wait
    for (auto it = S->pending.begin();
         it != S->pending.end() && (int)b->reqs.size() < S->max_batch;) {
      if (!S->busy[it->id]) {
        S->busy[it->id] = 1;
        b->reqs.push_back(std::move(*it));
        it = S->pending.erase(it);
      } else {
        ++it;
      }
    }
    ++S->inflight;
    ++S->nbatches;
    S->batched_requests += (long long)b->reqs.size();
    lk.unlock();
    launch(S, b);
    lk.lock();
  }
}

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  const size_t k = std::min(v.size() - 1, (size_t)std::floor(p * (v.size() - 1) + 0.5));
  std::nth_element(v.begin(), v.begin() + k, v.end());
  return v[k];
}

void free_batch(Batch& b) {
  if (b.stream) (void)hipStreamDestroy(b.stream);
  if (b.aux) (void)hipStreamDestroy(b.aux);
  if (b.fork_ev) (void)hipEventDestroy(b.fork_ev);
  if (b.join_ev) (void)hipEventDestroy(b.join_ev);
  if (b.h_dq) (void)hipHostFree(b.h_dq);
  if (b.h_shift) (void)hipHostFree(b.h_shift);
  if (b.h_energy) (void)hipHostFree(b.h_energy);
  if (b.d_dq) (void)hipFree(b.d_dq);
  if (b.d_shift) (void)hipFree(b.d_shift);
  if (b.d_energy) (void)hipFree(b.d_energy);
  if (b.staging) (void)hipFree(b.staging);
}

}  // namespace

extern "C" int ew_server_create(ew_pool* pool, const ew_server_options* opts, ew_server** out) {
  if (!pool || !out) return set_global_error(EW_ERR_ARGUMENT, "ew_server_create: NULL argument");
  *out = nullptr;
  ew_server_options o{};
  if (opts) o = *opts;
  if (o.max_batch < 0 || o.max_inflight < 0)
    return set_global_error(EW_ERR_ARGUMENT, "ew_server_create: negative limits");
  if (hipSetDevice(pool_device(pool)) != hipSuccess)
    return set_global_error(EW_ERR_DEVICE, "ew_server_create: hipSetDevice failed");
  ew_server* S = new ew_server();
  S->pool = pool;
  S->max_batch = o.max_batch > 0 ? o.max_batch : 64;
  S->max_inflight = o.max_inflight > 0 ? o.max_inflight : 2;
  const int nslot = pool_slots(pool);
  S->busy.assign(nslot, 0);
  size_t big = 0;
  std::vector<size_t> sizes;
  for (int i = 0; i < nslot; ++i) {
    const int n = pool_atoms(pool, i);
    if (n > 0) sizes.push_back((size_t)n);
  }
  if (sizes.empty()) {
    delete S;
    return set_global_error(EW_ERR_NO_GEOMETRY, "ew_server_create: the pool holds no structure");
  }
  // staging for the max_batch largest structures
  std::sort(sizes.rbegin(), sizes.rend());
  for (size_t k = 0; k < sizes.size() && (int)k < S->max_batch; ++k) big += sizes[k];
  S->max_atoms = sizes[0];
  S->batches.resize(S->max_inflight);
  for (Batch& b : S->batches) {
    b.server = S;
    const bool ok = hipStreamCreate(&b.stream) == hipSuccess &&
                    hipStreamCreate(&b.aux) == hipSuccess &&
                    hipEventCreateWithFlags(&b.fork_ev, kSyncEventFlags) == hipSuccess &&
                    hipEventCreateWithFlags(&b.join_ev, kSyncEventFlags) == hipSuccess &&
                    hipHostMalloc(&b.h_dq, sizeof(double) * big) == hipSuccess &&
                    hipHostMalloc(&b.h_shift, sizeof(double) * big) == hipSuccess &&
                    hipHostMalloc(&b.h_energy, sizeof(double) * S->max_batch) == hipSuccess &&
                    hipMalloc(&b.d_dq, sizeof(double) * big) == hipSuccess &&
                    hipMalloc(&b.d_shift, sizeof(double) * big) == hipSuccess &&
                    hipMalloc(&b.d_energy, sizeof(double) * S->max_batch) == hipSuccess;
    if (!ok) {
      (void)hipGetLastError();
      for (Batch& c : S->batches) free_batch(c);
      delete S;
      return set_global_error(EW_ERR_MEMORY, "ew_server_create: staging allocation failed");
    }
  }
  for (Batch& b : S->batches) S->free_batches.push_back(&b);
  S->worker = std::thread(worker_loop, S);
  *out = S;
  return EW_OK;
}

extern "C" int ew_server_destroy(ew_server* S) {
  if (!S) return EW_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (Batch& b : S->batches) {
    if (b.stream) (void)hipStreamSynchronize(b.stream);
    if (b.aux) (void)hipStreamSynchronize(b.aux);
  }
  {
    // host functions of the last batches may still be returning their slots
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [S] { return S->inflight == 0; });
  }
  for (Batch& b : S->batches) free_batch(b);
  delete S;
  return EW_OK;
}

extern "C" const char* ew_server_last_error(const ew_server* S) {
  return S ? S->err.c_str() : global_error().c_str();
}

extern "C" int ew_server_submit(ew_server* S, int id, const double* dq, double* shift,
                                double* energy, long long* ticket) {
  if (!S || !dq || !shift || !ticket)
    return set_global_error(EW_ERR_ARGUMENT, "ew_server_submit: NULL argument");
  const int n = id >= 0 && id < (int)S->busy.size() ? pool_atoms(S->pool, id) : -1;
  if (n <= 0) return S->fail(EW_ERR_ARGUMENT, "unknown structure id " + std::to_string(id));
  double net = 0.0, absq = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!std::isfinite(dq[i])) return S->fail(EW_ERR_NOT_FINITE, "non-finite charge");
    if (std::fabs(dq[i]) > kMaxCharge) return S->fail(EW_ERR_ARGUMENT, "|charge| above 1e4");
    net += dq[i];
    absq += std::fabs(dq[i]);
  }
  if (!pool_charged(S->pool) && std::fabs(net) > kNeutralityTol * absq)
    return S->fail(EW_ERR_NOT_NEUTRAL, "the cell must be charge neutral");
  Request r;
  r.id = id;
  r.n = n;
  r.dq.assign(dq, dq + n);
  r.shift = shift;
  r.energy = energy;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return S->fail(EW_ERR_ARGUMENT, "server is shutting down");
    r.ticket = S->next_ticket++;
    r.t_submit = Clock::now();
    S->status[r.ticket] = kPending;
    *ticket = r.ticket;
    ++S->submitted;
    S->pending.push_back(std::move(r));
  }
  S->cv_work.notify_one();
  return EW_OK;
}

extern "C" int ew_server_wait(ew_server* S, long long ticket, double timeout_s) {
  if (!S) return set_global_error(EW_ERR_ARGUMENT, "ew_server_wait: server is NULL");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->status.find(ticket);
  if (it == S->status.end()) return S->fail(EW_ERR_ARGUMENT, "unknown ticket");
  auto done = [&] { return S->status[ticket] != kPending; };
  if (timeout_s < 0.0)
    S->cv_done.wait(lk, done);
  else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
    return EW_ERR_PENDING;
  const int st = S->status[ticket];
  S->status.erase(ticket);
  return st;
}

extern "C" int ew_server_drain(ew_server* S) {
  if (!S) return set_global_error(EW_ERR_ARGUMENT, "ew_server_drain: server is NULL");
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait(lk, [S] { return S->pending.empty() && S->inflight == 0; });
  return EW_OK;
}

extern "C" int ew_server_stats(ew_server* S, ew_server_report* out, int reset) {
  if (!S || !out) return set_global_error(EW_ERR_ARGUMENT, "ew_server_stats: NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = ew_server_report{};
  out->submitted = S->submitted;
  out->completed = S->completed;
  out->failed = S->failed;
  out->batches = S->nbatches;
  out->mean_batch = S->nbatches ? double(S->batched_requests) / S->nbatches : 0.0;
  out->p50_ms = percentile(S->latency_ms, 0.50);
  out->p99_ms = percentile(S->latency_ms, 0.99);
  out->max_ms = S->latency_ms.empty() ? 0.0
                                      : *std::max_element(S->latency_ms.begin(), S->latency_ms.end());
  if (reset) {
    S->submitted = S->completed = S->failed = S->nbatches = S->batched_requests = 0;
    S->latency_ms.clear();
  }
  return EW_OK;
}
