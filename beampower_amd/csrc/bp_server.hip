// bp_server.hip -- serving form: resident networks, HBM-budget admission and
// continuous batching of window requests (no counterpart in the reference,
// whose beamform() serves one window per call: beampower/beampower.py:10-227).
//
// A request is one window of one resident network: waveform features
// (S, C, T) in host memory, reduce='max'.  Two native threads run the loop:
//   loader     admits requests in arrival order while their device memory
//              fits the budget (a first-fit arena over one allocation),
//              copies the features in, prestacks them (the network's phase
//              weights) and packs them into rows padded for the network in
//              the server's storage format;
//   scheduler  forms a batch whenever one of the `in_flight` slots is free:
//              ready requests in earliest-deadline order (a request's deadline
//              is its ready time plus kSlack times its estimated GPU time, so
//              short windows go ahead of long ones and a long one still gets
//              its turn), up to max_batch windows and about kBatchSeconds of
//              estimated GPU work (at least one request); launches it on the
//              slot's stream (bp_batch_*), completes finished batches
//              (outputs back to the caller's host buffers) and returns their
//              memory to the arena.
// Requests arriving while batches run are admitted as memory frees up and
// join the next launch; up to in_flight batches overlap on the device.  A
// slot's batch tables live in device memory allocated once with the server,
// so forming and retiring batches allocates and frees no device memory
// (hipFree would wait for every batch in flight).
// A request's result does not depend on what it is batched with: the batch
// runs the same tile code on the same tiles and reduces with an
// order-independent packed-key maximum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "beamform_gpu.h"
#include "bp_internal.h"

namespace {

// Scheduling: estimated GPU time of a request = its gather-FMAs
// (sources x stations x phases x samples) / kFmaRate, the stack kernel's
// measured rate on these windows (2.6e13/s); a batch takes about
// kBatchSeconds of such work; a request's deadline is its ready time plus
// kSlack times its estimated time.
constexpr double kFmaRate = 2.6e13;
constexpr double kBatchSeconds = 0.010;
constexpr double kSlack = 4.0;

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device allocation (offsets; free blocks
// coalesce with their neighbours).
class Arena {
 public:
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  hipError_t init(size_t bytes) {
    size_ = bytes;
    free_.clear();
    free_[0] = bytes;
    return hipMalloc(&base_, bytes);
  }
  // -1 if no free block is large enough
  long long alloc(size_t bytes) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return (long long)off;
    }
    return -1;
  }
  // Never throws: a block whose free-list node cannot be allocated stays
  // counted as used (lost to the budget) rather than ending the process.
  void release(long long off, size_t bytes) noexcept {
    if (off < 0) return;
    bytes = align_up(bytes);
    std::map<size_t, size_t>::iterator it;
    try {
      it = free_.emplace((size_t)off, bytes).first;
    } catch (...) {
      return;
    }
    used_ -= bytes;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  char *ptr(long long off) const { return static_cast<char *>(base_) + off; }
  size_t size() const { return size_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  void *base_ = nullptr;
  size_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

struct Network {
  bp_network *net = nullptr;
  size_t K = 0, S = 0, C = 0, P = 0, pad_lo = 0, pad_hi = 0;
  float *alpha = nullptr;  // device (S, C, P)
};

struct Request {
  long long id = 0;
  int net = 0;
  const float *u_host = nullptr;
  size_t T = 0;
  int oob = 0;
  float *bmax_host = nullptr;
  int *barg_host = nullptr;
  bp_request_info info{};
  double work = 0.0;      // gather-FMAs
  double deadline = 0.0;  // scheduling key (steady clock)
  // device memory: temporaries (freed once packed) and what the launch reads
  long long off_tmp = -1, off_keep = -1;
  size_t bytes_tmp = 0, bytes_keep = 0;
  bp_feature_view view{};
  std::vector<float> scale;
  float *d_bmax = nullptr;
  int *d_barg = nullptr;
  unsigned long long *d_keys = nullptr;
  std::string err;  // message of a failed request, for the waiting thread
  bool done = false;
};

struct Flight {
  bp_batch *batch = nullptr;
  hipEvent_t ev = nullptr;
  std::vector<Request *> reqs;
  int slot = -1;  // its stream and batch-table memory
};

thread_local std::string t_err;

int fail(int code, const char *msg) noexcept {
  try {
    t_err = msg;
  } catch (...) {
    t_err.clear();  // no memory for the message: the status still tells
  }
  return code;
}

int fail(int code, const std::string &msg) noexcept { return fail(code, msg.c_str()); }

// Handlers of the function-try-block of every int entry point: a host
// exception never crosses the C ABI, it becomes a status and a message.
#define BP_SERVER_CATCH                                                 \
  catch (const std::bad_alloc &) {                                      \
    return fail(BP_ENOMEM, "host memory allocation failed");            \
  }                                                                     \
  catch (const std::exception &e_) {                                    \
    return fail(BP_EINVAL, e_.what());                                  \
  }                                                                     \
  catch (...) {                                                         \
    return fail(BP_EINVAL, "unknown host exception");                   \
  }

}  // namespace

struct bp_server {
  bp_server_config cfg{};
  int device = 0;
  std::vector<Network> nets;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv_load, cv_sched, cv_done;
  std::deque<Request *> pending, ready;
  std::deque<Flight> flights;
  std::map<long long, Request *> reqs;
  hipStream_t load_stream = nullptr;
  std::vector<hipStream_t> compute;  // one stream per slot
  void *tables = nullptr;            // batch tables, table_bytes per slot
  size_t table_bytes = 0;
  std::vector<int> free_slots;
  std::thread loader, scheduler;
  bool stop = false;
  bool broken = false;  // a worker thread failed: submits are refused
  long long next_id = 1;
  long long n_batches = 0, n_batched_windows = 0, n_completed = 0;

  // memory a request needs: temporaries (features in, prestacked planar)
  // and what stays until completion (packed rows, outputs, keys)
  void sizes(const Network &n, size_t T, size_t &tmp, size_t &keep, size_t &stride) const {
    const size_t es = cfg.fmt == BP_F32 ? 4 : 2, planes = cfg.fmt == BP_F32 ? 1 : 2;
    stride = n.pad_lo + T + n.pad_hi;
    stride += stride & 1;
    tmp = align_up(n.S * n.C * T * 4) + align_up(n.S * n.P * T * 4);
    keep = align_up(n.S * n.P * planes * stride * es) + align_up(T * 4) + align_up(T * 4) +
           align_up(T * 8);
  }

  void loader_loop();
  void scheduler_loop();
  void guarded(void (bp_server::*loop)());
  void fail_request(Request *r, int status, const char *msg) noexcept;
  int admit(Request *r);
  void launch(std::vector<Request *> batch, int slot);
  void complete(Flight &f);
};

// Called with the lock held: request r ends with `status`, its device
// memory back in the arena.
void bp_server::fail_request(Request *r, int status, const char *msg) noexcept {
  r->info.status = status;
  try {
    r->err = msg;
  } catch (...) {
  }
  arena.release(r->off_tmp, r->bytes_tmp);
  arena.release(r->off_keep, r->bytes_keep);
  r->off_tmp = r->off_keep = -1;
  r->done = true;
}

// Runs a worker loop.  A host exception that leaves it (a container that
// cannot grow) would end the process; instead the server is marked broken,
// every request not yet launched fails with BP_ENOMEM, the batches in
// flight are completed, and later submits are refused.
void bp_server::guarded(void (bp_server::*loop)()) {
  try {
    (this->*loop)();
    return;
  } catch (...) {
  }
  std::deque<Flight> left;
  {
    std::lock_guard<std::mutex> lock(mu);
    broken = true;
    for (std::deque<Request *> *q : {&pending, &ready}) {
      for (Request *r : *q) fail_request(r, BP_ENOMEM, "server worker failed: host memory");
      q->clear();
    }
    left.swap(flights);
  }
  for (Flight &f : left) complete(f);
  cv_done.notify_all();
  cv_load.notify_all();
  cv_sched.notify_all();
}

int bp_server::admit(Request *r) {
  const Network &n = nets[r->net];
  size_t tmp, keep, stride;
  sizes(n, r->T, tmp, keep, stride);
  char *t = arena.ptr(r->off_tmp);
  char *k = arena.ptr(r->off_keep);
  float *d_u = reinterpret_cast<float *>(t);
  float *d_F = reinterpret_cast<float *>(t + align_up(n.S * n.C * r->T * 4));
  const size_t es = cfg.fmt == BP_F32 ? 4 : 2, planes = cfg.fmt == BP_F32 ? 1 : 2;
  void *packed = k;
  char *q = k + align_up(n.S * n.P * planes * stride * es);
  r->d_bmax = reinterpret_cast<float *>(q);
  r->d_barg = reinterpret_cast<int *>(q + align_up(r->T * 4));
  r->d_keys = reinterpret_cast<unsigned long long *>(q + 2 * align_up(r->T * 4));
  if (hipMemcpyAsync(d_u, r->u_host, n.S * n.C * r->T * 4, hipMemcpyHostToDevice,
                     load_stream) != hipSuccess)
    return fail(BP_EHIP, "feature upload failed");
  if (hipMemsetAsync(r->d_bmax, 0, r->T * 4, load_stream) != hipSuccess)
    return fail(BP_EHIP, "output initialisation failed");
  int rc = bp_prestack_planar(d_u, n.alpha, n.S, n.C, r->T, n.P, d_F, load_stream);
  if (rc != BP_OK) return fail(rc, bp_last_error());
  r->scale.assign(n.S, 1.0f);
  rc = bp_pack_features(d_F, n.S, n.P, r->T, cfg.fmt, n.pad_lo, stride, r->scale.data(), packed,
                        load_stream);
  if (rc != BP_OK) return fail(rc, bp_last_error());
  if (hipStreamSynchronize(load_stream) != hipSuccess)
    return fail(BP_EHIP, "admission stream failed");
  r->view.data = packed;
  r->view.fmt = cfg.fmt;
  r->view.n_samples = r->T;
  r->view.row_stride = stride;
  r->view.pad_before = n.pad_lo;
  r->view.station_scale = cfg.fmt == BP_F32 ? nullptr : r->scale.data();
  return BP_OK;
}

void bp_server::loader_loop() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    Request *r = nullptr;
    cv_load.wait(lock, [&] {
      if (stop) return true;
      if (pending.empty()) return false;
      Request *h = pending.front();
      const Network &n = nets[h->net];
      size_t tmp, keep, stride;
      sizes(n, h->T, tmp, keep, stride);
      if (tmp + keep > arena.size()) return true;  // can never fit: fail it
      h->off_tmp = arena.alloc(tmp);
      if (h->off_tmp < 0) return false;
      h->off_keep = arena.alloc(keep);
      if (h->off_keep < 0) {
        arena.release(h->off_tmp, tmp);
        h->off_tmp = -1;
        return false;
      }
      h->bytes_tmp = tmp;
      h->bytes_keep = keep;
      return true;
    });
    if (stop) break;
    r = pending.front();
    pending.pop_front();
    if (r->off_tmp < 0) {
      fail_request(r, BP_ENOMEM, "window needs more device memory than the budget");
      cv_done.notify_all();
      continue;
    }
    r->info.admit_s = now_s();
    lock.unlock();
    int rc;
    try {
      rc = admit(r);
    } catch (...) {
      rc = fail(BP_ENOMEM, "admission: host memory allocation failed");
    }
    lock.lock();
    arena.release(r->off_tmp, r->bytes_tmp);
    r->off_tmp = -1;
    r->info.ready_s = now_s();
    r->deadline = r->info.ready_s + kSlack * r->work / kFmaRate;
    if (rc != BP_OK) {
      fail_request(r, rc, t_err.c_str());
      cv_done.notify_all();
      cv_load.notify_all();
      continue;
    }
    try {
      ready.push_back(r);
    } catch (...) {
      fail_request(r, BP_ENOMEM, "admission: host memory allocation failed");
      cv_done.notify_all();
      cv_load.notify_all();
      continue;
    }
    cv_sched.notify_all();
  }
}

void bp_server::launch(std::vector<Request *> batch, int slot) {
  const int n = (int)batch.size();
  Flight f;
  f.slot = slot;
  int rc = BP_OK;
  double t = now_s();
  try {
    std::vector<bp_network *> nv(n);
    std::vector<bp_feature_view> views(n);
    std::vector<int> oob(n);
    std::vector<float *> bmax(n);
    std::vector<int *> barg(n);
    std::vector<unsigned long long *> keys(n);
    for (int i = 0; i < n; ++i) {
      Request *r = batch[i];
      nv[i] = nets[r->net].net;
      views[i] = r->view;
      oob[i] = r->oob;
      bmax[i] = r->d_bmax;
      barg[i] = r->d_barg;
      keys[i] = r->d_keys;
    }
    f.reqs = batch;
    hipStream_t st = compute[slot];
    // Grid: windows that fill the GPU alone in launches of their own, the
    // others one work-group per item.  A persistent grid holds every CU
    // slot until its queue is empty, so a batch launched on the other slot
    // would wait for all of it; one work-group per item lets the hardware
    // interleave the two (p50 at 50 req/s: 6 against 12 ms persistent).
    const bp_batch_placement pl{static_cast<char *>(tables) + (size_t)slot * table_bytes,
                                table_bytes, st, -1};
    rc = bp_batch_create_placed(n, nv.data(), views.data(), oob.data(), bmax.data(),
                                barg.data(), keys.data(), pl, &f.batch);
    if (rc != BP_OK) fail(rc, bp_last_error());
    t = now_s();
    if (rc == BP_OK && (rc = bp_batch_launch(f.batch, st)) != BP_OK) fail(rc, bp_last_error());
    if (rc == BP_OK && hipEventCreateWithFlags(&f.ev, hipEventDisableTiming) != hipSuccess)
      rc = fail(BP_EHIP, "hipEventCreate failed");
    if (rc == BP_OK && hipEventRecord(f.ev, st) != hipSuccess)
      rc = fail(BP_EHIP, "hipEventRecord failed");
  } catch (...) {
    rc = fail(BP_ENOMEM, "batch set-up: host memory allocation failed");
  }
  std::unique_lock<std::mutex> lock(mu);
  for (Request *r : batch) {
    r->info.launch_s = t;
    r->info.batch_windows = n;
    if (rc != BP_OK) {
      r->info.status = rc;
      r->err = t_err;
    }
  }
  ++n_batches;
  n_batched_windows += n;
  if (rc != BP_OK) {
    for (Request *r : batch) {
      arena.release(r->off_keep, r->bytes_keep);
      r->off_keep = -1;
      r->done = true;
    }
    if (f.batch) {
      // its table uploads may still be queued on the slot's stream
      (void)hipStreamSynchronize(compute[slot]);
      bp_batch_destroy(f.batch);
    }
    if (f.ev) (void)hipEventDestroy(f.ev);
    free_slots.push_back(slot);
    cv_done.notify_all();
    cv_load.notify_all();
    return;
  }
  try {
    flights.push_back(std::move(f));
  } catch (...) {
    lock.unlock();
    complete(f);  // no room to keep it in flight: finish it here
  }
}

// Called with the lock released; the flight has been removed from the list.
void bp_server::complete(Flight &f) {
  int rc = hipEventSynchronize(f.ev) == hipSuccess ? BP_OK : fail(BP_EHIP, "batch failed");
  const double t = now_s();
  for (Request *r : f.reqs) {
    if (rc == BP_OK &&
        (hipMemcpy(r->bmax_host, r->d_bmax, r->T * 4, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(r->barg_host, r->d_barg, r->T * 4, hipMemcpyDeviceToHost) != hipSuccess))
      rc = fail(BP_EHIP, "result download failed");
  }
  const double t_out = now_s();
  bp_batch_destroy(f.batch);  // its tables are the slot's: nothing is freed on the device
  (void)hipEventDestroy(f.ev);
  std::lock_guard<std::mutex> lock(mu);
  free_slots.push_back(f.slot);
  for (Request *r : f.reqs) {
    r->info.gpu_done_s = t;
    r->info.done_s = t_out;
    if (r->info.status == BP_OK && rc != BP_OK) {
      r->info.status = rc;
      r->err = t_err;
    }
    arena.release(r->off_keep, r->bytes_keep);
    r->off_keep = -1;
    r->done = true;
    ++n_completed;
  }
  cv_done.notify_all();
  cv_load.notify_all();
  cv_sched.notify_all();
}

void bp_server::scheduler_loop() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    if (!ready.empty() && !free_slots.empty()) {
      // earliest deadline first, up to max_batch windows and about
      // kBatchSeconds of estimated work (the first request always goes)
      std::stable_sort(ready.begin(), ready.end(),
                       [](const Request *a, const Request *b) { return a->deadline < b->deadline; });
      const double cap = kBatchSeconds * kFmaRate;
      std::vector<Request *> batch;
      // reserved before any request leaves `ready`, so a failed allocation
      // leaves them all there
      batch.reserve(std::min<size_t>(ready.size(), (size_t)cfg.max_batch));
      double work = 0.0;
      while (!ready.empty() && (int)batch.size() < cfg.max_batch &&
             (batch.empty() || work + ready.front()->work <= cap)) {
        work += ready.front()->work;
        batch.push_back(ready.front());
        ready.pop_front();
      }
      const int slot = free_slots.back();
      free_slots.pop_back();
      lock.unlock();
      launch(std::move(batch), slot);
      lock.lock();
      continue;
    }
    if (!flights.empty()) {
      // complete the batches that are done (in any order); if every slot
      // is taken, or the server stops, wait for the oldest
      bool done_any = false;
      for (size_t i = 0; i < flights.size(); ++i) {
        if (hipEventQuery(flights[i].ev) != hipSuccess) continue;
        Flight f = std::move(flights[i]);
        flights.erase(flights.begin() + (long)i);
        lock.unlock();
        complete(f);
        lock.lock();
        done_any = true;
        break;
      }
      if (done_any) continue;
      if (stop) {
        Flight f = std::move(flights.front());
        flights.pop_front();
        lock.unlock();
        complete(f);
        lock.lock();
      } else {
        cv_sched.wait_for(lock, std::chrono::microseconds(50));
      }
      continue;
    }
    if (stop) break;
    cv_sched.wait(lock);
  }
}

extern "C" {

const char *bp_server_last_error(void) { return t_err.c_str(); }

int bp_server_create(const bp_server_config *cfg, bp_server **out) try {
  if (!out) return fail(BP_EINVAL, "out is NULL");
  *out = nullptr;
  if (!cfg) return fail(BP_EINVAL, "NULL configuration");
  if (cfg->fmt != BP_F32 && cfg->fmt != BP_BF16)
    return fail(BP_EINVAL, "the server stores windows as BP_F32 or BP_BF16 (BP_F16 scales are "
                           "per window and would rebuild a network's schedule per batch)");
  if (cfg->hbm_budget < (1u << 20) || cfg->max_batch < 1 || cfg->max_batch > 4096 ||
      cfg->in_flight < 1 || cfg->in_flight > 4)
    return fail(BP_EINVAL, "hbm_budget >= 1 MiB, max_batch in [1, 4096], in_flight in [1, 4]");
  bp_server *s = new bp_server();  // a failure here reaches BP_SERVER_CATCH: nothing to free
  hipError_t e = hipSuccess;
  try {
    s->cfg = *cfg;
    s->compute.reserve((size_t)cfg->in_flight);
    e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = s->arena.init(cfg->hbm_budget);
    s->table_bytes = bp_batch_table_bytes(cfg->max_batch);
    if (e == hipSuccess) e = hipMalloc(&s->tables, s->table_bytes * (size_t)cfg->in_flight);
    for (int i = cfg->in_flight - 1; i >= 0; --i) s->free_slots.push_back(i);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->load_stream, hipStreamNonBlocking);
    for (int i = 0; e == hipSuccess && i < cfg->in_flight; ++i) {
      hipStream_t st = nullptr;
      e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
      if (e == hipSuccess) s->compute.push_back(st);
    }
    if (e == hipSuccess) {
      s->loader = std::thread([s] { s->guarded(&bp_server::loader_loop); });
      s->scheduler = std::thread([s] { s->guarded(&bp_server::scheduler_loop); });
    }
  } catch (...) {
    bp_server_destroy(s);  // stops and joins a thread that did start
    return fail(BP_ENOMEM, "server set-up: host memory or worker threads unavailable");
  }
  if (e != hipSuccess) {
    bp_server_destroy(s);
    return fail(e == hipErrorOutOfMemory ? BP_ENOMEM : BP_EHIP,
                std::string("server set-up: ") + hipGetErrorString(e));
  }
  *out = s;
  return BP_OK;
} BP_SERVER_CATCH

int bp_server_add_network(bp_server *s, const int *time_delays, const float *weights_sources,
                          const float *weights_phases, size_t n_sources, size_t n_stations,
                          size_t n_channels, size_t n_phases, int *network_id) try {
  if (!s || !network_id || !weights_phases) return fail(BP_EINVAL, "NULL argument");
  Network n;
  int rc = bp_network_create(time_delays, weights_sources, n_sources, n_stations, n_phases,
                             &n.net);
  if (rc != BP_OK) return fail(rc, bp_last_error());
  rc = bp_network_prepare(n.net, s->cfg.fmt, nullptr);
  if (rc == BP_OK) rc = bp_network_padding(n.net, &n.pad_lo, &n.pad_hi);
  if (rc != BP_OK) {
    bp_network_destroy(n.net);
    return fail(rc, bp_last_error());
  }
  n.K = n_sources;
  n.S = n_stations;
  n.C = n_channels;
  n.P = n_phases;
  const size_t na = n_stations * n_channels * n_phases * sizeof(float);
  if (hipMalloc(&n.alpha, std::max<size_t>(na, 16)) != hipSuccess ||
      hipMemcpy(n.alpha, weights_phases, na, hipMemcpyHostToDevice) != hipSuccess) {
    if (n.alpha) (void)hipFree(n.alpha);
    bp_network_destroy(n.net);
    return fail(BP_ENOMEM, "phase weights upload failed");
  }
  std::lock_guard<std::mutex> lock(s->mu);
  try {
    s->nets.push_back(n);
  } catch (...) {
    (void)hipFree(n.alpha);
    bp_network_destroy(n.net);
    throw;
  }
  *network_id = (int)s->nets.size() - 1;
  return BP_OK;
} BP_SERVER_CATCH

int bp_server_submit(bp_server *s, int network_id, const float *waveform_features,
                     size_t n_samples, int out_of_bounds, float *beam_max, int *beam_argmax,
                     long long *request_id) try {
  if (!s || !waveform_features || !beam_max || !beam_argmax || !request_id)
    return fail(BP_EINVAL, "NULL argument");
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 or 1");
  if (n_samples == 0 || n_samples > (1u << 30)) return fail(BP_EINVAL, "bad n_samples");
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->broken) return fail(BP_ENOMEM, "server stopped after a worker failed (host memory)");
  if (network_id < 0 || network_id >= (int)s->nets.size())
    return fail(BP_EINVAL, "unknown network " + std::to_string(network_id));
  std::unique_ptr<Request> owned(new Request());
  Request *r = owned.get();
  r->id = s->next_id++;
  r->net = network_id;
  r->u_host = waveform_features;
  r->T = n_samples;
  r->oob = out_of_bounds;
  {
    const Network &n = s->nets[network_id];
    r->work = (double)n.K * (double)n.S * (double)n.P * (double)n_samples;
  }
  r->bmax_host = beam_max;
  r->barg_host = beam_argmax;
  r->info.submit_s = now_s();
  s->reqs[r->id] = r;
  try {
    s->pending.push_back(r);
  } catch (...) {
    s->reqs.erase(r->id);
    throw;
  }
  owned.release();  // owned by the server until bp_server_wait
  *request_id = r->id;
  s->cv_load.notify_all();
  return BP_OK;
} BP_SERVER_CATCH

int bp_server_wait(bp_server *s, long long request_id, double timeout_s, bp_request_info *info) try {
  if (!s) return fail(BP_EINVAL, "NULL server");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->reqs.find(request_id);
  if (it == s->reqs.end()) return fail(BP_EINVAL, "unknown request " + std::to_string(request_id));
  Request *r = it->second;
  const auto pred = [r] { return r->done; };
  if (timeout_s < 0)
    s->cv_done.wait(lock, pred);
  else if (!s->cv_done.wait_for(lock, std::chrono::duration<double>(timeout_s), pred))
    return 1;  // not done yet
  if (info) *info = r->info;
  const int status = r->info.status;
  if (status != BP_OK) fail(status, r->err);
  s->reqs.erase(it);
  delete r;
  return status;
} BP_SERVER_CATCH

int bp_server_stats(bp_server *s, bp_server_counters *st) try {
  if (!s || !st) return fail(BP_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lock(s->mu);
  st->pending = (long long)s->pending.size();
  st->ready = (long long)s->ready.size();
  st->in_flight_batches = (long long)s->flights.size();
  st->completed = s->n_completed;
  st->batches = s->n_batches;
  st->batched_windows = s->n_batched_windows;
  st->arena_bytes = s->arena.size();
  st->arena_used = s->arena.used();
  st->arena_peak = s->arena.peak();
  return BP_OK;
} BP_SERVER_CATCH

int bp_server_destroy(bp_server *s) try {
  if (!s) return BP_OK;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
  }
  s->cv_load.notify_all();
  s->cv_sched.notify_all();
  if (s->loader.joinable()) s->loader.join();
  if (s->scheduler.joinable()) s->scheduler.join();  // completes the batches in flight
  for (auto &kv : s->reqs) delete kv.second;
  for (hipStream_t st : s->compute) (void)hipStreamDestroy(st);
  if (s->load_stream) (void)hipStreamDestroy(s->load_stream);
  if (s->tables) (void)hipFree(s->tables);
  for (Network &n : s->nets) {
    if (n.alpha) (void)hipFree(n.alpha);
    bp_network_destroy(n.net);
  }
  delete s;
  return BP_OK;
} BP_SERVER_CATCH

}  // extern "C"
