// beamform_gpu.hip -- C ABI of the gfx950 beamforming library.
//
// Drop-in symbols beamform / beamform_max replace the reference's GPU
// library (beampower/src/beamform_gpu.h:1-2, bound by beampower/core.py:65-87);
// prestack_waveform_features restates the CPU library's symbol
// (beamform_cpu.h:4) on the GPU.  The bp_* entry points expose the same
// engine on device-resident buffers.
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "beamform_gpu.h"
#include "bp_internal.h"
#include "bp_kernels.h"
#include "bp_host_pool.h"
#include "bp_order.h"

#define BP_VERSION "0.1.0"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) noexcept {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  try {
    g_err = buf;
  } catch (...) {
    g_err.clear();  // no memory for the message: the status still tells
  }
  return code;
}

#define HIP_TRY(expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess)                                                        \
      return fail(e_ == hipErrorOutOfMemory ? BP_ENOMEM : BP_EHIP, "%s: %s (%s:%d)", \
                  #expr, hipGetErrorString(e_), __FILE__, __LINE__);             \
  } while (0)

#define BP_TRY(expr)          \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != BP_OK) return rc_; \
  } while (0)

// Handlers of the function-try-block of every int entry point: a host
// exception (a std::vector that cannot grow, a thread that cannot start)
// never crosses the C ABI, it becomes a status and a message.  The
// entry point's device buffers are released as its locals unwind.
#define BP_CATCH                                                        \
  catch (const std::bad_alloc &) {                                      \
    return fail(BP_ENOMEM, "host memory allocation failed");            \
  }                                                                     \
  catch (const std::length_error &) {                                   \
    return fail(BP_ENOMEM, "host allocation too large");                \
  }                                                                     \
  catch (const std::exception &e_) {                                    \
    return fail(BP_EINVAL, "host exception: %s", e_.what());            \
  }                                                                     \
  catch (...) {                                                         \
    return fail(BP_EINVAL, "unknown host exception");                   \
  }

// Kernel configuration of the LDS shift-and-stack kernel (overridable at
// build time for tuning sweeps).  6 sources per wave keep the accumulators
// (6 x 8 FP32 per lane) within the 80 VGPRs of 6 waves per SIMD; on the
// 50-station, 192k-source, 1 h workload 5 and 4 sources per wave measured
// 8 % and 17 % slower (more staging per FMA).
#ifndef BP_KW
#define BP_KW 6
#endif
#ifndef BP_W
#define BP_W 8
#endif
constexpr int kKW = BP_KW;  // sources per wave
constexpr int kW = BP_W;    // waves per work-group -> source tile KW*W
constexpr int kTT = bp::kTimeTile;  // time tile (samples)
constexpr int kTileSources = kKW * kW;
// Window classes: the small window covers BP_SMALL_SHARE_PCT % of the tiles
// (or as many as fit BP_SMALL_WINDOW samples), the large one the rest up to
// kMaxWindow; wider tiles go to the direct kernel.
#ifndef BP_SMALL_WINDOW
#define BP_SMALL_WINDOW (bp::kTimeTile + 512)
#endif
#ifndef BP_SMALL_SHARE_PCT
#define BP_SMALL_SHARE_PCT 97
#endif
constexpr int kSmallLw = BP_SMALL_WINDOW;  // largest window of the common tiles
constexpr int kLargeLw = bp::kMaxWindow;  // largest window
constexpr int kLdsPerCU = 160 * 1024;
constexpr int kNumCU = 256;        // MI355X
#ifndef BP_TIME_GROUP
#define BP_TIME_GROUP 8
#endif
constexpr int kTimeGroup = BP_TIME_GROUP;  // block order: groups of time tiles (bp_kernels.h, block_tile)
constexpr double kSmallShare = BP_SMALL_SHARE_PCT / 100.0;  // tiles the small window covers
constexpr int kDirectChunk = kTileSources;
// Work-groups per launch of the LDS kernels: a dispatch's grid holds fewer
// than 2^32 work-items; a multiple of 8 keeps the XCD ranges whole.
constexpr long long kMaxLaunchBlocks = ((0xffffffffLL / (64 * kW)) / 8) * 8;

template <int FMT, int P, int NB, bool NONE>
using Lds = bp::StackLds<FMT, P, kKW, kW, NB, NONE>;

// Device buffer owned by RAII, or borrowed from memory that someone else
// owns (borrow(): never freed here).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;  // as allocated
  bool own = true;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p && own) (void)hipFree(p);
  }
  void borrow(void *ptr, size_t n) {
    if (p && own) (void)hipFree(p);
    p = ptr;
    bytes = n;
    own = false;
  }
  int alloc(size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (p) {  // a buffer that is allocated again releases the old memory
      if (own) (void)hipFree(p);
      p = nullptr;
      this->bytes = 0;
    }
    own = true;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? BP_ENOMEM : BP_EHIP,
                  "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    this->bytes = bytes;
    return BP_OK;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

struct Stream {
  hipStream_t s = nullptr;
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
  int create() {
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return BP_OK;
  }
};

// The current device is a gfx950.  The answer is kept per device: reading
// the properties cost about a millisecond per host-pointer call.
int check_device() {
  constexpr int kMaxDevices = 64;
  static std::atomic<int> known[kMaxDevices];  // 0 unknown, 1 gfx950
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev >= 0 && dev < kMaxDevices && known[dev].load(std::memory_order_acquire) == 1)
    return BP_OK;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(BP_ENODEV, "device %d is %s; this library is built for gfx950", dev,
                prop.gcnArchName);
  if (dev >= 0 && dev < kMaxDevices) known[dev].store(1, std::memory_order_release);
  return BP_OK;
}

// Per-thread kernel timing (bp_profiling).
struct Profile {
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0.0;
  long long launches = 0;
};
thread_local Profile g_prof;

int prof_begin(hipStream_t st, hipEvent_t *ev) {
  if (!g_prof.on) return BP_OK;
  HIP_TRY(hipEventCreate(&ev[0]));
  HIP_TRY(hipEventCreate(&ev[1]));
  HIP_TRY(hipEventRecord(ev[0], st));
  return BP_OK;
}

int prof_end(hipStream_t st, hipEvent_t *ev) {
  if (!g_prof.on) return BP_OK;
  HIP_TRY(hipEventRecord(ev[1], st));
  g_prof.pending.emplace_back(ev[0], ev[1]);
  return BP_OK;
}

bool size_ok(size_t v) { return v <= (size_t)INT32_MAX - 4096; }

std::once_flag g_attr_once;
int g_attr_rc = BP_OK;

template <int FMT, int P, int NB, bool NONE>
hipError_t set_lds_attr() {
  return hipFuncSetAttribute(
      reinterpret_cast<const void *>(&bp::stack_lds_kernel<FMT, P, kKW, kW, NB, NONE>),
      hipFuncAttributeMaxDynamicSharedMemorySize,
      std::min(kLdsPerCU, Lds<FMT, P, NB, NONE>::lds_bytes(kLargeLw)));
}

template <int FMT, bool NONE>
hipError_t set_lds_attrs() {
  hipError_t r[] = {set_lds_attr<FMT, 1, 2, NONE>(), set_lds_attr<FMT, 2, 2, NONE>(),
                    set_lds_attr<FMT, 3, 2, NONE>(), set_lds_attr<FMT, 4, 2, NONE>(),
                    set_lds_attr<FMT, 1, 3, NONE>(), set_lds_attr<FMT, 2, 3, NONE>(),
                    set_lds_attr<FMT, 3, 3, NONE>(), set_lds_attr<FMT, 4, 3, NONE>()};
  for (hipError_t x : r)
    if (x != hipSuccess) return x;
  return hipSuccess;
}

int init_kernel_attrs() {
  std::call_once(g_attr_once, [] {
    hipError_t r[] = {set_lds_attrs<BP_F32, false>(), set_lds_attrs<BP_F32, true>(),
                      set_lds_attrs<BP_F16, false>(), set_lds_attrs<BP_BF16, false>(),
                      set_lds_attrs<BP_F16, true>(),  set_lds_attrs<BP_BF16, true>()};
    for (hipError_t e : r)
      if (e != hipSuccess) {
        g_attr_rc = fail(BP_EHIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
        break;
      }
  });
  return g_attr_rc;
}

// Row stride of the 16-bit parity planes: T + 1 rounded up to 8 elements
// (even, so every plane starts 4-byte aligned; 16-byte aligned in fact).
size_t parity_stride(size_t T) { return (T + 1 + 7) / 8 * 8; }

}  // namespace

// The LDS kernel's schedule for one feature format and one set of station
// scales: it encodes the copy layout of the format and the weights
// beta * scale.  A schedule is built once and never rewritten, so a batch
// or a call that holds one keeps its weights whatever scale sets the
// network is used with afterwards, and building a new one cannot race with
// kernels still reading another on some stream.
struct Schedule {
  std::vector<float> scales;  // host copy of the station scales (1 for BP_F32)
  DevBuf sched, scale;
};

// Source tiles are split in three classes by their largest per-(s,p) delay
// spread: class 0 uses the LDS kernel with the small window Lw[0], class 1
// the LDS kernel with Lw[1] <= kLargeLw, class 2 the direct kernel.  The
// LDS kernel's schedules are built per format and scale set on first use
// (see Schedule); the network keeps the most recently used ones.
struct bp_network {
  int K = 0, S = 0, P = 0;
  int n_tiles = 0;
  int max_spread = 0;
  int Lw[2] = {0, 0};
  int pad_lo = 0, pad_hi = 0;  // feature row padding that makes every LDS tile interior
  std::vector<int> tiles[3];  // ascending tile ids per class (host copy)
  DevBuf tau, beta, lohi, tile_lo, tile_all, tile_lw, dtiles[3];
  // per format, most recently used first (at most kSchedulesPerFormat)
  std::vector<std::shared_ptr<Schedule>> scheds[3];
  // differential mode: wave-transposed tables in the caller's source order,
  // each built on first use and kept: weights (float32, or used bits with
  // the one weight value diff_cw), int32 delays, paired int16 delays
  DevBuf diff_tau, diff_beta, diff_tau_pairs;
  bool diff_ok = false, diff_pairs_ok = false;
  int diff_wform = 0;  // weights: 0 not built, 1 float32 table, 2 used bits
  float diff_cw = 0.0f;
  // source order: tau, beta and the plan are in position order; perm maps a
  // position to the caller's index (device and host), inv the reverse (host)
  bool permuted = false;
  DevBuf perm;
  std::vector<int> order, inv;
  mutable std::mutex mu;  // schedule and table builds
};

namespace {

// Schedules of a network per format: a fifth scale set replaces the least
// recently used one.  Dropping a schedule that nothing else holds frees its
// device memory, and hipFree waits for the device to be idle first (it
// performs an implicit hipDeviceSynchronize), so no kernel still reads it.
constexpr size_t kSchedulesPerFormat = 4;

bp::StackArgs make_args(const bp_network *net, const void *F, int fmt, int T, int oob,
                        const Schedule &sc) {
  bp::StackArgs a{};
  a.F = F;
  a.ldf = fmt == BP_F32 ? (long long)T : (long long)parity_stride(T);
  a.scale = fmt == BP_F32 ? nullptr : sc.scale.as<float>();
  a.tau = net->tau.as<int>();
  a.beta = net->beta.as<float>();
  a.src_lohi = net->lohi.as<int>();
  a.perm = net->permuted ? net->perm.as<int>() : nullptr;
  a.tile_lo = net->tile_lo.as<int>();
  a.tile_all = net->tile_all.as<int>();
  a.sched = sc.sched.as<int>();
  a.T = T;
  a.K = net->K;
  a.S = net->S;
  a.P = net->P;
  a.oob = oob;
  a.time_group = kTimeGroup;
  return a;
}

// Arguments for a feature view: rows of row_stride elements, sample 0 at
// element pad_before, readable zeros around [0, T).
bp::StackArgs make_args_view(const bp_network *net, const bp_feature_view &v, int oob,
                             const Schedule &sc) {
  bp::StackArgs a = make_args(net, nullptr, v.fmt, (int)v.n_samples, oob, sc);
  const size_t es = v.fmt == BP_F32 ? 4 : 2;
  a.F = static_cast<const char *>(v.data) + v.pad_before * es;
  a.ldf = (long long)v.row_stride;
  a.pad_lo = (int)v.pad_before;
  a.pad_hi = (int)(v.row_stride - v.pad_before - v.n_samples);
  return a;
}

int check_shapes(size_t T, size_t K, size_t S, size_t P);

int check_view(const bp_network *net, const bp_feature_view &v) {
  if (!v.data) return fail(BP_EINVAL, "NULL features");
  if (v.fmt != BP_F32 && v.fmt != BP_F16 && v.fmt != BP_BF16)
    return fail(BP_EINVAL, "unknown feature format %d", v.fmt);
  if (v.fmt == BP_F32 && v.station_scale)
    return fail(BP_EINVAL, "station_scale applies to the 16-bit formats only");
  BP_TRY(check_shapes(v.n_samples, net->K, net->S, net->P));
  if (!size_ok(v.row_stride) || !size_ok(v.pad_before) ||
      v.row_stride < v.pad_before + v.n_samples)
    return fail(BP_EINVAL, "row_stride %zu cannot hold pad_before %zu + %zu samples",
                v.row_stride, v.pad_before, v.n_samples);
  if (v.fmt != BP_F32 && ((v.row_stride | v.pad_before) & 1))
    return fail(BP_EINVAL, "16-bit rows need an even row_stride and pad_before");
  return BP_OK;
}

bool has_lds_tiles(const bp_network *net) {
  return !net->tiles[0].empty() || !net->tiles[1].empty();
}

// The schedule of format `fmt` for the station scales `scales` (host
// pointer to n_stations floats, or NULL for 1): found among the network's
// recent ones or built now (synchronously on `st`, so that it is complete
// before any stream uses it).
int ensure_schedule(bp_network *net, int fmt, const float *scales, hipStream_t st,
                    std::shared_ptr<Schedule> &out) {
  std::lock_guard<std::mutex> lock(net->mu);
  std::vector<float> want(net->S, 1.0f);
  if (scales) want.assign(scales, scales + net->S);
  std::vector<std::shared_ptr<Schedule>> &list = net->scheds[fmt];
  for (size_t i = 0; i < list.size(); ++i)
    if (list[i]->scales == want) {
      std::rotate(list.begin(), list.begin() + i, list.begin() + i + 1);
      out = list[0];
      return BP_OK;
    }
  auto sc = std::make_shared<Schedule>();
  sc->scales = want;
  const long long n = (long long)net->n_tiles * net->S * kTileSources;
  BP_TRY(sc->scale.alloc((size_t)net->S * sizeof(float)));
  HIP_TRY(hipMemcpyAsync(sc->scale.p, want.data(), want.size() * sizeof(float),
                         hipMemcpyHostToDevice, st));
  if (has_lds_tiles(net)) {
    BP_TRY(sc->sched.alloc((size_t)n * (net->P + 1) * sizeof(int)));
    hipLaunchKernelGGL(bp::schedule_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       net->tau.as<int>(), net->beta.as<float>(), net->tile_lo.as<int>(),
                       net->tile_lw.as<int>(),
                       fmt == BP_F32 ? (const float *)nullptr : sc->scale.as<float>(), net->K,
                       net->S, net->P, kTileSources, kKW, fmt, n, sc->sched.as<int>());
    HIP_TRY(hipGetLastError());
  }
  // the host vector `want` is read by the copy above: finish before returning
  HIP_TRY(hipStreamSynchronize(st));
  if (list.size() >= kSchedulesPerFormat) list.pop_back();
  list.insert(list.begin(), sc);
  out = std::move(sc);
  return BP_OK;
}

// Staging buffers of the LDS kernel: 2 (copies of station s + 1 in flight
// while station s is consumed) measured faster than 3 on the 50-station,
// 192k-source, 1 h workload (66.2 against 85.6 ms: three buffers leave room
// for two work-groups per CU instead of three); 3 is taken only when asked
// for at build time and two work-groups still fit a CU's LDS.
#ifndef BP_STAGE_BUFFERS
#define BP_STAGE_BUFFERS 2
#endif
// `blocks` work-groups in launches of at most kMaxLaunchBlocks each.
template <int FMT, int P, int NB, bool NONE>
void launch_lds_chunks(bp::StackArgs a, long long blocks, hipStream_t st) {
  for (long long b0 = 0; b0 < blocks; b0 += kMaxLaunchBlocks) {
    a.block_base = (int)b0;
    const unsigned n = (unsigned)std::min(kMaxLaunchBlocks, blocks - b0);
    hipLaunchKernelGGL((bp::stack_lds_kernel<FMT, P, kKW, kW, NB, NONE>), dim3(n),
                       dim3(64 * kW), (Lds<FMT, P, NB, NONE>::lds_bytes(a.Lw)), st, a);
  }
}

template <int FMT, int P, bool NONE>
void launch_lds_p(const bp::StackArgs &a, long long blocks, hipStream_t st) {
  if (BP_STAGE_BUFFERS == 3 && Lds<FMT, P, 3, NONE>::lds_bytes(a.Lw) <= kLdsPerCU / 2)
    launch_lds_chunks<FMT, P, 3, NONE>(a, blocks, st);
  else
    launch_lds_chunks<FMT, P, 2, NONE>(a, blocks, st);
}

template <int FMT, bool NONE>
void launch_lds(int P, const bp::StackArgs &a, long long blocks, hipStream_t st) {
  switch (P) {
    case 1: launch_lds_p<FMT, 1, NONE>(a, blocks, st); break;
    case 2: launch_lds_p<FMT, 2, NONE>(a, blocks, st); break;
    case 3: launch_lds_p<FMT, 3, NONE>(a, blocks, st); break;
    default: launch_lds_p<FMT, 4, NONE>(a, blocks, st);
  }
}

// Launch the stack kernels for sources [k_begin, k_end): every class's
// tiles overlapping the range.
template <int FMT, bool NONE>
int launch_stack(const bp_network *net, bp::StackArgs a, int k_begin, int k_end,
                 hipStream_t st, bool profile = true) {
  a.k_begin = k_begin;
  a.k_end = k_end;
  if (k_end <= k_begin || a.T <= 0) return BP_OK;
  // tiles holding the sources [k_begin, k_end): with a source order, the
  // range of their positions (tiles in between write nothing)
  int tile0 = k_begin / kTileSources;
  int tile1 = (k_end - 1) / kTileSources + 1;
  if (net->permuted && (k_begin > 0 || k_end < net->K)) {
    int lo = INT_MAX, hi = -1;
    for (int k = k_begin; k < k_end; ++k) {
      lo = std::min(lo, net->inv[k]);
      hi = std::max(hi, net->inv[k]);
    }
    tile0 = lo / kTileSources;
    tile1 = hi / kTileSources + 1;
  }
  hipEvent_t ev[2];
  if (profile) BP_TRY(prof_begin(st, ev));
  for (int cls = 0; cls < 3; ++cls) {
    const std::vector<int> &v = net->tiles[cls];
    const auto lo = std::lower_bound(v.begin(), v.end(), tile0);
    const auto hi = std::lower_bound(v.begin(), v.end(), tile1);
    const int n = (int)(hi - lo);
    if (n == 0) continue;
    a.tiles = net->dtiles[cls].as<int>() + (lo - v.begin());
    a.n_list = n;
    if (cls < 2) {
      // staging offsets within a station are 31-bit
      const long long planes = FMT == BP_F32 ? 1 : 2, es = FMT == BP_F32 ? 4 : 2;
      if ((net->P * planes * a.ldf + kLargeLw + 16) * es >= (1LL << 31))
        return fail(BP_EINVAL, "n_samples %d too large for the LDS kernel", a.T);
      BP_TRY(init_kernel_attrs());
      a.Lw = net->Lw[cls];
      a.n_time_tiles = (a.T + kTT - 1) / kTT;
      const long long blocks = (long long)a.n_time_tiles * n;
      if (blocks > INT32_MAX) return fail(BP_EINVAL, "grid too large (%lld blocks)", blocks);
      launch_lds<FMT, NONE>(net->P, a, blocks, st);
    } else {
      if (n > 65535) return fail(BP_EINVAL, "too many direct-kernel tiles (%d)", n);
      const unsigned gx = (unsigned)((a.T + 255) / 256);
      hipLaunchKernelGGL((bp::stack_direct_kernel<FMT, NONE>), dim3(gx, (unsigned)n), dim3(256),
                         0, st, a, kDirectChunk);
    }
    HIP_TRY(hipGetLastError());
  }
  if (profile) BP_TRY(prof_end(st, ev));
  return BP_OK;
}

// reduce='max' with schedule `sc` and key scratch `keys` (n_samples uint64):
// key fill, stack kernels, decode, all on `st`.  Allocates nothing (the
// kernel timer creates events when profiling is on and `profile` is set).
int network_max_with(bp_network *net, const bp_feature_view &v, int oob, float *bmax,
                     int *barg, unsigned long long *keys, hipStream_t st, const Schedule &sc,
                     bool profile) {
  const int T = (int)v.n_samples;
  const int fmt = v.fmt;
  const unsigned nb = (unsigned)((T + 255) / 256);
  hipLaunchKernelGGL(bp::fill_u64_kernel, dim3(nb), dim3(256), 0, st, keys, (long long)T,
                     0ull);
  HIP_TRY(hipGetLastError());
  bp::StackArgs a = make_args_view(net, v, oob, sc);
  a.keys = keys;
  if (fmt == BP_F32)
    BP_TRY((launch_stack<BP_F32, false>(net, a, 0, net->K, st, profile)));
  else if (fmt == BP_F16)
    BP_TRY((launch_stack<BP_F16, false>(net, a, 0, net->K, st, profile)));
  else
    BP_TRY((launch_stack<BP_BF16, false>(net, a, 0, net->K, st, profile)));
  hipLaunchKernelGGL(bp::finalize_max_kernel, dim3(nb), dim3(256), 0, st, keys, T, bmax,
                     barg);
  HIP_TRY(hipGetLastError());
  return BP_OK;
}

int network_max(bp_network *net, const bp_feature_view &v, int oob, float *bmax, int *barg,
                unsigned long long *keys, hipStream_t st) {
  std::shared_ptr<Schedule> sc;
  BP_TRY(ensure_schedule(net, v.fmt, v.fmt == BP_F32 ? nullptr : v.station_scale, st, sc));
  return network_max_with(net, v, oob, bmax, barg, keys, st, *sc, true);
}

int round_up8(long long v) { return (int)((v + 7) / 8 * 8); }

bp_feature_view plain_view(const void *F, int fmt, const float *scale, size_t T) {
  bp_feature_view v{};
  v.data = F;
  v.fmt = fmt;
  v.n_samples = T;
  v.row_stride = fmt == BP_F32 ? T : parity_stride(T);
  v.pad_before = 0;
  v.station_scale = scale;
  return v;
}

// Classify tiles by delay spread and choose the two LDS window lengths:
// multiples of 8 samples (whole 16-byte words of 16-bit samples) of at
// least TT + spread + 1 (the staging start of 16-bit samples is rounded
// down to even, one sample before the least delay at most).
// Largest LDS window of P phases whose staging buffers (float32, two
// buffers, the reduce='max' layout: the largest of the kernel's variants)
// fit one CU's LDS: 2,048 samples for one or two phases, less for more.
template <int P>
int lds_window_limit() {
  int lw = kLargeLw;
  while (lw > kTT && Lds<BP_F32, P, 2, false>::lds_bytes(lw) > kLdsPerCU) lw -= 64;
  return lw;
}

int lds_window_limit(int P) {
  switch (P) {
    case 1: return lds_window_limit<1>();
    case 2: return lds_window_limit<2>();
    case 3: return lds_window_limit<3>();
    case 4: return lds_window_limit<4>();
    default: return 0;
  }
}

void classify_tiles(bp_network *net, const std::vector<int> &tile_all) {
  const int n = net->n_tiles;
  std::vector<int> spreads;
  spreads.reserve(n);
  for (int j = 0; j < n; ++j) spreads.push_back(tile_all[8 * j + 3]);
  std::vector<int> sorted = spreads;
  std::sort(sorted.begin(), sorted.end());
  net->max_spread = sorted.empty() ? 0 : sorted.back();
  const int lw_max = lds_window_limit(net->P);  // 0: no LDS kernel for this phase count
  const bool lds = lw_max > 0;
  const int q = sorted.empty() ? 0 : sorted[std::min<size_t>(sorted.size() - 1,
                                                   (size_t)(kSmallShare * sorted.size()))];
  // the small window covers kSmallShare of the tiles, or as many as fit
  const int lw0 = std::min({kSmallLw, lw_max, round_up8((long long)kTT + q + 1)});
  int lw1 = round_up8((long long)kTT + net->max_spread + 1);
  lw1 = std::min(lw1, lw_max);
  if (lw0 > 0 && lw1 <= lw0) lw1 = 0;
  net->Lw[0] = lw0;
  net->Lw[1] = lw1;
  for (int c = 0; c < 3; ++c) net->tiles[c].clear();
  for (int j = 0; j < n; ++j) {
    const long long need = (long long)kTT + spreads[j] + 1;
    // a tile with an infinite or NaN weight goes to the direct kernel, which
    // drops out-of-range terms explicitly (b * 0 would be NaN)
    const bool ok = lds && tile_all[8 * j + 4] == 0;
    int cls = 2;
    if (ok && lw0 > 0 && need <= lw0)
      cls = 0;
    else if (ok && lw1 > 0 && need <= lw1)
      cls = 1;
    net->tiles[cls].push_back(j);
  }
  // Padding of the feature rows under which no LDS tile needs the guarded
  // edge path: a tile stages [t0 + lo, t0 + maxlo + Lw + 7) with lo rounded
  // down to even, t0 >= 0 and t0 <= T - 1.
  int lo_min = 0, hi_max = 0;
  for (int c = 0; c < 2; ++c)
    for (int j : net->tiles[c]) {
      lo_min = std::min(lo_min, tile_all[8 * j] & ~1);
      hi_max = std::max(hi_max, tile_all[8 * j + 2] + net->Lw[c] + 6);
    }
  net->pad_lo = round_up8(-(long long)lo_min);
  net->pad_hi = round_up8(hi_max);
}

// Bounds, tile classes and tile lists of a network whose delays and
// weights (in position order) are on the device.
int build_plan(bp_network *net, hipStream_t st) {
  const size_t K = net->K, S = net->S, P = net->P, tsp = (size_t)net->n_tiles * S * P;
  BP_TRY(net->lohi.alloc(2 * K * sizeof(int)));
  BP_TRY(net->tile_lo.alloc(tsp * sizeof(int)));
  BP_TRY(net->tile_all.alloc(8 * (size_t)net->n_tiles * sizeof(int)));
  hipLaunchKernelGGL(bp::source_bounds_kernel,
                     dim3((unsigned)((K + bp::kBoundsSources - 1) / bp::kBoundsSources)),
                     dim3(64 * bp::kBoundsSources), 0, st, net->tau.as<int>(),
                     net->beta.as<float>(), net->K, net->S, net->P, net->lohi.as<int>());
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(bp::tile_bounds_kernel, dim3((unsigned)net->n_tiles), dim3(256), 0, st,
                     net->tau.as<int>(), net->beta.as<float>(), net->K, net->S, net->P,
                     kTileSources, net->tile_lo.as<int>(), net->tile_all.as<int>());
  HIP_TRY(hipGetLastError());
  std::vector<int> tile_all(8 * (size_t)net->n_tiles);
  HIP_TRY(hipMemcpyAsync(tile_all.data(), net->tile_all.p, tile_all.size() * sizeof(int),
                         hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  classify_tiles(net, tile_all);

  std::vector<int> tile_lw(net->n_tiles, 0);
  for (int c = 0; c < 2; ++c)
    for (int j : net->tiles[c]) tile_lw[j] = net->Lw[c];
  BP_TRY(net->tile_lw.alloc(tile_lw.size() * sizeof(int)));
  HIP_TRY(hipMemcpyAsync(net->tile_lw.p, tile_lw.data(), tile_lw.size() * sizeof(int),
                         hipMemcpyHostToDevice, st));
  for (int c = 0; c < 3; ++c) {
    BP_TRY(net->dtiles[c].alloc(std::max<size_t>(1, net->tiles[c].size()) * sizeof(int)));
    if (!net->tiles[c].empty())
      HIP_TRY(hipMemcpyAsync(net->dtiles[c].p, net->tiles[c].data(),
                             net->tiles[c].size() * sizeof(int), hipMemcpyHostToDevice, st));
  }
  HIP_TRY(hipStreamSynchronize(st));  // the host vectors are read by the copies
  return BP_OK;
}

// Relative cost of a plan's stack launches: a tile of an LDS class costs
// 440 + its window length (a linear fit of the measured time per tile of
// 592- and 1,200-sample windows on the 50-station, 192k-source, 1 h
// workload: the staging grows with the window, the gathers do not), a
// direct-kernel tile ten 1,024-sample tiles.
double plan_cost(const bp_network *net) {
  double c = 0.0;
  for (int cls = 0; cls < 2; ++cls) c += (double)net->tiles[cls].size() * (440.0 + net->Lw[cls]);
  return c + (double)net->tiles[2].size() * 10.0 * (440.0 + 1024.0);
}

// Order policy of the per-call entry points, which plan a network for one
// call of n_samples: the similarity order is computed only when the most it
// can save on that call exceeds what computing it costs; otherwise as
// BP_ORDER_NATURAL.
constexpr int kOrderAutoWide = 3;

// Seconds per plan-cost unit, sample and station-phase of the stack
// kernels: 65.7 ms for 4,000 tiles of 440 + 592 units, 100 station-phases
// and 90,000 samples (50 stations, 192,000 sources, 1 h at 25 Hz).
constexpr double kUnitSeconds = 1.75e-15;

// The order cannot make a tile's window narrower than TT + 8 samples, so
// the natural plan's cost above that floor bounds the saving (53 ms at
// 90,000 samples for a shuffled 192,000-source grid, whose similarity plan
// saves about 50).  Computing the order takes about 10 ns per delay entry
// divided by the host threads (19 ms for 192,000 x 100 entries and 48 ms
// for 300,000 x 240 on 16 threads, 100 and 322 ms on one), plus a second
// plan (a few ms).
bool order_pays(const bp_network *net, size_t n_samples) {
  if (net->K <= kTileSources || !has_lds_tiles(net)) return false;
  const double floor = (double)net->n_tiles * (440.0 + kTT + 8);
  const double sp = (double)net->S * net->P;
  const double saving = (plan_cost(net) - floor) * sp * (double)n_samples * kUnitSeconds;
  const double entries = (double)net->K * sp;
  const double threads = (double)bp::order_detail::order_threads((size_t)entries);
  const double cost = entries * 1e-8 / threads + 5e-3;
  return saving > cost;
}

// ---- host-only planning (bp_plan) -------------------------------------------

// Host twin of tile_bounds_kernel's per-tile summary (bp_kernels.h): for
// source tile j, tile_all[8j..8j+4] = least delay, greatest delay, greatest
// per-(s,p) least delay, largest per-(s,p) spread over the tile's users of
// each (s,p), and 1 if a used weight is infinite or NaN; a tile using no
// station gets 0, 0, 0, 0.  Rows are in position order.
std::vector<int> host_tile_all(const int *tau, const float *beta, size_t K, size_t S, size_t P,
                               int n_tiles) {
  std::vector<int> all(8 * (size_t)n_tiles, 0);
  const size_t SP = S * P;
  for (int j = 0; j < n_tiles; ++j) {
    const size_t k0 = (size_t)j * kTileSources, k1 = std::min(K, k0 + kTileSources);
    int t_lo = INT_MAX, t_hi = INT_MIN, t_maxlo = INT_MIN, t_spread = 0, nonfinite = 0;
    for (size_t sp = 0; sp < SP; ++sp) {
      const size_t st = sp / P;
      int lo = INT_MAX, hi = INT_MIN;
      for (size_t k = k0; k < k1; ++k) {
        const float b = beta[k * S + st];
        if (b == 0.0f) continue;
        if (!std::isfinite(b)) nonfinite = 1;
        const int d = tau[k * SP + sp];
        lo = std::min(lo, d);
        hi = std::max(hi, d);
      }
      if (lo <= hi) {
        const long long spread = (long long)hi - (long long)lo;
        t_lo = std::min(t_lo, lo);
        t_hi = std::max(t_hi, hi);
        t_maxlo = std::max(t_maxlo, lo);
        t_spread = std::max(t_spread, spread > INT_MAX ? INT_MAX : (int)spread);
      }
    }
    int *o = all.data() + 8 * (size_t)j;
    const bool unused = t_lo == INT_MAX;
    o[0] = unused ? 0 : t_lo;
    o[1] = unused ? 0 : t_hi;
    o[2] = unused ? 0 : t_maxlo;
    o[3] = t_spread;
    o[4] = nonfinite;
  }
  return all;
}

// The plan bp_network_create_ordered would make for these host moveouts
// and weights, in a host-only network object (no device buffers): tiles,
// windows, padding and the source-order decision, by the same
// classify_tiles and plan_cost.
int host_network_plan(const int *tau, const float *beta, size_t K, size_t S, size_t P,
                      int order, std::unique_ptr<bp_network> &out) {
  auto make = [&]() {
    auto n = std::make_unique<bp_network>();
    n->K = (int)K;
    n->S = (int)S;
    n->P = (int)P;
    n->n_tiles = (int)((K + kTileSources - 1) / kTileSources);
    return n;
  };
  auto net = make();
  classify_tiles(net.get(), host_tile_all(tau, beta, K, S, P, net->n_tiles));
  const bool try_order =
      order == BP_ORDER_SIMILARITY || (order == BP_ORDER_AUTO && K > (size_t)kTileSources);
  if (try_order) {
    std::vector<int> ord = bp::similarity_order(tau, K, S * P, kTileSources);
    std::vector<int> tau_p(K * S * P);
    std::vector<float> beta_p(K * S);
    for (size_t i = 0; i < K; ++i) {
      std::memcpy(&tau_p[i * S * P], tau + (size_t)ord[i] * S * P, S * P * sizeof(int));
      std::memcpy(&beta_p[i * S], beta + (size_t)ord[i] * S, S * sizeof(float));
    }
    auto alt = make();
    classify_tiles(alt.get(), host_tile_all(tau_p.data(), beta_p.data(), K, S, P, alt->n_tiles));
    if (order == BP_ORDER_SIMILARITY || plan_cost(alt.get()) < plan_cost(net.get())) {
      alt->permuted = true;
      net = std::move(alt);
    }
  }
  out = std::move(net);
  return BP_OK;
}

// LDS bytes per work-group of the stack kernel for a window of Lw samples
// (the staging buffers launch_lds_p would pick).
template <int FMT, int P, bool NONE>
int lds_bytes_of(int Lw) {
  if (BP_STAGE_BUFFERS == 3 && Lds<FMT, P, 3, NONE>::lds_bytes(Lw) <= kLdsPerCU / 2)
    return Lds<FMT, P, 3, NONE>::lds_bytes(Lw);
  return Lds<FMT, P, 2, NONE>::lds_bytes(Lw);
}

template <int FMT, bool NONE>
int lds_bytes_of(int P, int Lw) {
  switch (P) {
    case 1: return lds_bytes_of<FMT, 1, NONE>(Lw);
    case 2: return lds_bytes_of<FMT, 2, NONE>(Lw);
    case 3: return lds_bytes_of<FMT, 3, NONE>(Lw);
    case 4: return lds_bytes_of<FMT, 4, NONE>(Lw);
    default: return 0;
  }
}

int lds_bytes_of(int fmt, int P, bool none, int Lw) {
  if (Lw <= 0) return 0;
  if (fmt == BP_F32) return none ? lds_bytes_of<BP_F32, true>(P, Lw) : lds_bytes_of<BP_F32, false>(P, Lw);
  if (fmt == BP_F16) return none ? lds_bytes_of<BP_F16, true>(P, Lw) : lds_bytes_of<BP_F16, false>(P, Lw);
  return none ? lds_bytes_of<BP_BF16, true>(P, Lw) : lds_bytes_of<BP_BF16, false>(P, Lw);
}

// Runs f(begin, end) over [0, n) bytes split across up to 16 host threads
// of the library's worker pool (pieces of at least `piece` bytes; small
// inputs stay on the calling thread).
template <class F>
void parallel_bytes(size_t n, F f, size_t piece = (size_t)8 << 20) {
  const size_t nt = std::min<size_t>(16, (n + piece - 1) / piece);
  if (nt <= 1) {
    f((size_t)0, n);
    return;
  }
  const size_t step = (n + nt - 1) / nt;
  bp::HostPool::shared().run(nt, [&](size_t i) {
    const size_t b = std::min(n, i * step), e = std::min(n, b + step);
    if (b < e) f(b, e);
  });
}

bool same_bytes(const void *a, const void *b, size_t n) {
  std::atomic<bool> same{true};
  parallel_bytes(n, [&](size_t lo, size_t hi) {
    if (std::memcmp(static_cast<const char *>(a) + lo, static_cast<const char *>(b) + lo,
                    hi - lo) != 0)
      same = false;
  });
  return same;
}

void copy_bytes(void *dst, const void *src, size_t n, size_t piece = (size_t)8 << 20) {
  parallel_bytes(
      n,
      [&](size_t lo, size_t hi) {
        std::memcpy(static_cast<char *>(dst) + lo, static_cast<const char *>(src) + lo,
                    hi - lo);
      },
      piece);
}

// Uploads from host memory the caller owns (pageable: the runtime's bounce
// buffers reach about 20 GB/s).  Here the data is copied by several host
// threads into one of two pinned 16 MiB buffers while the other one's DMA
// runs (about 50 GB/s over the link).  Two such pairs are allocated once per
// process, so two uploads can run at once (a call's features on one host
// thread, its moveouts and weights on another); an upload that finds both
// in use, or that is smaller than two buffers, copies from pageable memory.
// upload_host returns when its last DMA has completed (the buffers are then
// free for the next upload).
struct PinnedStage {
  std::mutex mu;
  bool tried = false;  // one attempt to allocate (a failed one is not repeated)
  int device = -1;
  void *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};
constexpr size_t kPinnedBytes = (size_t)16 << 20;
constexpr size_t kPinnedPiece = (size_t)2 << 20;  // per host thread of a chunk's copy

constexpr int kPinnedStages = 2;

PinnedStage *pinned_stages() {
  static PinnedStage *p = new PinnedStage[kPinnedStages];  // never freed: no device call at exit
  return p;
}

// A pinned stage for this thread (held by `lock`), or nullptr when both
// are in use or could not be allocated on the current device.
PinnedStage *acquire_pinned(std::unique_lock<std::mutex> &lock) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  PinnedStage *stages = pinned_stages();
  int k = 0;
  for (; k < kPinnedStages; ++k) {
    lock = std::unique_lock<std::mutex>(stages[k].mu, std::try_to_lock);
    if (lock.owns_lock()) break;
  }
  if (k == kPinnedStages) return nullptr;
  PinnedStage &ps = stages[k];
  if (!ps.tried) {
    ps.tried = true;
    if (hipHostMalloc(&ps.buf[0], kPinnedBytes, hipHostMallocDefault) == hipSuccess &&
        hipHostMalloc(&ps.buf[1], kPinnedBytes, hipHostMallocDefault) == hipSuccess &&
        hipEventCreateWithFlags(&ps.ev[0], hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&ps.ev[1], hipEventDisableTiming) == hipSuccess) {
      ps.device = dev;
    }
  }
  if (ps.device != dev) {
    lock = std::unique_lock<std::mutex>();
    return nullptr;
  }
  return &ps;
}

// n bytes to dst through the stage ps: fill(pinned, offset, len) writes
// bytes [offset, offset + len) of the data into a pinned buffer.  Chunks
// of 2, 2, 4, 8 and then 16 MiB (the first DMA starts after 2 MiB are
// staged instead of 16); every chunk but the last is a multiple of 2 MiB.
// Declared after the stage's lock, so that it runs first: on an early
// return (an error, an exception of a host thread) DMAs of the stage's
// buffers may still be queued on the stream, and they must end before the
// stage goes to its next user.
struct DrainOnExit {
  hipStream_t st;
  bool armed = true;
  ~DrainOnExit() {
    if (armed) (void)hipStreamSynchronize(st);
  }
};

template <class Fill>
int upload_pinned(void *dst, size_t n, hipStream_t st, PinnedStage *ps, Fill fill) {
  size_t done = 0;
  for (int i = 0; done < n; ++i) {
    const int j = i & 1;
    const size_t len = std::min({kPinnedBytes, n - done, std::max(kPinnedPiece, done)});
    if (i >= 2) HIP_TRY(hipEventSynchronize(ps->ev[j]));  // buffer j's previous DMA is done
    fill(ps->buf[j], done, len);
    HIP_TRY(hipMemcpyAsync(static_cast<char *>(dst) + done, ps->buf[j], len,
                           hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(ps->ev[j], st));
    done += len;
  }
  HIP_TRY(hipEventSynchronize(ps->ev[0]));
  HIP_TRY(hipEventSynchronize(ps->ev[1]));
  return BP_OK;
}

int upload_host(void *dst, const void *src, size_t n, hipStream_t st) {
  std::unique_lock<std::mutex> lock;
  PinnedStage *ps = n < 2 * kPinnedBytes ? nullptr : acquire_pinned(lock);
  if (!ps) {
    HIP_TRY(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st));
    return BP_OK;
  }
  DrainOnExit drain{st};
  const int rc = upload_pinned(dst, n, st, ps, [&](void *pin, size_t off, size_t len) {
    copy_bytes(pin, static_cast<const char *>(src) + off, len, kPinnedPiece);
  });
  drain.armed = rc != BP_OK;
  return rc;
}

// The two phases' delays of each (source, station pair) as one word of two
// int16 (phase 0 in the low half), each saturated to the int16 range: the
// words diff_pack_pairs_kernel makes on the device, made here by the host
// threads that stage the upload, so half the delay bytes cross PCIe.
// (K, S, 2) int32 -> (K, S) words on the device.  Returns false, having
// done nothing, when the delays are under 32 MiB or no pinned stage is
// free (the caller then uploads them as they are).
inline unsigned sat16_host(int v) {
  return (unsigned)(v < -32768 ? -32768 : v > 32767 ? 32767 : v) & 0xffffu;
}
int upload_packed_pairs(unsigned *dst, const int *tau, size_t n_words, hipStream_t st,
                        bool &packed) {
  packed = false;
  std::unique_lock<std::mutex> lock;
  PinnedStage *ps = 2 * n_words * sizeof(int) < 2 * kPinnedBytes ? nullptr : acquire_pinned(lock);
  if (!ps) return BP_OK;
  DrainOnExit drain{st};
  BP_TRY(upload_pinned(dst, n_words * sizeof(unsigned), st, ps,
                       [&](void *pin, size_t off, size_t len) {
                         unsigned *out = static_cast<unsigned *>(pin);
                         const int *in = tau + 2 * (off / sizeof(unsigned));
                         parallel_bytes(
                             len / sizeof(unsigned),
                             [&](size_t lo, size_t hi) {
                               for (size_t i = lo; i < hi; ++i)
                                 out[i] = sat16_host(in[2 * i]) | (sat16_host(in[2 * i + 1]) << 16);
                             },
                             kPinnedPiece / sizeof(unsigned));
                       }));
  drain.armed = false;
  packed = true;
  return BP_OK;
}

// The used-bit weight table of the differential kernel, built by host
// threads from (K, S) weights when every nonzero weight has one bit
// pattern (a weight is nonzero when it compares unequal to 0.0f: -0.0 is
// zero, NaN is not): the words diff_weight_bits_kernel and diff_mask_kernel
// make on the device, in their layout -- word w of source k holds bit s % 32
// for station s = 32 w + (s % 32), at ((k / 64) * W + w) * 64 + k % 64, zero
// for padded sources and stations.  cw: the weight (1.0 when none is
// nonzero).  Returns false as soon as two nonzero patterns differ.
bool host_used_bits(const float *beta, size_t K, size_t S, int W,
                    std::unique_ptr<unsigned[]> &words, size_t &n_words, float &cw) {
  const size_t nblk = (K + 63) / 64;
  n_words = nblk * (size_t)W * 64;
  words.reset(new unsigned[n_words]);  // every word is written below (not zeroed first)
  std::atomic<bool> differ{false};
  std::mutex mu;
  bool any = false;
  unsigned lo = 0u, hi = 0u;
  parallel_bytes(
      nblk,
      [&](size_t b0, size_t b1) {
        bool tany = false;
        unsigned tlo = 0u, thi = 0u;
        for (size_t b = b0; b < b1 && !differ.load(std::memory_order_relaxed); ++b) {
          for (size_t r = 0; r < 64; ++r) {
            const size_t k = b * 64 + r;
            for (int w = 0; w < W; ++w) {
              unsigned m = 0u;
              const size_t s0 = (size_t)w * 32;
              if (k < K && s0 < S) {
                const float *row = beta + k * S;
                const size_t s1 = std::min(S, s0 + 32);
                for (size_t s = s0; s < s1; ++s) {
                  const float v = row[s];
                  if (v != 0.0f) {
                    unsigned u;
                    std::memcpy(&u, &v, 4);
                    tlo = tany ? std::min(tlo, u) : u;
                    thi = tany ? std::max(thi, u) : u;
                    tany = true;
                    m |= 1u << (s - s0);
                  }
                }
              }
              words[(b * (size_t)W + (size_t)w) * 64 + r] = m;
            }
          }
          if (tany && tlo != thi) differ = true;
        }
        std::lock_guard<std::mutex> lock(mu);
        if (tany) {
          lo = any ? std::min(lo, tlo) : tlo;
          hi = any ? std::max(hi, thi) : thi;
          any = true;
        }
      },
      256);
  if (differ || lo != hi) return false;
  if (!any) {
    cw = 1.0f;  // no nonzero weight: the bits are all clear, cw is not applied
  } else {
    std::memcpy(&cw, &lo, 4);
  }
  return true;
}

// Device -> host memory the caller owns, after the work queued on `st`:
// 16 MiB chunks go by DMA into one pinned buffer while host threads copy
// the previous chunk out of the other (the runtime's pageable path, one
// thread behind bounce buffers, moved a 1.47 GB reduce='none' result at
// about 10 GB/s).  Returns when the data is in `dst`.
int download_host(void *dst, const void *src, size_t n, hipStream_t st) {
  std::unique_lock<std::mutex> lock;
  PinnedStage *ps = n < 2 * kPinnedBytes ? nullptr : acquire_pinned(lock);
  if (!ps) {
    HIP_TRY(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return BP_OK;
  }
  DrainOnExit drain{st};
  size_t prev_off = 0, prev_len = 0;
  for (size_t done = 0, i = 0; done < n || prev_len; ++i) {
    const int j = (int)(i & 1);
    const size_t len = std::min(kPinnedBytes, n - done);
    if (len) {
      // buffer j was last copied out two iterations ago, on this thread
      HIP_TRY(hipMemcpyAsync(ps->buf[j], static_cast<const char *>(src) + done, len,
                             hipMemcpyDeviceToHost, st));
      HIP_TRY(hipEventRecord(ps->ev[j], st));
    }
    if (prev_len) {
      const int pj = j ^ 1;
      HIP_TRY(hipEventSynchronize(ps->ev[pj]));
      copy_bytes(static_cast<char *>(dst) + prev_off, ps->buf[pj], prev_len, kPinnedPiece);
    }
    prev_off = done;
    prev_len = len;
    done += len;
  }
  drain.armed = false;
  return BP_OK;
}

// n_samples: the samples of the one call a kOrderAutoWide plan is made for.
int network_create(const int *tau, const float *beta, size_t K, size_t S, size_t P,
                   int order, hipStream_t st, std::unique_ptr<bp_network> &out,
                   size_t n_samples = 0) {
  auto make = [&]() {
    auto n = std::make_unique<bp_network>();
    n->K = (int)K;
    n->S = (int)S;
    n->P = (int)P;
    n->n_tiles = (int)((K + kTileSources - 1) / kTileSources);
    return n;
  };
  const size_t ksp = K * S * P, ks = K * S;
  auto net = make();
  BP_TRY(net->tau.alloc(ksp * sizeof(int)));
  BP_TRY(net->beta.alloc(ks * sizeof(float)));
  BP_TRY(upload_host(net->tau.p, tau, ksp * sizeof(int), st));
  BP_TRY(upload_host(net->beta.p, beta, ks * sizeof(float), st));
  BP_TRY(build_plan(net.get(), st));

  bool try_order = order == BP_ORDER_SIMILARITY || (order == BP_ORDER_AUTO && K > kTileSources);
  if (order == kOrderAutoWide) try_order = order_pays(net.get(), n_samples);
  if (try_order) {
    std::vector<int> ord = bp::similarity_order(tau, K, S * P, kTileSources);
    auto alt = make();
    BP_TRY(alt->perm.alloc(K * sizeof(int)));
    BP_TRY(alt->tau.alloc(ksp * sizeof(int)));
    BP_TRY(alt->beta.alloc(ks * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(alt->perm.p, ord.data(), K * sizeof(int), hipMemcpyHostToDevice, st));
    const unsigned gb = (unsigned)std::min<size_t>(4096, (ksp + 255) / 256);
    hipLaunchKernelGGL(bp::gather_rows_kernel, dim3(gb), dim3(256), 0, st,
                       net->tau.as<unsigned>(), alt->perm.as<int>(), (long long)K, (int)(S * P),
                       alt->tau.as<unsigned>());
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(bp::gather_rows_kernel, dim3(gb), dim3(256), 0, st,
                       net->beta.as<unsigned>(), alt->perm.as<int>(), (long long)K, (int)S,
                       alt->beta.as<unsigned>());
    HIP_TRY(hipGetLastError());
    BP_TRY(build_plan(alt.get(), st));  // synchronizes: `ord` stays alive until then
    if (order == BP_ORDER_SIMILARITY || plan_cost(alt.get()) < plan_cost(net.get())) {
      alt->permuted = true;
      alt->inv.assign(K, 0);
      for (size_t i = 0; i < K; ++i) alt->inv[ord[i]] = (int)i;
      alt->order = std::move(ord);
      net = std::move(alt);
    }
  }
  std::shared_ptr<Schedule> sc;
  BP_TRY(ensure_schedule(net.get(), BP_F32, nullptr, st, sc));
  out = std::move(net);
  return BP_OK;
}

int check_shapes(size_t T, size_t K, size_t S, size_t P) {
  if (!size_ok(T) || !size_ok(K) || !size_ok(S) || !size_ok(P))
    return fail(BP_EINVAL, "dimension too large (T=%zu K=%zu S=%zu P=%zu)", T, K, S, P);
  if (S * P > (size_t)INT32_MAX - 4096)
    return fail(BP_EINVAL, "n_stations * n_phases = %zu is too large", S * P);
  if (K > 0 && S > 0 && P > 0 && (K * S * P) / (S * P) != K)
    return fail(BP_EINVAL, "K*S*P overflows");
  return BP_OK;
}

// Shapes with no station, phase or source: every beam is empty.  Strict
// mode has no valid source (lo/hi stay at their sentinels); flexible mode
// gives beam 0 for every source (beamform.c:74-105 with no terms).
void degenerate_max(size_t T, size_t K, int oob, float *bmax, int *barg) {
  for (size_t t = 0; t < T; ++t) {
    if (oob == 1 && K > 0) bmax[t] = 0.0f;
    barg[t] = 0;
  }
}

// ---- plan reuse of the host-pointer entry points ---------------------------
//
// A caller of the reference's interface passes the same moveouts and
// weights on every call of a detection loop (one call per window of data).
// The network built for them (device tables, plan, schedules) is kept with a
// host copy of those inputs, and a later call on the same device whose
// moveouts and weights are equal byte for byte (compared in full, on several
// host threads) reuses it instead of uploading and planning them again; on
// its first reuse the network is re-planned once with the similarity source
// order where that is cheaper for the kernels.  Results do not depend on
// any of this: the plan is a function of the inputs, and the kernels'
// results do not depend on the plan's source order.  One network is kept
// (the last one built); bp_dropin_cache_clear() releases it, and
// BP_DROPIN_CACHE=0 in the environment turns the reuse off.

struct DropinCache {
  std::mutex mu;
  int device = -1;
  size_t K = 0, S = 0, P = 0;
  std::unique_ptr<int[]> tau;  // host copies (not zero-filled first: copied in parallel)
  std::unique_ptr<float[]> beta;
  std::shared_ptr<bp_network> net;
  long long reuses = 0;
};

// Never destroyed: its device memory must not be freed by a static
// destructor after the runtime has shut down.
DropinCache &dropin_cache() {
  static DropinCache *c = new DropinCache;
  return *c;
}

bool dropin_cache_on() {
  const char *e = std::getenv("BP_DROPIN_CACHE");
  return !(e && e[0] == '0');
}

// The network of one host-pointer call.  keep: built by this call, to be
// kept; replan: reused for the first time, to be re-planned with the
// similarity order.  Both are done by dropin_finish while the call's
// kernels run.
struct DropinUse {
  std::shared_ptr<bp_network> net;
  bool keep = false, replan = false;
  int device = 0;
};

// The kept network when the inputs match, else a new one (planned for this
// call's n_samples).
int dropin_network(const int *tau, const float *beta, size_t K, size_t S, size_t P, size_t T,
                   hipStream_t st, DropinUse &u) {
  const bool on = dropin_cache_on();
  HIP_TRY(hipGetDevice(&u.device));
  DropinCache &c = dropin_cache();
  if (on) {
    std::lock_guard<std::mutex> lock(c.mu);
    if (c.net && c.device == u.device && c.K == K && c.S == S && c.P == P &&
        same_bytes(c.tau.get(), tau, K * S * P * sizeof(int)) &&
        same_bytes(c.beta.get(), beta, K * S * sizeof(float))) {
      u.net = c.net;
      u.replan = c.reuses++ == 0 && !c.net->permuted && K > (size_t)kTileSources;
      return BP_OK;
    }
  }
  std::unique_ptr<bp_network> net;
  BP_TRY(network_create(tau, beta, K, S, P, kOrderAutoWide, st, net, T));
  u.net = std::shared_ptr<bp_network>(std::move(net));
  u.keep = on;
  return BP_OK;
}

// After the call's kernels are launched on `st`: keep the new network with
// host copies of its inputs (published once `st` has finished, so that
// other threads find its tables complete), or re-plan the reused one on a
// stream of its own.  The call's inputs are still the caller's arrays.
// Keeping and re-planning are best effort: when memory for them is short
// the call still succeeds, and only an error of the call's own stream is
// returned.
int dropin_finish(const int *tau, const float *beta, size_t K, size_t S, size_t P,
                  hipStream_t st, const DropinUse &u) {
  DropinCache &c = dropin_cache();
  if (u.keep) {
    std::unique_ptr<int[]> tau_h(new (std::nothrow) int[K * S * P]);
    std::unique_ptr<float[]> beta_h(new (std::nothrow) float[K * S]);
    if (tau_h && beta_h) {
      copy_bytes(tau_h.get(), tau, K * S * P * sizeof(int));
      copy_bytes(beta_h.get(), beta, K * S * sizeof(float));
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (!tau_h || !beta_h) return BP_OK;
    std::lock_guard<std::mutex> lock(c.mu);
    c.tau = std::move(tau_h);
    c.beta = std::move(beta_h);
    c.device = u.device;
    c.K = K;
    c.S = S;
    c.P = P;
    c.reuses = 0;
    c.net = u.net;
  } else if (u.replan) {
    try {
      Stream rs;
      std::unique_ptr<bp_network> net;
      if (rs.create() != BP_OK ||
          network_create(tau, beta, K, S, P, BP_ORDER_AUTO, rs.s, net) != BP_OK ||
          hipStreamSynchronize(rs.s) != hipSuccess)
        return BP_OK;  // the reused network stays
      std::lock_guard<std::mutex> lock(c.mu);
      if (c.net == u.net) c.net = std::shared_ptr<bp_network>(std::move(net));
    } catch (...) {
      return BP_OK;
    }
  }
  return BP_OK;
}

// The network and the raw (S, T, P) features of a host-pointer call: the
// features' upload runs on a second host thread while the moveouts and
// weights are compared with the kept ones (or their network is built).
int dropin_inputs(const float *host_stp, size_t nf, const int *tau, const float *beta,
                  size_t K, size_t S, size_t P, size_t T, hipStream_t st, DevBuf &f_stp,
                  DropinUse &use) {
  BP_TRY(f_stp.alloc(nf * sizeof(float)));
  if (nf == 0) return dropin_network(tau, beta, K, S, P, T, st, use);
  int up_rc = BP_OK, dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::string up_err;  // the upload thread's message, reported by this thread
  std::thread up([&] {
    try {
      up_rc = hipSetDevice(dev) == hipSuccess  // a new thread starts on device 0
                  ? upload_host(f_stp.p, host_stp, nf * sizeof(float), st)
                  : fail(BP_EHIP, "hipSetDevice(%d) failed", dev);
    } catch (...) {
      up_rc = fail(BP_EHIP, "feature upload failed");
    }
    if (up_rc != BP_OK) up_err = g_err;
  });
  int rc = BP_OK;
  try {
    rc = dropin_network(tau, beta, K, S, P, T, st, use);
  } catch (...) {
    up.join();
    throw;
  }
  up.join();
  BP_TRY(rc);
  if (up_rc != BP_OK) return fail(up_rc, "%s", up_err.c_str());
  return BP_OK;
}

// Raw (S, T, P) features on the device -> planar rows padded for the
// network (no LDS tile takes the guarded edge path).
int stage_dropin_features(const bp_network *net, const DevBuf &f_stp, size_t S, size_t T,
                          size_t P, DevBuf &f_pad, bp_feature_view &v, hipStream_t st) {
  const size_t ld = (size_t)net->pad_lo + T + (size_t)net->pad_hi;
  BP_TRY(f_pad.alloc(S * P * ld * sizeof(float)));
  hipLaunchKernelGGL(bp::interleaved_to_planar_kernel,
                     dim3((unsigned)((ld + 255) / 256), (unsigned)S), dim3(256), 0, st,
                     f_stp.as<float>(), (int)S, (int)T, (int)P, (long long)ld,
                     (long long)net->pad_lo, f_pad.as<float>());
  HIP_TRY(hipGetLastError());
  v = bp_feature_view{};
  v.data = f_pad.p;
  v.fmt = BP_F32;
  v.n_samples = T;
  v.row_stride = ld;
  v.pad_before = (size_t)net->pad_lo;
  return BP_OK;
}

// Features of a host-pointer call: prestacked (S, T, P), as beamform_max
// and beamform take them, or raw (S, C, T) with their (S, C, P) phase
// weights, as the fused entry points take them (raw; with C = 0 neither
// array is read).
struct HostFeatures {
  const float *data = nullptr;
  const float *alpha = nullptr;
  size_t C = 0;
  bool raw = false;
};

// Network and padded FP32 rows of a host-pointer call.  Raw features are
// prestacked on the device straight into the padded rows
// (beamform.c:108-152, the FMA chain of prestack_kernel).
int dropin_stage(const HostFeatures &hf, const int *tau, const float *beta, size_t K,
                 size_t S, size_t P, size_t T, hipStream_t st, DropinUse &use, DevBuf &f_in,
                 DevBuf &d_alpha, DevBuf &f_pad, bp_feature_view &v) {
  if (!hf.raw) {
    BP_TRY(dropin_inputs(hf.data, S * T * P, tau, beta, K, S, P, T, st, f_in, use));
    return stage_dropin_features(use.net.get(), f_in, S, T, P, f_pad, v, st);
  }
  const size_t C = hf.C;
  BP_TRY(d_alpha.alloc(S * C * P * sizeof(float)));
  if (C > 0)
    HIP_TRY(hipMemcpyAsync(d_alpha.p, hf.alpha, S * C * P * sizeof(float),
                           hipMemcpyHostToDevice, st));
  BP_TRY(dropin_inputs(hf.data, S * C * T, tau, beta, K, S, P, T, st, f_in, use));
  const bp_network *net = use.net.get();
  const size_t ld = (size_t)net->pad_lo + T + (size_t)net->pad_hi;
  BP_TRY(f_pad.alloc(S * P * ld * sizeof(float)));
  hipLaunchKernelGGL(bp::prestack_rows_kernel, dim3((unsigned)((ld + 255) / 256), (unsigned)S),
                     dim3(256), 0, st, f_in.as<float>(), d_alpha.as<float>(), (int)S, (int)C,
                     (int)T, (int)P, (long long)ld, (long long)net->pad_lo, f_pad.as<float>());
  HIP_TRY(hipGetLastError());
  v = bp_feature_view{};
  v.data = f_pad.p;
  v.fmt = BP_F32;
  v.n_samples = T;
  v.row_stride = ld;
  v.pad_before = (size_t)net->pad_lo;
  return BP_OK;
}

// beamform_max (beamform.c:230-336) of a host-pointer call.
int dropin_max(const HostFeatures &hf, const int *time_delays, const float *weights_sources,
               size_t n_samples, size_t n_sources, size_t n_stations, size_t n_phases,
               int out_of_bounds, float *beam_max, int *beam_argmax) {
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 (strict) or 1 (flexible), not %d",
                out_of_bounds);
  BP_TRY(check_shapes(n_samples, n_sources, n_stations, n_phases));
  if (n_samples == 0) return BP_OK;
  if (!beam_max || !beam_argmax) return fail(BP_EINVAL, "NULL output");
  if (n_sources == 0 || n_stations == 0 || n_phases == 0) {
    degenerate_max(n_samples, n_sources, out_of_bounds, beam_max, beam_argmax);
    return BP_OK;
  }
  if ((!hf.data && !(hf.raw && hf.C == 0)) || (hf.raw && hf.C > 0 && !hf.alpha) ||
      !time_delays || !weights_sources)
    return fail(BP_EINVAL, "NULL input");
  BP_TRY(check_device());
  Stream st;
  BP_TRY(st.create());
  DropinUse use;
  const size_t T = n_samples;
  bp_feature_view v{};
  DevBuf f_in, d_alpha, f_pad, keys, bmax, barg;
  BP_TRY(dropin_stage(hf, time_delays, weights_sources, n_sources, n_stations, n_phases, T,
                      st.s, use, f_in, d_alpha, f_pad, v));
  bp_network *net = use.net.get();
  BP_TRY(keys.alloc(T * sizeof(unsigned long long)));
  BP_TRY(bmax.alloc(T * sizeof(float)));
  BP_TRY(barg.alloc(T * sizeof(int)));
  // the caller's beam_max is kept where no source is valid (beamform.c:332-333)
  HIP_TRY(hipMemcpyAsync(bmax.p, beam_max, T * sizeof(float), hipMemcpyHostToDevice, st.s));
  BP_TRY(network_max(net, v, out_of_bounds, bmax.as<float>(), barg.as<int>(),
                     keys.as<unsigned long long>(), st.s));
  BP_TRY(dropin_finish(time_delays, weights_sources, n_sources, n_stations, n_phases, st.s, use));
  HIP_TRY(hipMemcpyAsync(beam_max, bmax.p, T * sizeof(float), hipMemcpyDeviceToHost, st.s));
  HIP_TRY(hipMemcpyAsync(beam_argmax, barg.p, T * sizeof(int), hipMemcpyDeviceToHost, st.s));
  HIP_TRY(hipStreamSynchronize(st.s));
  return BP_OK;
}

// beamform (reduce='none', beamform.c:154-228) of a host-pointer call.
int dropin_none(const HostFeatures &hf, const int *time_delays, const float *weights_sources,
                size_t n_samples, size_t n_sources, size_t n_stations, size_t n_phases,
                int out_of_bounds, float *beam) {
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 (strict) or 1 (flexible), not %d",
                out_of_bounds);
  BP_TRY(check_shapes(n_samples, n_sources, n_stations, n_phases));
  if (n_samples == 0 || n_sources == 0) return BP_OK;
  if (!beam) return fail(BP_EINVAL, "NULL output");
  const size_t T = n_samples, K = n_sources;
  if (n_stations == 0 || n_phases == 0) {
    // no terms: strict leaves every cell invalid (0), flexible gives 0 beams
    std::memset(beam, 0, K * T * sizeof(float));
    return BP_OK;
  }
  if ((!hf.data && !(hf.raw && hf.C == 0)) || (hf.raw && hf.C > 0 && !hf.alpha) ||
      !time_delays || !weights_sources)
    return fail(BP_EINVAL, "NULL input");
  BP_TRY(check_device());
  Stream st;
  BP_TRY(st.create());
  DropinUse use;
  bp_feature_view v{};
  DevBuf f_in, d_alpha, f_pad, dbeam;
  BP_TRY(dropin_stage(hf, time_delays, weights_sources, n_sources, n_stations, n_phases, T,
                      st.s, use, f_in, d_alpha, f_pad, v));
  bp_network *net = use.net.get();
  BP_TRY(dbeam.alloc(K * T * sizeof(float)));
  // invalid strict cells read back as 0, as the wrapper's zero-filled buffer
  const long long n = (long long)(K * T);
  const unsigned fill_blocks = (unsigned)std::min<long long>(65536, (n + 255) / 256);
  hipLaunchKernelGGL(bp::fill_f32_kernel, dim3(fill_blocks), dim3(256), 0, st.s,
                     dbeam.as<float>(), n, 0.0f);
  HIP_TRY(hipGetLastError());
  std::shared_ptr<Schedule> sc;
  BP_TRY(ensure_schedule(net, BP_F32, nullptr, st.s, sc));
  bp::StackArgs a = make_args_view(net, v, out_of_bounds, *sc);
  a.beam = dbeam.as<float>();
  BP_TRY((launch_stack<BP_F32, true>(net, a, 0, (int)K, st.s)));
  BP_TRY(dropin_finish(time_delays, weights_sources, n_sources, n_stations, n_phases, st.s, use));
  return download_host(beam, dbeam.p, K * T * sizeof(float), st.s);
}

// Channels of a fused call: C fits the prestack kernel and S * C * T
// elements are addressable.
int check_channels(size_t T, size_t S, size_t C) {
  if (!size_ok(C) || (S > 0 && C > 0 && T > 0 && (S * C * T) / (S * C) != T))
    return fail(BP_EINVAL, "n_stations * n_channels * n_samples overflows (C=%zu)", C);
  return BP_OK;
}

// ---------------------------------------------------------------------------
// Batched form: many windows (network, features, length, outputs) in one
// persistent launch per LDS window class.
// ---------------------------------------------------------------------------

template <int FMT, int P, int NB>
hipError_t set_batch_attr() {
  return hipFuncSetAttribute(
      reinterpret_cast<const void *>(&bp::stack_batch_kernel<FMT, P, kKW, kW, NB>),
      hipFuncAttributeMaxDynamicSharedMemorySize,
      std::min(kLdsPerCU, Lds<FMT, P, NB, false>::lds_bytes(kLargeLw)));
}

std::once_flag g_batch_attr_once;
int g_batch_attr_rc = BP_OK;

int init_batch_attrs() {
  std::call_once(g_batch_attr_once, [] {
    hipError_t r[] = {
        set_batch_attr<BP_F32, 1, 2>(),  set_batch_attr<BP_F32, 2, 2>(),
        set_batch_attr<BP_F32, 3, 2>(),  set_batch_attr<BP_F32, 4, 2>(),
        set_batch_attr<BP_F32, 1, 3>(),  set_batch_attr<BP_F32, 2, 3>(),
        set_batch_attr<BP_F32, 3, 3>(),  set_batch_attr<BP_F32, 4, 3>(),
        set_batch_attr<BP_F16, 1, 2>(),  set_batch_attr<BP_F16, 2, 2>(),
        set_batch_attr<BP_F16, 3, 2>(),  set_batch_attr<BP_F16, 4, 2>(),
        set_batch_attr<BP_F16, 1, 3>(),  set_batch_attr<BP_F16, 2, 3>(),
        set_batch_attr<BP_F16, 3, 3>(),  set_batch_attr<BP_F16, 4, 3>(),
        set_batch_attr<BP_BF16, 1, 2>(), set_batch_attr<BP_BF16, 2, 2>(),
        set_batch_attr<BP_BF16, 3, 2>(), set_batch_attr<BP_BF16, 4, 2>(),
        set_batch_attr<BP_BF16, 1, 3>(), set_batch_attr<BP_BF16, 2, 3>(),
        set_batch_attr<BP_BF16, 3, 3>(), set_batch_attr<BP_BF16, 4, 3>()};
    for (hipError_t e : r)
      if (e != hipSuccess) {
        g_batch_attr_rc = fail(BP_EHIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
        break;
      }
  });
  return g_batch_attr_rc;
}

}  // namespace

// Windows of one LDS class run by one fused launch: descriptors (host copy
// and device), their work items numbered consecutively, the XCD ranges.
// Default grid of a batch: one persistent launch per window class, three
// work-groups per CU (the occupancy of the stack kernel's one-per-item
// grid) over a work queue.  On the 256-window batched workload it equals
// the windows launched one by one; on a serving-like mix it is 3 % faster.
constexpr int kBatchBlocksPerCU = 3;

struct FusedSet {
  std::vector<bp::BatchWindow> wins;
  DevBuf dwins;
  DevBuf dsplit;            // 9 item indices: XCD x walks [split[x], split[x+1])
  DevBuf dqueue;            // 8 work-queue counters of a persistent launch
  long long split_h[9] = {};  // host copy (the source of an asynchronous upload)
  long long n_items = 0;
  long long max_range = 0;  // largest XCD range (items)
  int max_lw = 0;
  void add(const bp::StackArgs &a) {
    wins.push_back(bp::BatchWindow{a, n_items});
    n_items += (long long)a.n_time_tiles * a.n_list;
    max_lw = std::max(max_lw, a.Lw);
  }
};

struct bp_batch {
  int fmt = BP_F32, P = 0, n = 0, max_T = 0;
  bp_batch_config cfg{kTimeGroup, 2, kBatchBlocksPerCU, 0.0f};
  // per LDS class: every window fused (blocks_per_cu >= 0), or the windows
  // that fill the GPU alone in launches of their own and the others fused
  // (blocks_per_cu < 0)
  FusedSet all[2], small[2];
  std::vector<bp::StackArgs> large[2];
  std::vector<bp::StackArgs> direct;  // per window with direct-kernel tiles
  std::vector<int> direct_n;
  // the windows' schedules: held, so that later calls of their networks
  // with other station scales leave this batch's weights as they were
  std::vector<std::shared_ptr<Schedule>> scheds;
  DevBuf dkeys;
  std::vector<bp::BatchKeys> keys_h;  // host copy of dkeys
  // Where the device tables live: hipMalloc'd (ws == NULL), or carved from a
  // caller's workspace and uploaded asynchronously on `upload` (a batch of
  // the serving loop: no allocation, no synchronisation, no hipFree).
  char *ws = nullptr;
  size_t ws_bytes = 0, ws_used = 0;
  hipStream_t upload = nullptr;
  hipStream_t cap = nullptr;  // private stream for graph capture
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~bp_batch() { drop_graph(); if (cap) (void)hipStreamDestroy(cap); }
  // device memory of `bytes` for a table
  int table(DevBuf &d, size_t bytes) {
    if (!ws) return d.alloc(bytes);
    const size_t off = (ws_used + 255) / 256 * 256;
    if (off + bytes > ws_bytes)
      return fail(BP_EINVAL, "batch tables need more than the %zu-byte workspace", ws_bytes);
    d.borrow(ws + off, bytes);
    ws_used = off + bytes;
    return BP_OK;
  }
  // host -> device copy of a table (asynchronous on `upload` when placed)
  int put(DevBuf &d, const void *src, size_t bytes) {
    if (upload) {
      HIP_TRY(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, upload));
    } else {
      HIP_TRY(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
    }
    return BP_OK;
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

namespace {

// A window of at least this many work items (source tile x time tile)
// fills the GPU's 512 work-group slots 8 times over.
constexpr long long kLargeWindowItems = 4096;

// Tuning results of this process, keyed by the batch's shape.
std::mutex g_tune_mu;
std::map<std::tuple<int, int, long long, long long, int, int>, bp_batch_config> g_tune_cache;

std::tuple<int, int, long long, long long, int, int> tune_key(const bp_batch *b) {
  return std::make_tuple(b->fmt, b->P, b->all[0].n_items, b->all[1].n_items, b->all[0].max_lw,
                         b->all[1].max_lw);
}

// Stage buffers usable for class c: 3 only when two work-groups still fit a CU.
template <int FMT, int P>
int class_nb(const bp_batch *b, int c) {
  return (b->cfg.stage_buffers == 3 &&
          Lds<FMT, P, 3, false>::lds_bytes(b->all[c].max_lw) <= kLdsPerCU / 2) ? 3 : 2;
}

// Upload a fused set's descriptors and cut its items into 8 XCD ranges at
// equal shares of the cost (items x stations): windows differ 8x in
// stations, and equal item counts made the fused launch 23 % slower than
// the windows one by one.
int finish_fused(bp_batch *b, FusedSet &f) {
  if (f.wins.empty()) return BP_OK;
  BP_TRY(b->table(f.dwins, f.wins.size() * sizeof(bp::BatchWindow)));
  const std::vector<bp::BatchWindow> &ws = f.wins;
  auto items_of = [&](size_t w) {
    return (w + 1 < ws.size() ? ws[w + 1].item_begin : f.n_items) - ws[w].item_begin;
  };
  double total = 0.0;
  for (size_t w = 0; w < ws.size(); ++w) total += (double)items_of(w) * ws[w].a.S;
  long long *split = f.split_h;
  split[0] = 0;
  split[8] = f.n_items;
  size_t w = 0;
  double before = 0.0;  // cost of the windows before w
  for (int x = 1; x < 8; ++x) {
    const double target = total * x / 8.0;
    for (;;) {
      const long long items = items_of(w);
      const double cost = (double)items * ws[w].a.S;
      if (before + cost >= target || w + 1 == ws.size()) {
        const long long k = (long long)std::ceil((target - before) / ws[w].a.S);
        split[x] = std::min(f.n_items, ws[w].item_begin + std::max(0LL, std::min(items, k)));
        break;
      }
      before += cost;
      ++w;
    }
    split[x] = std::max(split[x], split[x - 1]);
  }
  for (int x = 0; x < 8; ++x) f.max_range = std::max(f.max_range, split[x + 1] - split[x]);
  BP_TRY(b->table(f.dsplit, sizeof(f.split_h)));
  BP_TRY(b->table(f.dqueue, 8 * sizeof(unsigned long long)));
  return b->put(f.dsplit, f.split_h, sizeof(f.split_h));
}

// One window's tiles of class c in a launch of its own: the kernel a window
// alone runs.
template <int FMT, int P>
int launch_window_alone(const bp::StackArgs &a, int nb, hipStream_t st) {
  const long long blocks = (long long)a.n_time_tiles * a.n_list;
  if (blocks > INT32_MAX) return fail(BP_EINVAL, "grid too large (%lld blocks)", blocks);
  if (nb == 3)
    launch_lds_chunks<FMT, P, 3, false>(a, blocks, st);
  else
    launch_lds_chunks<FMT, P, 2, false>(a, blocks, st);
  HIP_TRY(hipGetLastError());
  return BP_OK;
}

// A fused launch over the windows of `f`.  blocks_per_cu > 0: a persistent
// grid of that many work-groups per CU over a work queue; otherwise one
// work-group per item (8 x the largest XCD range).
template <int FMT, int P>
int launch_fused(const FusedSet &f, int nb, int blocks_per_cu, hipStream_t st) {
  if (f.n_items == 0) return BP_OK;
  const int lds = nb == 3 ? Lds<FMT, P, 3, false>::lds_bytes(f.max_lw)
                          : Lds<FMT, P, 2, false>::lds_bytes(f.max_lw);
  long long want;
  if (blocks_per_cu <= 0) {
    want = 8 * f.max_range;
  } else {
    const int per_cu = std::max(1, std::min(blocks_per_cu, kLdsPerCU / lds));
    want = std::min((long long)kNumCU * per_cu, (f.n_items + 7) / 8 * 8);
  }
  // the kernel walks its XCD's items in a loop: any grid covers them.  A
  // persistent grid takes its items from a work queue, cleared here by a
  // kernel (no memset node when the batch is captured in a graph).
  const unsigned grid = (unsigned)std::max<long long>(8, std::min(want, kMaxLaunchBlocks));
  unsigned long long *queue = nullptr;
  if (blocks_per_cu > 0) {
    queue = f.dqueue.as<unsigned long long>();
    hipLaunchKernelGGL(bp::fill_u64_kernel, dim3(1), dim3(64), 0, st, queue, 8LL, 0ull);
    HIP_TRY(hipGetLastError());
  }
  if (nb == 3)
    hipLaunchKernelGGL((bp::stack_batch_kernel<FMT, P, kKW, kW, 3>), dim3(grid), dim3(64 * kW),
                       lds, st, f.dwins.as<bp::BatchWindow>(), (int)f.wins.size(),
                       f.dsplit.as<long long>(), queue);
  else
    hipLaunchKernelGGL((bp::stack_batch_kernel<FMT, P, kKW, kW, 2>), dim3(grid), dim3(64 * kW),
                       lds, st, f.dwins.as<bp::BatchWindow>(), (int)f.wins.size(),
                       f.dsplit.as<long long>(), queue);
  HIP_TRY(hipGetLastError());
  return BP_OK;
}

template <int FMT, int P>
int launch_batch_class(bp_batch *b, int c, hipStream_t st) {
  if (b->all[c].n_items == 0) return BP_OK;
  const int nb = class_nb<FMT, P>(b, c);
  if (b->cfg.blocks_per_cu >= 0) return launch_fused<FMT, P>(b->all[c], nb, b->cfg.blocks_per_cu, st);
  for (const bp::StackArgs &a : b->large[c]) BP_TRY((launch_window_alone<FMT, P>(a, nb, st)));
  return launch_fused<FMT, P>(b->small[c], nb, 0, st);
}

template <int FMT>
int launch_batch_fmt(bp_batch *b, hipStream_t st) {
  for (int c = 0; c < 2; ++c) {
    switch (b->P) {
      case 1: BP_TRY((launch_batch_class<FMT, 1>(b, c, st))); break;
      case 2: BP_TRY((launch_batch_class<FMT, 2>(b, c, st))); break;
      case 3: BP_TRY((launch_batch_class<FMT, 3>(b, c, st))); break;
      case 4: BP_TRY((launch_batch_class<FMT, 4>(b, c, st))); break;
      default: break;  // no LDS tiles: every tile is a direct one
    }
  }
  for (size_t i = 0; i < b->direct.size(); ++i) {
    const bp::StackArgs &a = b->direct[i];
    const unsigned gx = (unsigned)((a.T + 255) / 256);
    hipLaunchKernelGGL((bp::stack_direct_kernel<FMT, false>), dim3(gx, (unsigned)b->direct_n[i]),
                       dim3(256), 0, st, a, kDirectChunk);
    HIP_TRY(hipGetLastError());
  }
  return BP_OK;
}

// Enqueue the whole batch: keys, stack launches, decode.  No host
// synchronisation and no allocation, so it can be captured in a graph.
int batch_enqueue(bp_batch *b, hipStream_t st, bool profile) {
  const unsigned gx = (unsigned)std::min(1024, (b->max_T + 255) / 256);
  hipLaunchKernelGGL(bp::fill_keys_batch_kernel, dim3(std::max(1u, gx), (unsigned)b->n),
                     dim3(256), 0, st, b->dkeys.as<bp::BatchKeys>());
  HIP_TRY(hipGetLastError());
  hipEvent_t ev[2];
  if (profile) BP_TRY(prof_begin(st, ev));
  if (b->fmt == BP_F32)
    BP_TRY(launch_batch_fmt<BP_F32>(b, st));
  else if (b->fmt == BP_F16)
    BP_TRY(launch_batch_fmt<BP_F16>(b, st));
  else
    BP_TRY(launch_batch_fmt<BP_BF16>(b, st));
  if (profile) BP_TRY(prof_end(st, ev));
  hipLaunchKernelGGL(bp::finalize_max_batch_kernel, dim3(std::max(1u, gx), (unsigned)b->n),
                     dim3(256), 0, st, b->dkeys.as<bp::BatchKeys>());
  HIP_TRY(hipGetLastError());
  return BP_OK;
}

// Apply a configuration: the block order lives in the window descriptors.
int batch_configure(bp_batch *b, const bp_batch_config &c) {
  if (c.time_group < 1 || c.time_group > 64 || (c.stage_buffers != 2 && c.stage_buffers != 3) ||
      c.blocks_per_cu < -1 || c.blocks_per_cu > 4)
    return fail(BP_EINVAL, "invalid batch configuration (%d, %d, %d)", c.time_group,
                c.stage_buffers, c.blocks_per_cu);
  b->cfg = c;
  for (int k = 0; k < 2; ++k) {
    for (FusedSet *f : {&b->all[k], &b->small[k]}) {
      if (f->wins.empty()) continue;
      for (auto &w : f->wins) w.a.time_group = c.time_group;
      BP_TRY(b->put(f->dwins, f->wins.data(), f->wins.size() * sizeof(bp::BatchWindow)));
    }
    for (auto &a : b->large[k]) a.time_group = c.time_group;
  }
  b->drop_graph();
  return BP_OK;
}

// Differential mode (beamform.c:338-394), one thread per source.
// (K, S, P) delays and (K, S) weights on the device -> the wave-transposed
// tables of the differential kernel (64-source blocks, entry-major), padded
// to bp::diff_padded_stations(S, P) stations.
// (K, E) row-major 4-byte table -> wave-transposed (ceil(K/64), E_pad, 64).
int transpose_diff_table(const void *in, long long K, int E, int E_pad, DevBuf &out,
                         hipStream_t st) {
  const long long nblk = (K + bp::kDiffBlock - 1) / bp::kDiffBlock;
  BP_TRY(out.alloc((size_t)(nblk * E_pad * bp::kDiffBlock) * 4));
  const unsigned tiles =
      (unsigned)std::min<long long>((E_pad + bp::kDiffBlock - 1) / bp::kDiffBlock, 1024);
  hipLaunchKernelGGL(bp::diff_tables_kernel, dim3((unsigned)nblk, tiles), dim3(256), 0, st,
                     static_cast<const unsigned *>(in), K, E, E_pad, out.as<unsigned>());
  HIP_TRY(hipGetLastError());
  return BP_OK;
}

// Tables of the differential kernel: weights (ceil(K/64), S_pad, 64) float32
// and delays either (ceil(K/64), S_pad * P, 64) int32 or, for two phases,
// (ceil(K/64), S_pad, 64) words of two saturated int16 (`pairs`).
int build_diff_tau(const int *tau, long long K, int S, int P, bool pairs, DevBuf &tauT,
                   hipStream_t st) {
  const int S_pad = bp::diff_padded_stations(S, P);
  if (!pairs) return transpose_diff_table(tau, K, S * P, S_pad * P, tauT, st);
  DevBuf wide;
  BP_TRY(transpose_diff_table(tau, K, S * 2, S_pad * 2, wide, st));
  const long long nblk = (K + bp::kDiffBlock - 1) / bp::kDiffBlock;
  const long long n = nblk * S_pad * bp::kDiffBlock;
  BP_TRY(tauT.alloc((size_t)n * 4));
  hipLaunchKernelGGL(bp::diff_pack_pairs_kernel,
                     dim3((unsigned)std::min<long long>(65536, (n + 255) / 256)), dim3(256), 0,
                     st, wide.as<unsigned>(), nblk, S_pad, tauT.as<unsigned>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));  // before `wide` is freed
  return BP_OK;
}

// Weights of the differential kernel.  For 1-4 phases, when every nonzero
// weight has the same bits (wform 2, that value in cw; no nonzero weight at
// all counts too), the used-bit table (ceil(K/64), diff_mask_words, 64);
// otherwise the float32 table (ceil(K/64), S_pad, 64) (wform 1).
int build_diff_weights(const float *beta, long long K, int S, int P, DevBuf &out, int &wform,
                       float &cw, hipStream_t st) {
  const long long n = K * S;
  if (P >= 1 && P <= 4 && n > 0) {
    DevBuf bits;
    BP_TRY(bits.alloc(2 * sizeof(unsigned)));
    HIP_TRY(hipMemsetAsync(bits.p, 0xff, sizeof(unsigned), st));
    HIP_TRY(hipMemsetAsync(bits.as<unsigned>() + 1, 0, sizeof(unsigned), st));
    hipLaunchKernelGGL(bp::diff_weight_bits_kernel,
                       dim3((unsigned)std::min<long long>(2048, (n + 255) / 256)), dim3(256), 0,
                       st, reinterpret_cast<const unsigned *>(beta), n, bits.as<unsigned>());
    HIP_TRY(hipGetLastError());
    unsigned h[2] = {0u, 0u};
    HIP_TRY(hipMemcpyAsync(h, bits.p, sizeof h, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (h[0] >= h[1]) {  // one nonzero weight value, or none
      cw = h[0] == h[1] ? __builtin_bit_cast(float, h[0]) : 1.0f;
      const int W = bp::diff_mask_words(S, P);
      const long long nw = (K + bp::kDiffBlock - 1) / bp::kDiffBlock * W * bp::kDiffBlock;
      BP_TRY(out.alloc((size_t)nw * sizeof(unsigned)));
      // one wave per (two sources, word) item: nw / 128 items, 4 waves a group
      hipLaunchKernelGGL(bp::diff_mask_kernel,
                         dim3((unsigned)std::min<long long>(65536, (nw / 128 + 3) / 4)),
                         dim3(256), 0, st, beta, K, S, W, out.as<unsigned>());
      HIP_TRY(hipGetLastError());
      wform = 2;
      return BP_OK;
    }
  }
  BP_TRY(transpose_diff_table(beta, K, S, bp::diff_padded_stations(S, P), out, st));
  wform = 1;
  cw = 0.0f;
  return BP_OK;
}

// Two-phase delays fit the paired int16 tables for windows of at most 32768
// samples (the saturated delays then stay outside the window).
bool diff_pairs_ok(long long P, long long T) { return P == 2 && T <= 32768; }

template <int FMT, int PT, bool PAIRS>
void launch_diff_form(const bp::StackArgs &a, const int *tauT, const float *wT, float *beam,
                      bool mask, float cw, hipStream_t st) {
  const dim3 grid((unsigned)((a.K + bp::kDiffSources - 1) / bp::kDiffSources));
  const dim3 block(bp::kDiffSources);
  const int S_pad = bp::diff_padded_stations(a.S, a.P);
  if constexpr (PT > 0) {
    if (mask) {
      hipLaunchKernelGGL((bp::differential_kernel<FMT, PT, PAIRS, true>), grid, block, 0, st, a,
                         tauT, wT, S_pad, beam, cw);
      return;
    }
  }
  hipLaunchKernelGGL((bp::differential_kernel<FMT, PT, PAIRS, false>), grid, block, 0, st, a,
                     tauT, wT, S_pad, beam, 0.0f);
}

// wT: the weight table of build_diff_weights, used bits when `mask`.
template <int FMT>
void launch_differential_p(const bp::StackArgs &a, const int *tauT, const float *wT,
                           float *beam, bool pairs, bool mask, float cw, hipStream_t st) {
  switch (a.P) {
    case 1: launch_diff_form<FMT, 1, false>(a, tauT, wT, beam, mask, cw, st); break;
    case 2:
      if (pairs)
        launch_diff_form<FMT, 2, true>(a, tauT, wT, beam, mask, cw, st);
      else
        launch_diff_form<FMT, 2, false>(a, tauT, wT, beam, mask, cw, st);
      break;
    case 3: launch_diff_form<FMT, 3, false>(a, tauT, wT, beam, mask, cw, st); break;
    case 4: launch_diff_form<FMT, 4, false>(a, tauT, wT, beam, mask, cw, st); break;
    default: launch_diff_form<FMT, 0, false>(a, tauT, wT, beam, false, 0.0f, st);
  }
}

// Size limits of the kernel's 32-bit offsets.
int check_differential(int fmt, long long S, long long P, long long ldf) {
  const long long sp = S * P, rows = sp * (fmt == BP_F32 ? 1 : 2);
  // offsets inside a wave's block of the padded tables
  if (S >= (1ll << 24) || P >= (1ll << 24) ||
      (long long)bp::diff_padded_stations((int)S, (int)P) * P * bp::kDiffBlock >= (1ll << 30))
    return fail(BP_EINVAL, "n_stations * n_phases = %lld is too large for differential mode", sp);
  if (rows * ldf >= (1ll << 30))
    return fail(BP_EINVAL, "differential mode takes features under 2^30 elements (%lld rows of "
                "%lld)", rows, ldf);
  return BP_OK;
}

int launch_differential(int fmt, const bp::StackArgs &a, const int *tauT, const float *wT,
                        int wform, float cw, float *beam, bool pairs, hipStream_t st) {
  if (pairs && a.P != 2) return fail(BP_EINVAL, "paired delay tables need two phases");
  const bool mask = wform == 2;
  if (mask && (a.P < 1 || a.P > 4)) return fail(BP_EINVAL, "used-bit weights need 1-4 phases");
  if (fmt == BP_F32)
    launch_differential_p<BP_F32>(a, tauT, wT, beam, pairs, mask, cw, st);
  else if (fmt == BP_F16)
    launch_differential_p<BP_F16>(a, tauT, wT, beam, pairs, mask, cw, st);
  else
    launch_differential_p<BP_BF16>(a, tauT, wT, beam, pairs, mask, cw, st);
  HIP_TRY(hipGetLastError());
  return BP_OK;
}

}  // namespace

extern "C" {

const char *bp_last_error(void) { return g_err.c_str(); }

const char *bp_version(void) { return BP_VERSION; }

int bp_profiling(int enable) try {
  g_prof.on = enable != 0;
  return BP_OK;
} BP_CATCH

int bp_profiling_read(double *stack_ms_total, long long *stack_launches) try {
  for (auto &e : g_prof.pending) {
    HIP_TRY(hipEventSynchronize(e.second));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
    g_prof.ms += ms;
    g_prof.launches += 1;
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  g_prof.pending.clear();
  if (stack_ms_total) *stack_ms_total = g_prof.ms;
  if (stack_launches) *stack_launches = g_prof.launches;
  g_prof.ms = 0.0;
  g_prof.launches = 0;
  return BP_OK;
} BP_CATCH

int bp_network_create(const int *time_delays, const float *weights_sources, size_t n_sources,
                      size_t n_stations, size_t n_phases, bp_network **out) try {
  return bp_network_create_ordered(time_delays, weights_sources, n_sources, n_stations,
                                    n_phases, BP_ORDER_AUTO, out);
} BP_CATCH

int bp_network_create_ordered(const int *time_delays, const float *weights_sources,
                              size_t n_sources, size_t n_stations, size_t n_phases, int order,
                              bp_network **out) try {
  if (!out) return fail(BP_EINVAL, "out is NULL");
  if (order != BP_ORDER_NATURAL && order != BP_ORDER_SIMILARITY && order != BP_ORDER_AUTO)
    return fail(BP_EINVAL, "unknown source order policy %d", order);
  *out = nullptr;
  if (n_sources == 0 || n_stations == 0 || n_phases == 0)
    return fail(BP_EINVAL, "network needs n_sources, n_stations, n_phases > 0");
  if (!time_delays || !weights_sources) return fail(BP_EINVAL, "NULL input");
  BP_TRY(check_shapes(0, n_sources, n_stations, n_phases));
  BP_TRY(check_device());
  Stream st;
  BP_TRY(st.create());
  std::unique_ptr<bp_network> net;
  BP_TRY(network_create(time_delays, weights_sources, n_sources, n_stations, n_phases, order,
                        st.s, net));
  *out = net.release();
  return BP_OK;
} BP_CATCH

int bp_network_destroy(bp_network *net) try {
  delete net;
  return BP_OK;
} BP_CATCH

int bp_dropin_cache_clear(void) try {
  DropinCache &c = dropin_cache();
  std::lock_guard<std::mutex> lock(c.mu);
  c.net.reset();
  c.tau.reset();
  c.beta.reset();
  c.device = -1;
  c.K = c.S = c.P = 0;
  c.reuses = 0;
  return BP_OK;
} BP_CATCH

int bp_network_source_order(const bp_network *net, int *permuted, int *order) try {
  if (!net) return fail(BP_EINVAL, "NULL network");
  if (permuted) *permuted = net->permuted ? 1 : 0;
  if (order)
    for (int i = 0; i < net->K; ++i) order[i] = net->permuted ? net->order[i] : i;
  return BP_OK;
} BP_CATCH

int bp_kernel_shape(int *source_tile, int *time_tile, int *waves_per_simd) try {
  if (source_tile) *source_tile = kTileSources;
  if (time_tile) *time_tile = kTT;
  if (waves_per_simd) *waves_per_simd = BP_WAVES_PER_EU;
  return BP_OK;
} BP_CATCH

int bp_network_info(const bp_network *net, int *source_tile, int *max_tile_spread,
                    int *n_tiles) try {
  if (!net) return fail(BP_EINVAL, "NULL network");
  if (source_tile) *source_tile = kTileSources;
  if (max_tile_spread) *max_tile_spread = net->max_spread;
  if (n_tiles) *n_tiles = net->n_tiles;
  return BP_OK;
} BP_CATCH

int bp_network_classes(const bp_network *net, int *tiles_per_class, int *window_per_class) try {
  if (!net) return fail(BP_EINVAL, "NULL network");
  for (int c = 0; c < 3; ++c) {
    if (tiles_per_class) tiles_per_class[c] = (int)net->tiles[c].size();
    if (window_per_class) window_per_class[c] = c < 2 ? net->Lw[c] : 0;
  }
  return BP_OK;
} BP_CATCH

int bp_network_beamform_max(bp_network *net, const void *features, int fmt,
                            const float *station_scale, size_t n_samples, int out_of_bounds,
                            float *beam_max, int *beam_argmax, unsigned long long *keys,
                            void *stream) try {
  if (!net || !features || !beam_max || !beam_argmax || !keys)
    return fail(BP_EINVAL, "NULL argument");
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 (strict) or 1 (flexible), not %d",
                out_of_bounds);
  if (fmt != BP_F32 && fmt != BP_F16 && fmt != BP_BF16)
    return fail(BP_EINVAL, "unknown feature format %d", fmt);
  if (fmt == BP_F32 && station_scale)
    return fail(BP_EINVAL, "station_scale applies to the 16-bit formats only");
  BP_TRY(check_shapes(n_samples, net->K, net->S, net->P));
  if (n_samples == 0) return BP_OK;
  return network_max(net, plain_view(features, fmt, station_scale, n_samples), out_of_bounds,
                     beam_max, beam_argmax, keys, static_cast<hipStream_t>(stream));
} BP_CATCH

int bp_network_beamform_max_view(bp_network *net, const bp_feature_view *features,
                                 int out_of_bounds, float *beam_max, int *beam_argmax,
                                 unsigned long long *keys, void *stream) try {
  if (!net || !features || !beam_max || !beam_argmax || !keys)
    return fail(BP_EINVAL, "NULL argument");
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 (strict) or 1 (flexible), not %d",
                out_of_bounds);
  BP_TRY(check_view(net, *features));
  if (features->n_samples == 0) return BP_OK;
  return network_max(net, *features, out_of_bounds, beam_max, beam_argmax, keys,
                     static_cast<hipStream_t>(stream));
} BP_CATCH

}  // extern "C"

// ---- reduce='max' without allocation ---------------------------------------

// Key scratch of bp_network_beamform_max_noalloc: one packed (value, source)
// key per sample.  A __device__ array: every GPU the library runs on holds
// its own copy, at its own address, until the process ends.  In host code
// its name is only a stand-in for the symbol; the device address comes from
// hipGetSymbolAddress (scratch_for_current_device).
static __device__ __align__(256) unsigned long long scratch[BP_NOALLOC_MAX_SAMPLES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long *g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch, looked up on the first
// call on that device (with that device current) and kept.
int scratch_for_current_device(unsigned long long **out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(BP_ENODEV, "device %d: the key scratch serves devices 0..%d", dev,
                kMaxDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<unsigned long long *>(p);
  }
  *out = g_scratch_ptr[dev];
  return BP_OK;
}

// An existing schedule of the network for (fmt, scales), or NULL; builds
// nothing and allocates nothing.
std::shared_ptr<Schedule> find_schedule(bp_network *net, int fmt, const float *scales) {
  std::lock_guard<std::mutex> lock(net->mu);
  for (const std::shared_ptr<Schedule> &sc : net->scheds[fmt]) {
    bool same = true;
    for (int s = 0; s < net->S && same; ++s)
      same = sc->scales[s] == (scales ? scales[s] : 1.0f);
    if (same) return sc;
  }
  return nullptr;
}

}  // namespace

extern "C" {

int bp_network_beamform_max_noalloc(bp_network *net, const bp_feature_view *features,
                                    int out_of_bounds, float *beam_max, int *beam_argmax,
                                    void *stream) try {
  if (!net || !features || !beam_max || !beam_argmax) return fail(BP_EINVAL, "NULL argument");
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 (strict) or 1 (flexible), not %d",
                out_of_bounds);
  BP_TRY(check_view(net, *features));
  if (features->n_samples > (size_t)BP_NOALLOC_MAX_SAMPLES)
    return fail(BP_EINVAL,
                "n_samples %zu exceeds the %d samples of the no-allocation key scratch "
                "(use bp_network_beamform_max_view with a workspace)",
                features->n_samples, BP_NOALLOC_MAX_SAMPLES);
  if (features->n_samples == 0) return BP_OK;
  const int fmt = features->fmt;
  std::shared_ptr<Schedule> sc =
      find_schedule(net, fmt, fmt == BP_F32 ? nullptr : features->station_scale);
  if (!sc)
    return fail(BP_EINVAL,
                "no schedule for feature format %d with these station scales: build it first "
                "(bp_network_prepare)", fmt);
  unsigned long long *keys = nullptr;
  BP_TRY(scratch_for_current_device(&keys));
  return network_max_with(net, *features, out_of_bounds, beam_max, beam_argmax, keys,
                          static_cast<hipStream_t>(stream), *sc, false);
} BP_CATCH

int bp_network_prepare(bp_network *net, int fmt, const float *station_scale) try {
  if (!net) return fail(BP_EINVAL, "NULL network");
  if (fmt != BP_F32 && fmt != BP_F16 && fmt != BP_BF16)
    return fail(BP_EINVAL, "unknown feature format %d", fmt);
  if (fmt == BP_F32 && station_scale)
    return fail(BP_EINVAL, "station_scale applies to the 16-bit formats only");
  BP_TRY(check_device());
  Stream st;
  BP_TRY(st.create());
  std::shared_ptr<Schedule> sc;
  return ensure_schedule(net, fmt, station_scale, st.s, sc);
} BP_CATCH

int bp_plan(const bp_plan_request *req, bp_plan_result *res) try {
  if (!req || !res) return fail(BP_EINVAL, "NULL argument");
  *res = bp_plan_result{};
  const size_t K = req->n_sources, S = req->n_stations, P = req->n_phases, T = req->n_samples;
  const int fmt = req->fmt;
  if (fmt != BP_F32 && fmt != BP_F16 && fmt != BP_BF16)
    return fail(BP_EINVAL, "unknown feature format %d", fmt);
  if (req->reduce != 0 && req->reduce != 1)
    return fail(BP_EINVAL, "reduce must be 0 ('max') or 1 ('none'), not %d", req->reduce);
  if (req->order != BP_ORDER_NATURAL && req->order != BP_ORDER_SIMILARITY &&
      req->order != BP_ORDER_AUTO)
    return fail(BP_EINVAL, "unknown source order %d", req->order);
  BP_TRY(check_shapes(T, K, S, P));
  if ((req->time_delays == nullptr) != (req->weights_sources == nullptr))
    return fail(BP_EINVAL, "time_delays and weights_sources come together");
  const bool none = req->reduce == 1;
  res->source_tile = kTileSources;
  res->time_tile = kTT;
  res->lds_window_limit = lds_window_limit((int)P);
  for (int c = 0; c < 3; ++c) res->tiles_per_class[c] = -1;
  res->window_per_class[0] = res->window_per_class[1] = -1;
  res->similarity_order = -1;
  res->pad_before = res->pad_after = -1;
  const int n_tiles = (int)((K + kTileSources - 1) / kTileSources);
  bool lds_tiles = res->lds_window_limit > 0 && K > 0;
  long long pad_lo = 0, pad_hi = 0;
  bool permuted = false;
  if (req->time_delays && K > 0 && S > 0 && P > 0) {
    std::unique_ptr<bp_network> net;
    BP_TRY(host_network_plan(req->time_delays, req->weights_sources, K, S, P, req->order, net));
    for (int c = 0; c < 3; ++c) res->tiles_per_class[c] = (int)net->tiles[c].size();
    res->window_per_class[0] = net->Lw[0];
    res->window_per_class[1] = net->Lw[1];
    res->similarity_order = net->permuted ? 1 : 0;
    res->pad_before = pad_lo = net->pad_lo;
    res->pad_after = pad_hi = net->pad_hi;
    permuted = net->permuted;
    lds_tiles = has_lds_tiles(net.get());
    const int wave_slots = 4 * BP_WAVES_PER_EU / kW;  // work-groups per CU by waves
    for (int c = 0; c < 2; ++c) {
      res->lds_bytes[c] = lds_bytes_of(fmt, (int)P, none, net->Lw[c]);
      res->workgroups_per_cu[c] =
          res->lds_bytes[c] > 0 ? std::min(kLdsPerCU / res->lds_bytes[c], wave_slots) : 0;
    }
  }
  // resident network: moveouts, weights, per-source bounds, tile tables,
  // source order, one schedule of the format with its station scales
  const size_t sched = lds_tiles ? (size_t)n_tiles * S * kTileSources * (P + 1) * 4 : 0;
  res->network_bytes = K * S * P * 4 + K * S * 4 + 2 * K * 4 + (size_t)n_tiles * S * P * 4 +
                       (size_t)n_tiles * (8 + 1 + 1) * 4 + (permuted ? K * 4 : 0) + sched +
                       S * 4;
  // the window in rows padded for the network (16-bit: two parity planes)
  const size_t rows = fmt == BP_F32 ? S * P : 2 * S * P, es = fmt == BP_F32 ? 4 : 2;
  size_t stride = (size_t)pad_lo + T + (size_t)pad_hi;
  if (fmt != BP_F32) stride += stride & 1;
  res->feature_bytes = rows * stride * es;
  res->output_bytes = none ? K * T * 4 : T * (4 + 4 + 8);
  // a host-pointer call also uploads the interleaved (S, T, P) features
  res->call_bytes = res->network_bytes + res->feature_bytes + S * T * P * 4 + res->output_bytes;
  res->fits = req->hbm_budget == 0 || res->call_bytes <= req->hbm_budget;
  return BP_OK;
} BP_CATCH

int bp_network_padding(const bp_network *net, size_t *pad_before, size_t *pad_after) try {
  if (!net) return fail(BP_EINVAL, "NULL network");
  if (pad_before) *pad_before = (size_t)net->pad_lo;
  if (pad_after) *pad_after = (size_t)net->pad_hi;
  return BP_OK;
} BP_CATCH

size_t bp_parity_stride(size_t n_samples) { return parity_stride(n_samples); }

int bp_pack_features(const float *F_spt, size_t n_stations, size_t n_phases, size_t n_samples,
                     int fmt, size_t pad_before, size_t row_stride, float *station_scale,
                     void *out, void *stream) try {
  if (!F_spt || !out) return fail(BP_EINVAL, "NULL argument");
  if (fmt != BP_F32 && fmt != BP_F16 && fmt != BP_BF16)
    return fail(BP_EINVAL, "unknown feature format %d", fmt);
  if (fmt != BP_F32 && !station_scale)
    return fail(BP_EINVAL, "16-bit packing returns station scales: station_scale is NULL");
  BP_TRY(check_shapes(n_samples, 0, n_stations, n_phases));
  if (!size_ok(row_stride) || !size_ok(pad_before) || row_stride < pad_before + n_samples)
    return fail(BP_EINVAL, "row_stride %zu cannot hold pad_before %zu + %zu samples", row_stride,
                pad_before, n_samples);
  if (fmt != BP_F32 && ((row_stride | pad_before) & 1))
    return fail(BP_EINVAL, "16-bit rows need an even row_stride and pad_before");
  const size_t S = n_stations, P = n_phases, T = n_samples;
  if (S == 0 || P == 0 || row_stride == 0) return BP_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DevBuf dscale;
  if (fmt == BP_BF16) {
    // bfloat16 keeps the float32 exponent range: scale 1 (a NULL scale
    // table), no host round trip, asynchronous
    for (size_t s = 0; s < S; ++s) station_scale[s] = 1.0f;
  } else if (fmt == BP_F16) {
    DevBuf maxbits;
    BP_TRY(maxbits.alloc(S * sizeof(unsigned)));
    BP_TRY(dscale.alloc(S * sizeof(float)));
    std::vector<unsigned> mb(S, 0u);
    HIP_TRY(hipMemcpyAsync(maxbits.p, mb.data(), S * sizeof(unsigned), hipMemcpyHostToDevice,
                           st));
    if (T > 0) {
      hipLaunchKernelGGL(bp::station_maxabs_kernel, dim3(64, (unsigned)S), dim3(256), 0, st,
                         F_spt, (int)P, (long long)T, maxbits.as<unsigned>());
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(mb.data(), maxbits.p, S * sizeof(unsigned), hipMemcpyDeviceToHost,
                           st));
    HIP_TRY(hipStreamSynchronize(st));
    // power-of-two scales: FP16 maps the station's max |F| into [2^14, 2^15)
    // (headroom below 65504); BF16 keeps the float32 exponent range: scale 1.
    // The caller gets the weight multipliers 1/scale (exact).
    std::vector<float> sigma(S, 1.0f);
    for (size_t s = 0; s < S; ++s) {
      const float m = __builtin_bit_cast(float, mb[s]);
      if (fmt == BP_F16 && m > 0.0f && m <= 3.0e38f) {
        int e = 0;
        std::frexp(m, &e);  // m = f * 2^e, f in [0.5, 1)
        sigma[s] = std::ldexp(1.0f, 15 - e);
      }
      station_scale[s] = 1.0f / sigma[s];
    }
    HIP_TRY(hipMemcpyAsync(dscale.p, sigma.data(), S * sizeof(float), hipMemcpyHostToDevice, st));
  }
  hipLaunchKernelGGL(bp::pack_features_kernel,
                     dim3((unsigned)((row_stride + 255) / 256), (unsigned)(S * P)), dim3(256), 0,
                     st, F_spt, (int)P, (long long)T, dscale.as<float>(), fmt,
                     (long long)row_stride, (long long)pad_before, out);
  HIP_TRY(hipGetLastError());
  if (fmt == BP_F16) HIP_TRY(hipStreamSynchronize(st));  // dscale is freed on return
  return BP_OK;
} BP_CATCH

int bp_quantize_planar(const float *F_spt, size_t n_stations, size_t n_phases, size_t n_samples,
                       int fmt, float *station_scale, void *out, void *stream) try {
  if (!F_spt || !station_scale || !out) return fail(BP_EINVAL, "NULL argument");
  if (fmt != BP_F16 && fmt != BP_BF16)
    return fail(BP_EINVAL, "quantization targets BP_F16 or BP_BF16, not %d", fmt);
  BP_TRY(check_shapes(n_samples, 0, n_stations, n_phases));
  if (n_stations == 0 || n_phases == 0 || n_samples == 0) return BP_OK;
  return bp_pack_features(F_spt, n_stations, n_phases, n_samples, fmt, 0,
                          parity_stride(n_samples), station_scale, out, stream);
} BP_CATCH

int bp_network_beamform(bp_network *net, const float *features, size_t n_samples,
                        int out_of_bounds, size_t source_begin, size_t source_end,
                        float *beam, void *stream) try {
  if (!net || !features || !beam) return fail(BP_EINVAL, "NULL argument");
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 (strict) or 1 (flexible), not %d",
                out_of_bounds);
  if (source_begin > source_end || source_end > (size_t)net->K)
    return fail(BP_EINVAL, "bad source range [%zu, %zu) for %d sources", source_begin,
                source_end, net->K);
  BP_TRY(check_shapes(n_samples, net->K, net->S, net->P));
  if (n_samples == 0) return BP_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::shared_ptr<Schedule> sc;
  BP_TRY(ensure_schedule(net, BP_F32, nullptr, st, sc));
  bp::StackArgs a = make_args(net, features, BP_F32, (int)n_samples, out_of_bounds, *sc);
  a.beam = beam;
  return launch_stack<BP_F32, true>(net, a, (int)source_begin, (int)source_end, st);
} BP_CATCH

int bp_network_beamform_view(bp_network *net, const bp_feature_view *features,
                             int out_of_bounds, size_t source_begin, size_t source_end,
                             float *beam, void *stream) try {
  if (!net || !features || !beam) return fail(BP_EINVAL, "NULL argument");
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 (strict) or 1 (flexible), not %d",
                out_of_bounds);
  if (source_begin > source_end || source_end > (size_t)net->K)
    return fail(BP_EINVAL, "bad source range [%zu, %zu) for %d sources", source_begin,
                source_end, net->K);
  BP_TRY(check_view(net, *features));
  if (features->n_samples == 0) return BP_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int fmt = features->fmt;
  std::shared_ptr<Schedule> sc;
  BP_TRY(ensure_schedule(net, fmt, fmt == BP_F32 ? nullptr : features->station_scale, st, sc));
  bp::StackArgs a = make_args_view(net, *features, out_of_bounds, *sc);
  a.beam = beam;
  if (fmt == BP_F32)
    return launch_stack<BP_F32, true>(net, a, (int)source_begin, (int)source_end, st);
  if (fmt == BP_F16)
    return launch_stack<BP_F16, true>(net, a, (int)source_begin, (int)source_end, st);
  return launch_stack<BP_BF16, true>(net, a, (int)source_begin, (int)source_end, st);
} BP_CATCH

// ---- differential mode (beamform.c:338-394) -------------------------------

int bp_network_diff_table_bytes(const bp_network *net, size_t *delay_bytes,
                                size_t *weight_bytes, int *weight_form) try {
  if (!net) return fail(BP_EINVAL, "NULL network");
  std::lock_guard<std::mutex> lock(net->mu);
  if (delay_bytes) *delay_bytes = net->diff_tau.bytes + net->diff_tau_pairs.bytes;
  if (weight_bytes) *weight_bytes = net->diff_wform ? net->diff_beta.bytes : 0;
  if (weight_form) *weight_form = net->diff_wform;
  return BP_OK;
} BP_CATCH

int bp_network_beamform_differential(bp_network *net, const bp_feature_view *features,
                                     float *beam, void *stream) try {
  if (!net || !features || !beam) return fail(BP_EINVAL, "NULL argument");
  BP_TRY(check_view(net, *features));
  if (features->n_samples == 0 || net->K == 0) return BP_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int fmt = features->fmt;
  BP_TRY(check_differential(fmt, net->S, net->P, (long long)features->row_stride));
  std::shared_ptr<Schedule> sc;
  BP_TRY(ensure_schedule(net, fmt, fmt == BP_F32 ? nullptr : features->station_scale, st, sc));
  // The tables are in the caller's source order even when the stack plan
  // orders sources by similarity: a wave's 64 lanes are 64 sources, and on
  // a grid 64 consecutive indices gather from fewer cache lines than a
  // similarity tile (0.51 against 0.59 ms at 1,225 pairs x 192,000 sources).
  const bool pairs = diff_pairs_ok(net->P, (long long)features->n_samples);
  {
    std::lock_guard<std::mutex> lock(net->mu);
    bool &tau_ok = pairs ? net->diff_pairs_ok : net->diff_ok;
    if (!tau_ok || net->diff_wform == 0) {
      const size_t K = net->K, S = net->S, P = net->P;
      const int *tau_src = net->tau.as<int>();
      const float *beta_src = net->beta.as<float>();
      DevBuf inv, tau, beta;
      if (net->permuted) {
        BP_TRY(inv.alloc(K * sizeof(int)));
        BP_TRY(tau.alloc(K * S * P * sizeof(int)));
        BP_TRY(beta.alloc(K * S * sizeof(float)));
        HIP_TRY(hipMemcpyAsync(inv.p, net->inv.data(), K * sizeof(int), hipMemcpyHostToDevice,
                               st));
        const unsigned gb = (unsigned)std::min<size_t>(4096, (K * S * P + 255) / 256);
        hipLaunchKernelGGL(bp::gather_rows_kernel, dim3(gb), dim3(256), 0, st,
                           net->tau.as<unsigned>(), inv.as<int>(), (long long)K, (int)(S * P),
                           tau.as<unsigned>());
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(bp::gather_rows_kernel, dim3(gb), dim3(256), 0, st,
                           net->beta.as<unsigned>(), inv.as<int>(), (long long)K, (int)S,
                           beta.as<unsigned>());
        HIP_TRY(hipGetLastError());
        tau_src = tau.as<int>();
        beta_src = beta.as<float>();
      }
      if (net->diff_wform == 0) {
        int wform = 0;
        float cw = 0.0f;
        BP_TRY(build_diff_weights(beta_src, net->K, net->S, net->P, net->diff_beta, wform, cw,
                                  st));
        net->diff_wform = wform;
        net->diff_cw = cw;
      }
      if (!tau_ok)
        BP_TRY(build_diff_tau(tau_src, net->K, net->S, net->P, pairs,
                              pairs ? net->diff_tau_pairs : net->diff_tau, st));
      // ready for calls on any stream, and before the temporaries are freed
      HIP_TRY(hipStreamSynchronize(st));
      tau_ok = true;
    }
  }
  bp::StackArgs a = make_args_view(net, *features, 0, *sc);
  a.perm = nullptr;  // tables in the caller's order
  const DevBuf &tauT = pairs ? net->diff_tau_pairs : net->diff_tau;
  return launch_differential(fmt, a, tauT.as<int>(), net->diff_beta.as<float>(), net->diff_wform,
                             net->diff_cw, beam, pairs, st);
} BP_CATCH

// The CPU library's entry point (beamform_cpu.h:6, beamform.c:338-394) on
// the GPU, host pointers, synchronous; num_threads is accepted and unused.
// Sources left out keep the caller's beam values (the reference wrapper
// passes zeros, beampower.py:198).
int beamform_differential(float *waveform_features, int *time_delays, float *weights,
                          size_t n_samples, size_t n_sources, size_t n_stations,
                          size_t n_phases, int num_threads, float *beam) try {
  (void)num_threads;
  BP_TRY(check_shapes(n_samples, n_sources, n_stations, n_phases));
  // no sample, no source, or no station/phase (no source uses a station)
  if (n_samples == 0 || n_sources == 0 || n_stations == 0 || n_phases == 0) return BP_OK;
  if (!beam) return fail(BP_EINVAL, "NULL output");
  if (!waveform_features || !time_delays || !weights) return fail(BP_EINVAL, "NULL input");
  BP_TRY(check_device());
  Stream st;
  BP_TRY(st.create());
  const size_t T = n_samples, K = n_sources, S = n_stations, P = n_phases;
  BP_TRY(check_differential(BP_F32, (long long)S, (long long)P, (long long)T));
  DevBuf tau, beta, tauT, betaT, f_stp, f_pl, dbeam;
  const bool pairs = diff_pairs_ok((long long)P, (long long)T);
  BP_TRY(tau.alloc((pairs ? K * S : K * S * P) * sizeof(int)));
  int wform = 0;  // 2: used-bit weights built on the host (betaT holds them)
  float cw = 0.0f;
  std::unique_ptr<unsigned[]> bits_h;
  size_t n_bits = 0;
  BP_TRY(f_stp.alloc(S * T * P * sizeof(float)));
  BP_TRY(f_pl.alloc(S * T * P * sizeof(float)));
  BP_TRY(dbeam.alloc(K * sizeof(float)));
  // the delay and weight tables (2.8 GB at 1,225 station pairs x 192,000
  // sources) go up through the two pinned stages at once, the weights on a
  // second host thread; two phases' delays go up as packed int16 pairs
  // where the kernel reads them so (windows up to 32,768 samples), and
  // weights with one nonzero value as their used bits
  bool packed = false;
  HIP_TRY(hipMemcpyAsync(f_stp.p, waveform_features, S * T * P * sizeof(float),
                         hipMemcpyHostToDevice, st.s));
  HIP_TRY(hipMemcpyAsync(dbeam.p, beam, K * sizeof(float), hipMemcpyHostToDevice, st.s));
  {
    int dev = 0, up_rc = BP_OK;
    HIP_TRY(hipGetDevice(&dev));
    std::string up_err;
    std::thread up([&] {
      try {
        if (hipSetDevice(dev) != hipSuccess) {  // a new thread starts on device 0
          up_rc = fail(BP_EHIP, "hipSetDevice(%d) failed", dev);
        } else if (P >= 1 && P <= 4 && K * S * sizeof(float) >= 2 * kPinnedBytes &&
                   host_used_bits(weights, K, S, bp::diff_mask_words((int)S, (int)P), bits_h,
                                  n_bits, cw)) {
          // one nonzero weight value: its used bits instead of the weights
          up_rc = betaT.alloc(n_bits * sizeof(unsigned));
          if (up_rc == BP_OK)
            up_rc = upload_host(betaT.p, bits_h.get(), n_bits * sizeof(unsigned), st.s);
          if (up_rc == BP_OK) wform = 2;
        } else {
          up_rc = beta.alloc(K * S * sizeof(float));
          if (up_rc == BP_OK) up_rc = upload_host(beta.p, weights, K * S * sizeof(float), st.s);
        }
      } catch (...) {
        up_rc = fail(BP_EHIP, "weight upload failed");
      }
      if (up_rc != BP_OK) up_err = g_err;
    });
    int rc = BP_OK;
    try {
      if (pairs) rc = upload_packed_pairs(tau.as<unsigned>(), time_delays, K * S, st.s, packed);
      if (rc == BP_OK && !packed) {
        if (pairs) rc = tau.alloc(K * S * P * sizeof(int));
        if (rc == BP_OK) rc = upload_host(tau.p, time_delays, K * S * P * sizeof(int), st.s);
      }
    } catch (...) {
      up.join();
      throw;
    }
    up.join();
    BP_TRY(rc);
    if (up_rc != BP_OK) return fail(up_rc, "%s", up_err.c_str());
  }
  hipLaunchKernelGGL(bp::interleaved_to_planar_kernel,
                     dim3((unsigned)((T + 255) / 256), (unsigned)S), dim3(256), 0, st.s,
                     f_stp.as<float>(), (int)S, (int)T, (int)P, (long long)T, 0ll,
                     f_pl.as<float>());
  HIP_TRY(hipGetLastError());
  bp::StackArgs a{};
  a.F = f_pl.p;
  a.ldf = (long long)T;
  a.tau = packed ? nullptr : tau.as<int>();  // the kernel reads only the tables
  a.beta = beta.as<float>();
  a.T = (int)T;
  a.K = (int)K;
  a.S = (int)S;
  a.P = (int)P;
  if (wform != 2)
    BP_TRY(build_diff_weights(beta.as<float>(), (long long)K, (int)S, (int)P, betaT, wform, cw,
                              st.s));
  if (packed)  // the pairs, transposed as build_diff_tau transposes and packs the int32 delays
    BP_TRY(transpose_diff_table(tau.p, (long long)K, (int)S,
                                bp::diff_padded_stations((int)S, (int)P), tauT, st.s));
  else
    BP_TRY(build_diff_tau(tau.as<int>(), (long long)K, (int)S, (int)P, pairs, tauT, st.s));
  BP_TRY(launch_differential(BP_F32, a, tauT.as<int>(), betaT.as<float>(), wform, cw,
                             dbeam.as<float>(), pairs, st.s));
  HIP_TRY(hipMemcpyAsync(beam, dbeam.p, K * sizeof(float), hipMemcpyDeviceToHost, st.s));
  HIP_TRY(hipStreamSynchronize(st.s));
  return BP_OK;
} BP_CATCH

// ---- detection consumer (notebooks/3_beampower.ipynb:497-626) ------------

int bp_detect_threshold_type(const float *beam_max, size_t n_samples, size_t window,
                             double overlap, double num_dev, int threshold_type,
                             float *threshold, void *stream) try {
  if (!beam_max || !threshold) return fail(BP_EINVAL, "NULL argument");
  if (threshold_type != BP_THRESHOLD_RMS && threshold_type != BP_THRESHOLD_MAD)
    return fail(BP_EINVAL, "threshold_type must be BP_THRESHOLD_RMS or BP_THRESHOLD_MAD, not %d",
                threshold_type);
  const long long n = (long long)n_samples, W = (long long)window;
  const long long shift = (long long)((1.0 - overlap) * (double)window);
  if (!size_ok(n_samples) || W < 1 || n <= W || shift < 1)
    return fail(BP_EINVAL, "need n_samples (%zu) > window (%zu) and a shift of >= 1 sample",
                n_samples, window);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long nw = (n - W) / shift + 1;
  // per window: center and deviation (rms: mean and population standard
  // deviation in double; mad: median and median absolute deviation, float)
  std::vector<float> center((size_t)nw), deviation((size_t)nw);
  DevBuf stats, dthr;
  if (threshold_type == BP_THRESHOLD_RMS) {
    BP_TRY(stats.alloc((size_t)nw * 2 * sizeof(double)));
    hipLaunchKernelGGL(bp::window_stats_kernel, dim3((unsigned)nw), dim3(256), 0, st, beam_max,
                       W, shift, stats.as<double>());
    HIP_TRY(hipGetLastError());
    std::vector<double> h((size_t)nw * 2);
    HIP_TRY(hipMemcpyAsync(h.data(), stats.p, h.size() * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (long long j = 0; j < nw; ++j) {
      center[j] = (float)h[2 * j];
      deviation[j] = (float)h[2 * j + 1];
    }
  } else {
    BP_TRY(stats.alloc((size_t)nw * 2 * sizeof(float)));
    hipLaunchKernelGGL(bp::window_median_kernel, dim3((unsigned)nw), dim3(256), 0, st,
                       beam_max, W, shift, stats.as<float>());
    HIP_TRY(hipGetLastError());
    std::vector<float> h((size_t)nw * 2);
    HIP_TRY(hipMemcpyAsync(h.data(), stats.p, h.size() * sizeof(float), hipMemcpyDeviceToHost,
                           st));
    HIP_TRY(hipStreamSynchronize(st));
    for (long long j = 0; j < nw; ++j) {
      center[j] = h[2 * j];
      deviation[j] = h[2 * j + 1];
    }
  }
  BP_TRY(dthr.alloc((size_t)nw * sizeof(float)));
  // center + num_dev * deviation in float32, then the one-step maximum with
  // the previous and (after that) the next window (:556-560)
  std::vector<float> thr((size_t)nw);
  const float nd = (float)num_dev;
  for (long long j = 0; j < nw; ++j) {
    const float dev = nd * deviation[j];
    thr[j] = center[j] + dev;
  }
  std::vector<float> a = thr;
  for (long long j = 1; j < nw; ++j) thr[j] = std::max(a[j - 1], a[j]);
  for (long long j = 0; j + 1 < nw; ++j) thr[j] = std::max(thr[j], thr[j + 1]);
  HIP_TRY(hipMemcpyAsync(dthr.p, thr.data(), thr.size() * sizeof(float), hipMemcpyHostToDevice,
                         st));
  hipLaunchKernelGGL(bp::expand_threshold_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, st, dthr.as<float>(), (int)nw, n, W, shift, threshold);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));  // thr and dthr are released on return
  return BP_OK;
} BP_CATCH

int bp_detect_threshold(const float *beam_max, size_t n_samples, size_t window, double overlap,
                        double num_dev, float *threshold, void *stream) try {
  return bp_detect_threshold_type(beam_max, n_samples, window, overlap, num_dev,
                                  BP_THRESHOLD_RMS, threshold, stream);
} BP_CATCH

int bp_detect_peaks(const float *beam_max, const float *threshold, size_t n_samples,
                    size_t distance, int *peaks, size_t capacity, size_t *n_peaks,
                    void *stream) try {
  if (!beam_max || !n_peaks || (capacity && !peaks)) return fail(BP_EINVAL, "NULL argument");
  *n_peaks = 0;
  if (!size_ok(n_samples) || distance < 1)
    return fail(BP_EINVAL, "bad n_samples (%zu) or distance (%zu)", n_samples, distance);
  const long long n = (long long)n_samples;
  if (n < 3) return BP_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  DevBuf flags;
  BP_TRY(flags.alloc((size_t)n));
  HIP_TRY(hipMemsetAsync(flags.p, 0, (size_t)n, st));
  hipLaunchKernelGGL(bp::local_max_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     beam_max, n, flags.as<unsigned char>());
  HIP_TRY(hipGetLastError());
  std::vector<unsigned char> f((size_t)n);
  std::vector<float> x((size_t)n), thr;
  HIP_TRY(hipMemcpyAsync(f.data(), flags.p, (size_t)n, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(x.data(), beam_max, (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                         st));
  if (threshold) {
    thr.resize((size_t)n);
    HIP_TRY(hipMemcpyAsync(thr.data(), threshold, (size_t)n * sizeof(float),
                           hipMemcpyDeviceToHost, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  std::vector<long long> cand;
  for (long long i = 0; i < n; ++i)
    if (f[(size_t)i]) cand.push_back(i);
  // scipy.signal.find_peaks(distance=...): visit peaks by decreasing height
  // (equal heights: later index first); a visited, kept peak removes every
  // peak closer than `distance`
  std::vector<size_t> order(cand.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    const float ha = x[(size_t)cand[a]], hb = x[(size_t)cand[b]];
    return ha != hb ? ha > hb : cand[a] > cand[b];
  });
  std::vector<unsigned char> keep(cand.size(), 1);
  const long long dist = (long long)distance;
  for (size_t o : order) {
    if (!keep[o]) continue;
    for (long long k = (long long)o - 1; k >= 0 && cand[o] - cand[(size_t)k] < dist; --k)
      keep[(size_t)k] = 0;
    for (size_t k = o + 1; k < cand.size() && cand[k] - cand[o] < dist; ++k) keep[k] = 0;
  }
  size_t m = 0;
  for (size_t i = 0; i < cand.size(); ++i) {
    if (!keep[i]) continue;
    const size_t t = (size_t)cand[i];
    if (threshold && !(x[t] > thr[t])) continue;
    if (m < capacity) peaks[m] = (int)t;
    ++m;
  }
  *n_peaks = m;
  return BP_OK;
} BP_CATCH

int bp_interleaved_to_planar(const float *F_stp, size_t n_stations, size_t n_samples,
                             size_t n_phases, float *F_spt, void *stream) try {
  if (!F_stp || !F_spt) return fail(BP_EINVAL, "NULL argument");
  BP_TRY(check_shapes(n_samples, 0, n_stations, n_phases));
  if (n_samples == 0 || n_stations == 0 || n_phases == 0) return BP_OK;
  hipLaunchKernelGGL(bp::interleaved_to_planar_kernel,
                     dim3((unsigned)((n_samples + 255) / 256), (unsigned)n_stations),
                     dim3(256), 0, static_cast<hipStream_t>(stream), F_stp, (int)n_stations,
                     (int)n_samples, (int)n_phases, (long long)n_samples, 0LL, F_spt);
  HIP_TRY(hipGetLastError());
  return BP_OK;
} BP_CATCH

int bp_prestack_planar(const float *u_sct, const float *weights_phases, size_t n_stations,
                       size_t n_channels, size_t n_samples, size_t n_phases, float *F_spt,
                       void *stream) try {
  if (!u_sct || !weights_phases || !F_spt) return fail(BP_EINVAL, "NULL argument");
  BP_TRY(check_shapes(n_samples, n_channels, n_stations, n_phases));
  if (n_samples == 0 || n_stations == 0 || n_phases == 0) return BP_OK;
  hipLaunchKernelGGL(bp::prestack_kernel,
                     dim3((unsigned)((n_samples + 255) / 256), (unsigned)n_stations),
                     dim3(256), 0, static_cast<hipStream_t>(stream), u_sct, weights_phases,
                     (int)n_stations, (int)n_channels, (int)n_samples, (int)n_phases, 1,
                     F_spt);
  HIP_TRY(hipGetLastError());
  return BP_OK;
} BP_CATCH

// ---- drop-in entry points ------------------------------------------------

int beamform_max(float *waveform_features, int *time_delays, float *weights_sources,
                 size_t n_samples, size_t n_sources, size_t n_stations, size_t n_phases,
                 int out_of_bounds, float *beam_max, int *beam_argmax) try {
  HostFeatures hf;
  hf.data = waveform_features;
  return dropin_max(hf, time_delays, weights_sources, n_samples, n_sources, n_stations,
                    n_phases, out_of_bounds, beam_max, beam_argmax);
} BP_CATCH

int beamform(float *waveform_features, int *time_delays, float *weights_sources,
             size_t n_samples, size_t n_sources, size_t n_stations, size_t n_phases,
             int out_of_bounds, float *beam) try {
  HostFeatures hf;
  hf.data = waveform_features;
  return dropin_none(hf, time_delays, weights_sources, n_samples, n_sources, n_stations,
                     n_phases, out_of_bounds, beam);
} BP_CATCH

int bp_beamform_max_fused(const float *waveform_features, const float *weights_phases,
                          const int *time_delays, const float *weights_sources,
                          size_t n_samples, size_t n_sources, size_t n_stations,
                          size_t n_channels, size_t n_phases, int out_of_bounds,
                          float *beam_max, int *beam_argmax) try {
  BP_TRY(check_channels(n_samples, n_stations, n_channels));
  HostFeatures hf;
  hf.data = waveform_features;
  hf.alpha = weights_phases;
  hf.C = n_channels;
  hf.raw = true;
  return dropin_max(hf, time_delays, weights_sources, n_samples, n_sources, n_stations,
                    n_phases, out_of_bounds, beam_max, beam_argmax);
} BP_CATCH

int bp_beamform_fused(const float *waveform_features, const float *weights_phases,
                      const int *time_delays, const float *weights_sources, size_t n_samples,
                      size_t n_sources, size_t n_stations, size_t n_channels,
                      size_t n_phases, int out_of_bounds, float *beam) try {
  BP_TRY(check_channels(n_samples, n_stations, n_channels));
  HostFeatures hf;
  hf.data = waveform_features;
  hf.alpha = weights_phases;
  hf.C = n_channels;
  hf.raw = true;
  return dropin_none(hf, time_delays, weights_sources, n_samples, n_sources, n_stations,
                     n_phases, out_of_bounds, beam);
} BP_CATCH

int prestack_waveform_features(float *waveform_features, float *weights_phases,
                               size_t n_samples, size_t n_stations, size_t n_channels,
                               size_t n_phases, int num_threads, float *prestacked) try {
  (void)num_threads;
  BP_TRY(check_shapes(n_samples, n_channels, n_stations, n_phases));
  const size_t S = n_stations, C = n_channels, T = n_samples, P = n_phases;
  if (S == 0 || T == 0 || P == 0) return BP_OK;
  if (!prestacked) return fail(BP_EINVAL, "NULL output");
  if (C == 0) {
    std::memset(prestacked, 0, S * T * P * sizeof(float));
    return BP_OK;
  }
  if (!waveform_features || !weights_phases) return fail(BP_EINVAL, "NULL input");
  BP_TRY(check_device());
  Stream st;
  BP_TRY(st.create());
  DevBuf du, dw, dout;
  BP_TRY(du.alloc(S * C * T * sizeof(float)));
  BP_TRY(dw.alloc(S * C * P * sizeof(float)));
  BP_TRY(dout.alloc(S * T * P * sizeof(float)));
  HIP_TRY(hipMemcpyAsync(dw.p, weights_phases, S * C * P * sizeof(float),
                         hipMemcpyHostToDevice, st.s));
  BP_TRY(upload_host(du.p, waveform_features, S * C * T * sizeof(float), st.s));
  hipLaunchKernelGGL(bp::prestack_kernel, dim3((unsigned)((T + 255) / 256), (unsigned)S),
                     dim3(256), 0, st.s, du.as<float>(), dw.as<float>(), (int)S, (int)C,
                     (int)T, (int)P, 0, dout.as<float>());
  HIP_TRY(hipGetLastError());
  BP_TRY(download_host(prestacked, dout.p, S * T * P * sizeof(float), st.s));
  return BP_OK;
} BP_CATCH

}  // extern "C"

namespace {

// Shared by bp_batch_create (pl == NULL: tables hipMalloc'd, synchronous
// uploads) and bp_batch_create_placed.
int batch_create(int n_windows, bp_network *const *networks, const bp_feature_view *features,
                 const int *out_of_bounds, float *const *beam_max, int *const *beam_argmax,
                 unsigned long long *const *keys, const bp_batch_placement *pl,
                 bp_batch **out) {
  if (!out) return fail(BP_EINVAL, "out is NULL");
  *out = nullptr;
  if (n_windows <= 0 || !networks || !features || !out_of_bounds || !beam_max ||
      !beam_argmax || !keys)
    return fail(BP_EINVAL, "a batch needs n_windows > 0 and non-NULL arrays");
  if (n_windows > 65535) return fail(BP_EINVAL, "too many windows (%d)", n_windows);
  const int fmt = features[0].fmt;
  auto b = std::make_unique<bp_batch>();
  b->fmt = fmt;
  b->n = n_windows;
  Stream own;
  hipStream_t st = nullptr;
  if (pl) {
    if (!pl->ws || pl->ws_bytes < bp_batch_table_bytes(n_windows))
      return fail(BP_EINVAL, "batch table workspace of %zu bytes, %zu needed", pl->ws_bytes,
                  bp_batch_table_bytes(n_windows));
    b->ws = static_cast<char *>(pl->ws);
    b->ws_bytes = pl->ws_bytes;
    b->upload = pl->stream;
    st = pl->stream;
  } else {
    BP_TRY(check_device());
    BP_TRY(own.create());
    st = own.s;
  }
  std::vector<bp::BatchKeys> &kd = b->keys_h;
  for (int i = 0; i < n_windows; ++i) {
    bp_network *net = networks[i];
    if (!net || !beam_max[i] || !beam_argmax[i] || !keys[i])
      return fail(BP_EINVAL, "window %d: NULL network or output", i);
    const bp_feature_view &fv = features[i];
    if (fv.fmt != fmt)
      return fail(BP_EINVAL, "window %d: format %d, window 0 has %d (one format per batch)", i,
                  fv.fmt, fmt);
    BP_TRY(check_view(net, fv));
    if (i == 0) b->P = net->P;
    if (net->P != b->P)
      return fail(BP_EINVAL, "window %d has %d phases, window 0 has %d", i, net->P, b->P);
    const int oob = out_of_bounds[i];
    if (oob != 0 && oob != 1)
      return fail(BP_EINVAL, "window %d: out_of_bounds must be 0 or 1, not %d", i, oob);
    std::shared_ptr<Schedule> sched;
    const float *sc = fmt == BP_F32 ? nullptr : fv.station_scale;
    if (pl) {
      sched = find_schedule(net, fmt, sc);
      if (!sched)
        return fail(BP_EINVAL, "window %d: its network has no schedule for format %d and these "
                    "station scales (bp_network_prepare)", i, fmt);
    } else {
      BP_TRY(ensure_schedule(net, fmt, sc, st, sched));
    }
    b->scheds.push_back(sched);
    const int T = (int)fv.n_samples;
    kd.push_back(bp::BatchKeys{keys[i], beam_max[i], beam_argmax[i], T});
    b->max_T = std::max(b->max_T, T);
    if (T == 0) continue;
    bp::StackArgs a = make_args_view(net, fv, oob, *sched);
    a.keys = keys[i];
    a.k_begin = 0;
    a.k_end = net->K;
    a.n_time_tiles = (T + kTT - 1) / kTT;
    const long long planes = fmt == BP_F32 ? 1 : 2, es = fmt == BP_F32 ? 4 : 2;
    const bool lds_ok = (net->P * planes * a.ldf + kLargeLw + 16) * es < (1LL << 31);
    for (int c = 0; c < 2; ++c) {
      if (net->tiles[c].empty()) continue;
      if (!lds_ok) return fail(BP_EINVAL, "window %d: n_samples %d too large for the LDS kernel", i, T);
      a.tiles = net->dtiles[c].as<int>();
      a.n_list = (int)net->tiles[c].size();
      a.Lw = net->Lw[c];
      b->all[c].add(a);
      if ((long long)a.n_time_tiles * a.n_list >= kLargeWindowItems)
        b->large[c].push_back(a);
      else
        b->small[c].add(a);
    }
    if (!net->tiles[2].empty()) {
      const int nd = (int)net->tiles[2].size();
      if (nd > 65535) return fail(BP_EINVAL, "window %d: too many direct-kernel tiles (%d)", i, nd);
      a.tiles = net->dtiles[2].as<int>();
      a.n_list = nd;
      b->direct.push_back(a);
      b->direct_n.push_back(nd);
    }
  }
  for (int c = 0; c < 2; ++c) {
    BP_TRY(finish_fused(b.get(), b->all[c]));
    BP_TRY(finish_fused(b.get(), b->small[c]));
  }
  BP_TRY(b->table(b->dkeys, kd.size() * sizeof(bp::BatchKeys)));
  BP_TRY(b->put(b->dkeys, kd.data(), kd.size() * sizeof(bp::BatchKeys)));
  BP_TRY(init_batch_attrs());
  BP_TRY(init_kernel_attrs());
  bp_batch_config cfg = b->cfg;
  if (pl) cfg.blocks_per_cu = pl->blocks_per_cu;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune_cache.find(tune_key(b.get()));
    if (it != g_tune_cache.end()) cfg = it->second;
  }
  BP_TRY(batch_configure(b.get(), cfg));
  if (!pl) HIP_TRY(hipStreamSynchronize(st));
  *out = b.release();
  return BP_OK;
}

}  // namespace

size_t bp_batch_table_bytes(int n_windows) {
  const size_t n = (size_t)std::max(0, n_windows);
  auto up = [](size_t v) { return (v + 255) / 256 * 256; };
  // four fused sets (two classes, all windows and the small ones): window
  // descriptors, XCD split and work-queue counters each; the keys table
  return 4 * (up(n * sizeof(bp::BatchWindow)) + up(9 * sizeof(long long)) +
              up(8 * sizeof(unsigned long long))) +
         up(n * sizeof(bp::BatchKeys)) + 256;
}

int bp_batch_create_placed(int n_windows, bp_network *const *networks,
                           const bp_feature_view *features, const int *out_of_bounds,
                           float *const *beam_max, int *const *beam_argmax,
                           unsigned long long *const *keys, const bp_batch_placement &pl,
                           bp_batch **out) try {
  return batch_create(n_windows, networks, features, out_of_bounds, beam_max, beam_argmax, keys,
                      &pl, out);
} BP_CATCH

extern "C" {

int bp_batch_create(int n_windows, bp_network *const *networks,
                    const bp_feature_view *features, const int *out_of_bounds,
                    float *const *beam_max, int *const *beam_argmax,
                    unsigned long long *const *keys, bp_batch **out) try {
  return batch_create(n_windows, networks, features, out_of_bounds, beam_max, beam_argmax, keys,
                      nullptr, out);
} BP_CATCH

int bp_batch_launch(bp_batch *b, void *stream) try {
  if (!b) return fail(BP_EINVAL, "NULL batch");
  return batch_enqueue(b, static_cast<hipStream_t>(stream), true);
} BP_CATCH

int bp_batch_graph_launch(bp_batch *b, void *stream) try {
  if (!b) return fail(BP_EINVAL, "NULL batch");
  if (!b->exec) {
    if (!b->cap) HIP_TRY(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    const int rc = batch_enqueue(b, b->cap, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc != BP_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIP_TRY(e);
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  }
  HIP_TRY(hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)));
  return BP_OK;
} BP_CATCH

int bp_batch_get_config(const bp_batch *b, bp_batch_config *cfg) try {
  if (!b || !cfg) return fail(BP_EINVAL, "NULL argument");
  *cfg = b->cfg;
  return BP_OK;
} BP_CATCH

int bp_batch_set_config(bp_batch *b, const bp_batch_config *cfg) try {
  if (!b || !cfg) return fail(BP_EINVAL, "NULL argument");
  return batch_configure(b, *cfg);
} BP_CATCH

int bp_batch_tune(bp_batch *b, int reps, bp_batch_config *tried, int max_tried, int *n_tried,
                  bp_batch_config *best, void *stream) try {
  if (!b) return fail(BP_EINVAL, "NULL batch");
  if (reps < 1) reps = 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<bp_batch_config> cands;
  for (int tg : {1, 2, 4, 8})
    for (int nb : {2, 3})
      for (int bpc : {-1, 0, 1, 2, 3}) cands.push_back(bp_batch_config{tg, nb, bpc, 0.0f});
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int count = 0, rc = BP_OK;
  bp_batch_config win{};
  win.ms = -1.0f;
  for (bp_batch_config &c : cands) {
    if ((rc = batch_configure(b, c)) != BP_OK) break;
    if ((rc = batch_enqueue(b, st, false)) != BP_OK) break;  // warm-up
    if (hipEventRecord(e0, st) != hipSuccess) { rc = fail(BP_EHIP, "hipEventRecord"); break; }
    for (int r = 0; r < reps && rc == BP_OK; ++r) rc = batch_enqueue(b, st, false);
    if (rc != BP_OK) break;
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      rc = fail(BP_EHIP, "timing events failed");
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    c.ms = ms / reps;
    if (tried && count < max_tried) tried[count] = c;
    ++count;
    if (win.ms < 0 || c.ms < win.ms) win = c;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (n_tried) *n_tried = count;
  if (rc != BP_OK) return rc;
  BP_TRY(batch_configure(b, win));
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_cache[tune_key(b)] = win;
  }
  if (best) *best = win;
  return BP_OK;
} BP_CATCH

int bp_batch_destroy(bp_batch *b) try {
  delete b;
  return BP_OK;
} BP_CATCH

}  // extern "C"
