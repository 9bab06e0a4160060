// beamform_gpu.hip -- C ABI of the gfx950 beamforming library.
//
// Drop-in symbols beamform / beamform_max replace the reference's GPU
// library (beampower/src/beamform_gpu.h:1-2, bound by beampower/core.py:65-87);
// prestack_waveform_features restates the CPU library's symbol
// (beamform_cpu.h:4) on the GPU.  The bp_* entry points expose the same
// engine on device-resident buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "beamform_gpu.h"
#include "bp_kernels.h"

#define BP_VERSION "0.1.0"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
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

// Kernel configuration of the LDS shift-and-stack kernel.
constexpr int kR = 8;    // samples per lane -> time tile 512
constexpr int kKW = 8;   // sources per wave
constexpr int kW = 8;    // waves per work-group -> source tile 64
constexpr int kNJ = 4;   // staging registers per slot -> Lmax <= 2048
constexpr int kTileSources = kKW * kW;
constexpr int kDirectChunk = 64;

template <int P, bool NONE>
using Lds = bp::StackLds<P, 4 / P, kR, kKW, kW, kNJ, NONE>;

// Device buffer owned by RAII.
struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? BP_ENOMEM : BP_EHIP,
                  "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
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

int check_device() {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(BP_ENODEV, "device %d is %s; this library is built for gfx950", dev,
                prop.gcnArchName);
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

template <int P, bool NONE>
hipError_t set_lds_attr() {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(
                                 &bp::stack_lds_kernel<P, 4 / P, kR, kKW, kW, kNJ, NONE>),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)Lds<P, NONE>::lds_bytes(Lds<P, NONE>::MAX_LMAX));
}

int init_kernel_attrs() {
  std::call_once(g_attr_once, [] {
    hipError_t e = set_lds_attr<1, false>();
    if (e == hipSuccess) e = set_lds_attr<2, false>();
    if (e == hipSuccess) e = set_lds_attr<1, true>();
    if (e == hipSuccess) e = set_lds_attr<2, true>();
    if (e != hipSuccess)
      g_attr_rc = fail(BP_EHIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
  });
  return g_attr_rc;
}

}  // namespace

struct bp_network {
  int K = 0, S = 0, P = 0;
  int n_tiles = 0;
  int max_spread = 0;
  DevBuf tau, beta, lohi, tile_lo, tile_hi, tile_all;
};

namespace {

bp::StackArgs make_args(const bp_network *net, const float *F, int T, int oob) {
  bp::StackArgs a{};
  a.F = F;
  a.tau = net->tau.as<int>();
  a.beta = net->beta.as<float>();
  a.src_lohi = net->lohi.as<int>();
  a.tile_lo = net->tile_lo.as<int>();
  a.tile_hi = net->tile_hi.as<int>();
  a.tile_all = net->tile_all.as<int>();
  a.T = T;
  a.K = net->K;
  a.S = net->S;
  a.P = net->P;
  a.oob = oob;
  return a;
}

// Launch the stack kernel (LDS or direct) for sources [k_begin, k_end).
template <bool NONE>
int launch_stack(const bp_network *net, bp::StackArgs a, int k_begin, int k_end,
                 hipStream_t st) {
  a.k_begin = k_begin;
  a.k_end = k_end;
  if (k_end <= k_begin || a.T <= 0) return BP_OK;
  const int P = net->P;
  const long long Lmax = (long long)kR * 64 + net->max_spread;
  const bool lds_ok = (P == 1 || P == 2) && Lmax <= (long long)kNJ * kW * 64;
  hipEvent_t ev[2];
  BP_TRY(prof_begin(st, ev));
  if (lds_ok) {
    BP_TRY(init_kernel_attrs());
    a.Lmax = (int)Lmax;
    a.n_time_tiles = (a.T + kR * 64 - 1) / (kR * 64);
    const int tile0 = k_begin / kTileSources;
    const int tile1 = (k_end - 1) / kTileSources + 1;
    a.tile_begin = tile0;
    const long long blocks = (long long)a.n_time_tiles * (tile1 - tile0);
    if (blocks > INT32_MAX) return fail(BP_EINVAL, "grid too large (%lld blocks)", blocks);
    size_t lds = P == 1 ? Lds<1, NONE>::lds_bytes(a.Lmax) : Lds<2, NONE>::lds_bytes(a.Lmax);
    if (P == 1)
      hipLaunchKernelGGL((bp::stack_lds_kernel<1, 4, kR, kKW, kW, kNJ, NONE>),
                         dim3((unsigned)blocks), dim3(64 * kW), lds, st, a);
    else
      hipLaunchKernelGGL((bp::stack_lds_kernel<2, 2, kR, kKW, kW, kNJ, NONE>),
                         dim3((unsigned)blocks), dim3(64 * kW), lds, st, a);
  } else {
    // sources per block row: at least kDirectChunk, enough to keep gridDim.y < 65536
    const long long n = (long long)k_end - k_begin;
    const int chunk = (int)std::max<long long>(kDirectChunk, (n + 65534) / 65535);
    const unsigned gx = (unsigned)((a.T + 255) / 256);
    const unsigned gy = (unsigned)((n + chunk - 1) / chunk);
    hipLaunchKernelGGL((bp::stack_direct_kernel<NONE>), dim3(gx, gy), dim3(256), 0, st, a,
                       chunk);
  }
  HIP_TRY(hipGetLastError());
  BP_TRY(prof_end(st, ev));
  return BP_OK;
}

int network_max(bp_network *net, const float *F, size_t n_samples, int oob, float *bmax,
                int *barg, unsigned long long *keys, hipStream_t st) {
  const int T = (int)n_samples;
  const unsigned nb = (unsigned)((T + 255) / 256);
  hipLaunchKernelGGL(bp::fill_u64_kernel, dim3(nb), dim3(256), 0, st, keys, (long long)T,
                     0ull);
  HIP_TRY(hipGetLastError());
  bp::StackArgs a = make_args(net, F, T, oob);
  a.keys = keys;
  BP_TRY(launch_stack<false>(net, a, 0, net->K, st));
  hipLaunchKernelGGL(bp::finalize_max_kernel, dim3(nb), dim3(256), 0, st, keys, T, bmax,
                     barg);
  HIP_TRY(hipGetLastError());
  return BP_OK;
}

int network_create(const int *tau, const float *beta, size_t K, size_t S, size_t P,
                   hipStream_t st, std::unique_ptr<bp_network> &out) {
  auto net = std::make_unique<bp_network>();
  net->K = (int)K;
  net->S = (int)S;
  net->P = (int)P;
  net->n_tiles = (int)((K + kTileSources - 1) / kTileSources);
  const size_t ksp = K * S * P, ks = K * S, tsp = (size_t)net->n_tiles * S * P;
  BP_TRY(net->tau.alloc(ksp * sizeof(int)));
  BP_TRY(net->beta.alloc(ks * sizeof(float)));
  BP_TRY(net->lohi.alloc(2 * K * sizeof(int)));
  BP_TRY(net->tile_lo.alloc(tsp * sizeof(int)));
  BP_TRY(net->tile_hi.alloc(tsp * sizeof(int)));
  BP_TRY(net->tile_all.alloc(2 * (size_t)net->n_tiles * sizeof(int)));
  DevBuf spread;
  BP_TRY(spread.alloc(sizeof(int)));
  HIP_TRY(hipMemcpyAsync(net->tau.p, tau, ksp * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(net->beta.p, beta, ks * sizeof(float), hipMemcpyHostToDevice, st));
  int zero = 0;
  HIP_TRY(hipMemcpyAsync(spread.p, &zero, sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(bp::source_bounds_kernel, dim3((unsigned)((K + 255) / 256)), dim3(256),
                     0, st, net->tau.as<int>(), net->beta.as<float>(), net->K, net->S,
                     net->P, net->lohi.as<int>());
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(bp::tile_bounds_kernel, dim3((unsigned)net->n_tiles), dim3(256), 0, st,
                     net->tau.as<int>(), net->beta.as<float>(), net->K, net->S, net->P,
                     kTileSources, net->tile_lo.as<int>(), net->tile_hi.as<int>(),
                     net->tile_all.as<int>(), spread.as<int>());
  HIP_TRY(hipGetLastError());
  int ms = 0;
  HIP_TRY(hipMemcpyAsync(&ms, spread.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  net->max_spread = ms;
  out = std::move(net);
  return BP_OK;
}

int check_shapes(size_t T, size_t K, size_t S, size_t P) {
  if (!size_ok(T) || !size_ok(K) || !size_ok(S) || !size_ok(P))
    return fail(BP_EINVAL, "dimension too large (T=%zu K=%zu S=%zu P=%zu)", T, K, S, P);
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

}  // namespace

extern "C" {

const char *bp_last_error(void) { return g_err.c_str(); }

const char *bp_version(void) { return BP_VERSION; }

int bp_profiling(int enable) {
  g_prof.on = enable != 0;
  return BP_OK;
}

int bp_profiling_read(double *stack_ms_total, long long *stack_launches) {
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
}

int bp_network_create(const int *time_delays, const float *weights_sources, size_t n_sources,
                      size_t n_stations, size_t n_phases, bp_network **out) {
  if (!out) return fail(BP_EINVAL, "out is NULL");
  *out = nullptr;
  if (n_sources == 0 || n_stations == 0 || n_phases == 0)
    return fail(BP_EINVAL, "network needs n_sources, n_stations, n_phases > 0");
  if (!time_delays || !weights_sources) return fail(BP_EINVAL, "NULL input");
  BP_TRY(check_shapes(0, n_sources, n_stations, n_phases));
  BP_TRY(check_device());
  Stream st;
  BP_TRY(st.create());
  std::unique_ptr<bp_network> net;
  BP_TRY(network_create(time_delays, weights_sources, n_sources, n_stations, n_phases, st.s,
                        net));
  *out = net.release();
  return BP_OK;
}

int bp_network_destroy(bp_network *net) {
  delete net;
  return BP_OK;
}

int bp_network_info(const bp_network *net, int *source_tile, int *max_tile_spread,
                    int *n_tiles) {
  if (!net) return fail(BP_EINVAL, "NULL network");
  if (source_tile) *source_tile = kTileSources;
  if (max_tile_spread) *max_tile_spread = net->max_spread;
  if (n_tiles) *n_tiles = net->n_tiles;
  return BP_OK;
}

int bp_network_beamform_max(bp_network *net, const void *features, int fmt,
                            const float *station_scale, size_t n_samples, int out_of_bounds,
                            float *beam_max, int *beam_argmax, unsigned long long *keys,
                            void *stream) {
  if (!net || !features || !beam_max || !beam_argmax || !keys)
    return fail(BP_EINVAL, "NULL argument");
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 (strict) or 1 (flexible), not %d",
                out_of_bounds);
  if (fmt != BP_F32) return fail(BP_EINVAL, "feature format %d not supported", fmt);
  if (station_scale) return fail(BP_EINVAL, "station_scale applies to narrow formats only");
  BP_TRY(check_shapes(n_samples, net->K, net->S, net->P));
  if (n_samples == 0) return BP_OK;
  return network_max(net, static_cast<const float *>(features), n_samples, out_of_bounds,
                     beam_max, beam_argmax, keys, static_cast<hipStream_t>(stream));
}

int bp_network_beamform(bp_network *net, const float *features, size_t n_samples,
                        int out_of_bounds, size_t source_begin, size_t source_end,
                        float *beam, void *stream) {
  if (!net || !features || !beam) return fail(BP_EINVAL, "NULL argument");
  if (out_of_bounds != 0 && out_of_bounds != 1)
    return fail(BP_EINVAL, "out_of_bounds must be 0 (strict) or 1 (flexible), not %d",
                out_of_bounds);
  if (source_begin > source_end || source_end > (size_t)net->K)
    return fail(BP_EINVAL, "bad source range [%zu, %zu) for %d sources", source_begin,
                source_end, net->K);
  BP_TRY(check_shapes(n_samples, net->K, net->S, net->P));
  if (n_samples == 0) return BP_OK;
  bp::StackArgs a = make_args(net, features, (int)n_samples, out_of_bounds);
  a.beam = beam;
  return launch_stack<true>(net, a, (int)source_begin, (int)source_end,
                            static_cast<hipStream_t>(stream));
}

int bp_interleaved_to_planar(const float *F_stp, size_t n_stations, size_t n_samples,
                             size_t n_phases, float *F_spt, void *stream) {
  if (!F_stp || !F_spt) return fail(BP_EINVAL, "NULL argument");
  BP_TRY(check_shapes(n_samples, 0, n_stations, n_phases));
  if (n_samples == 0 || n_stations == 0 || n_phases == 0) return BP_OK;
  hipLaunchKernelGGL(bp::interleaved_to_planar_kernel,
                     dim3((unsigned)((n_samples + 255) / 256), (unsigned)n_stations),
                     dim3(256), 0, static_cast<hipStream_t>(stream), F_stp, (int)n_stations,
                     (int)n_samples, (int)n_phases, F_spt);
  HIP_TRY(hipGetLastError());
  return BP_OK;
}

int bp_prestack_planar(const float *u_sct, const float *weights_phases, size_t n_stations,
                       size_t n_channels, size_t n_samples, size_t n_phases, float *F_spt,
                       void *stream) {
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
}

// ---- drop-in entry points ------------------------------------------------

int beamform_max(float *waveform_features, int *time_delays, float *weights_sources,
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
  if (!waveform_features || !time_delays || !weights_sources)
    return fail(BP_EINVAL, "NULL input");
  BP_TRY(check_device());
  Stream st;
  BP_TRY(st.create());
  std::unique_ptr<bp_network> net;
  BP_TRY(network_create(time_delays, weights_sources, n_sources, n_stations, n_phases, st.s,
                        net));
  const size_t T = n_samples, nf = n_stations * n_samples * n_phases;
  DevBuf f_stp, f_spt, keys, bmax, barg;
  BP_TRY(f_stp.alloc(nf * sizeof(float)));
  BP_TRY(f_spt.alloc(nf * sizeof(float)));
  BP_TRY(keys.alloc(T * sizeof(unsigned long long)));
  BP_TRY(bmax.alloc(T * sizeof(float)));
  BP_TRY(barg.alloc(T * sizeof(int)));
  HIP_TRY(hipMemcpyAsync(f_stp.p, waveform_features, nf * sizeof(float),
                         hipMemcpyHostToDevice, st.s));
  // the caller's beam_max is kept where no source is valid (beamform.c:332-333)
  HIP_TRY(hipMemcpyAsync(bmax.p, beam_max, T * sizeof(float), hipMemcpyHostToDevice, st.s));
  BP_TRY(bp_interleaved_to_planar(f_stp.as<float>(), n_stations, n_samples, n_phases,
                                  f_spt.as<float>(), st.s));
  BP_TRY(network_max(net.get(), f_spt.as<float>(), T, out_of_bounds, bmax.as<float>(),
                     barg.as<int>(), keys.as<unsigned long long>(), st.s));
  HIP_TRY(hipMemcpyAsync(beam_max, bmax.p, T * sizeof(float), hipMemcpyDeviceToHost, st.s));
  HIP_TRY(hipMemcpyAsync(beam_argmax, barg.p, T * sizeof(int), hipMemcpyDeviceToHost, st.s));
  HIP_TRY(hipStreamSynchronize(st.s));
  return BP_OK;
}

int beamform(float *waveform_features, int *time_delays, float *weights_sources,
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
  if (!waveform_features || !time_delays || !weights_sources)
    return fail(BP_EINVAL, "NULL input");
  BP_TRY(check_device());
  Stream st;
  BP_TRY(st.create());
  std::unique_ptr<bp_network> net;
  BP_TRY(network_create(time_delays, weights_sources, n_sources, n_stations, n_phases, st.s,
                        net));
  const size_t nf = n_stations * n_samples * n_phases;
  DevBuf f_stp, f_spt, dbeam;
  BP_TRY(f_stp.alloc(nf * sizeof(float)));
  BP_TRY(f_spt.alloc(nf * sizeof(float)));
  BP_TRY(dbeam.alloc(K * T * sizeof(float)));
  HIP_TRY(hipMemcpyAsync(f_stp.p, waveform_features, nf * sizeof(float),
                         hipMemcpyHostToDevice, st.s));
  BP_TRY(bp_interleaved_to_planar(f_stp.as<float>(), n_stations, n_samples, n_phases,
                                  f_spt.as<float>(), st.s));
  // invalid strict cells read back as 0, as the wrapper's zero-filled buffer
  const long long n = (long long)(K * T);
  hipLaunchKernelGGL(bp::fill_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     st.s, dbeam.as<float>(), n, 0.0f);
  HIP_TRY(hipGetLastError());
  bp::StackArgs a = make_args(net.get(), f_spt.as<float>(), (int)T, out_of_bounds);
  a.beam = dbeam.as<float>();
  BP_TRY(launch_stack<true>(net.get(), a, 0, (int)K, st.s));
  HIP_TRY(hipMemcpyAsync(beam, dbeam.p, K * T * sizeof(float), hipMemcpyDeviceToHost, st.s));
  HIP_TRY(hipStreamSynchronize(st.s));
  return BP_OK;
}

int prestack_waveform_features(float *waveform_features, float *weights_phases,
                               size_t n_samples, size_t n_stations, size_t n_channels,
                               size_t n_phases, int num_threads, float *prestacked) {
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
  HIP_TRY(hipMemcpyAsync(du.p, waveform_features, S * C * T * sizeof(float),
                         hipMemcpyHostToDevice, st.s));
  HIP_TRY(hipMemcpyAsync(dw.p, weights_phases, S * C * P * sizeof(float),
                         hipMemcpyHostToDevice, st.s));
  hipLaunchKernelGGL(bp::prestack_kernel, dim3((unsigned)((T + 255) / 256), (unsigned)S),
                     dim3(256), 0, st.s, du.as<float>(), dw.as<float>(), (int)S, (int)C,
                     (int)T, (int)P, 0, dout.as<float>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(prestacked, dout.p, S * T * P * sizeof(float),
                         hipMemcpyDeviceToHost, st.s));
  HIP_TRY(hipStreamSynchronize(st.s));
  return BP_OK;
}

}  // extern "C"
