/*
 * beamform_gpu.h -- C ABI of beampower_amd/lib/beamform_gpu.so (gfx950).
 *
 * Drop-in entry points.  Their names, argument order, types and meaning are
 * those the reference's Python layer binds for device="gpu":
 *   reference prototypes  beampower/src/beamform_gpu.h:1-2
 *   ctypes argtypes       beampower/core.py:65-87
 *   call sites            beampower/beampower.py:156-190
 * All pointers are HOST pointers (NumPy buffers); each call is synchronous.
 * The reference returns void and its wrapper never reads the result (ctypes
 * restype stays c_int), so returning an int status is ABI-compatible:
 * 0 on success, a negative BP_E* code otherwise (never exit()).
 * Calls are re-entrant: each call uses its own stream and buffers.
 */
#ifndef BEAMFORM_GPU_H
#define BEAMFORM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BP_OK 0
#define BP_EINVAL (-22)   /* bad argument (shape, mode, null pointer) */
#define BP_ENOMEM (-12)   /* device allocation failed */
#define BP_EHIP (-5)      /* HIP runtime error */
#define BP_ENODEV (-19)   /* no usable gfx950 device */

/* Replaces beamform(...) of beamform.cu:462-659 (reduce='none').
 * F: prestacked features (n_stations, n_samples, n_phases) float32
 * time_delays: (n_sources, n_stations, n_phases) int32 moveouts in samples
 * weights_sources: (n_sources, n_stations) float32
 * out_of_bounds: 0 = strict, 1 = flexible
 * beam: (n_sources, n_samples) float32 output; cells that are invalid in
 *       strict mode are set to 0 (the wrapper passes a zero-filled buffer). */
int beamform(float *waveform_features, int *time_delays, float *weights_sources,
             size_t n_samples, size_t n_sources, size_t n_stations,
             size_t n_phases, int out_of_bounds, float *beam);

/* Replaces beamform_max(...) of beamform.cu:207-460 (reduce='max').
 * beam_max: (n_samples) float32, written only where some source is valid
 *           (the caller's value is kept elsewhere, beamform.c:332-333)
 * beam_argmax: (n_samples) int32, lowest source index attaining the max,
 *           0 where no source is valid (beamform.c:269-270,334). */
int beamform_max(float *waveform_features, int *time_delays, float *weights_sources,
                 size_t n_samples, size_t n_sources, size_t n_stations,
                 size_t n_phases, int out_of_bounds, float *beam_max,
                 int *beam_argmax);

/* beamform and beamform_max keep the network they build for time_delays
 * and weights_sources (device tables, plan, schedules) with a host copy of
 * those two arrays; a later call on the same device whose arrays are equal
 * byte for byte (compared in full) reuses it instead of uploading and
 * planning again, and re-plans it once with the similarity source order on
 * its first reuse.  Results are the same either way.  One network is kept
 * (the last one built); BP_DROPIN_CACHE=0 in the environment turns the reuse
 * off.  bp_dropin_cache_clear releases the kept network (calls in progress
 * keep theirs until they return).  Extension, no reference counterpart. */
int bp_dropin_cache_clear(void);

/* Replaces the CPU library's prestack_waveform_features (beamform_cpu.h:4,
 * beamform.c:108-152; argtypes core.py:40-49), computed on the GPU.
 * waveform_features: (n_stations, n_channels, n_samples) float32
 * weights_phases: (n_stations, n_channels, n_phases) float32
 * prestacked: (n_stations, n_samples, n_phases) float32 output.
 * num_threads is accepted for signature compatibility and ignored. */
int prestack_waveform_features(float *waveform_features, float *weights_phases,
                               size_t n_samples, size_t n_stations,
                               size_t n_channels, size_t n_phases,
                               int num_threads, float *prestacked);

/* Replaces the CPU library's beamform_differential (beamform_cpu.h:6,
 * beamform.c:338-394; argtypes core.py:26-35), which the reference's GPU
 * path lacks (beampower.py:220-223).  The time delays are relative to a
 * zero lag at the centre sample t = (n_samples - 1) / 2.
 * waveform_features: (n_stations, n_samples, n_phases) float32 prestacked
 * time_delays: (n_sources, n_stations, n_phases) int32
 * weights: (n_sources, n_stations) float32
 * beam: (n_sources) float32; beam[k] = the beam of source k at t when all
 *       of its used delays point inside [0, n_samples), else left as passed
 *       (the wrapper passes zeros, beampower.py:198).
 * num_threads is accepted for signature compatibility and ignored. */
int beamform_differential(float *waveform_features, int *time_delays, float *weights,
                          size_t n_samples, size_t n_sources, size_t n_stations,
                          size_t n_phases, int num_threads, float *beam);

/* The reference wrapper's two library calls for device="gpu" in one:
 * prestack_waveform_features (beamform_cpu.h:4, beamform.c:108-152) followed
 * by beamform_max (beamform_gpu.h:2) or beamform (beamform_gpu.h:1), as
 * beampower.py:103-105,156-190 makes them.  The (n_stations, n_channels,
 * n_samples) features and the (n_stations, n_channels, n_phases) phase
 * weights go to the device as they are and are prestacked there, straight
 * into the rows the stack kernels read, so the prestacked array never
 * crosses PCIe.  Results are bit-identical to the two calls; the other
 * arguments, the outputs and the reuse of the network for equal
 * time_delays and weights_sources are those of beamform_max / beamform.
 * Host pointers, synchronous.  Extension: beampower_amd.beamform uses them
 * for device="gpu". */
int bp_beamform_max_fused(const float *waveform_features, const float *weights_phases,
                          const int *time_delays, const float *weights_sources,
                          size_t n_samples, size_t n_sources, size_t n_stations,
                          size_t n_channels, size_t n_phases, int out_of_bounds,
                          float *beam_max, int *beam_argmax);
int bp_beamform_fused(const float *waveform_features, const float *weights_phases,
                      const int *time_delays, const float *weights_sources, size_t n_samples,
                      size_t n_sources, size_t n_stations, size_t n_channels,
                      size_t n_phases, int out_of_bounds, float *beam);

/* ---- extension API (no counterpart in the reference) -------------------- */

/* Thread-local message for the last non-zero status of this thread. */
const char *bp_last_error(void);

/* Library version string. */
const char *bp_version(void);

/* Feature storage formats for the device-resident entry points. */
#define BP_F32 0
#define BP_BF16 1
#define BP_F16 2

/* A network: moveouts and source weights resident on the device, with the
 * per-source and per-source-tile delay bounds precomputed once.
 * tau, beta are HOST pointers with the layouts of beamform_max.
 * bp_network_create plans with BP_ORDER_AUTO. */
typedef struct bp_network bp_network;
int bp_network_create(const int *time_delays, const float *weights_sources,
                      size_t n_sources, size_t n_stations, size_t n_phases,
                      bp_network **out);
int bp_network_destroy(bp_network *net);

/* Source order of a network's plan (no counterpart in the reference, which
 * scans sources in the caller's order, beamform.c:255-300).  Source tiles
 * hold consecutive positions; a tile's LDS window grows with the spread of
 * its sources' delays.
 *   BP_ORDER_NATURAL     positions = the caller's source indices;
 *   BP_ORDER_SIMILARITY  positions from a recursive bisection of the delay
 *                        vectors, so that a tile holds sources of similar
 *                        delays (always applied);
 *   BP_ORDER_AUTO        the similarity order when its plan is cheaper than
 *                        the natural one (the natural order is kept on a tie).
 * Results are the same in every order: beams are indexed, and max ties
 * broken, by the caller's source index.  Creating a network with the
 * similarity order costs a host pass over the delays (about 70 ms at
 * 192,000 sources x 100 delay entries). */
#define BP_ORDER_NATURAL 0
#define BP_ORDER_SIMILARITY 1
#define BP_ORDER_AUTO 2
int bp_network_create_ordered(const int *time_delays, const float *weights_sources,
                              size_t n_sources, size_t n_stations, size_t n_phases,
                              int order, bp_network **out);
/* *permuted = 1 when the plan's positions differ from the caller's indices;
 * order (n_sources ints, may be NULL) receives the caller's source index at
 * each plan position. */
int bp_network_source_order(const bp_network *net, int *permuted, int *order);
/* Largest per-(station, phase) delay spread inside one source tile. */
int bp_network_info(const bp_network *net, int *source_tile, int *max_tile_spread,
                    int *n_tiles);
/* Compile-time shape of the LDS kernel of this build (no counterpart in the
 * reference): sources per tile, samples per time tile, and the waves per SIMD
 * it is compiled for (beampower_amd.tune chooses among builds). */
int bp_kernel_shape(int *source_tile, int *time_tile, int *waves_per_simd);
/* Source tiles per kernel class (small LDS window, large LDS window,
 * direct gathers) and the LDS window length of each class (0 = direct). */
int bp_network_classes(const bp_network *net, int *tiles_per_class, int *window_per_class);

/* reduce='max' on device-resident features.
 * features: DEVICE pointer in the format `fmt`:
 *   BP_F32  planar float32 (n_stations, n_phases, n_samples);
 *   BP_F16 / BP_BF16  16-bit parity planes as written by bp_quantize_planar:
 *     (n_stations, n_phases, 2, bp_parity_stride(n_samples)), plane 0 holding
 *     x[t] and plane 1 x[t+1].
 * station_scale: HOST pointer (n_stations), 16-bit formats only: multiplier
 *   of station s's weights, i.e. the inverse of the scale its features were
 *   stored with (bp_quantize_planar returns it).  NULL = 1.  Powers of two
 *   keep the folding exact.  Must be NULL for BP_F32.
 * beam_max, beam_argmax: DEVICE pointers (n_samples); beam_max keeps its
 *   previous value where no source is valid.
 * keys: DEVICE scratch of n_samples uint64 (caller-provided workspace).
 * stream: hipStream_t (NULL = default stream).  Asynchronous, except that
 *   the first call with a format, or with new station_scale values, builds
 *   the network's schedule for it synchronously.  A schedule, once built,
 *   is never rewritten; a network keeps those of its four most recently
 *   used scale sets per format, and dropping an older one frees device
 *   memory, which waits for the device to be idle (hipFree). */
int bp_network_beamform_max(bp_network *net, const void *features, int fmt,
                            const float *station_scale, size_t n_samples,
                            int out_of_bounds, float *beam_max, int *beam_argmax,
                            unsigned long long *keys, void *stream);

/* A device feature buffer in a storage format, with optional zero padding
 * around every row.  Row r (r = s*P + p for BP_F32; r = (s*P + p)*2 + plane
 * for the 16-bit parity planes) starts at element r * row_stride of `data`;
 * its sample 0 is element pad_before of the row.  The elements before
 * sample 0 and from sample n_samples to the end of the row must read as
 * zero features (bp_pack_features writes them so).  Padding at least
 * bp_network_padding() on both sides lets every LDS tile of the network
 * stage its windows without the guarded edge path.  16-bit rows need even
 * row_stride and pad_before.  station_scale: HOST (n_stations) weight
 * multipliers of 16-bit features (from bp_pack_features), or NULL. */
typedef struct {
  const void *data;
  int fmt;
  size_t n_samples;
  size_t row_stride;
  size_t pad_before;
  const float *station_scale;
} bp_feature_view;

/* reduce='max' on a feature view; otherwise as bp_network_beamform_max. */
int bp_network_beamform_max_view(bp_network *net, const bp_feature_view *features,
                                 int out_of_bounds, float *beam_max, int *beam_argmax,
                                 unsigned long long *keys, void *stream);

/* reduce='max' on a feature view without allocating anything (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  As bp_network_beamform_max_view, but the key scratch
 * is the library's own: a device array of BP_NOALLOC_MAX_SAMPLES packed
 * keys (32 MiB) that each GPU holds from the first call on it until the
 * process ends.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (on different streams or from different threads) race
 * and must not be made.  Fails (BP_EINVAL, nothing launched) when n_samples
 * exceeds BP_NOALLOC_MAX_SAMPLES or when the network has no schedule yet
 * for the view's format and station scales (build it with
 * bp_network_prepare or any allocating call first).  Allocates no host or
 * device memory and creates no events; asynchronous on `stream`. */
#define BP_NOALLOC_MAX_SAMPLES 4194304
int bp_network_beamform_max_noalloc(bp_network *net, const bp_feature_view *features,
                                    int out_of_bounds, float *beam_max, int *beam_argmax,
                                    void *stream);

/* reduce='none' on a feature view (any format) for the sources
 * [source_begin, source_end): beam DEVICE (n, n_samples) float32; cells
 * invalid in strict mode are left untouched. */
int bp_network_beamform_view(bp_network *net, const bp_feature_view *features,
                             int out_of_bounds, size_t source_begin, size_t source_end,
                             float *beam, void *stream);

/* Differential mode on a resident network (same rule as
 * beamform_differential): beam DEVICE (n_sources) float32, entries of
 * sources out of range at the centre sample of the view are left
 * untouched.  Any storage format.  Asynchronous on `stream`. */
int bp_network_beamform_differential(bp_network *net, const bp_feature_view *features,
                                     float *beam, void *stream);

/* Device bytes of a network's differential tables, built on the first
 * differential call (0 before): delays (int32 and/or paired int16) and
 * weights.  weight_form: 0 not built, 1 float32 weights, 2 used bits with
 * one weight value (every nonzero weight of the network has the same bits;
 * 1-4 phases).  Any output pointer may be NULL.  Extension, no reference
 * counterpart. */
int bp_network_diff_table_bytes(const bp_network *net, size_t *delay_bytes,
                                size_t *weight_bytes, int *weight_form);

/* Host-only plan of a call (no counterpart in the reference; no device is
 * touched): what the engine would allocate and how it would tile a window,
 * so that admission and kernel-fit decisions can be made before any
 * upload.  Request: the shape, the feature storage format, the reduction
 * (0 'max', 1 'none'), an HBM budget in bytes (0: not checked) and
 * optionally the moveouts and weights (HOST, the layouts of beamform_max)
 * with a source order (BP_ORDER_*, as bp_network_create_ordered).  With
 * them the plan is the one bp_network_create_ordered would make: tile
 * classes, LDS windows, row padding, source order; without them those
 * fields are -1 and the sizes assume every tile in the LDS kernel.
 * Result:
 *   source_tile, time_tile      kernel shape of this build
 *   lds_window_limit            largest LDS window (samples) for n_phases,
 *                               0 when the LDS kernel does not take n_phases
 *   lds_bytes[2], workgroups_per_cu[2]  per LDS class (small, large window)
 *   network_bytes               resident moveouts, weights, bounds, tile
 *                               tables and one schedule of the format
 *   feature_bytes               the window as the engine stores it (rows
 *                               padded for the network, in the format)
 *   output_bytes                'max': beam_max, argmax and key scratch;
 *                               'none': the (n_sources, n_samples) beams
 *   call_bytes                  what a host-pointer call (beamform /
 *                               beamform_max) allocates in all: the above
 *                               plus the interleaved upload of the features
 *   fits                        call_bytes <= hbm_budget (1 without budget) */
typedef struct {
  size_t n_sources, n_stations, n_phases, n_samples;
  int fmt;
  int reduce;
  size_t hbm_budget;
  const int *time_delays;
  const float *weights_sources;
  int order;
} bp_plan_request;
typedef struct {
  int source_tile, time_tile;
  int lds_window_limit;
  int tiles_per_class[3];
  int window_per_class[2];
  int similarity_order;
  long long pad_before, pad_after;
  int lds_bytes[2];
  int workgroups_per_cu[2];
  size_t network_bytes, feature_bytes, output_bytes, call_bytes;
  int fits;
} bp_plan_result;
int bp_plan(const bp_plan_request *request, bp_plan_result *result);

/* Zero padding (samples) a feature view of this network needs before and
 * after every row so that no LDS tile takes the guarded edge path. */
int bp_network_padding(const bp_network *net, size_t *pad_before, size_t *pad_after);

/* Planar float32 features (n_stations, n_phases, n_samples) on the device ->
 * a feature view's buffer `out` (DEVICE, rows of row_stride elements: S*P
 * rows for BP_F32, S*P*2 for the 16-bit formats) with sample 0 at element
 * pad_before and zeros elsewhere.  16-bit formats choose per-station
 * power-of-two scales as bp_quantize_planar does and write their weight
 * multipliers to station_scale (HOST, n_stations): BP_F16 measures the
 * stations (a host round trip) and returns after the packing; BP_BF16
 * (scale 1) and BP_F32 (station_scale ignored) are asynchronous. */
int bp_pack_features(const float *F_spt, size_t n_stations, size_t n_phases, size_t n_samples,
                     int fmt, size_t pad_before, size_t row_stride, float *station_scale,
                     void *out, void *stream);

/* reduce='none' on device-resident planar f32 features for the sources
 * [source_begin, source_end) of the network.  beam: DEVICE (n, n_samples);
 * invalid cells are left untouched. */
int bp_network_beamform(bp_network *net, const float *features, size_t n_samples,
                        int out_of_bounds, size_t source_begin, size_t source_end,
                        float *beam, void *stream);

/* 16-bit feature storage.  bp_parity_stride(T) is the row stride (elements)
 * of the parity planes.  bp_quantize_planar reads planar float32 features
 * (n_stations, n_phases, n_samples) on the device, chooses per-station
 * power-of-two scales (BP_F16: the station's max |F| maps into
 * [2^14, 2^15); BP_BF16: 1), writes the scaled, rounded parity planes to
 * `out` (DEVICE, 2 * n_stations * n_phases * bp_parity_stride(T) uint16) and
 * the weight multipliers 1/scale to station_scale (HOST, n_stations).
 * Non-finite features saturate to the format's largest finite value (NaN to
 * 0).  Synchronous. */
size_t bp_parity_stride(size_t n_samples);
int bp_quantize_planar(const float *F_spt, size_t n_stations, size_t n_phases,
                       size_t n_samples, int fmt, float *station_scale, void *out,
                       void *stream);

/* Layout / prestack kernels on device pointers (asynchronous on `stream`). */
int bp_interleaved_to_planar(const float *F_stp, size_t n_stations, size_t n_samples,
                             size_t n_phases, float *F_spt, void *stream);
int bp_prestack_planar(const float *u_sct, const float *weights_phases,
                       size_t n_stations, size_t n_channels, size_t n_samples,
                       size_t n_phases, float *F_spt, void *stream);

/* Batched reduce='max' (no counterpart in the reference, which takes one
 * window per call): n_windows windows, each with its own network, feature
 * view, bounds mode and DEVICE outputs (beam_max, beam_argmax and keys
 * scratch of n_samples elements, as for bp_network_beamform_max), all in one
 * feature format and with one number of phases.  By default the LDS-class
 * tiles of all windows run in one persistent launch per window class, three
 * work-groups per CU taking work items from a queue (see bp_batch_config);
 * results are bit-identical to bp_network_beamform_max_view on each window
 * alone.
 * Windows of one network may have different station scales.  Creation
 * builds missing schedules synchronously, and the batch holds the schedules
 * of its windows: later calls or batches of the same networks with other
 * station scales do not change its results
 * and uploads the window table; the networks, features and outputs must
 * outlive the batch.  bp_batch_launch enqueues the batch on `stream`
 * (no allocation, no host synchronisation); bp_batch_graph_launch captures
 * that sequence into a hipGraph on first use and replays the graph after.
 * bp_batch_tune times every configuration of the search space (time tiles
 * per block-order group 1/2/4/8, staging buffers 2/3, and the five grid
 * policies of blocks_per_cu -1..3) on this batch, keeps the fastest and
 * remembers it for batches of the same shape in this process. */
typedef struct bp_batch bp_batch;
typedef struct {
  int time_group;     /* time tiles per block-order group (default 8; 1 = time-major) */
  int stage_buffers;  /* 2 or 3 (3 only where two work-groups still fit a CU) */
  int blocks_per_cu;  /* 1-4 (default 3): all windows in one persistent launch of that many
                         work-groups per CU (fewer where their LDS does not fit), which
                         take work items from a queue, their XCD's first; 0: all fused,
                         one work-group per work item; -1: windows of >= 4096 work items
                         in launches of their own, the others fused with one work-group
                         per work item */
  float ms;           /* measured time of the whole batch (tuning output) */
} bp_batch_config;
int bp_batch_create(int n_windows, bp_network *const *networks,
                    const bp_feature_view *features, const int *out_of_bounds,
                    float *const *beam_max, int *const *beam_argmax,
                    unsigned long long *const *keys, bp_batch **out);
int bp_batch_launch(bp_batch *batch, void *stream);
int bp_batch_graph_launch(bp_batch *batch, void *stream);
int bp_batch_tune(bp_batch *batch, int reps, bp_batch_config *tried, int max_tried,
                  int *n_tried, bp_batch_config *best, void *stream);
int bp_batch_get_config(const bp_batch *batch, bp_batch_config *config);
int bp_batch_set_config(bp_batch *batch, const bp_batch_config *config);
int bp_batch_destroy(bp_batch *batch);

/* Build the network's schedule for a feature format now (otherwise built
 * on first use); station_scale as for bp_network_beamform_max.
 * Synchronous. */
int bp_network_prepare(bp_network *net, int fmt, const float *station_scale);

/* Serving form (no counterpart in the reference): resident networks,
 * admission by a device-memory budget and continuous batching of window
 * requests (reduce='max').  A request is one window of one network:
 * waveform features (n_stations, n_channels, n_samples) float32 in HOST
 * memory, prestacked on the device with the network's phase weights, stored
 * in `fmt` rows padded for the network, beamformed in a batched launch with
 * whatever other requests are ready, and its beam_max / beam_argmax (HOST,
 * n_samples; beam_max 0 where no source is valid) written back.  Requests
 * are admitted in arrival order while their device memory fits
 * `hbm_budget`.  Whenever one of the `in_flight` batch slots (each its own
 * stream) is free, the ready requests are taken in earliest-deadline order
 * (deadline = ready time + 4 x the request's estimated GPU time, from its
 * sources x stations x phases x samples), up to max_batch windows and about
 * 10 ms of estimated work per batch (at least one request), so short
 * windows do not wait behind long ones.  Batch tables live in memory
 * allocated with the server: forming and retiring batches allocates and
 * frees no device memory.  A result does not depend on the requests it is
 * batched with.  The caller keeps the
 * feature and output buffers alive until bp_server_wait returns for the
 * request.  Thread-safe; bp_server_last_error() has this thread's message. */
typedef struct bp_server bp_server;
typedef struct {
  size_t hbm_budget;  /* bytes of device memory for admitted windows */
  int max_batch;      /* most windows per batched launch */
  int fmt;            /* BP_BF16 or BP_F32 storage of admitted windows */
  int in_flight;      /* batches on the device at once (1..4) */
} bp_server_config;
typedef struct {
  int status;         /* BP_OK or the error of this request */
  int batch_windows;  /* windows in the launch that served it */
  double submit_s, admit_s, ready_s, launch_s, gpu_done_s, done_s;  /* steady clock */
} bp_request_info;
typedef struct {
  long long pending, ready, in_flight_batches, completed, batches, batched_windows;
  size_t arena_bytes, arena_used, arena_peak;
} bp_server_counters;
int bp_server_create(const bp_server_config *config, bp_server **out);
int bp_server_add_network(bp_server *server, const int *time_delays,
                          const float *weights_sources, const float *weights_phases,
                          size_t n_sources, size_t n_stations, size_t n_channels,
                          size_t n_phases, int *network_id);
int bp_server_submit(bp_server *server, int network_id, const float *waveform_features,
                     size_t n_samples, int out_of_bounds, float *beam_max, int *beam_argmax,
                     long long *request_id);
/* Waits for the request (timeout_s < 0: no limit); returns its status, or 1
 * if the timeout passed first (the request stays pending). */
int bp_server_wait(bp_server *server, long long request_id, double timeout_s,
                   bp_request_info *info);
int bp_server_stats(bp_server *server, bp_server_counters *stats);
int bp_server_destroy(bp_server *server);
const char *bp_server_last_error(void);

/* Detection consumer (no counterpart in the library; the tutorial's step
 * after the beamforming, notebooks/3_beampower.ipynb:497-626).
 * bp_detect_threshold: time-dependent threshold of beam_max (DEVICE,
 *   n_samples) into threshold (DEVICE, n_samples): per window of `window`
 *   samples starting every int((1 - overlap) * window) samples, the mean
 *   plus num_dev population standard deviations (the tutorial's "rms"
 *   flavour; exact zeros are left out of the statistics where the tutorial
 *   fills them with unseeded noise), the maximum with each neighbouring
 *   window, each sample taking its window's value.  Needs n_samples > window.
 *   Synchronous.
 * bp_detect_peaks: local maxima of beam_max (plateaus at their midpoint) at
 *   least `distance` samples apart, higher peaks first (equal heights: the
 *   later one), as scipy.signal.find_peaks(distance=...); with threshold
 *   (DEVICE, or NULL) only those strictly above it.  Writes up to `capacity`
 *   ascending sample indices to peaks (HOST) and the total count to
 *   n_peaks.  Synchronous. */
int bp_detect_threshold(const float *beam_max, size_t n_samples, size_t window, double overlap,
                        double num_dev, float *threshold, void *stream);
/* The threshold in either of the tutorial's flavours
 * (notebooks/3_beampower.ipynb:497-578, threshold_type "rms" or "mad"):
 * BP_THRESHOLD_RMS is bp_detect_threshold; BP_THRESHOLD_MAD takes each
 * window's median as center and the median of the absolute deviations from
 * it as deviation (numpy's median in float32: the mean of the two middle
 * values for an even count), exact zeros again left out.  Synchronous. */
#define BP_THRESHOLD_RMS 0
#define BP_THRESHOLD_MAD 1
int bp_detect_threshold_type(const float *beam_max, size_t n_samples, size_t window,
                             double overlap, double num_dev, int threshold_type,
                             float *threshold, void *stream);
int bp_detect_peaks(const float *beam_max, const float *threshold, size_t n_samples,
                    size_t distance, int *peaks, size_t capacity, size_t *n_peaks,
                    void *stream);

/* Kernel timing: when enabled, the dominant (stack) kernel of every
 * bp_network_* call is bracketed with HIP events on its own stream and the
 * elapsed milliseconds are accumulated per thread. */
int bp_profiling(int enable);
int bp_profiling_read(double *stack_ms_total, long long *stack_launches);

#ifdef __cplusplus
}
#endif
#endif
