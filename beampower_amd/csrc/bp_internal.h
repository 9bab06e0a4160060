// bp_internal.h -- interfaces shared by the library's own translation units
// (beamform_gpu.hip, bp_server.hip); not part of the C ABI and not exported.
#pragma once
#include <hip/hip_runtime.h>

#include "beamform_gpu.h"

#define BP_INTERNAL __attribute__((visibility("hidden")))

// Where a batch's device tables go: `ws_bytes` of device memory at `ws`
// that the caller owns and keeps until the batch is destroyed, uploaded
// asynchronously on `stream` (launch the batch on the same stream).  Such a
// batch allocates no device memory and synchronises nothing, so creating
// and destroying it cannot stall work on other streams.  blocks_per_cu is
// the batch's grid policy (bp_batch_config), unless a tuned configuration
// for its shape exists.
struct bp_batch_placement {
  void *ws;
  size_t ws_bytes;
  hipStream_t stream;
  int blocks_per_cu;
};

// Device bytes the tables of a batch of n_windows windows need at most.
BP_INTERNAL size_t bp_batch_table_bytes(int n_windows);

// bp_batch_create with the tables placed as `pl` says.  The windows'
// networks must already hold their schedules for the batch's format and
// station scales (bp_network_prepare): the schedule lookup then allocates
// nothing.
BP_INTERNAL int bp_batch_create_placed(int n_windows, bp_network *const *networks,
                                       const bp_feature_view *features, const int *out_of_bounds,
                                       float *const *beam_max, int *const *beam_argmax,
                                       unsigned long long *const *keys,
                                       const bp_batch_placement &pl, bp_batch **out);
