// bp_device.h -- device-side helpers shared by the beamforming kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <limits.h>

namespace bp {

constexpr int kWave = 64;

// Order-preserving map of a float onto uint32 (larger float -> larger key).
// -0.0 maps to the key of +0.0: the two compare equal under the
// reference's '>' (beamform.c:326), so a tie between them must go to the
// lower source index like any other tie.  (A beam can be -0.0: an
// underflowing negative product, fmaf(-1e-30f, 1e-30f, +0.0f), gives it.)
__device__ __forceinline__ uint32_t float_order(float v) {
  uint32_t b = __float_as_uint(v);
  b = b == 0x80000000u ? 0u : b;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float float_unorder(uint32_t o) {
  uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __uint_as_float(b);
}

// 64-bit reduction key: value order in the high word; in the low word the
// inverted source index (bits 31..1) and whether the value is -0.0 (bit 0).
// An unsigned max picks the largest value and, among equal values (+0.0 and
// -0.0 included), the lowest source index -- the same winner as the
// reference's ascending scan with a strict '>' (beamform.c:320-331) -- and
// bit 0 then restores the winner's own zero sign.  Source indices are below
// 2^31.  Key 0 means "no candidate" (candidates satisfy v > -FLT_MAX, whose
// order word is >= 0x00800000).
__device__ __forceinline__ unsigned long long pack_key(float v, int k) {
  const uint32_t negz = __float_as_uint(v) == 0x80000000u ? 1u : 0u;
  return ((unsigned long long)float_order(v) << 32) |
         (unsigned long long)(((0x7fffffffu - (uint32_t)k) << 1) | negz);
}

__device__ __forceinline__ float key_value(unsigned long long key) {
  return (key & 1ull) ? -0.0f : float_unorder((uint32_t)(key >> 32));
}

__device__ __forceinline__ int key_index(unsigned long long key) {
  return (int)(0x7fffffffu - ((uint32_t)key >> 1));
}

__device__ __forceinline__ int uniform(int x) {
  return __builtin_amdgcn_readfirstlane(x);
}

// Valid time range [vlo, vhi) of a source in strict mode: t + lo >= 0 and
// t + hi < T (beamform.c:196-201); empty when no station is used.
__device__ __forceinline__ void strict_range(int lo, int hi, int T, int &vlo, int &vhi) {
  if (lo > hi) {  // sentinels INT_MAX / INT_MIN: no used station
    vlo = 1;
    vhi = 0;
    return;
  }
  long long a = -(long long)lo;
  long long b = (long long)T - (long long)hi;
  vlo = (int)(a < 0 ? 0 : (a > T ? T : a));
  vhi = (int)(b < 0 ? 0 : (b > T ? T : b));
}

}  // namespace bp
