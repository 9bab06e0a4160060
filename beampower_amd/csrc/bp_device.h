// bp_device.h -- device-side helpers shared by the beamforming kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <limits.h>

namespace bp {

constexpr int kWave = 64;

// Order-preserving map of a float onto uint32 (larger float -> larger key).
// Beams reaching the reduction are never -0.0 (an FP32 sum started at +0.0
// cannot become -0.0 under round-to-nearest), so no zero canonicalisation
// is needed.
__device__ __forceinline__ uint32_t float_order(float v) {
  uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float float_unorder(uint32_t o) {
  uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __uint_as_float(b);
}

// 64-bit reduction key: value in the high word, inverted source index in the
// low word, so an unsigned max picks the largest value and, among equal
// values, the lowest source index -- the same winner as the reference's
// ascending scan with a strict '>' (beamform.c:320-331).  Key 0 means
// "no candidate" (candidates satisfy v > -FLT_MAX, whose order word is
// >= 0x00800000).
__device__ __forceinline__ unsigned long long pack_key(float v, int k) {
  return ((unsigned long long)float_order(v) << 32) |
         (unsigned long long)(0xffffffffu - (uint32_t)k);
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
