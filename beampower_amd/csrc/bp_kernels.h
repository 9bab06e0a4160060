// bp_kernels.h -- gfx950 kernels of the beamforming hot path.
//
// Algorithm (reference: beampower/src/beamform.c):
//   F[s,p,t]  = sum_c alpha[s,c,p] * u[s,c,t]                 (prestack, :108-152)
//   b_k(t)    = sum_{s: beta[k,s]!=0} sum_p beta[k,s] * F[s,p,t+tau[k,s,p]]   (:49-105)
//   reduce=max: max over valid k, lowest k on ties            (:230-336)
// Every beam is accumulated as a chain of fused multiply-adds starting at
// +0.0f, stations ascending then phases ascending -- the exact sequence of
// the reference's pip/CMake build -- so results are bit-identical to it.
#pragma once
#include "bp_device.h"

namespace bp {

struct StackArgs {
  const void *F;          // features at sample 0 of row 0: float32 rows (S, P) or for
                          // the 16-bit formats parity-plane rows (S, P, 2), x[t] and x[t+1]
  long long ldf;          // feature row stride in elements
  int pad_lo, pad_hi;     // readable zero samples before 0 and from T on, in every row
  const float *scale;     // per-station weight multiplier of the 16-bit formats (or NULL)
  const int *tau;         // (K, S, P) moveouts
  const float *beta;      // (K, S) source weights
  const int *src_lohi;    // (K, 2) per-source min/max delay over used stations
  const int *perm;        // (K) original index of the source at each position, or NULL
                          // (identity); tau, beta, src_lohi and the tiles are in position order
  const int *tile_lo;     // (n_tiles, S, P) min delay over the tile's users of (s,p)
  const int *tile_all;    // (n_tiles, 8) min lo, max hi, max lo, max (hi - lo), non-finite weight flag
  const int *sched;       // (n_tiles, S, KT, P+1) LDS byte offsets and weight bits
  const int *tiles;       // source tiles processed by this launch
  int n_list;             // number of entries of `tiles`
  int T, K, S, P;
  int oob;                // 0 strict, 1 flexible
  int Lw;                 // LDS window length (samples), multiple of 8
  int n_time_tiles;
  int time_group;         // block order: time tiles per group (>= 1)
  int block_base;         // first block of this launch in the launch order (a class's blocks
                          // may need several launches: see run())
  int k_begin, k_end;     // sources written, by original index: [k_begin, k_end)
  unsigned long long *keys;  // reduce=max: (T) packed keys of (value, original index)
  float *beam;               // reduce=none: beam[(original index - k_begin) * T + t]
};

// ---------------------------------------------------------------------------
// Plan kernels (run once per network).
// ---------------------------------------------------------------------------

// beamform.c:10-47 -- per-source min/max over used stations (beta != 0).
// One wave per source (kBoundsSources per work-group): the lanes stride
// over the source's S*P delay entries, so every load instruction reads
// consecutive words, then a wave reduction.
constexpr int kBoundsSources = 4;
__global__ __launch_bounds__(64 * kBoundsSources) void source_bounds_kernel(
    const int *__restrict__ tau, const float *__restrict__ beta, int K, int S, int P,
    int *__restrict__ lohi) {
  const long long k = (long long)blockIdx.x * kBoundsSources + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (k >= K) return;  // whole wave
  const int SP = S * P;
  const int *tk = tau + k * SP;
  const float *bk = beta + k * S;
  int lo = INT_MAX, hi = INT_MIN;
  for (int j = lane; j < SP; j += 64) {
    if (bk[j / P] == 0.0f) continue;  // the station is not used
    const int d = tk[j];
    lo = d < lo ? d : lo;
    hi = d > hi ? d : hi;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  if (lane == 0) {
    lohi[2 * k] = lo;
    lohi[2 * k + 1] = hi;
  }
}

// Per source tile [tile*KT, tile*KT+KT): min delay of every (s,p) over the
// tile's sources that use station s (INT_MAX when none does).  tile_all
// (8 ints per tile) gets the tile's min lo, max hi, max lo, its largest
// per-(s,p) spread hi - lo, and 1 if some weight of the tile is infinite or
// NaN.  The tile's window of (s,p) for time tile t0 is [t0 + lo, t0 + TT - 1 + hi].
__global__ void tile_bounds_kernel(const int *__restrict__ tau,
                                   const float *__restrict__ beta, int K, int S,
                                   int P, int KT, int *__restrict__ tile_lo,
                                   int *__restrict__ tile_all) {
  __shared__ int s_lo, s_hi, s_maxlo, s_spread, s_nonfinite;
  const int tile = blockIdx.x;
  const int SP = S * P;
  if (threadIdx.x == 0) {
    s_lo = INT_MAX;
    s_hi = INT_MIN;
    s_maxlo = INT_MIN;
    s_spread = 0;
    s_nonfinite = 0;
  }
  __syncthreads();
  const int k0 = tile * KT;
  const int k1 = min(K, k0 + KT);
  for (int sp = threadIdx.x; sp < SP; sp += blockDim.x) {
    const int s = sp / P;
    int lo = INT_MAX, hi = INT_MIN;
    for (int k = k0; k < k1; ++k) {
      const float b = beta[(size_t)k * S + s];
      if (b == 0.0f) continue;
      if (!__builtin_isfinite(b)) s_nonfinite = 1;
      int d = tau[(size_t)k * SP + sp];
      lo = d < lo ? d : lo;
      hi = d > hi ? d : hi;
    }
    tile_lo[(size_t)tile * SP + sp] = lo;
    if (lo <= hi) {
      long long spread = (long long)hi - (long long)lo;
      atomicMin(&s_lo, lo);
      atomicMax(&s_hi, hi);
      atomicMax(&s_maxlo, lo);
      atomicMax(&s_spread, spread > INT_MAX ? INT_MAX : (int)spread);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int *o = tile_all + 8 * tile;
    // A tile none of whose sources uses a station reads no feature (every
    // weight is 0 and every gather reads the zero block), but the LDS kernel
    // still places its staging window at t0 + the tile's least delay: that
    // window goes to delay 0, where the staging addresses stay in the rows
    // (INT_MAX there would put the buffer base gigabytes past them).
    const bool unused = s_lo == INT_MAX;
    o[0] = unused ? 0 : s_lo;
    o[1] = unused ? 0 : s_hi;
    o[2] = unused ? 0 : s_maxlo;
    o[3] = s_spread;
    o[4] = s_nonfinite;
    o[5] = o[6] = o[7] = 0;
  }
}

// Round a float to the nearest bfloat16 (ties to even); returns its bits.
__host__ __device__ __forceinline__ unsigned bf16_bits(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

// LDS layout of a staged window of Lw samples (Lw a multiple of 8) of one
// (station, phase): SPA copies shifted by 0..SPA-1 samples (SPA = 4 for
// float32 features, 8 for 16-bit ones), back to back, each Lw/SPA 16-byte
// words long: word m of copy c holds the SPA samples starting at
// w0 + c + SPA*m, w0 = t0 + the staging start.  The Lw words of a phase are
// rounded up to whole 1 KiB rows (64 lanes x 16 bytes, one staging
// instruction per wave).
// Source order of a plan: out[i, :] = in[perm[i], :] over rows of D 32-bit
// words (delays, weights).
__global__ void gather_rows_kernel(const unsigned *__restrict__ in, const int *__restrict__ perm,
                                   long long K, int D, unsigned *__restrict__ out) {
  const long long n = K * D;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long k = i / D;
    out[i] = in[(long long)perm[k] * D + (i - k * D)];
  }
}

// Samples per lane and source of the LDS kernel (8 or 16): its time tile is
// 64 * BP_NACC samples.
#ifndef BP_NACC
#define BP_NACC 8
#endif
constexpr int kTimeTile = 64 * BP_NACC;
constexpr int kMaxWindow = 2048;  // largest LDS window (samples)
__host__ __device__ __forceinline__ int window_slot_bytes(int Lw) { return (Lw + 63) / 64 * 1024; }

// Per-tile schedule of the LDS kernel.  For tile j and station s, one row
// of KW*(P+1) words per group of KW source slots (one wave's sources,
// kk = g*KW + q, k = j*KT + kk): the KW*P LDS byte offsets in (source,
// phase) order, then the KW weight words.  The
// moveout d of phase p lies e = d - lo samples into the window (lo: the
// staging start, tile_lo[j,s,p] rounded down to even for 16-bit samples),
// so the SPA samples starting at (lane sample + e) sit 16-byte aligned in
// copy e % SPA, e / SPA words past the lane's own word.  The weight word is
// beta[k,s] * scale[s] as float32 bits, or for bfloat16 features its bf16
// rounding in the high half.  Sources past K and unused stations get weight
// -0.0 and the offset of the all-zero block that follows the phase slots, so
// the kernel needs no branch for them: fma(-0, +0, acc) == acc.  Tiles served
// by the direct kernel (Lw = 0) get offsets 0.
__global__ void schedule_kernel(const int *__restrict__ tau, const float *__restrict__ beta,
                                const int *__restrict__ tile_lo,
                                const int *__restrict__ tile_lw,
                                const float *__restrict__ scale, int K, int S, int P, int KT,
                                int KW, int fmt, long long n, int *__restrict__ sched) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int kk = (int)(idx % KT);
  const long long js = idx / KT;
  const int s = (int)(js % S);
  const int tile = (int)(js / S);
  const int k = tile * KT + kk;
  const int Lw = tile_lw[tile];
  const int spa = fmt == BP_F32 ? 4 : 8;
  const float b = k < K ? beta[(size_t)k * S + s] : 0.0f;
  const int q = kk % KW;
  int *row = sched + (idx - q) * (P + 1);  // the row of slots idx - q .. idx - q + KW - 1
  for (int p = 0; p < P; ++p) {
    int off = 0;
    if (Lw > 0) {
      const int SB = window_slot_bytes(Lw);
      off = P * SB;
      if (b != 0.0f) {
        int lo = tile_lo[((size_t)tile * S + s) * P + p];
        lo = fmt == BP_F32 ? lo : (lo & ~1);
        const int e = tau[((size_t)k * S + s) * P + p] - lo;
        off = p * SB + 16 * ((e % spa) * (Lw / spa) + e / spa);
      }
    }
    row[q * P + p] = off;
  }
  // An unused term gets weight -0.0: fma(-0, +0, acc) is acc bit for bit,
  // a zero acc keeping its sign, as when the reference skips the station
  // (beamform.c:61-62); with weight +0.0 an acc of -0.0 would become +0.0.
  const float w = b != 0.0f ? b * (scale ? scale[s] : 1.0f) : -0.0f;
  row[KW * P + q] = fmt == BP_BF16 ? (int)(bf16_bits(w) << 16) : __float_as_int(w);
}

// ---------------------------------------------------------------------------
// Reduced-precision feature storage.
// ---------------------------------------------------------------------------

// Per-station max |F| over phases and time (float bits are order-preserving
// for non-negative values, so an integer max of |x| bits is the max).
__global__ void station_maxabs_kernel(const float *__restrict__ F, int P, long long T,
                                      unsigned *__restrict__ maxbits) {
  const int s = blockIdx.y;
  unsigned m = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)P * T;
       i += (long long)gridDim.x * blockDim.x) {
    const unsigned u = __float_as_uint(F[(size_t)s * P * T + i]) & 0x7fffffffu;
    if (u <= 0x7f800000u) m = u > m ? u : m;  // NaN excluded
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned v = __shfl_xor(m, o);
    m = v > m ? v : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(maxbits + s, m);
}

// Planar float32 features (S*P rows of T) -> rows of `ldout` elements in
// the storage format, sample 0 at element pad_out, zeros outside [0, T):
//   F32   : row (s,p) holds x[t];
//   16-bit: parity planes, row (s,p,0) holds q(x[t] * scale[s]) and row
//           (s,p,1) q(x[t+1] * scale[s]).
// Non-finite inputs saturate to the format's largest finite value (NaN to 0).
__device__ __forceinline__ unsigned short quantize16(float x, int fmt) {
  if (x != x) x = 0.0f;
  if (fmt == BP_F16) {
    x = fminf(fmaxf(x, -65504.0f), 65504.0f);
    return __builtin_bit_cast(unsigned short, (_Float16)x);
  }
  x = fminf(fmaxf(x, -3.3895313892515355e38f), 3.3895313892515355e38f);
  return (unsigned short)bf16_bits(x);
}

__global__ void pack_features_kernel(const float *__restrict__ F, int P, long long T,
                                     const float *__restrict__ scale, int fmt, long long ldout,
                                     long long pad_out, void *__restrict__ out) {
  const int sp = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ldout) return;
  const long long t = i - pad_out;
  const float *row = F + (size_t)sp * T;
  if (fmt == BP_F32) {
    static_cast<float *>(out)[(size_t)sp * ldout + i] = (t >= 0 && t < T) ? row[t] : 0.0f;
    return;
  }
  const float sc = scale ? scale[sp / P] : 1.0f;
  unsigned short *o = static_cast<unsigned short *>(out);
  for (int par = 0; par < 2; ++par) {
    const long long tt = t + par;
    const float x = (tt >= 0 && tt < T) ? row[tt] * sc : 0.0f;
    o[((size_t)sp * 2 + par) * ldout + i] = quantize16(x, fmt);
  }
}

// ---------------------------------------------------------------------------
// Detection consumer (notebooks/3_beampower.ipynb:497-626): time-dependent
// threshold and local maxima of beam_max.
// ---------------------------------------------------------------------------

// Mean and population standard deviation of x over the window
// [j*shift, j*shift + window) of block j, exact zeros excluded, two passes
// accumulated in double with a fixed reduction tree (deterministic).
__global__ __launch_bounds__(256) void window_stats_kernel(const float *__restrict__ x,
                                                           long long window, long long shift,
                                                           double *__restrict__ stats) {
  __shared__ double s_sum[256], s_cnt[256];
  const float *w = x + (long long)blockIdx.x * shift;
  const int tid = threadIdx.x;
  double s = 0.0, c = 0.0;
  for (long long i = tid; i < window; i += 256) {
    const float v = w[i];
    if (v != 0.0f) {
      s += v;
      c += 1.0;
    }
  }
  s_sum[tid] = s;
  s_cnt[tid] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) {
      s_sum[tid] += s_sum[tid + o];
      s_cnt[tid] += s_cnt[tid + o];
    }
    __syncthreads();
  }
  const double n = s_cnt[0];
  const double mean = n > 0.0 ? s_sum[0] / n : 0.0;
  __syncthreads();
  double q = 0.0;
  for (long long i = tid; i < window; i += 256) {
    const float v = w[i];
    if (v != 0.0f) q += ((double)v - mean) * ((double)v - mean);
  }
  s_sum[tid] = q;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) s_sum[tid] += s_sum[tid + o];
    __syncthreads();
  }
  if (tid == 0) {
    stats[2 * blockIdx.x] = mean;
    stats[2 * blockIdx.x + 1] = n > 0.0 ? sqrt(s_sum[0] / n) : 0.0;
  }
}

// "mad" flavour: per window j (samples [j*shift, j*shift + window)), the
// median of the non-zero samples and the median of their absolute
// deviations from it, as numpy's median computes them in float32: the middle
// value of n, or for even n the float32 mean of the two middle values
// ((a + b) / 2).  stats[2j] = center, stats[2j+1] = deviation (0 when the
// window holds no non-zero sample).  Each median is an exact radix
// selection over order-preserving keys: four passes of 8-bit digits, a
// 256-bin histogram in LDS per pass.
struct MedianScratch {
  unsigned hist[256];
  unsigned digit, rank;
  unsigned count;
};

// Key of sample v for the selection: its value (mode 0) or |v - center|
// (mode 1); exact zeros take no part (they return false).
__device__ __forceinline__ bool median_key(float v, int mode, float center, unsigned &key) {
  if (v == 0.0f) return false;
  key = float_order(mode ? fabsf(v - center) : v);
  return true;
}

// The rank-th smallest key (0-based) among the window's selected samples.
__device__ unsigned select_rank(const float *__restrict__ w, long long n, unsigned rank,
                                int mode, float center, MedianScratch &sm) {
  const int tid = threadIdx.x;
  unsigned prefix = 0u, mask = 0u;
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (int b = tid; b < 256; b += blockDim.x) sm.hist[b] = 0u;
    __syncthreads();
    for (long long i = tid; i < n; i += blockDim.x) {
      unsigned key;
      if (median_key(w[i], mode, center, key) && (key & mask) == prefix)
        atomicAdd(&sm.hist[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      unsigned r = rank, b = 0;
      while (b < 255u && sm.hist[b] <= r) r -= sm.hist[b++];
      sm.digit = b;
      sm.rank = r;
    }
    __syncthreads();
    prefix |= sm.digit << shift;
    mask |= 255u << shift;
    rank = sm.rank;
    __syncthreads();  // the histogram is cleared next pass
  }
  return prefix;
}

__device__ float median_of(const float *__restrict__ w, long long n, unsigned count, int mode,
                           float center, MedianScratch &sm) {
  const float a = float_unorder(select_rank(w, n, (count - 1) / 2, mode, center, sm));
  if (count & 1u) return a;
  const float b = float_unorder(select_rank(w, n, count / 2, mode, center, sm));
  return (a + b) / 2.0f;
}

__global__ __launch_bounds__(256) void window_median_kernel(const float *__restrict__ x,
                                                            long long window, long long shift,
                                                            float *__restrict__ stats) {
  __shared__ MedianScratch sm;
  const float *w = x + (long long)blockIdx.x * shift;
  if (threadIdx.x == 0) sm.count = 0u;
  __syncthreads();
  unsigned c = 0;
  for (long long i = threadIdx.x; i < window; i += blockDim.x) c += w[i] != 0.0f;
  atomicAdd(&sm.count, c);
  __syncthreads();
  const unsigned count = sm.count;
  float center = 0.0f, deviation = 0.0f;
  if (count > 0) {
    center = median_of(w, window, count, 0, 0.0f, sm);
    deviation = median_of(w, window, count, 1, center, sm);
  }
  if (threadIdx.x == 0) {
    stats[2 * blockIdx.x] = center;
    stats[2 * blockIdx.x + 1] = deviation;
  }
}

// Per-window thresholds -> every sample: sample t takes window t / shift
// (clamped to the last), samples before window/2 take the value of sample
// window/2 and samples from n - (window - window/2) on the value of the
// sample before.
__global__ void expand_threshold_kernel(const float *__restrict__ thr, int nw, long long n,
                                        long long window, long long shift,
                                        float *__restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long half = window / 2, end = n - (window - half);
  const long long tt = t < half ? half : (t >= end ? end - 1 : t);
  long long j = tt / shift;
  if (j >= nw) j = nw - 1;
  out[t] = thr[j];
}

// Local maxima of x, plateaus at their midpoint (the left one of two):
// the thread at the left edge of a rise scans the plateau and flags its
// midpoint when the series falls after it.  The first and last samples are
// never maxima.
__global__ void local_max_kernel(const float *__restrict__ x, long long n,
                                 unsigned char *__restrict__ flag) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i >= n - 1) return;
  const float xi = x[i];
  if (!(x[i - 1] < xi)) return;
  long long j = i + 1;
  while (j < n - 1 && x[j] == xi) ++j;
  if (x[j] < xi) flag[(i + j - 1) / 2] = 1;
}

// ---------------------------------------------------------------------------
// Layout and prestack.
// ---------------------------------------------------------------------------

// (S,T,P) phase-interleaved -> planar rows (S,P) of `ld` elements, sample 0
// at element `pad`, zeros outside [0, T).
__global__ void interleaved_to_planar_kernel(const float *__restrict__ in, int S, int T,
                                             int P, long long ld, long long pad,
                                             float *__restrict__ out) {
  const int s = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ld) return;
  const long long t = i - pad;
  const bool in_range = t >= 0 && t < T;
  const float *src = in + ((size_t)s * T + (in_range ? t : 0)) * P;
  for (int p = 0; p < P; ++p) out[((size_t)s * P + p) * ld + i] = in_range ? src[p] : 0.0f;
}

// beamform.c:108-152: F[s,.,p] = sum_c alpha[s,c,p] * u[s,c,t], accumulated
// from +0.0f in ascending c with fused multiply-adds.  planar selects the
// output layout (S,P,T) instead of the reference's (S,T,P).
__global__ void prestack_kernel(const float *__restrict__ u,
                                const float *__restrict__ alpha, int S, int C, int T,
                                int P, int planar, float *__restrict__ out) {
  const int s = blockIdx.y;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  for (int p = 0; p < P; ++p) {
    float acc = 0.0f;
    for (int c = 0; c < C; ++c)
      acc = fmaf(alpha[((size_t)s * C + c) * P + p], u[((size_t)s * C + c) * T + t], acc);
    if (planar)
      out[((size_t)s * P + p) * T + t] = acc;
    else
      out[((size_t)s * T + t) * P + p] = acc;
  }
}

// The same prestack written straight into planar rows of stride ld that
// hold sample t at t + pad, with zeros in the padding (the rows a network's
// LDS tiles stage without edge guards).
__global__ void prestack_rows_kernel(const float *__restrict__ u,
                                     const float *__restrict__ alpha, int S, int C, int T,
                                     int P, long long ld, long long pad,
                                     float *__restrict__ out) {
  const int s = blockIdx.y;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ld) return;
  const long long t = i - pad;
  const bool in_range = t >= 0 && t < T;
  for (int p = 0; p < P; ++p) {
    float acc = 0.0f;
    if (in_range)
      for (int c = 0; c < C; ++c)
        acc = fmaf(alpha[((size_t)s * C + c) * P + p], u[((size_t)s * C + c) * T + t], acc);
    out[((size_t)s * P + p) * ld + i] = acc;
  }
}

// ---------------------------------------------------------------------------
// Reduction helpers.
// ---------------------------------------------------------------------------

__global__ void fill_u64_kernel(unsigned long long *__restrict__ p, long long n,
                                unsigned long long v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Grid-stride: an output of K x T floats can hold more elements than a
// dispatch has work-items (2^32).
__global__ void fill_f32_kernel(float *__restrict__ p, long long n, float v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// keys -> (beam_max, beam_argmax): beam_max written only where a candidate
// exists, argmax always written, 0 when none (beamform.c:332-334).
__global__ void finalize_max_kernel(const unsigned long long *__restrict__ keys, int T,
                                    float *__restrict__ beam_max,
                                    int *__restrict__ beam_argmax) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  unsigned long long key = keys[t];
  if (key) {
    beam_max[t] = key_value(key);
    beam_argmax[t] = key_index(key);
  } else {
    beam_argmax[t] = 0;
  }
}

// ---------------------------------------------------------------------------
// Shift-and-stack, LDS-staged (the hot kernel).
//
// A work-group of W waves covers one source tile (KT = W*KW sources; wave w
// owns KW of them) and one time tile of TT = 64 * NACC samples (512 by
// default: NACC = 8 accumulators per lane and source).  A feature sample
// is ES bytes (4 for float32, 2 for the 16-bit storage formats), so one
// 16-byte LDS access holds SPA = 16/ES consecutive samples.  Lane l owns
// R = NACC/SPA runs of SPA consecutive samples, starting at
// t0 + SPA*l + 64*SPA*r, so a wave's gather of one (source, station, phase)
// term is R ds_read_b128 of 64 consecutive 16-byte words: conflict-free, and
// SPA samples per LDS lane-access, which is what the gfx950 LDS pipe needs
// to reach its bandwidth (it serves 16 lanes per clock at any width up to
// 16 bytes).
//
// Stations are visited one at a time (ascending, so every (k,t) FMA chain
// runs in ascending (s, p) order).  For each of the station's P phases the
// window [t0 + lo, t0 + lo + Lw + SPA - 1) of planar features is staged into
// LDS as SPA copies shifted by 0..SPA-1 samples, in a ring of NB buffers:
// the global->LDS copies of station s + NB - 1 are in flight while station
// s is consumed.  Positions outside [0, T) are staged as exact zeros.
// Accumulators acc[KW][NACC] stay in registers for the whole tile.  The
// per-(source, station) LDS offsets and weights come from the precomputed
// schedule as wave-uniform scalar loads.
// ---------------------------------------------------------------------------

// Block -> (time tile, listed tile) with XCD and L2 locality.  Work-groups
// are dispatched round-robin over the 8 XCDs, so block b runs on XCD b % 8;
// XCD x is given the contiguous range [x*N/8, (x+1)*N/8) of an order that
// walks groups of TG consecutive time tiles, all listed source tiles per
// group and the group's time tiles innermost.  The work-groups running
// together on an XCD then share TG time tiles, whose feature windows stay in
// that XCD's L2, and read each source tile's schedule once per TG time
// tiles instead of once per time tile.  TG = 8 at the 1 h, 192,000-source
// drop-in shape: 1.9 MB of windows per XCD, 5.3 GB of HBM reads per launch
// (41 GB with TG = 1, where the schedule stream evicts the windows).
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int n8 = n & ~7;
  if (b >= n8) return b;
  return (b & 7) * (n8 >> 3) + (b >> 3);
}

__device__ __forceinline__ void block_tile(int L, int n_time_tiles, int n_list, int TG, int &tt,
                                           int &li) {
  const int g = L / (TG * n_list);
  const int gsize = min(TG, n_time_tiles - g * TG);  // last group may be short
  const int rem = L - g * TG * n_list;
  li = rem / gsize;
  tt = g * TG + (rem - li * gsize);
}

// (source, phase) gather steps in flight for FP32 samples.  One keeps the
// registers within the 6-waves-per-SIMD budget below (two spill there,
// three ran 1.5 % slower at 4 waves per SIMD).
#ifndef BP_PIPE_DEPTH
#define BP_PIPE_DEPTH 1
#endif
// 16-bit samples: one 16-byte read per step, so two steps ahead fit the
// registers (1.3 % faster than one step ahead on the drop-in workload in
// BF16; two steps ahead made FP32 4 % slower).
#ifndef BP_PIPE_DEPTH_16
#define BP_PIPE_DEPTH_16 2
#endif
// Wave priority (s_setprio) over the first BP_PRIO_PCT % of a station's
// steps: its first reads and, at step BP_DMA_AT, a later station's copies.
// Waves that have just passed the station's barrier then issue ahead of
// the waves of the other work-groups of the CU that are finishing their
// stations, which keeps the work-groups out of step with one another.
// Priority 2 over the first 6 of 12 steps measured 3-4 % faster than
// none in every format on the drop-in workload (FP32 61.69 -> 59.66 ms);
// 3 of 12 steps gave nothing, 5-9 the same, and priority 1 or 3 about the
// same as 2; raising it only around the copies gave 0.7 %.
#ifndef BP_PRIO
#define BP_PRIO 2
#endif
#ifndef BP_PRIO_PCT
#define BP_PRIO_PCT 50
#endif
#ifndef BP_BATCH_PRIO
#define BP_BATCH_PRIO BP_PRIO
#endif
// Steps (of a station's KW*P gather steps) at which the copies of a later
// station's phase p are issued: BP_DMA_AT + p * BP_DMA_GAP.  All phases at
// step 4 measured 1.2 % faster than at step 0 (FP32, drop-in workload);
// spreading the phases over the station was slower.
#ifndef BP_DMA_AT
#define BP_DMA_AT 4
#endif
#ifndef BP_DMA_GAP
#define BP_DMA_GAP 0
#endif
// Where in a step's products the next step's LDS reads are issued, in
// 16-byte words of the step (FP32: 8 words, read 0 and read 1; 16-bit: 4
// words, one read): a wave then alternates products and reads instead of
// issuing its reads in one burst before its products.
// Measured on the drop-in workload (HIP events, builds interleaved in one
// process): FP32 reads before words 3 and 5 of 8 run 1-2 % faster than both
// reads before the products (0, 0); 16-bit samples are fastest with the
// read first.
#ifndef BP_GATHER_AT_F32_0
#define BP_GATHER_AT_F32_0 3
#endif
#ifndef BP_GATHER_AT_F32_1
#define BP_GATHER_AT_F32_1 5
#endif
#ifndef BP_GATHER_AT_16
#define BP_GATHER_AT_16 0
#endif

// Work-group barrier without the memory-model fence of __syncthreads(),
// which waits for every outstanding global->LDS copy of the wave (the
// fence makes those copies' LDS writes visible).  Callers order their own
// LDS traffic with explicit s_waitcnt; the empty asm statements keep the
// compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void barrier_no_fence() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// s_waitcnt vmcnt(n) (expcnt and lgkmcnt left alone) for a wave-uniform n
// in [0, 15]: the counter field is an immediate, so one branch per value.
__device__ __forceinline__ void wait_vmem(int n) {
  switch (n) {
#define BP_WAIT_CASE(k) \
  case k:               \
    __builtin_amdgcn_s_waitcnt(0x0F70 | k); \
    break;
    BP_WAIT_CASE(1) BP_WAIT_CASE(2) BP_WAIT_CASE(3) BP_WAIT_CASE(4) BP_WAIT_CASE(5)
    BP_WAIT_CASE(6) BP_WAIT_CASE(7) BP_WAIT_CASE(8) BP_WAIT_CASE(9) BP_WAIT_CASE(10)
    BP_WAIT_CASE(11) BP_WAIT_CASE(12) BP_WAIT_CASE(13) BP_WAIT_CASE(14) BP_WAIT_CASE(15)
#undef BP_WAIT_CASE
    default:
      __builtin_amdgcn_s_waitcnt(0x0F70);
      break;
  }
}

typedef __attribute__((address_space(1))) char global_char;

// Raw buffer descriptor (stride 0, dword data format) based at `base`,
// covering 2 GiB: staging windows lie within it by construction.  The base
// is wave-uniform; passing it through readfirstlane states so, which keeps
// the descriptor in scalar registers (a descriptor the compiler believes
// divergent costs a loop over lanes at every load).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t window_rsrc(const global_char *base) {
  const unsigned long long b = (unsigned long long)base;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __builtin_amdgcn_make_buffer_rsrc((void *)(((unsigned long long)hi << 32) | lo),
                                           (short)0, 0x7fffffff, 0x00020000);
}

typedef __attribute__((address_space(3))) void lds_void;
// Read-only per-network tables are read through the constant address space:
// wave-uniform loads from it are always scalar (s_load), whatever the
// compiler can prove about intervening memory operations.
typedef const __attribute__((address_space(4))) int const_int;

__device__ __forceinline__ const_int *as_const(const int *p) {
  return (const_int *)(p);
}

// Original index of the source at position k of a network's plan.
template <class A>
__device__ __forceinline__ int source_index(const A &a, int k) {
  return a.perm ? as_const(a.perm)[k] : k;
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// Storage formats of the staged features.  Each provides the element size
// and the accumulation of one 16-byte LDS word (SPA samples) with weight
// word `bw` from the schedule:
//   F32  : 4 samples, v_fma_f32 with the float weight;
//   F16  : 8 samples, v_fma_mix_f32 (f16 sample x f32 weight + f32 acc, one
//          rounding, no conversion instruction);
//   BF16 : 8 samples, v_dot2_f32_bf16 against (w, 0) and (0, w): the product
//          of two bf16 values is exact in f32, so each is one FMA with one
//          rounding; the weight word holds the bf16 weight in its high half.
template <int FMT>
struct Fmt;

template <>
struct Fmt<BP_F32> {
  static constexpr int ES = 4, SPA = 4;
  // word q (0..3) of a 16-byte LDS word: sample q
  __device__ __forceinline__ static void mac_word(float *acc, unsigned w, int q, int bw) {
    acc[q] = fmaf(__int_as_float(bw), __uint_as_float(w), acc[q]);
  }
  __device__ __forceinline__ static void mac(float *acc, const uint4 &v, int bw) {
    mac_word(acc, v.x, 0, bw);
    mac_word(acc, v.y, 1, bw);
    mac_word(acc, v.z, 2, bw);
    mac_word(acc, v.w, 3, bw);
  }
};

template <>
struct Fmt<BP_F16> {
  static constexpr int ES = 2, SPA = 8;
  __device__ __forceinline__ static float lo(unsigned w) {
    return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu));
  }
  __device__ __forceinline__ static float hi(unsigned w) {
    return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16));
  }
  // word q (0..3) of a 16-byte LDS word: samples 2q and 2q + 1
  __device__ __forceinline__ static void mac_word(float *acc, unsigned w, int q, int bw) {
    const float b = __int_as_float(bw);
    acc[2 * q] = fmaf(lo(w), b, acc[2 * q]);
    acc[2 * q + 1] = fmaf(hi(w), b, acc[2 * q + 1]);
  }
  __device__ __forceinline__ static void mac(float *acc, const uint4 &v, int bw) {
    mac_word(acc, v.x, 0, bw);
    mac_word(acc, v.y, 1, bw);
    mac_word(acc, v.z, 2, bw);
    mac_word(acc, v.w, 3, bw);
  }
};

template <>
struct Fmt<BP_BF16> {
  static constexpr int ES = 2, SPA = 8;
  // word q (0..3) of a 16-byte LDS word: samples 2q and 2q + 1
  __device__ __forceinline__ static void mac_word(float *acc, unsigned w, int q, int bw) {
    const bf16x2_t whi = __builtin_bit_cast(bf16x2_t, (unsigned)bw & 0xffff0000u);  // (0, w)
    const bf16x2_t wlo = __builtin_bit_cast(bf16x2_t, (unsigned)bw >> 16);          // (w, 0)
    const bf16x2_t x = __builtin_bit_cast(bf16x2_t, w);
    acc[2 * q] = __builtin_amdgcn_fdot2_f32_bf16(x, wlo, acc[2 * q], false);
    acc[2 * q + 1] = __builtin_amdgcn_fdot2_f32_bf16(x, whi, acc[2 * q + 1], false);
  }
  __device__ __forceinline__ static void mac(float *acc, const uint4 &v, int bw) {
    mac_word(acc, v.x, 0, bw);
    mac_word(acc, v.y, 1, bw);
    mac_word(acc, v.z, 2, bw);
    mac_word(acc, v.w, 3, bw);
  }
};

// The reference's flexible chain for one (source position k, sample t)
// (beamform.c:74-105): used stations ascending, phases ascending, terms
// outside [0, T) skipped, features read from HBM.  The LDS kernel instead
// accumulates an out-of-range term as fma(w, 0, acc); that equals acc except
// for the sign of an acc that is exactly zero (w * 0 is a signed zero), and
// a sign difference of a zero acc disappears at the next nonzero term.  So
// the two chains differ only in the sign of a beam that ends at zero, and
// such beams are recomputed here.  The weight and feature are those the LDS
// kernel applies (station scale; 16-bit samples from parity plane 0).
template <int FMT, int P, class A>
__device__ __forceinline__ float flexible_chain(const A &a, int k, int t) {
  const int S = a.S;
  const float *beta = a.beta + (size_t)k * S;
  const int *tau = a.tau + (size_t)k * S * P;
  float acc = 0.0f;
  for (int s = 0; s < S; ++s) {
    const float b0 = beta[s];
    if (b0 == 0.0f) continue;
    const float w = b0 * (a.scale ? a.scale[s] : 1.0f);
    for (int p = 0; p < P; ++p) {
      const long long tg = (long long)t + tau[s * P + p];
      if (tg < 0 || tg >= a.T) continue;
      const size_t row = (size_t)s * P + p;
      float x;
      if (FMT == BP_F32) {
        x = static_cast<const float *>(a.F)[row * a.ldf + tg];
      } else {
        const unsigned short h = static_cast<const unsigned short *>(a.F)[row * 2 * a.ldf + tg];
        x = (float)__builtin_bit_cast(_Float16, h);
      }
      acc = fmaf(w, x, acc);
    }
  }
  return acc;
}

template <int FMT, int P, int KW, int W, int NB, bool NONE, int PRIO = BP_PRIO>
struct StackLds {
  static_assert(P >= 1 && P <= 4, "the LDS kernel takes one to four phases");
  static_assert(NB == 2 || NB == 3, "two or three staging buffers");
  using F = Fmt<FMT>;
  static constexpr int ES = F::ES;
  static constexpr int SPA = F::SPA;
  static constexpr int NC = SPA;          // shifted copies of a window
  static constexpr int NACC = BP_NACC;    // accumulators per lane and source
  static_assert(NACC == 8 || NACC == 16, "8 or 16 samples per lane");
  static_assert(KW * NACC <= 64 && KW <= 8, "zero-beam masks: one bit per accumulator");
  static constexpr int R = NACC / SPA;    // 16-byte words per lane per term
  static constexpr int PL = ES == 4 ? 1 : 2;  // feature planes per (s, p) in HBM
  static constexpr int TT = kTimeTile;
  static constexpr int KT = KW * W;
  static constexpr int NT = 64 * W;
  static constexpr int ZERO_BYTES = 1024 * R;  // all-zero block read by unused terms

  // LDS of one buffer: P phase slots (window_slot_bytes each), then the
  // zero block.  NB buffers form a ring.
  __host__ __device__ static int buf_bytes(int Lw) {
    return P * window_slot_bytes(Lw) + ZERO_BYTES;
  }
  __host__ __device__ static int lds_bytes(int Lw) {
    return (NONE ? 0 : TT * 8) + NB * buf_bytes(Lw);
  }

  // Global->LDS staging of interior tiles (buffer_load_dwordx4 ... lds, no
  // registers).  A phase slot is ceil(Lw / 64) rows of 64 lanes x 16 bytes;
  // row j is slot words [64j, 64j + 64), and lane l moves word
  // u = 64j + l (clamped to the last word Lw - 1, so the slot padding gets
  // a harmless duplicate).  Wave w moves rows j = w, w + W, ... of every
  // phase.  A row's lane offsets from the window start do not depend on
  // the phase or the station, so they are computed once; a row then costs
  // one instruction with the station's descriptor, the phase's scalar
  // offset and the lane's offset.  For 16-bit samples the features are kept
  // twice in HBM, as x[t] and as x[t+1] ("parity planes"), and the window
  // start is even: copy c then reads plane (c & 1) at the even element
  // c - (c & 1) + SPA*m, so every row is 4-byte aligned.
  static constexpr int NRMAX = (kMaxWindow / 64 + W - 1) / W;  // rows per phase and wave

  struct RowPlan {
    int n;                 // rows of this wave per phase
    unsigned voff[NRMAX];  // this lane's byte offset from the window start, row wave + i*W
  };

  __device__ __forceinline__ static void plan_rows(int Lw, long long ldf, int wave, int lane,
                                                   RowPlan &rp) {
    const int nrp = (Lw + 63) / 64;
    const int nw = Lw / SPA;
    rp.n = (nrp - wave + W - 1) / W;
    // c = u / nw for u < Lw <= 2048 without an integer division:
    // (u + 1/2) / nw lies at least 1/(2 nw) >= 2^-11 away from an integer,
    // far beyond the float rounding of the reciprocal and the product
    const float inv_nw = __builtin_amdgcn_rcpf((float)nw);
#pragma unroll
    for (int i = 0; i < NRMAX; ++i) {
      const int j = wave + i * W;
      const int u = min(64 * j + lane, Lw - 1);
      const int c = (int)(((float)u + 0.5f) * inv_nw), m = u - c * nw;
      if (ES == 4) {
        rp.voff[i] = 4u * (unsigned)(c + SPA * m);
      } else {
        const int par = c & 1;
        rp.voff[i] = (unsigned)par * 2u * (unsigned)ldf + 2u * (unsigned)(c - par + SPA * m);
      }
    }
  }

  // Staging start of (s, p) relative to t0: the least delay of the tile's
  // users of the phase at the station (the tile's least delay when no
  // source uses the station, which keeps the copies in range), rounded down
  // to even for 16-bit samples.
  __device__ __forceinline__ static int stage_lo(const_int *tlo, int tall0, int s, int p) {
    int lo = tlo[s * P + p];
    lo = lo == INT_MAX ? tall0 : lo;
    return ES == 4 ? lo : (lo & ~1);
  }

  // A station's staging sources: a buffer descriptor based at plane 0 of
  // phase 0 at t0 + the tile's least staging start, and each phase's byte
  // offset from there (non-negative, below 2^31 as the host checks).
  struct StationSrc {
    __amdgpu_buffer_rsrc_t rs;
    unsigned soff[P];
  };

  template <class A>
  __device__ __forceinline__ static void station_src(const A &a, const_int *tlo,
                                                     int tall0, int s, int t0,
                                                     StationSrc &src) {
    const int lo0 = ES == 4 ? tall0 : (tall0 & ~1);
    src.rs = window_rsrc((const global_char *)(a.F) +
                         (((long long)s * P * PL) * a.ldf + t0 + lo0) * ES);
#pragma unroll
    for (int p = 0; p < P; ++p)
      src.soff[p] = __builtin_amdgcn_readfirstlane(
          (unsigned)(((long long)p * PL * a.ldf + stage_lo(tlo, tall0, s, p) - lo0) * ES));
  }

  // The copies are buffer loads to LDS (MUBUF): unlike the FLAT-encoded
  // global load to LDS, the compiler then keeps its waits for the gathers
  // exact (lgkmcnt(n)) while copies are in flight instead of draining all
  // outstanding LDS reads at every wait.
  __device__ __forceinline__ static void dma_phase(const StationSrc &src, const RowPlan &rp,
                                                   int SB, int wave, int p,
                                                   unsigned char *__restrict__ buf) {
#pragma unroll
    for (int i = 0; i < NRMAX; ++i) {
      if (i >= rp.n) break;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          src.rs, (lds_void *)(buf + p * SB + 1024 * (wave + i * W)), 16, rp.voff[i],
          src.soff[p], 0, 0);
    }
  }
  __device__ __forceinline__ static void dma_station(const StationSrc &src, const RowPlan &rp,
                                                     int SB, int wave,
                                                     unsigned char *__restrict__ buf) {
#pragma unroll
    for (int p = 0; p < P; ++p) dma_phase(src, rp, SB, wave, p, buf);
  }
  // Step of a station's gathers at which the copies of phase p of a later
  // station are issued: spread through the station rather than in one burst
  // before it.
  __host__ __device__ static constexpr int dma_step(int p) {
    return BP_DMA_AT + p * BP_DMA_GAP < STEPS ? BP_DMA_AT + p * BP_DMA_GAP : STEPS - 1;
  }

  // Edge tiles (a window crosses [0, T)): element-wise guarded copies
  // through registers, zeros outside [0, T).  Slot word u = c*(Lw/SPA) + m
  // of phase p holds x[w0 + c + SPA*m .. + SPA-1].
  template <class A>
  __device__ __forceinline__ static unsigned load_elem(const A &a, int s, int p,
                                                       long long t) {
    if (t < 0 || t >= a.T) return 0u;
    const global_char *base = (const global_char *)(a.F);
    if (ES == 4)
      return *(const __attribute__((address_space(1))) unsigned *)(
          base + (((size_t)s * P + p) * a.ldf + t) * 4);
    return *(const __attribute__((address_space(1))) unsigned short *)(
        base + (((size_t)s * P + p) * 2 * a.ldf + t) * 2);
  }

  template <class A>
  __device__ __forceinline__ static void stage_station_guarded(const A &a,
                                                               const_int *tlo, int s, int t0,
                                                               int tid, unsigned char *buf) {
    const int Lw = a.Lw, nw = Lw / SPA, SB = window_slot_bytes(Lw);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int lo0 = tlo[s * P + p];
      if (lo0 == INT_MAX) continue;  // no source of the tile uses (s, p)
      const long long w0 = (long long)t0 + (ES == 4 ? lo0 : (lo0 & ~1));
      for (int u = tid; u < Lw; u += NT) {
        const int c = u / nw, m = u - c * nw;
        const long long t = w0 + c + (long long)SPA * m;
        unsigned q[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          if (ES == 4) {
            q[d] = load_elem(a, s, p, t + d);
          } else {
            q[d] = load_elem(a, s, p, t + 2 * d) | (load_elem(a, s, p, t + 2 * d + 1) << 16);
          }
        }
        *reinterpret_cast<uint4 *>(buf + p * SB + 16 * u) = make_uint4(q[0], q[1], q[2], q[3]);
      }
    }
  }

  // Operands of one station for this wave's KW sources: KW*(P+1)
  // wave-uniform words (the KW*P byte offsets in step order, then the KW
  // weight words).
  __device__ __forceinline__ static void load_sched(const_int *sg, int (&sv)[KW * (P + 1)]) {
#pragma unroll
    for (int i = 0; i < KW * (P + 1); ++i) sv[i] = sg[i];
  }

  // One gather step is one (source, phase) term: R ds_read_b128 at LDS
  // address `addr` (an absolute LDS address, so that no base is added).
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(3))) u32x4 lds_u32x4;
  __device__ __forceinline__ static void gather_one(unsigned addr, int r, uint4 (&v)[R]) {
    const u32x4 w = *(lds_u32x4 *)(size_t)(addr + 1024u * r);
    v[r] = make_uint4(w.x, w.y, w.z, w.w);
  }
  __device__ __forceinline__ static void gather(unsigned addr, uint4 (&v)[R]) {
#pragma unroll
    for (int r = 0; r < R; ++r) gather_one(addr, r, v);
  }
  __device__ __forceinline__ static unsigned lds_address(const unsigned char *p) {
    return (unsigned)(size_t)(const __attribute__((address_space(3))) unsigned char *)p;
  }

  // Lane addresses of two steps in one instruction: a 64-bit add of the
  // lane's offset, held in both halves of a register pair, and the two
  // steps' (buffer + schedule) offsets as a scalar pair.  Neither half
  // carries (LDS offsets are below 2^18), so the halves are the two 32-bit
  // sums; one v_lshl_add_u64 replaces two v_add_u32.
  __device__ __forceinline__ static unsigned long long pair_addr(unsigned long long lb2,
                                                                 unsigned o0, unsigned o1) {
    const unsigned long long o = ((unsigned long long)o1 << 32) | o0;
    unsigned long long a2;
    asm("v_lshl_add_u64 %0, %1, 0, %2" : "=v"(a2) : "v"(lb2), "s"(o));
    return a2;
  }

  // One station: straight-line gathers and FMAs over the KW*P (source,
  // phase) steps, software-pipelined DEPTH steps ahead (the gathers of
  // steps i+1..i+DEPTH are in flight during the FMAs of step i).  A fence
  // after each step keeps the compiler from hoisting every gather of the
  // station at once, which would exhaust the registers: the empty asm
  // statements take the step's accumulators (its FMAs complete there) and
  // the lane's LDS offset (later gathers depend on it, so they stay behind).
  // boff: byte offset of the station's staging buffer in the staging area.
  static constexpr int STEPS = KW * P;
  static constexpr int DEPTH_FMT = ES == 4 ? BP_PIPE_DEPTH : BP_PIPE_DEPTH_16;
  static constexpr int DEPTH = DEPTH_FMT < STEPS ? DEPTH_FMT : STEPS - 1;
  static constexpr int PRIO_STEPS = STEPS * BP_PRIO_PCT / 100;
  // where in a step's 4R words the next step's reads are issued: 32-bit
  // samples read 1 KiB twice per step, 16-bit samples once
  __host__ __device__ static constexpr int gather_at(int r) {
    return ES == 4 ? (r == 0 ? BP_GATHER_AT_F32_0 : BP_GATHER_AT_F32_1) : BP_GATHER_AT_16;
  }

  // hook(i) runs at the start of step i (the copies of a later station).
  // lb2: the lane's byte offset (16 * lane) in both halves; the caller
  // keeps it in one register pair for the whole tile (the fences below
  // pass it through, so no station copies it).
  template <class Hook>
  __device__ __forceinline__ static void consume_station(const int (&sv)[KW * (P + 1)],
                                                         const unsigned char *stage,
                                                         unsigned boff,
                                                         unsigned long long &lb2,
                                                         float (&acc)[KW][NACC], Hook hook) {
    const unsigned lb = (unsigned)lb2;
    const unsigned base = lds_address(stage) + boff;
    uint4 v[DEPTH + 1][R];
    unsigned addr[STEPS];
    // address of step n: computed with step n + 1 for even n
    auto step_addr = [&](int n) {
      if (n % 2 == 0 && n + 1 < STEPS) {
        const unsigned long long a2 =
            pair_addr(lb2, (unsigned)sv[n] + base, (unsigned)sv[n + 1] + base);
        addr[n] = (unsigned)a2;
        addr[n + 1] = (unsigned)(a2 >> 32);
      } else if (n % 2 == 0) {
        addr[n] = lb + (unsigned)sv[n] + base;
      }
      return addr[n];
    };
#pragma unroll
    for (int i = 0; i < DEPTH; ++i) gather(step_addr(i), v[i % (DEPTH + 1)]);
#pragma unroll
    for (int i = 0; i < STEPS; ++i) {
      hook(i);
      // raised priority over the first PRIO_STEPS steps of the station
      if (PRIO > 0 && i == 0) __builtin_amdgcn_s_setprio(PRIO);
      if (PRIO > 0 && i == PRIO_STEPS) __builtin_amdgcn_s_setprio(0);
      const int kk = i / P;
      const int n = i + DEPTH;
      const int bw = sv[STEPS + kk];
      // the step's 4R words in order (each word is one or two products),
      // with read r of step n's gather issued before word gather_at(r)
      const uint4(&x)[R] = v[i % (DEPTH + 1)];
      unsigned an = 0;
#pragma unroll
      for (int u = 0; u < 4 * R; ++u) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (u == gather_at(r) && n < STEPS) {
            __builtin_amdgcn_sched_barrier(0);
            if (r == 0) an = step_addr(n);
            gather_one(an, r, v[n % (DEPTH + 1)]);
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        const uint4 &w = x[u / 4];
        const unsigned word = u % 4 == 0 ? w.x : u % 4 == 1 ? w.y : u % 4 == 2 ? w.z : w.w;
        F::mac_word(&acc[kk][SPA * (u / 4)], word, u % 4, bw);
      }
#pragma unroll
      for (int j = 0; j < NACC; ++j) asm volatile("" : "+v"(acc[kk][j]));
      asm volatile("" : "+v"(lb2));
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  // One pipeline step: start a later station's global->LDS copies into
  // `nxt`, then consume the current station from buffer `boff` of the
  // staging area.  The staging area is restrict-qualified against the copy
  // target in one scope, so the compiler knows the copies cannot alias the
  // gathers and does not wait for them before reading.
  __device__ __forceinline__ static void station_step(bool dma, const StationSrc &src,
                                                      const RowPlan &rp, int SB, int wave,
                                                      const int (&sv)[KW * (P + 1)],
                                                      const unsigned char *__restrict__ stage,
                                                      unsigned boff,
                                                      unsigned char *__restrict__ nxt,
                                                      unsigned long long &lb2,
                                                      float (&acc)[KW][NACC]) {
    consume_station(sv, stage, boff, lb2, acc, [&](int i) {
      bool here = false;  // i is a constant of the unrolled steps
#pragma unroll
      for (int p = 0; p < P; ++p) here = here || i == dma_step(p);
      if (!here) return;
      __builtin_amdgcn_sched_barrier(0);
      if (dma) {
#pragma unroll
        for (int p = 0; p < P; ++p)
          if (i == dma_step(p)) dma_phase(src, rp, SB, wave, p, nxt);
      }
      __builtin_amdgcn_sched_barrier(0);
    });
  }

  // Sample of accumulator j of a lane: run j / SPA, position j % SPA.
  __device__ __forceinline__ static int sample_of(int t0, int lane, int j) {
    return t0 + SPA * lane + 64 * SPA * (j / SPA) + (j % SPA);
  }

  // One work-group's tile: time tile tt (samples [512 tt, 512 tt + 512)) of
  // source tile `tile`, all stations.
  template <class A>
  __device__ __forceinline__ static void tile_job(const A &a, int tt, int tile,
                                                  unsigned char *smem) {
    const int Lw = a.Lw;
    const int SB = window_slot_bytes(Lw);
    unsigned long long *red = reinterpret_cast<unsigned long long *>(smem);
    unsigned char *stage = smem + (NONE ? 0 : TT * 8);
    const int bbytes = buf_bytes(Lw);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uniform(tid >> 6);
    const int t0 = tt * TT;
    const int S = a.S;
    const_int *tlo = as_const(a.tile_lo) + (size_t)tile * S * P;
    const_int *tall = as_const(a.tile_all) + 8 * tile;
    const int tall0 = tall[0];
    const int kbase = tile * KT + wave * KW;
    // staged range of every slot: [t0 + lo, t0 + lo + Lw + SPA - 1), with
    // lo rounded down to even for 16-bit samples; rows are readable (and
    // zero) from -pad_lo to T + pad_hi
    const bool edge = (long long)t0 + (ES == 4 ? tall0 : (tall0 & ~1)) < -(long long)a.pad_lo ||
                      (long long)t0 + tall[2] + Lw + SPA - 1 > (long long)a.T + a.pad_hi;

    if (!NONE)
      for (int i = tid; i < TT; i += NT) red[i] = 0ull;
    for (int i = tid; i < NB * (ZERO_BYTES / 16); i += NT) {
      const int b = i / (ZERO_BYTES / 16), w = i - b * (ZERO_BYTES / 16);
      *reinterpret_cast<uint4 *>(stage + b * bbytes + P * SB + 16 * w) = make_uint4(0u, 0u, 0u, 0u);
    }

    float acc[KW][NACC];
#pragma unroll
    for (int kk = 0; kk < KW; ++kk)
#pragma unroll
      for (int j = 0; j < NACC; ++j) acc[kk][j] = 0.0f;

    // schedule rows of this wave: station s at sched_w + s * KT * (P + 1)
    const_int *sched_w = as_const(a.sched) + ((size_t)tile * S * KT + wave * KW) * (P + 1);
    int sv[KW * (P + 1)];
    unsigned long long lb2 = ((unsigned long long)(16u * lane) << 32) | (16u * lane);
    if (edge) {
      // synchronous: stage, barrier, consume, barrier
      for (int s = 0; s < S; ++s) {
        load_sched(sched_w + (size_t)s * KT * (P + 1), sv);
        stage_station_guarded(a, tlo, s, t0, tid, stage);
        __syncthreads();
        consume_station(sv, stage, 0u, lb2, acc, [](int) {});
        __syncthreads();
      }
    } else {
      // Ring of NB buffers: the copies of station s + NB - 1 are issued
      // while station s is consumed.  Every wave issues P * rp.n rows per
      // station, so at the end of station s waiting until at most that
      // many rows per station issued after s + 1 are in flight means the
      // wave's rows of station s + 1 have landed; the barrier then makes
      // every wave's rows visible and frees buffer s for reuse.
      RowPlan rp;
      plan_rows(Lw, a.ldf, wave, lane, rp);
      const int rows = P * rp.n;
      StationSrc src;
      const int npro = min(NB - 1, S);
      __syncthreads();  // zero blocks written
      for (int s = 0; s < npro; ++s) {
        station_src(a, tlo, tall0, s, t0, src);
        dma_station(src, rp, SB, wave, stage + s * bbytes);
      }
      wait_vmem(rows * (npro - 1));
      barrier_no_fence();
      load_sched(sched_w, sv);
      if (NB - 1 < S) station_src(a, tlo, tall0, NB - 1, t0, src);
      int ci = 0;       // ring index of station s
      int di = NB - 1;  // ring index of station s + NB - 1 (kept as its own scalar: derived
                        // from ci, the select went through the VALU)
      for (int s = 0; s < S; ++s) {
        const int sd = s + NB - 1;
        station_step(sd < S, src, rp, SB, wave, sv, stage, (unsigned)(ci * bbytes),
                     stage + di * bbytes, lb2, acc);
        if (s + 1 < S) {
          // next station's operands and staging sources: the scalar loads
          // overlap the barrier
          load_sched(sched_w + (size_t)(s + 1) * KT * (P + 1), sv);
          if (sd + 1 < S) station_src(a, tlo, tall0, sd + 1, t0, src);
        }
        // every gather of station s has been consumed by its FMAs, so only
        // the copies need waiting for
        if (NB == 2)
          wait_vmem(0);
        else
          wait_vmem(rows * max(0, min(sd, S - 1) - (s + 1)));
        barrier_no_fence();
        di = ci;
        ci = ci + 1 == NB ? 0 : ci + 1;
      }
    }

    // Flexible mode, time tiles where a term can fall outside [0, T): beams
    // that end at zero take the reference's chain (flexible_chain).  The
    // bfloat16 form is held to a tolerance, not to the bits, and skips this.
    const bool flex_edge =
        FMT != BP_BF16 && a.oob &&
        ((long long)t0 + tall[0] < 0 || (long long)t0 + TT - 1 + tall[1] >= a.T);

    if (NONE) {
      // bit kk * NACC + j: the beam of (kbase + kk, sample j) ended at zero
      unsigned long long zero_beams = 0ull;
      if (flex_edge) {
#pragma unroll
        for (int kk = 0; kk < KW; ++kk)
#pragma unroll
          for (int j = 0; j < NACC; ++j)
            if (acc[kk][j] == 0.0f) zero_beams |= 1ull << (kk * NACC + j);
      }
#pragma unroll
      for (int kk = 0; kk < KW; ++kk) {
        const int k = kbase + kk;
        if (k >= a.K) continue;
        const int ko = source_index(a, k);
        if (ko < a.k_begin || ko >= a.k_end) continue;
        int vlo = 0, vhi = a.T;
        if (!a.oob)
          strict_range(as_const(a.src_lohi)[2 * k], as_const(a.src_lohi)[2 * k + 1], a.T, vlo,
                       vhi);
        float *row = a.beam + (size_t)(ko - a.k_begin) * a.T;
        if (vlo <= t0 && t0 + TT <= vhi && ((size_t)row & 15) == 0) {
          // the whole tile is valid for source k and the row 16-byte aligned
          // (wave-uniform): each run of SPA samples as 16-byte stores
#pragma unroll
          for (int j = 0; j < NACC; j += 4)
            *reinterpret_cast<float4 *>(row + sample_of(t0, lane, j)) =
                make_float4(acc[kk][j], acc[kk][j + 1], acc[kk][j + 2], acc[kk][j + 3]);
        } else {
#pragma unroll
          for (int j = 0; j < NACC; ++j) {
            const int t = sample_of(t0, lane, j);
            if (t >= vlo && t < vhi) row[t] = acc[kk][j];
          }
        }
      }
      // flexible edge tiles: zero beams again from the reference's chain (the
      // same lane stored the value above, so this store replaces it)
      while (zero_beams) {
        const int i = __builtin_ctzll(zero_beams);
        zero_beams &= zero_beams - 1;
        const int kk = i / NACC, j = i - kk * NACC;
        const int k = kbase + kk, t = sample_of(t0, lane, j);
        if (k >= a.K || t >= a.T) continue;
        const int ko = a.perm ? a.perm[k] : k;
        if (ko < a.k_begin || ko >= a.k_end) continue;
        a.beam[(size_t)(ko - a.k_begin) * a.T + t] = flexible_chain<FMT, P>(a, k, t);
      }
    } else {
      float bv[NACC];
      int bk[NACC];
#pragma unroll
      for (int j = 0; j < NACC; ++j) {
        bv[j] = -FLT_MAX;
        bk[j] = 0;
      }
      // The common case, tested once per wave: every source of the wave
      // exists, is written, and is valid over the whole time tile.  Then the
      // running max needs no per-source or per-sample test.
      bool whole = kbase + KW <= a.K && a.k_begin == 0 && a.k_end == a.K;
      if (whole && !a.oob) {
#pragma unroll
        for (int kk = 0; kk < KW; ++kk) {
          const int k = kbase + kk;
          int vlo, vhi;
          strict_range(as_const(a.src_lohi)[2 * k], as_const(a.src_lohi)[2 * k + 1], a.T, vlo,
                       vhi);
          whole = whole && vlo <= t0 && t0 + TT <= vhi;
        }
      }
      if (whole) {
#pragma unroll
        for (int kk = 0; kk < KW; ++kk) {
          const int ko = source_index(a, kbase + kk);
#pragma unroll
          for (int j = 0; j < NACC; ++j) {
            const float v = acc[kk][j];
            if (v > bv[j]) {
              bv[j] = v;
              bk[j] = ko;
            }
          }
        }
      } else {
        // Candidates in position order, which within a tile is ascending
        // original index: with a strict '>' every lane keeps, of its sources,
        // the largest value and among equal values the lowest original index
        // (the reference's ascending scan).
#pragma unroll
        for (int kk = 0; kk < KW; ++kk) {
          const int k = kbase + kk;
          if (k >= a.K) continue;
          const int ko = source_index(a, k);
          if (ko < a.k_begin || ko >= a.k_end) continue;
          int vlo = 0, vhi = a.T;
          if (!a.oob)
            strict_range(as_const(a.src_lohi)[2 * k], as_const(a.src_lohi)[2 * k + 1], a.T, vlo,
                         vhi);
          if (vlo <= t0 && t0 + TT <= vhi) {
            // every sample of the tile is valid for source k (wave-uniform)
#pragma unroll
            for (int j = 0; j < NACC; ++j) {
              const float v = acc[kk][j];
              if (v > bv[j]) {
                bv[j] = v;
                bk[j] = ko;
              }
            }
          } else {
#pragma unroll
            for (int j = 0; j < NACC; ++j) {
              const int t = sample_of(t0, lane, j);
              const float v = acc[kk][j];
              if (t >= vlo && t < vhi && v > bv[j]) {
                bv[j] = v;
                bk[j] = ko;
              }
            }
          }
        }
      }
      if (flex_edge) {
        // Where a lane's maximum is a zero, the winner is its first candidate
        // (ascending position) whose beam is zero, and the winner's zero sign
        // comes from the reference's chain.  Nibble j of `win`: 8 | kk of the
        // winner at sample j.
        unsigned long long win = 0ull;
#pragma unroll
        for (int j = 0; j < NACC; ++j) {
          if (bv[j] != 0.0f) continue;
          int w = -1;
#pragma unroll
          for (int kk = KW - 1; kk >= 0; --kk) {
            const int k = kbase + kk;
            if (k >= a.K) continue;
            const int ko = source_index(a, k);
            if (ko >= a.k_begin && ko < a.k_end && acc[kk][j] == 0.0f) w = kk;
          }
          if (w >= 0) win |= (unsigned long long)(8 | w) << (4 * j);
        }
        while (win) {
          const int i = __builtin_ctzll(win) / 4;
          const int kk = (int)((win >> (4 * i)) & 7ull);
          win &= ~(15ull << (4 * i));
          const float v = flexible_chain<FMT, P>(a, kbase + kk, sample_of(t0, lane, i));
#pragma unroll
          for (int j = 0; j < NACC; ++j)
            if (j == i) bv[j] = v;
        }
      }
#pragma unroll
      for (int j = 0; j < NACC; ++j)
        if (bv[j] > -FLT_MAX) atomicMax(&red[sample_of(0, lane, j)], pack_key(bv[j], bk[j]));
      __syncthreads();
      for (int i = tid; i < TT; i += NT) {
        const unsigned long long key = red[i];
        if (key && t0 + i < a.T) atomicMax(a.keys + t0 + i, key);
      }
    }
  }

  // One launch over the listed tiles of one network: block -> (time tile,
  // listed tile) with XCD and L2 locality.  A dispatch holds fewer than 2^32
  // work-items, so a class of more blocks than that allows runs as several
  // launches of consecutive block ranges [block_base, block_base + grid).
  __device__ __forceinline__ static void run(const StackArgs &a, unsigned char *smem) {
    const int L = a.block_base + xcd_remap(blockIdx.x, gridDim.x);
    int tt, li;
    block_tile(L, a.n_time_tiles, a.n_list, a.time_group, tt, li);
    tile_job(a, tt, as_const(a.tiles)[li], smem);
  }
};

// 6 waves per SIMD (VGPRs <= 80): three work-groups of 8 waves per CU when
// their LDS fits (FP32 windows up to 704 samples), so two compute while one
// waits at its per-station barrier.  With 6 sources per wave this measured
// 6.9 % faster than 8 sources per wave at 4 waves per SIMD (two work-groups).
#ifndef BP_WAVES_PER_EU
#define BP_WAVES_PER_EU 6
#endif
template <int FMT, int P, int KW, int W, int NB, bool NONE>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(BP_WAVES_PER_EU)))
void stack_lds_kernel(StackArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  StackLds<FMT, P, KW, W, NB, NONE>::run(a, smem);
}

// ---------------------------------------------------------------------------
// Batched form: one launch over the LDS-class tiles of many windows (each
// its own network, length and feature buffer), reduce='max'.  Window w
// contributes n_time_tiles(w) * n_list(w) work items, numbered from
// item_begin(w).  The grid is a multiple of 8 work-groups.  Work-group b
// runs on XCD b % 8, and XCD x walks the contiguous item range
// [split[x], split[x+1]), so that within an XCD consecutive items are
// consecutive time tiles of one window (the same L2 locality as one launch
// per window).  Items cost different amounts (S varies per window), so the
// host places the XCD boundaries `split` at equal shares of the total cost
// (items x stations) rather than of the item count.
//   queue == NULL: one work-group per item (8 x the largest XCD range;
//     surplus work-groups exit at once), or a grid that strides over its
//     XCD's range;
//   queue != NULL: a persistent grid over a work queue.  queue[x] counts
//     the items of XCD x's range handed out so far (zero at launch).  A
//     work-group takes the next item of its own XCD's range with an atomic
//     add until the range is exhausted, then takes items of the other
//     XCDs' ranges the same way, so no XCD idles while another still has
//     work.  Counters only grow, so every work-group ends after at most
//     eight failed takes.
// ---------------------------------------------------------------------------
struct BatchWindow {
  StackArgs a;           // the window's arguments: its tiles of one LDS class
  long long item_begin;  // first work item of the window
};

typedef const __attribute__((address_space(4))) BatchWindow const_batch_window;

template <int FMT, int P, int KW, int W, int NB>
__device__ __forceinline__ void batch_item(const_batch_window *cw, int n_wins, long long L,
                                           unsigned char *smem) {
  // the window holding item L: the last with item_begin <= L
  int lo = 0, hi = n_wins - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cw[mid].item_begin <= L)
      lo = mid;
    else
      hi = mid - 1;
  }
  // the window's arguments stay in constant memory: fields are scalar
  // loads where used, like a kernel argument's
  typedef const __attribute__((address_space(4))) StackArgs const_args;
  const_args &a = cw[lo].a;
  int tt, li;
  block_tile((int)(L - cw[lo].item_begin), a.n_time_tiles, a.n_list, a.time_group, tt, li);
  StackLds<FMT, P, KW, W, NB, false, BP_BATCH_PRIO>::tile_job(a, tt, as_const(a.tiles)[li], smem);
  __syncthreads();  // the LDS is reused by the next item
}

template <int FMT, int P, int KW, int W, int NB>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(BP_WAVES_PER_EU)))
void stack_batch_kernel(const BatchWindow *wins, int n_wins, const long long *split,
                        unsigned long long *queue) {
  extern __shared__ __align__(16) unsigned char smem[];
  const_batch_window *cw = (const_batch_window *)wins;
  typedef const __attribute__((address_space(4))) long long const_ll;
  const int x = blockIdx.x & 7;
  if (!queue) {
    const long long end = ((const_ll *)split)[x + 1];
    const int step = gridDim.x >> 3;
    for (long long L = ((const_ll *)split)[x] + (blockIdx.x >> 3); L < end; L += step)
      batch_item<FMT, P, KW, W, NB>(cw, n_wins, L, smem);
    return;
  }
  // The taken item is broadcast through the first word of the LDS, which
  // tile_job clears before it uses it (the reduction keys).
  unsigned long long *slot = reinterpret_cast<unsigned long long *>(smem);
  for (int r = 0; r < 8; ++r) {
    const int q = (x + r) & 7;
    const long long b0 = ((const_ll *)split)[q];
    const long long n = ((const_ll *)split)[q + 1] - b0;
    for (;;) {
      if (threadIdx.x == 0) *slot = atomicAdd(queue + q, 1ull);
      __syncthreads();
      const unsigned long long g = *slot;
      __syncthreads();  // every thread has read the slot before it is reused
      const long long i = (long long)(((unsigned long long)(unsigned)uniform((int)(g >> 32)) << 32) |
                                      (unsigned)uniform((int)g));
      if (i >= n) break;
      batch_item<FMT, P, KW, W, NB>(cw, n_wins, b0 + i, smem);
    }
  }
}

// Batched key initialisation and decoding: blockIdx.y selects the window.
struct BatchKeys {
  unsigned long long *keys;
  float *beam_max;
  int *beam_argmax;
  int T;
};

__global__ void fill_keys_batch_kernel(const BatchKeys *w) {
  const BatchKeys b = w[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < b.T;
       i += (long long)gridDim.x * blockDim.x)
    b.keys[i] = 0ull;
}

__global__ void finalize_max_batch_kernel(const BatchKeys *w) {
  const BatchKeys b = w[blockIdx.y];
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < b.T;
       t += (long long)gridDim.x * blockDim.x) {
    const unsigned long long key = b.keys[t];
    if (key) {
      b.beam_max[t] = key_value(key);
      b.beam_argmax[t] = key_index(key);
    } else {
      b.beam_argmax[t] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// Shift-and-stack, direct gathers from global memory.  Used for source tiles
// the LDS kernel does not take (n_phases outside 1..4, or a tile whose
// delay spread exceeds the largest LDS window).  Same FMA order, same bits.
// One thread per time sample; blockIdx.y selects a listed source tile of
// `chunk` sources.
// ---------------------------------------------------------------------------
// Feature element t of (s, p) as float: float32 planar, or the 16-bit
// parity plane 0.
template <int FMT>
__device__ __forceinline__ float feature(const StackArgs &a, int s, int p, long long t) {
  if (FMT == BP_F32) return static_cast<const float *>(a.F)[((size_t)s * a.P + p) * a.ldf + t];
  const unsigned short h =
      static_cast<const unsigned short *>(a.F)[((size_t)s * a.P + p) * 2 * a.ldf + t];
  if (FMT == BP_F16) return (float)__builtin_bit_cast(_Float16, h);
  return __uint_as_float((unsigned)h << 16);
}

// Weight of a used station as the kernels apply it: times the station's
// 16-bit storage scale, and for BF16 rounded to bfloat16 like the features.
template <int FMT>
__device__ __forceinline__ float applied_weight(float b0, float scale) {
  float b = b0 * scale;
  if (FMT == BP_BF16) b = __uint_as_float(bf16_bits(b) << 16);
  return b;
}

template <int FMT>
__device__ __forceinline__ float station_weight(const StackArgs &a, float b0, int s) {
  return applied_weight<FMT>(b0, a.scale ? a.scale[s] : 1.0f);
}

template <int FMT, bool NONE>
__global__ __launch_bounds__(256) void stack_direct_kernel(StackArgs a, int chunk) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long tile = a.tiles[blockIdx.y];
  const int k0 = (int)(tile * chunk);
  const int k1 = (int)min((long long)a.K, tile * chunk + chunk);
  const int S = a.S, P = a.P, T = a.T;
  float bv = -FLT_MAX;
  int bk = 0;
  for (int k = k0; k < k1; ++k) {
    const int ko = a.perm ? a.perm[k] : k;
    if (ko < a.k_begin || ko >= a.k_end) continue;
    int vlo = 0, vhi = T;
    if (!a.oob) strict_range(a.src_lohi[2 * k], a.src_lohi[2 * k + 1], T, vlo, vhi);
    if (t >= T || t < vlo || t >= vhi) continue;
    float acc = 0.0f;
    for (int s = 0; s < S; ++s) {
      const float b0 = a.beta[(size_t)k * S + s];
      if (b0 == 0.0f) continue;
      const float b = station_weight<FMT>(a, b0, s);
      for (int p = 0; p < P; ++p) {
        const long long tg = t + a.tau[((size_t)k * S + s) * P + p];
        if (tg < 0 || tg >= T) continue;  // only reachable in flexible mode
        acc = fmaf(b, feature<FMT>(a, s, p, tg), acc);
      }
    }
    if (NONE) {
      a.beam[(size_t)(ko - a.k_begin) * T + t] = acc;
    } else if (acc > bv) {
      bv = acc;
      bk = ko;
    }
  }
  if (!NONE && t < T && bv > -FLT_MAX) atomicMax(a.keys + t, pack_key(bv, bk));
}

// ---------------------------------------------------------------------------
// Differential mode (beamform.c:338-394): the time delays are relative to a
// zero lag at the centre sample t = (T - 1) / 2 of a short window (the
// features are inter-station cross-correlations), and each source gets one
// beam, at t, if it uses a station and every used delay points inside
// [0, T); other entries of `beam` are left as they are.  The FMA chain is
// the one of every other beam.
//
// One thread per source, 64 sources per wave.  The tables are read in a
// wave-transposed layout built once per network (diff_tables_kernel): for
// the 64 sources of wave block b, delay entry j = s * P + p of all 64 is
// one contiguous 256-byte line, tauT[(b * S * P + j) * 64 + lane], and so
// is weight s, betaT[(b * S + s) * 64 + lane].  Every table load of a wave
// is then one full line (row-major (K, S, P) rows would give each lane its
// own line, and the load instructions, not the bytes, would set the pace).
// ---------------------------------------------------------------------------
constexpr int kDiffSources = 256;  // per work-group: 4 independent waves
constexpr int kDiffBlock = 64;     // sources per wave block of the transposed tables

// (K, E_in) row-major 4-byte table -> (ceil(K / 64), E_out, 64) wave-transposed,
// E_out >= E_in, zero past the last source and past entry E_in.  Grid
// (ceil(K / 64), any), 256 threads, the y blocks striding over 64-entry
// tiles; a 64 x 64 tile goes through LDS so both sides are coalesced.
__global__ __launch_bounds__(256) void diff_tables_kernel(const unsigned *__restrict__ in,
                                                          long long K, int E_in, int E_out,
                                                          unsigned *__restrict__ out) {
  __shared__ unsigned tile[kDiffBlock][kDiffBlock + 1];
  const long long b = blockIdx.x;
  for (int e0 = blockIdx.y * kDiffBlock; e0 < E_out; e0 += gridDim.y * kDiffBlock) {
    __syncthreads();  // the previous tile is written out
    for (int i = threadIdx.x; i < kDiffBlock * kDiffBlock; i += blockDim.x) {
      const int r = i / kDiffBlock, c = i % kDiffBlock;
      const long long k = b * kDiffBlock + r;
      tile[r][c] = (k < K && e0 + c < E_in) ? in[k * E_in + e0 + c] : 0u;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kDiffBlock * kDiffBlock; i += blockDim.x) {
      const int c = i / kDiffBlock, r = i % kDiffBlock;
      if (e0 + c < E_out) out[(b * E_out + e0 + c) * kDiffBlock + r] = tile[r][c];
    }
  }
}

// Two phases' delays of a station as one word of int16 halves (phase 0 low,
// phase 1 high), saturated to the int16 range, from wave-transposed int32
// entries (ceil(K/64), 2 * S_pad, 64) into (ceil(K/64), S_pad, 64).  For
// windows of at most 32768 samples (centre t <= 16383) a saturated delay
// still points outside [0, T), and every delay inside it is exact, so the
// kernel's terms and bounds are unchanged.
__device__ __forceinline__ unsigned sat16(unsigned v) {
  const int d = (int)v;
  return (unsigned)(d < -32768 ? -32768 : d > 32767 ? 32767 : d) & 0xffffu;
}

__global__ void diff_pack_pairs_kernel(const unsigned *__restrict__ in, long long n_blk,
                                       int S_pad, unsigned *__restrict__ out) {
  const long long n = n_blk * S_pad * kDiffBlock;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long row = i / kDiffBlock, lane = i - row * kDiffBlock;  // row = b * S_pad + s
    const long long b = row / S_pad, s = row - b * S_pad;
    const long long src = ((b * S_pad + s) * 2) * kDiffBlock + lane;  // entry 2s of block b
    out[i] = sat16(in[src]) | (sat16(in[src + kDiffBlock]) << 16);
  }
}

// Uniform weights.  When every nonzero weight of a network has the same
// bits (all ones, or any 0/1 station mask), the weight of (source, station)
// is that value c where the station is used (beta != 0) and 0 elsewhere:
// one bit instead of a float32.  The kernel then applies the same weight c
// (the same FMA chain) from a bit table, 1/32 of the bytes.
//
// diff_weight_bits_kernel: min (out[0]) and max (out[1]) of the bit
// patterns of the nonzero weights (-0.0 is zero); the host initialises
// out to {0xffffffff, 0}.  Equal bounds: uniform.  No nonzero weight
// leaves out[0] > out[1].
__global__ __launch_bounds__(256) void diff_weight_bits_kernel(const unsigned *__restrict__ beta,
                                                               long long n,
                                                               unsigned *__restrict__ out) {
  unsigned lo = 0xffffffffu, hi = 0u;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const unsigned v = beta[i];
    if (__uint_as_float(v) != 0.0f) {
      lo = min(lo, v);
      hi = max(hi, v);
    }
  }
  // wave reduction, then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, (unsigned)__shfl_xor((int)lo, off, 64));
    hi = max(hi, (unsigned)__shfl_xor((int)hi, off, 64));
  }
  if (threadIdx.x % 64 == 0) {
    if (lo != 0xffffffffu) atomicMin(out, lo);
    if (hi != 0u) atomicMax(out + 1, hi);
  }
}

// (K, S) row-major weights -> wave-transposed used-bit words
// (ceil(K/64), W, 64): bit (s % 32) of word s / 32 of source k is
// beta[k, s] != 0; bits past station S and sources past K are 0.  Each
// wave item is two sources and one word: lanes 0-31 read 32 consecutive
// weights of the first source, lanes 32-63 of the second (coalesced row
// pieces), and a ballot forms both words.  Waves stride over the items.
__global__ __launch_bounds__(256) void diff_mask_kernel(const float *__restrict__ beta,
                                                        long long K, int S, int W,
                                                        unsigned *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const long long n_src = (K + 63) / 64 * 64;  // every word of the last block is written
  const long long n_items = n_src / 2 * W;
  const long long n_waves = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long it = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; it < n_items;
       it += n_waves) {
    const long long pair = it / W;
    const int w = (int)(it - pair * W);
    const long long k = 2 * pair + (lane >> 5);
    const int s = w * 32 + (lane & 31);
    const bool used = k < K && s < S && beta[k * S + s] != 0.0f;
    const unsigned long long m = __ballot(used);
    if ((lane & 31) == 0)
      out[((k >> 6) * W + w) * 64 + (k & 63)] = (unsigned)(m >> (lane & 32));
  }
}

// Stations per chunk of the differential kernel for P phases (1-4), and the
// station count of its tables: S rounded up to two chunks, plus one chunk
// that the last prefetch reads.  Padding stations have zero weights.  A
// chunk's 16 gathers are in flight together (32 measured 8 % slower at
// 1,225 pairs x 192,000 sources; so did an XCD-grouped block order, +0 %).
__host__ __device__ constexpr int diff_chunk_stations(int P) { return P >= 1 && P <= 4 ? 16 / P : 1; }
__host__ __device__ inline int diff_padded_stations(int S, int P) {
  const int cs = diff_chunk_stations(P);
  return (S + 2 * cs - 1) / (2 * cs) * (2 * cs) + cs;
}
// Words per source of the used-bit table: the padded stations, and one more
// word that a chunk straddling a word boundary reads (5 stations per chunk).
__host__ __device__ inline int diff_mask_words(int S, int P) {
  return (diff_padded_stations(S, P) + 31) / 32 + 1;
}

// Feature element t of planar row `row` (= s * P + p): float32 rows, or
// plane 0 of the 16-bit parity rows.  The element offset is 32-bit (the
// host keeps the view under 2^30 elements).
template <int FMT>
__device__ __forceinline__ float feature_row(const StackArgs &a, unsigned row, unsigned t) {
  if (FMT == BP_F32) return static_cast<const float *>(a.F)[row * (unsigned)a.ldf + t];
  const unsigned short h =
      static_cast<const unsigned short *>(a.F)[row * 2u * (unsigned)a.ldf + t];
  if (FMT == BP_F16) return (float)__builtin_bit_cast(_Float16, h);
  return __uint_as_float((unsigned)h << 16);
}

// Buffer descriptor over [base, base + bytes), base wave-uniform and bytes
// below 2^32: raw loads take a 32-bit per-lane offset and a scalar offset
// (no 64-bit address arithmetic), and a load past the range returns zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t range_rsrc(const void *base,
                                                             unsigned long long bytes) {
  const unsigned long long b = (unsigned long long)base;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __builtin_amdgcn_make_buffer_rsrc((void *)(((unsigned long long)hi << 32) | lo),
                                           (short)0, (int)(unsigned)bytes, 0x00020000);
}

// PT: the phase count when it is 1..4 (a chunk of CS = 16 / PT stations is
// loaded at once), 0 for any other count (station by station).  The
// per-source bounds of the reference (beamform.c:10-47) are taken in the
// same pass: a used delay outside [0, T) adds no term and marks the source
// out of range, and the beam is written only for a source that uses a
// station and never went out of range -- for it no term was left out, so
// the chain is the reference's.  The bounds are private to each source
// (the reference shares them between OpenMP threads,
// beamform.c:353-354,370-380; this is its one-thread behaviour).
//
// PAIRS (two phases only): the delays come as one word of two int16 per
// (source, station) (diff_pack_pairs_kernel), one load instead of two and
// 8 instead of 12 table bytes per (source, station) with the weight.
//
// MASK (1-4 phases): uniform weights; betaT is the used-bit table of
// diff_mask_kernel, and a used station's weight is cw.
template <int FMT, int PT, bool PAIRS = false, bool MASK = false>
__global__ __launch_bounds__(kDiffSources) void differential_kernel(
    StackArgs a, const int *__restrict__ tauT, const float *__restrict__ betaT, int S_pad,
    float *__restrict__ beam, float cw = 0.0f) {
  static_assert(!PAIRS || PT == 2, "paired delays are for two phases");
  static_assert(!MASK || PT > 0, "used-bit weights are for 1-4 phases");
  constexpr int CS = diff_chunk_stations(PT);  // stations per chunk (PT > 0)
  static_assert(!MASK || CS <= 32, "a chunk's used bits span two words at most");
  constexpr int NE = PT > 0 ? CS * PT : 1;
  const int lane = threadIdx.x % kDiffBlock;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kDiffBlock);
  const long long blk = (long long)blockIdx.x * (kDiffSources / kDiffBlock) + wv;
  const long long k0 = blk * kDiffBlock;
  if (k0 >= a.K) return;  // whole wave past the last source
  const int S = a.S, P = PT > 0 ? PT : a.P;
  const long long T = a.T, t = (T - 1) / 2;
  // 32-bit offsets inside the wave's block: 64 * S_pad * P < 2^30 (checked by the host)
  const int *tb = tauT + blk * S_pad * (PAIRS ? 1 : P) * kDiffBlock + lane;
  const float *wb = betaT + blk * S_pad * kDiffBlock + lane;
  float acc = 0.0f;
  bool used = false, inside = true;

  if (PT == 0) {
    // any phase count: one station at a time
    for (int s0 = 0; s0 < S; ++s0) {
      const float b0 = wb[(unsigned)s0 * kDiffBlock];
      const bool u = b0 != 0.0f;
      const float sc = (FMT == BP_F32 || !a.scale) ? 1.0f : a.scale[s0];
      const float wt = applied_weight<FMT>(b0, sc);
      used |= u;
      for (int p = 0; p < P; ++p) {
        const long long tg = t + tb[(unsigned)(s0 * P + p) * kDiffBlock];
        const bool in = tg >= 0 && tg < T;
        const bool live = u && in;
        const float xv = feature_row<FMT>(a, live ? (unsigned)(s0 * P + p) : 0u,
                                          live ? (unsigned)tg : 0u);
        inside &= !u || in;
        const float y = fmaf(wt, xv, acc);
        acc = live ? y : acc;
      }
    }
    if (k0 + lane < a.K && used && inside) beam[k0 + lane] = acc;
    return;
  }

  // 1-4 phases: chunks of CS stations.  The tables are padded to S_pad
  // stations (pairs of whole chunks and one more) with zero weights, so
  // every chunk is whole and the padding stations are simply unused.  The next chunk's
  // delays and weights are in flight while the current chunk's gathers and
  // FMAs run (two register sets used in turn; the scheduling barriers keep
  // the gathers ahead of those loads, so waiting for a gather never waits
  // for them).
  struct Chunk {
    int d[NE];
    float b[CS];
  };
  // Every load of a chunk is a buffer load: the wave's block of a table, or
  // the features, under a wave-uniform descriptor, with the lane's 32-bit
  // offset and a scalar offset per entry (a 256-byte table line per
  // station or (station, phase)).  The host keeps the features under 2^30
  // elements (2^32 bytes, 4 * ldf bytes per row in every format).
  const unsigned lane4 = 4u * (unsigned)lane;
  constexpr int DE = PAIRS ? 1 : (PT > 0 ? PT : 1);  // delay words per station
  const int MW = diff_mask_words(S, P);
  const __amdgpu_buffer_rsrc_t rs_tau =
      range_rsrc(tauT + blk * S_pad * DE * kDiffBlock, (unsigned long long)S_pad * DE * 256u);
  const __amdgpu_buffer_rsrc_t rs_w =
      MASK ? range_rsrc(betaT + blk * MW * kDiffBlock, (unsigned long long)MW * 256u)
           : range_rsrc(betaT + blk * S_pad * kDiffBlock, (unsigned long long)S_pad * 256u);
  const unsigned row_bytes = 4u * (unsigned)a.ldf;
  const __amdgpu_buffer_rsrc_t rs_f =
      range_rsrc(a.F, (unsigned long long)S * P * row_bytes);
  auto load = [&](int s0, Chunk &ch) {
    // used bits of the chunk's stations: one word, or two when the chunk
    // straddles a word boundary (chunks start at multiples of CS)
    unsigned m0 = 0u, m1 = 0u;
    if (MASK) {
      m0 = __builtin_amdgcn_raw_buffer_load_b32(rs_w, lane4, (unsigned)(s0 >> 5) * 256u, 0);
      if (32 % CS != 0)
        m1 = __builtin_amdgcn_raw_buffer_load_b32(rs_w, lane4, (unsigned)((s0 >> 5) + 1) * 256u,
                                                  0);
    }
#pragma unroll
    for (int c = 0; c < CS; ++c) {
      if (MASK) {
        const int sc = s0 + c;
        const unsigned m = (32 % CS == 0 || (sc >> 5) == (s0 >> 5)) ? m0 : m1;
        ch.b[c] = ((m >> (sc & 31)) & 1u) ? cw : 0.0f;
      } else {
        ch.b[c] = __uint_as_float(
            __builtin_amdgcn_raw_buffer_load_b32(rs_w, lane4, (unsigned)(s0 + c) * 256u, 0));
      }
      if (PAIRS) {
        const unsigned w =
            __builtin_amdgcn_raw_buffer_load_b32(rs_tau, lane4, (unsigned)(s0 + c) * 256u, 0);
        ch.d[c * PT] = (int)(short)(w & 0xffffu);
        ch.d[c * PT + 1] = (int)(short)(w >> 16);
      } else {
#pragma unroll
        for (int p = 0; p < PT; ++p)
          ch.d[c * PT + p] = (int)__builtin_amdgcn_raw_buffer_load_b32(
              rs_tau, lane4, (unsigned)((s0 + c) * PT + p) * 256u, 0);
      }
    }
  };
  // 32-bit sample arithmetic: t < 2^30, and t + d computed modulo 2^32 lies
  // in [0, T) exactly when the true sum does (a wrapped negative sum is at
  // least 2^31 > T).
  //
  // A live entry (used station, delay in range) loads its feature at byte
  // offset row * row_bytes + t_g * ESZ from the feature descriptor; every
  // other entry loads at offset 0xffffffff, past the descriptor's range
  // (below 2^32 bytes), which reads no memory and returns +0.  The chain
  // then takes every entry as fma(w', x, acc) with w' = -0.0 for an unused
  // station: fma(-0, +0, acc) is acc exactly, signed zeros and NaN
  // included.  A used entry out of range would add w * (+0), but it marks
  // the source out of range and its beam is never written; every written
  // beam is the reference's chain over its used stations.  The entry tests
  // are thus made once, at the gathers, and the chain is one FMA an entry.
  const unsigned tu = (unsigned)t, Tu = (unsigned)T;
  constexpr unsigned ESZ = FMT == BP_F32 ? 4u : 2u;
  auto gather = [&](int s0, const Chunk &ch, float (&x)[NE]) {
    // row offsets as a running scalar sum (an opaque stride keeps the
    // compiler from hoisting one constant per entry into scalar registers)
    unsigned rb = row_bytes;
    asm volatile("" : "+s"(rb));
    unsigned roff = (unsigned)(s0 * PT) * rb;
#pragma unroll
    for (int c = 0; c < CS; ++c) {
      const bool u = ch.b[c] != 0.0f;
      used |= u;
#pragma unroll
      for (int p = 0; p < PT; ++p) {
        const unsigned tg = tu + (unsigned)ch.d[c * PT + p];
        const bool in = tg < Tu;
        inside &= !u || in;
        const unsigned off = (u && in) ? roff + tg * ESZ : 0xffffffffu;
        if (FMT == BP_F32) {
          x[c * PT + p] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs_f, off, 0, 0));
        } else {
          const unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(rs_f, off, 0, 0);
          x[c * PT + p] = FMT == BP_F16 ? (float)__builtin_bit_cast(_Float16, h)
                                        : __uint_as_float((unsigned)h << 16);
        }
        roff += rb;
      }
    }
  };
  auto chain = [&](int s0, const Chunk &ch, const float (&x)[NE]) {
#pragma unroll
    for (int c = 0; c < CS; ++c) {
      const float sc =
          (FMT == BP_F32 || !a.scale) ? 1.0f : a.scale[s0 + c < S ? s0 + c : S - 1];
      const float wt = ch.b[c] != 0.0f ? applied_weight<FMT>(ch.b[c], sc) : -0.0f;
#pragma unroll
      for (int p = 0; p < PT; ++p) acc = fmaf(wt, x[c * PT + p], acc);
    }
  };
  Chunk c0, c1;
  float x[NE];
  const int s_end = S_pad - CS;  // whole pairs of chunks; the last chunk is read ahead only
  load(0, c0);
  for (int s0 = 0; s0 < s_end; s0 += 2 * CS) {
    gather(s0, c0, x);
    __builtin_amdgcn_sched_barrier(0);
    load(s0 + CS, c1);
    __builtin_amdgcn_sched_barrier(0);
    chain(s0, c0, x);
    gather(s0 + CS, c1, x);
    __builtin_amdgcn_sched_barrier(0);
    load(s0 + 2 * CS, c0);
    __builtin_amdgcn_sched_barrier(0);
    chain(s0 + CS, c1, x);
  }
  if (k0 + lane < a.K && used && inside) beam[k0 + lane] = acc;
}

}  // namespace bp
