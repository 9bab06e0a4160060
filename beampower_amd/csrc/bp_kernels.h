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
  const float *F;         // planar features (S, P, T)
  const int *tau;         // (K, S, P) moveouts
  const float *beta;      // (K, S) source weights
  const int *src_lohi;    // (K, 2) per-source min/max delay over used stations
  const int *tile_lo;     // (n_tiles, S, P) min delay over the tile's users of (s,p)
  const int *tile_hi;     // (n_tiles, S, P) max delay (INT_MIN when unused)
  const int *tile_all;    // (n_tiles, 2) min of tile_lo / max of tile_hi
  int T, K, S, P;
  int oob;                // 0 strict, 1 flexible
  int Lmax;               // LDS window length per (s,p) slot
  int n_time_tiles;
  int tile_begin;         // first source tile of the launch
  int k_begin, k_end;     // sources written: [k_begin, k_end)
  unsigned long long *keys;  // reduce=max: (T) packed keys
  float *beam;               // reduce=none: beam[(k - k_begin) * T + t]
};

// ---------------------------------------------------------------------------
// Plan kernels (run once per network).
// ---------------------------------------------------------------------------

// beamform.c:10-47 -- per-source min/max over used stations (beta != 0).
__global__ void source_bounds_kernel(const int *__restrict__ tau,
                                     const float *__restrict__ beta, int K, int S,
                                     int P, int *__restrict__ lohi) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  int lo = INT_MAX, hi = INT_MIN;
  const int *tk = tau + (size_t)k * S * P;
  const float *bk = beta + (size_t)k * S;
  for (int s = 0; s < S; ++s) {
    if (bk[s] == 0.0f) continue;
    for (int p = 0; p < P; ++p) {
      int d = tk[s * P + p];
      lo = d < lo ? d : lo;
      hi = d > hi ? d : hi;
    }
  }
  lohi[2 * k] = lo;
  lohi[2 * k + 1] = hi;
}

// Per source tile [tile*KT, tile*KT+KT): min/max delay of every (s,p) over
// the tile's sources that use station s.  These bound the feature window a
// tile needs: [t0 + lo, t0 + TT - 1 + hi].
__global__ void tile_bounds_kernel(const int *__restrict__ tau,
                                   const float *__restrict__ beta, int K, int S,
                                   int P, int KT, int *__restrict__ tile_lo,
                                   int *__restrict__ tile_hi, int *__restrict__ tile_all,
                                   int *__restrict__ max_spread) {
  __shared__ int s_lo, s_hi, s_spread;
  const int tile = blockIdx.x;
  const int SP = S * P;
  if (threadIdx.x == 0) {
    s_lo = INT_MAX;
    s_hi = INT_MIN;
    s_spread = 0;
  }
  __syncthreads();
  const int k0 = tile * KT;
  const int k1 = min(K, k0 + KT);
  for (int sp = threadIdx.x; sp < SP; sp += blockDim.x) {
    const int s = sp / P;
    int lo = INT_MAX, hi = INT_MIN;
    for (int k = k0; k < k1; ++k) {
      if (beta[(size_t)k * S + s] == 0.0f) continue;
      int d = tau[(size_t)k * SP + sp];
      lo = d < lo ? d : lo;
      hi = d > hi ? d : hi;
    }
    tile_lo[(size_t)tile * SP + sp] = lo;
    tile_hi[(size_t)tile * SP + sp] = hi;
    if (lo <= hi) {
      long long spread = (long long)hi - (long long)lo;
      atomicMin(&s_lo, lo);
      atomicMax(&s_hi, hi);
      atomicMax(&s_spread, spread > INT_MAX ? INT_MAX : (int)spread);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    tile_all[2 * tile] = s_lo;
    tile_all[2 * tile + 1] = s_hi;
    atomicMax(max_spread, s_spread);
  }
}

// ---------------------------------------------------------------------------
// Layout and prestack.
// ---------------------------------------------------------------------------

// (S,T,P) phase-interleaved -> (S,P,T) planar.
__global__ void interleaved_to_planar_kernel(const float *__restrict__ in, int S, int T,
                                             int P, float *__restrict__ out) {
  const int s = blockIdx.y;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const float *src = in + ((size_t)s * T + t) * P;
  for (int p = 0; p < P; ++p) out[((size_t)s * P + p) * T + t] = src[p];
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

// ---------------------------------------------------------------------------
// Reduction helpers.
// ---------------------------------------------------------------------------

__global__ void fill_u64_kernel(unsigned long long *__restrict__ p, long long n,
                                unsigned long long v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_f32_kernel(float *__restrict__ p, long long n, float v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
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
    beam_max[t] = float_unorder((uint32_t)(key >> 32));
    beam_argmax[t] = (int)(0xffffffffu - (uint32_t)key);
  } else {
    beam_argmax[t] = 0;
  }
}

// ---------------------------------------------------------------------------
// Shift-and-stack, LDS-staged (the hot kernel).
//
// Work-group = W waves over one source tile (KT = W*KW sources, wave w owns
// KW consecutive sources) and one time tile of TT = 64*R samples (lane l
// owns samples t0 + l + 64 r, r < R, so every LDS gather of a wave reads 64
// consecutive words: conflict-free).  Stations are visited in groups of G;
// the feature windows of a group's G*P (s,p) slots,
//   [t0 + tile_lo[s,p], t0 + TT - 1 + tile_hi[s,p]],
// are staged in LDS (double-buffered: the next group's global loads are in
// flight while the current group is consumed).  Accumulators acc[KW][R]
// live in registers across all groups, so the per-(k,t) FMA chain runs in
// ascending (s, p) order.  Moveouts and weights are wave-uniform scalar loads.
// ---------------------------------------------------------------------------

template <int P, int G, int R, int KW, int W, int NJ, bool NONE>
struct StackLds {
  static constexpr int TT = 64 * R;
  static constexpr int KT = KW * W;
  static constexpr int SLOTS = G * P;
  static constexpr int NT = 64 * W;
  static constexpr int MAX_LMAX = NJ * NT;

  static size_t lds_bytes(int Lmax) {
    return (NONE ? 0 : (size_t)TT * 8) + 2ull * SLOTS * Lmax * sizeof(float);
  }

  __device__ __forceinline__ static void load_group(const StackArgs &a, const int *tlo, const int *thi,
                                    int g, int t0, int tid, float (&pre)[SLOTS][NJ]) {
#pragma unroll
    for (int slot = 0; slot < SLOTS; ++slot) {
      const int si = slot / P, p = slot % P;
      const int s = g * G + si;
      int lo = INT_MAX, len = 0;
      if (s < a.S) {
        lo = tlo[s * P + p];
        if (lo != INT_MAX) len = TT + (thi[s * P + p] - lo);
      }
      const float *src = a.F + ((size_t)s * P + p) * a.T;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int e = tid + j * NT;
        const long long tg = (long long)t0 + lo + e;
        float v = 0.0f;
        if (e < len && tg >= 0 && tg < a.T) v = src[tg];
        pre[slot][j] = v;
      }
    }
  }

  __device__ __forceinline__ static void store_group(float *buf, int Lmax, int tid,
                                     const float (&pre)[SLOTS][NJ]) {
#pragma unroll
    for (int slot = 0; slot < SLOTS; ++slot)
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int e = tid + j * NT;
        if (e < Lmax) buf[slot * Lmax + e] = pre[slot][j];
      }
  }

  // Window positions outside [0, T) are staged as exact zeros, and
  // fma(b, 0, acc) == acc for every finite b (acc is never -0.0), so in
  // flexible mode an out-of-range term drops out exactly as the reference's
  // explicit skip (beamform.c:97-98).  Only a non-finite weight (b*0 = NaN)
  // needs the explicit per-term range test: `edge_flex` marks flexible-mode
  // tiles whose windows cross [0, T), where that slow path can be taken.
  __device__ __forceinline__ static void consume_group(const StackArgs &a, const int *tlo,
                                                       const float *buf, int g, int kbase,
                                                       int t0, int lane, bool edge_flex,
                                                       float (&acc)[KW][R]) {
    const int S = a.S, Lmax = a.Lmax;
#pragma unroll
    for (int kk = 0; kk < KW; ++kk) {
      const int k = kbase + kk;
      if (k < a.k_begin || k >= a.k_end) continue;
#pragma unroll
      for (int si = 0; si < G; ++si) {
        const int s = g * G + si;
        if (s >= S) continue;
        const float b = a.beta[(size_t)k * S + s];
        if (b == 0.0f) continue;  // unused station (beamform.c:61-62)
        if (edge_flex && !__builtin_isfinite(b)) {
#pragma unroll 1
          for (int p = 0; p < P; ++p) {
            const int d = a.tau[((size_t)k * S + s) * P + p];
            const float *x = buf + (si * P + p) * Lmax + (d - tlo[s * P + p]) + lane;
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int tg = t0 + lane + 64 * r + d;
              if (tg >= 0 && tg < a.T) acc[kk][r] = fmaf(b, x[64 * r], acc[kk][r]);
            }
          }
          continue;
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const int d = a.tau[((size_t)k * S + s) * P + p];
          const float *x = buf + (si * P + p) * Lmax + (d - tlo[s * P + p]) + lane;
#pragma unroll
          for (int r = 0; r < R; ++r) acc[kk][r] = fmaf(b, x[64 * r], acc[kk][r]);
        }
      }
    }
  }

  __device__ __forceinline__ static void run(const StackArgs &a, unsigned char *smem) {
    unsigned long long *red = reinterpret_cast<unsigned long long *>(smem);
    float *stage = reinterpret_cast<float *>(smem + (NONE ? 0 : (size_t)TT * 8));
    const int Lmax = a.Lmax;
    const int buf_elems = SLOTS * Lmax;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uniform(tid >> 6);
    const int tt = blockIdx.x % a.n_time_tiles;
    const int tile = a.tile_begin + blockIdx.x / a.n_time_tiles;
    const int t0 = tt * TT;
    const int SP = a.S * P;
    const int *tlo = a.tile_lo + (size_t)tile * SP;
    const int *thi = a.tile_hi + (size_t)tile * SP;
    const int kbase = tile * KT + wave * KW;
    const bool edge_flex =
        a.oob && ((long long)t0 + a.tile_all[2 * tile] < 0 ||
                  (long long)t0 + TT - 1 + a.tile_all[2 * tile + 1] >= a.T);

    if (!NONE)
      for (int i = tid; i < TT; i += NT) red[i] = 0ull;

    float acc[KW][R];
#pragma unroll
    for (int kk = 0; kk < KW; ++kk)
#pragma unroll
      for (int r = 0; r < R; ++r) acc[kk][r] = 0.0f;

    const int n_groups = (a.S + G - 1) / G;
    float pre[SLOTS][NJ];
    load_group(a, tlo, thi, 0, t0, tid, pre);
    store_group(stage, Lmax, tid, pre);
    __syncthreads();
    for (int g = 0; g < n_groups; ++g) {
      const float *cur = stage + (g & 1) * buf_elems;
      float *nxt = stage + ((g + 1) & 1) * buf_elems;
      const bool more = g + 1 < n_groups;
      if (more) load_group(a, tlo, thi, g + 1, t0, tid, pre);
      consume_group(a, tlo, cur, g, kbase, t0, lane, edge_flex, acc);
      if (more) store_group(nxt, Lmax, tid, pre);
      __syncthreads();
    }

    if (NONE) {
#pragma unroll
      for (int kk = 0; kk < KW; ++kk) {
        const int k = kbase + kk;
        if (k < a.k_begin || k >= a.k_end) continue;
        int vlo = 0, vhi = a.T;
        if (!a.oob) strict_range(a.src_lohi[2 * k], a.src_lohi[2 * k + 1], a.T, vlo, vhi);
        float *row = a.beam + (size_t)(k - a.k_begin) * a.T;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int t = t0 + lane + 64 * r;
          if (t >= vlo && t < vhi) row[t] = acc[kk][r];
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int t = t0 + lane + 64 * r;
        float bv = -FLT_MAX;
        int bk = 0;
#pragma unroll
        for (int kk = 0; kk < KW; ++kk) {
          const int k = kbase + kk;
          if (k < a.k_begin || k >= a.k_end) continue;
          int vlo = 0, vhi = a.T;
          if (!a.oob) strict_range(a.src_lohi[2 * k], a.src_lohi[2 * k + 1], a.T, vlo, vhi);
          const float v = acc[kk][r];
          if (t >= vlo && t < vhi && v > bv) {
            bv = v;
            bk = k;
          }
        }
        if (bv > -FLT_MAX) atomicMax(&red[lane + 64 * r], pack_key(bv, bk));
      }
      __syncthreads();
      for (int i = tid; i < TT; i += NT) {
        const unsigned long long key = red[i];
        if (key && t0 + i < a.T) atomicMax(a.keys + t0 + i, key);
      }
    }
  }
};

template <int P, int G, int R, int KW, int W, int NJ, bool NONE>
__global__ __launch_bounds__(64 * W) void stack_lds_kernel(StackArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  StackLds<P, G, R, KW, W, NJ, NONE>::run(a, smem);
}

// ---------------------------------------------------------------------------
// Shift-and-stack, direct gathers from global memory.  Used when the LDS
// kernel does not apply (n_phases outside {1,2}, or a source tile whose delay
// spread exceeds the LDS window).  Same FMA order, so same bits.
// One thread per time sample; blockIdx.y selects a chunk of sources.
// ---------------------------------------------------------------------------
template <bool NONE>
__global__ __launch_bounds__(256) void stack_direct_kernel(StackArgs a, int chunk) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k0 = a.k_begin + blockIdx.y * chunk;
  const int k1 = min(a.k_end, k0 + chunk);
  const int S = a.S, P = a.P, T = a.T;
  float bv = -FLT_MAX;
  int bk = 0;
  for (int k = k0; k < k1; ++k) {
    int vlo = 0, vhi = T;
    if (!a.oob) strict_range(a.src_lohi[2 * k], a.src_lohi[2 * k + 1], T, vlo, vhi);
    if (t >= T || t < vlo || t >= vhi) continue;
    float acc = 0.0f;
    for (int s = 0; s < S; ++s) {
      const float b = a.beta[(size_t)k * S + s];
      if (b == 0.0f) continue;
      for (int p = 0; p < P; ++p) {
        const long long tg = t + a.tau[((size_t)k * S + s) * P + p];
        if (tg < 0 || tg >= T) continue;  // only reachable in flexible mode
        acc = fmaf(b, a.F[((size_t)s * P + p) * T + tg], acc);
      }
    }
    if (NONE) {
      a.beam[(size_t)(k - a.k_begin) * T + t] = acc;
    } else if (acc > bv) {
      bv = acc;
      bk = k;
    }
  }
  if (!NONE && t < T && bv > -FLT_MAX) atomicMax(a.keys + t, pack_key(bv, bk));
}

}  // namespace bp
