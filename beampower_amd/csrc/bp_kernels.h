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
  const int *tile_all;    // (n_tiles, 8) min lo, max hi, max lo, max (hi - lo), non-finite weight flag
  const int *sched;       // (n_tiles, S, KT, P+1) LDS byte offsets and weight bits
  const int *tiles;       // source tiles processed by this launch
  int n_list;             // number of entries of `tiles`
  int T, K, S, P;
  int oob;                // 0 strict, 1 flexible
  int Lw;                 // LDS window length per (s,p) copy, multiple of 4
  int n_time_tiles;
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
    o[0] = s_lo;
    o[1] = s_hi;
    o[2] = s_maxlo;
    o[3] = s_spread;
    o[4] = s_nonfinite;
    o[5] = o[6] = o[7] = 0;
  }
}

// Per-tile schedule of the LDS kernel.  For tile j, station s and source
// slot kk (k = j*KT + kk): P LDS byte offsets, then the weight's bits.  The
// moveout d of phase p lies e = d - tile_lo[j,s,p] samples into the window;
// the window is kept as 4 copies shifted by 0..3 samples, so the 4 samples
// starting at (lane sample + e) sit 16-byte aligned in copy e&3, 16*(e>>2)
// bytes past the lane's own offset.  A copy holds CS = Lw rounded up to 256
// floats; slot p starts at p*16*CS bytes and copy c 4*CS*c bytes into it.
// Sources past K and unused stations get weight 0 and the offset P*16*CS of
// the all-zero block that follows the slots, so the kernel needs no branch
// for them: fma(0, 0, acc) == acc.
// Tiles served by the direct kernel (Lw = 0) get offsets 0.
__global__ void schedule_kernel(const int *__restrict__ tau, const float *__restrict__ beta,
                                const int *__restrict__ tile_lo,
                                const int *__restrict__ tile_lw, int K, int S, int P,
                                int KT, long long n, int *__restrict__ sched) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int kk = (int)(idx % KT);
  const long long js = idx / KT;
  const int s = (int)(js % S);
  const int tile = (int)(js / S);
  const int k = tile * KT + kk;
  const int Lw = tile_lw[tile];
  const float b = k < K ? beta[(size_t)k * S + s] : 0.0f;
  int *out = sched + idx * (P + 1);
  for (int p = 0; p < P; ++p) {
    int off = 0;
    if (Lw > 0) {
      const int CS = (Lw + 255) & ~255;  // copy stride, whole 1 KiB rows
      off = P * 16 * CS;
      if (b != 0.0f) {
        const int e =
            tau[((size_t)k * S + s) * P + p] - tile_lo[((size_t)tile * S + s) * P + p];
        off = p * 16 * CS + (e & 3) * 4 * CS + 16 * (e >> 2);
      }
    }
    out[p] = off;
  }
  out[P] = __float_as_int(b);
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
// A work-group of W waves covers one source tile (KT = W*KW sources; wave w
// owns KW of them) and one time tile of TT = 256*R4 samples.  Lane l owns
// the 4-sample runs starting at t0 + 4l + 256r (r < R4), so a wave's gather
// of one (source, station, phase) term is R4 ds_read_b128 of 64 consecutive
// 16-byte words: conflict-free, and 4 samples per LDS lane-access, which is
// what the gfx950 LDS pipe needs to reach its bandwidth (it serves 16 lanes
// per clock at any width up to 16 bytes).
//
// Stations are visited one at a time (ascending, so every (k,t) FMA chain
// runs in ascending (s, p) order).  For each of the station's P phases the
// window [t0 + lo, t0 + lo + Lw + 3) of planar features is staged into LDS
// as 4 copies shifted by 0..3 samples, double-buffered: the next station's
// global loads are in flight while the current one is consumed.  Positions
// outside [0, T) are staged as exact zeros.  Accumulators acc[KW][4*R4]
// stay in registers for the whole tile.  The per-(source, station) LDS
// offsets and weights come from the precomputed schedule as wave-uniform
// scalar loads.
// ---------------------------------------------------------------------------

// Block -> (time tile, listed tile) with XCD locality: work-groups are
// dispatched round-robin over the 8 XCDs, so block b runs on XCD b % 8;
// XCD x is given the contiguous range [x*N/8, (x+1)*N/8) of the time-major
// order, so the feature windows of a time tile stay in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int n8 = n & ~7;
  if (b >= n8) return b;
  return (b & 7) * (n8 >> 3) + (b >> 3);
}

typedef __attribute__((address_space(1))) void global_void;
typedef __attribute__((address_space(3))) void lds_void;
// Read-only per-network tables are read through the constant address space:
// wave-uniform loads from it are always scalar (s_load), whatever the
// compiler can prove about intervening memory operations.
typedef const __attribute__((address_space(4))) int const_int;

__device__ __forceinline__ const_int *as_const(const int *p) {
  return (const_int *)(p);
}

// Copy stride of a window of Lw samples: whole 1 KiB chunks (256 floats),
// so staging always moves complete 64-lane x 16-byte rows.
__host__ __device__ __forceinline__ int copy_stride(int Lw) { return (Lw + 255) & ~255; }

template <int P, int R4, int KW, int W, int MAXLW, bool NONE>
struct StackLds {
  static_assert(P == 1 || P == 2, "the LDS kernel takes one or two phases");
  static constexpr int TT = 256 * R4;
  static constexpr int KT = KW * W;
  static constexpr int NT = 64 * W;
  static constexpr int NACC = 4 * R4;
  static constexpr int MAX_LW = MAXLW;         // largest window this instantiation stages
  static constexpr int NU = (MAX_LW + NT - 1) / NT;  // edge-path staging units per thread
  static constexpr int ZERO_BYTES = 16 * 64 * R4;  // all-zero block read by unused terms

  // LDS of one buffer: P slots of 4 copies (stride copy_stride(Lw) floats),
  // then the zero block.
  __host__ __device__ static size_t buf_bytes(int Lw) {
    return (size_t)P * 16 * copy_stride(Lw) + ZERO_BYTES;
  }
  static size_t lds_bytes(int Lw) { return (NONE ? 0 : (size_t)TT * 8) + 2 * buf_bytes(Lw); }

  // Interior tiles: the station's windows go global -> LDS directly
  // (global_load_lds_dwordx4, no registers).  Row (p, c, j) of a station is
  // copy c of phase p, quads [64j, 64j + 64): lane l loads
  // x[w0 + 4m + c .. +3], m = 64j + l (clamped to the window's last quad, so
  // the padding past the window gets a harmless duplicate), into LDS at
  // slot p + copy c + 16m.  Row r belongs to wave r % W.  The rows of a wave
  // are the same for every station, so their LDS offsets and per-lane
  // element offsets are computed once (RowPlan).
  static constexpr int NR = (P * 4 * (((MAXLW + 255) & ~255) >> 8) + W - 1) / W;

  struct RowPlan {
    int n;          // rows of this wave
    int p[NR];      // phase of each row (wave-uniform)
    int lds[NR];    // LDS byte offset of each row in a buffer (wave-uniform)
    unsigned eoff[NR];  // this lane's byte offset 4*(4m + c) within the window
  };

  __device__ __forceinline__ static void plan_rows(int Lw, int wave, int lane, RowPlan &rp) {
    const int CS = copy_stride(Lw);
    const int nch = CS >> 8;
    const int last_quad = (Lw >> 2) - 1;
    const int rows = P * 4 * nch;
    rp.n = 0;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const int row = wave + i * W;
      const int rr = row < rows ? row : 0;
      const int p = rr / (4 * nch);
      const int rem = rr - p * 4 * nch;
      const int c = rem / nch;
      const int j = rem - c * nch;
      rp.p[i] = p;
      rp.lds[i] = p * 16 * CS + c * 4 * CS + 1024 * j;
      rp.eoff[i] = 4u * (unsigned)(4 * min(64 * j + lane, last_quad) + c);
      rp.n += row < rows ? 1 : 0;
    }
  }

  __device__ __forceinline__ static void dma_station(const StackArgs &a, const_int *tlo, int s,
                                                     int t0, const RowPlan &rp,
                                                     unsigned char *__restrict__ buf) {
    int lo[P];
    const float *plane[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      lo[p] = tlo[s * P + p];
      plane[p] = a.F + ((size_t)s * P + p) * a.T + (t0 + lo[p]);
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      if (i >= rp.n) break;
      const int p = rp.p[i];
      const int lop = P == 1 ? lo[0] : (p == 0 ? lo[0] : lo[P - 1]);
      if (lop == INT_MAX) continue;  // no source of the tile uses station s
      const char *base =
          reinterpret_cast<const char *>(P == 1 ? plane[0] : (p == 0 ? plane[0] : plane[P - 1]));
      __builtin_amdgcn_global_load_lds((global_void *)(base + rp.eoff[i]),
                                       (lds_void *)(buf + rp.lds[i]), 16, 0, 0);
    }
  }

  // Edge tiles (a window crosses [0, T)): guarded loads through registers,
  // zeros outside [0, T).  Unit u = c*(Lw/4) + m of slot p.
  __device__ __forceinline__ static void stage_station_guarded(const StackArgs &a,
                                                               const_int *tlo, int s, int t0,
                                                               int tid, unsigned char *buf) {
    const int Lw = a.Lw, CS = copy_stride(Lw), quads = Lw >> 2;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int lo = tlo[s * P + p];
      if (lo == INT_MAX) continue;
      const float *plane = a.F + ((size_t)s * P + p) * a.T;
      const long long w0 = (long long)t0 + lo;
      for (int u = tid; u < Lw; u += NT) {
        const int c = u / quads, m = u - c * quads;
        const long long t = w0 + 4 * m + c;
        float4 v;
        v.x = (t + 0 >= 0 && t + 0 < a.T) ? plane[t + 0] : 0.f;
        v.y = (t + 1 >= 0 && t + 1 < a.T) ? plane[t + 1] : 0.f;
        v.z = (t + 2 >= 0 && t + 2 < a.T) ? plane[t + 2] : 0.f;
        v.w = (t + 3 >= 0 && t + 3 < a.T) ? plane[t + 3] : 0.f;
        *reinterpret_cast<float4 *>(buf + p * 16 * CS + c * 4 * CS + 16 * m) = v;
      }
    }
  }

  // Operands of one station for this wave's KW sources: KW*(P+1)
  // wave-uniform words (P byte offsets and the weight per source).
  __device__ __forceinline__ static void load_sched(const_int *sg, int (&sv)[KW * (P + 1)]) {
#pragma unroll
    for (int i = 0; i < KW * (P + 1); ++i) sv[i] = sg[i];
  }

  __device__ __forceinline__ static void gather(const unsigned char *buf, unsigned lb,
                                                const int *svk, float4 (&v)[P][R4]) {
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int r = 0; r < R4; ++r)
        v[p][r] = *reinterpret_cast<const float4 *>(buf + (lb + (unsigned)svk[p]) + 1024 * r);
  }

  // One station: straight-line gathers and FMAs, software-pipelined by one
  // source (the gathers of source kk+1 are in flight during the FMAs of kk).
  // A fence after each source keeps the compiler from hoisting every gather
  // of the station at once, which would exhaust the registers: the empty asm
  // statements take the source's accumulators (its FMAs complete there) and
  // the lane's LDS offset (later gathers depend on it, so they stay behind).
  __device__ __forceinline__ static void consume_station(const int (&sv)[KW * (P + 1)],
                                                         const unsigned char *buf, int lane,
                                                         float (&acc)[KW][NACC]) {
    unsigned lb = 16u * (unsigned)lane;
    float4 v[2][P][R4];
    gather(buf, lb, sv, v[0]);
#pragma unroll
    for (int kk = 0; kk < KW; ++kk) {
      const int c = kk & 1;
      if (kk + 1 < KW) gather(buf, lb, sv + (kk + 1) * (P + 1), v[c ^ 1]);
      const float b = __int_as_float(sv[kk * (P + 1) + P]);
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int r = 0; r < R4; ++r) {
          acc[kk][4 * r + 0] = fmaf(b, v[c][p][r].x, acc[kk][4 * r + 0]);
          acc[kk][4 * r + 1] = fmaf(b, v[c][p][r].y, acc[kk][4 * r + 1]);
          acc[kk][4 * r + 2] = fmaf(b, v[c][p][r].z, acc[kk][4 * r + 2]);
          acc[kk][4 * r + 3] = fmaf(b, v[c][p][r].w, acc[kk][4 * r + 3]);
        }
#pragma unroll
      for (int j = 0; j < NACC; ++j) asm volatile("" : "+v"(acc[kk][j]));
      asm volatile("" : "+v"(lb));
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  // One pipeline step: start the next station's global->LDS copies into
  // `nxt`, then consume the current station from `cur`.  Both buffers are
  // restrict-qualified in one scope, so the compiler knows the copies cannot
  // alias the gathers and does not wait for them before reading `cur`.
  __device__ __forceinline__ static void station_step(const StackArgs &a, const_int *tlo,
                                                      int s_next, int t0, bool dma,
                                                      const RowPlan &rp,
                                                      const int (&sv)[KW * (P + 1)],
                                                      const unsigned char *__restrict__ cur,
                                                      unsigned char *__restrict__ nxt,
                                                      int lane, float (&acc)[KW][NACC]) {
    if (dma) dma_station(a, tlo, s_next, t0, rp, nxt);
    consume_station(sv, cur, lane, acc);
  }

  __device__ __forceinline__ static void run(const StackArgs &a, unsigned char *smem) {
    const int Lw = a.Lw;
    const int CS = copy_stride(Lw);
    unsigned long long *red = reinterpret_cast<unsigned long long *>(smem);
    unsigned char *stage = smem + (NONE ? 0 : TT * 8);
    const int bbytes = (int)buf_bytes(Lw);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uniform(tid >> 6);
    const int L = xcd_remap(blockIdx.x, a.n_time_tiles * a.n_list);
    const int tt = L / a.n_list;
    const int tile = as_const(a.tiles)[L - tt * a.n_list];
    const int t0 = tt * TT;
    const int S = a.S;
    const_int *tlo = as_const(a.tile_lo) + (size_t)tile * S * P;
    const_int *tall = as_const(a.tile_all) + 8 * tile;
    const int kbase = tile * KT + wave * KW;
    // staged range of every slot: [t0 + lo, t0 + lo + Lw + 3)
    const bool edge = (long long)t0 + tall[0] < 0 ||
                      (long long)t0 + tall[2] + Lw + 3 > (long long)a.T;

    if (!NONE)
      for (int i = tid; i < TT; i += NT) red[i] = 0ull;
    for (int i = tid; i < ZERO_BYTES / 16; i += NT) {
      const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
      *reinterpret_cast<float4 *>(stage + P * 16 * CS + 16 * i) = z;
      *reinterpret_cast<float4 *>(stage + bbytes + P * 16 * CS + 16 * i) = z;
    }

    float acc[KW][NACC];
#pragma unroll
    for (int kk = 0; kk < KW; ++kk)
#pragma unroll
      for (int j = 0; j < NACC; ++j) acc[kk][j] = 0.0f;

    // schedule rows of this wave: station s at sched_w + s * KT * (P + 1)
    const_int *sched_w = as_const(a.sched) + ((size_t)tile * S * KT + wave * KW) * (P + 1);
    int sv[KW * (P + 1)];
    load_sched(sched_w, sv);
    RowPlan rp;
    plan_rows(Lw, wave, lane, rp);
    if (edge)
      stage_station_guarded(a, tlo, 0, t0, tid, stage);
    else
      dma_station(a, tlo, 0, t0, rp, stage);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int s = 0; s < S; ++s) {
      const unsigned char *cur = stage + (s & 1) * bbytes;
      unsigned char *nxt = stage + ((s + 1) & 1) * bbytes;
      const bool more = s + 1 < S;
      station_step(a, tlo, s + 1, t0, more && !edge, rp, sv, cur, nxt, lane, acc);
      if (more) {
        // next station's operands: the scalar loads overlap the barrier
        load_sched(sched_w + (size_t)(s + 1) * KT * (P + 1), sv);
        if (edge) stage_station_guarded(a, tlo, s + 1, t0, tid, nxt);
      }
      // the station's LDS-DMA writes must land before the barrier
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
      __syncthreads();
    }

    if (NONE) {
#pragma unroll
      for (int kk = 0; kk < KW; ++kk) {
        const int k = kbase + kk;
        if (k < a.k_begin || k >= a.k_end) continue;
        int vlo = 0, vhi = a.T;
        if (!a.oob)
          strict_range(as_const(a.src_lohi)[2 * k], as_const(a.src_lohi)[2 * k + 1], a.T, vlo,
                       vhi);
        float *row = a.beam + (size_t)(k - a.k_begin) * a.T;
#pragma unroll
        for (int j = 0; j < NACC; ++j) {
          const int t = t0 + 4 * lane + 256 * (j >> 2) + (j & 3);
          if (t >= vlo && t < vhi) row[t] = acc[kk][j];
        }
      }
    } else {
      float bv[NACC];
      int bk[NACC];
#pragma unroll
      for (int j = 0; j < NACC; ++j) {
        bv[j] = -FLT_MAX;
        bk[j] = 0;
      }
#pragma unroll
      for (int kk = 0; kk < KW; ++kk) {
        const int k = kbase + kk;
        if (k < a.k_begin || k >= a.k_end) continue;
        int vlo = 0, vhi = a.T;
        if (!a.oob)
          strict_range(as_const(a.src_lohi)[2 * k], as_const(a.src_lohi)[2 * k + 1], a.T, vlo,
                       vhi);
#pragma unroll
        for (int j = 0; j < NACC; ++j) {
          const int t = t0 + 4 * lane + 256 * (j >> 2) + (j & 3);
          const float v = acc[kk][j];
          if (t >= vlo && t < vhi && v > bv[j]) {
            bv[j] = v;
            bk[j] = k;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < NACC; ++j)
        if (bv[j] > -FLT_MAX)
          atomicMax(&red[4 * lane + 256 * (j >> 2) + (j & 3)], pack_key(bv[j], bk[j]));
      __syncthreads();
      for (int i = tid; i < TT; i += NT) {
        const unsigned long long key = red[i];
        if (key && t0 + i < a.T) atomicMax(a.keys + t0 + i, key);
      }
    }
  }
};

// At least 4 waves per SIMD (VGPRs <= 128): two work-groups per CU, so one
// computes while the other waits at its per-station barrier.
template <int P, int R4, int KW, int W, int MAXLW, bool NONE>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4)))
void stack_lds_kernel(StackArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  StackLds<P, R4, KW, W, MAXLW, NONE>::run(a, smem);
}

// ---------------------------------------------------------------------------
// Shift-and-stack, direct gathers from global memory.  Used for source tiles
// the LDS kernel does not take (n_phases outside {1,2}, or a tile whose
// delay spread exceeds the largest LDS window).  Same FMA order, same bits.
// One thread per time sample; blockIdx.y selects a listed source tile of
// `chunk` sources.
// ---------------------------------------------------------------------------
template <bool NONE>
__global__ __launch_bounds__(256) void stack_direct_kernel(StackArgs a, int chunk) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long tile = a.tiles[blockIdx.y];
  const int k0 = (int)max((long long)a.k_begin, tile * chunk);
  const int k1 = (int)min((long long)a.k_end, tile * chunk + chunk);
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
