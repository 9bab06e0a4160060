// probe_peaks.hip -- measured ceilings of the two pipes the shift-and-stack
// kernel uses on gfx950: LDS ds_read_b128 (conflict-free, 16 KiB per block,
// 8 blocks of 4 waves per CU) and VALU v_fma_f32 (16 independent chains).
// Every LDS index is masked into s[]; the program prints one line per probe.
// `probe_peaks copy` instead runs three copies of 256 MiB (read and written
// once each), the known-size kernel that calibrates the FETCH_SIZE and
// WRITE_SIZE counters, then three streaming reads of 2 GiB: the HBM
// bandwidth a kernel can reach, beside the 8 TB/s of the data sheet.
// Build: python -m beampower_amd.build --probe  (-> beampower_amd/lib/probe_peaks)
#include <hip/hip_runtime.h>
#include <cstdio>
__global__ __launch_bounds__(256) void lds128(float *out, int iters) {
  __shared__ float4 s[1024];  // 16 KiB
  for (int i = threadIdx.x; i < 1024; i += 256) s[i] = make_float4(i & 3, 1, 2, 3);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  unsigned acc = 0;
  unsigned base = 16u * lane;  // + 1024*u below, u < 8 -> max byte 1008 + 7168 + 15 < 16384
  for (int it = 0; it < iters; ++it) {
    const unsigned o = (unsigned)__builtin_amdgcn_readfirstlane((it & 7) * 1024) & 8191u;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const float4 v = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(s) + ((base + o + 1024u * u) & 16383u));
      acc = acc | __float_as_uint(v.x) | __float_as_uint(v.y) | __float_as_uint(v.z) | __float_as_uint(v.w);
    }
  }
  out[blockIdx.x * 256 + threadIdx.x] = __uint_as_float(acc);
}
__global__ __launch_bounds__(256) void valu_fma(float *out, int iters, float b) {
  float a[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = threadIdx.x + i;
  for (int it = 0; it < iters; ++it)
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = fmaf(b, a[i], 0.5f);
  float r = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) r += a[i];
  out[blockIdx.x * 256 + threadIdx.x] = r;
}
__global__ void copy4(const float4 *__restrict__ a, float4 *__restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
// Streaming read: each thread sums `per` float4 a grid-stride apart and
// writes one float (reads dominate: 2 GiB in, 4 MiB out).
__global__ __launch_bounds__(256) void read4(const float4 *__restrict__ a, float *__restrict__ o,
                                             long n) {
  const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  float acc = 0.f;
  for (long i = i0; i < n; i += stride) {
    const float4 v = a[i];
    acc += v.x + v.y + v.z + v.w;
  }
  o[i0] = acc;
}
int copy_calibration() {
  const long bytes = 256l << 20, n = bytes / 16;
  float4 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  (void)hipMemset(a, 0, bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float ms = 0.f;
  for (int i = 0; i < 3; ++i) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(copy4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
  }
  printf("copy4: %ld bytes read, %ld bytes written per launch, %.3f ms, %.2f TB/s read+write (%s)\n",
         bytes, bytes, ms, 2.0 * bytes / ms / 1e9, hipGetErrorString(hipGetLastError()));
  (void)hipFree(a);
  (void)hipFree(b);
  // read bandwidth: 2 GiB, 8 blocks of 256 threads per CU
  const long rbytes = 2048l << 20, rn = rbytes / 16;
  const unsigned rblocks = 256 * 8;
  float *o;
  if (hipMalloc(&a, rbytes) != hipSuccess || hipMalloc(&o, (size_t)rblocks * 256 * 4) != hipSuccess)
    return 1;
  (void)hipMemset(a, 0, rbytes);
  for (int i = 0; i < 3; ++i) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(read4, dim3(rblocks), dim3(256), 0, 0, a, o, rn);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
  }
  printf("read4: %ld bytes read per launch, %.3f ms, %.2f TB/s (%s)\n", rbytes, ms,
         rbytes / ms / 1e9, hipGetErrorString(hipGetLastError()));
  (void)hipFree(a);
  (void)hipFree(o);
  return 0;
}
int main(int argc, char **argv) {
  if (argc > 1 && argv[1][0] == 'c') return copy_calibration();
  float *out; const int blocks = 256 * 8; (void)hipMalloc(&out, (size_t)blocks * 256 * 4);
  hipEvent_t a, b; (void)hipEventCreate(&a); (void)hipEventCreate(&b); float ms;
  for (int rep = 0; rep < 2; ++rep) {
    const int it = 2048;
    hipLaunchKernelGGL(lds128, dim3(blocks), dim3(256), 0, 0, out, 16);
    (void)hipEventRecord(a); hipLaunchKernelGGL(lds128, dim3(blocks), dim3(256), 0, 0, out, it); (void)hipEventRecord(b); (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(&ms, a, b);
    double bytes = (double)blocks * 256 * it * 8 * 16;
    printf("lds b128 (8 WG/CU x 4 waves): %.3f ms %.1f TB/s %.1f B/clk/CU (%s)\n", ms, bytes / ms / 1e9, bytes / (ms * 1e-3) / 2.4e9 / 256, hipGetErrorString(hipGetLastError()));
    const int vi = 8192;
    hipLaunchKernelGGL(valu_fma, dim3(blocks), dim3(256), 0, 0, out, 16, 0.999f);
    (void)hipEventRecord(a); hipLaunchKernelGGL(valu_fma, dim3(blocks), dim3(256), 0, 0, out, vi, 0.999f); (void)hipEventRecord(b); (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(&ms, a, b);
    double fmas = (double)blocks * 256 * vi * 16;
    printf("valu v_fma_f32: %.3f ms %.2f TFMA/s = %.1f TFLOP/s (%.1f FMA/clk/CU)\n", ms, fmas / ms / 1e9, 2 * fmas / ms / 1e9, fmas / (ms * 1e-3) / 2.4e9 / 256);
  }
  (void)hipFree(out);
  return 0;
}
