// Throughput calibration kernels: the sustained lane rate of the VALU operations the field
// arithmetic is built from, measured on the device that runs the prover. The INT32 roofline
// of the wide (128/256-bit) fields is priced against the measured v_mad_u64_u32 rate.
//
// Each thread runs CHAINS independent dependency chains for `iters` iterations, so the
// measured rate is issue-bound rather than latency-bound. Results are stored so no chain is
// dead code.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "calib.hpp"

namespace sc {

constexpr int CHAINS = 8;

__global__ void __launch_bounds__(256) calib_mad_u64(uint32_t b, int iters, uint64_t* out) {
  uint64_t x[CHAINS];
#pragma unroll
  for (int k = 0; k < CHAINS; ++k) x[k] = threadIdx.x + blockIdx.x * 7919u + k;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < CHAINS; ++k) x[k] = (uint64_t)(uint32_t)x[k] * b + x[k];
  }
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < CHAINS; ++k) s ^= x[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) calib_mul_hi_u32(uint32_t b, int iters, uint32_t* out) {
  uint32_t x[CHAINS];
#pragma unroll
  for (int k = 0; k < CHAINS; ++k) x[k] = threadIdx.x * 31u + k + 1u;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < CHAINS; ++k) x[k] = __umulhi(x[k], b) + x[k];
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < CHAINS; ++k) s ^= x[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) calib_add_co(uint32_t b, int iters, uint32_t* out) {
  uint64_t x[CHAINS];
#pragma unroll
  for (int k = 0; k < CHAINS; ++k) x[k] = threadIdx.x + k;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < CHAINS; ++k) x[k] = x[k] + x[(k + 3) % CHAINS] + b;  // 2x (add_co + addc)
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < CHAINS; ++k) s ^= (uint32_t)x[k] ^ (uint32_t)(x[k] >> 32);
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) calib_fma_f64(double b, int iters, double* out) {
  double x[CHAINS];
#pragma unroll
  for (int k = 0; k < CHAINS; ++k) x[k] = (double)(threadIdx.x + k);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < CHAINS; ++k) x[k] = __fma_rn(x[k], b, 0.5);
  }
  double s = 0;
#pragma unroll
  for (int k = 0; k < CHAINS; ++k) s += x[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Streaming HBM probes over `words` uint4 elements: a read-only XOR reduction and a copy.
__global__ void __launch_bounds__(256) calib_read(const uint4* __restrict__ src, uint64_t words, uint32_t* out) {
  uint32_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (uint64_t)gridDim.x * 256) {
    const uint4 v = src[i];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  out[blockIdx.x * 256 + threadIdx.x] = acc;
}

__global__ void __launch_bounds__(256) calib_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t words) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (uint64_t)gridDim.x * 256) dst[i] = src[i];
}

// Two 16-byte words read per 16-byte word written, contiguous per thread: the traffic shape
// of a fold round (4 entries read, 2 written, per pair and table).
__global__ void __launch_bounds__(256) calib_mix21(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                  uint64_t words_out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < words_out; i += (uint64_t)gridDim.x * 256) {
    const uint4 x = src[2 * i], y = src[2 * i + 1];
    dst[i] = make_uint4(x.x ^ y.x, x.y ^ y.y, x.z ^ y.z, x.w ^ y.w);
  }
}

hipError_t calibrate_bandwidth(hipStream_t s, double* rates, int n) {
  const uint64_t bytes = (uint64_t)1 << 31;  // 2 GiB per buffer: far beyond the 256 MB last-level cache
  const uint64_t words = bytes / 16;
  void* src = nullptr;
  void* dst = nullptr;
  hipError_t e = hipMalloc(&src, bytes);
  if (e == hipSuccess) e = hipMalloc(&dst, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(src, 1, bytes, s);
  hipEvent_t a = nullptr, b = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  // The best over launch shapes: on MI355X a 1024-block grid streams copies and 2:1 mixes
  // about 15% faster than 4096 blocks (5.5 vs 4.7 TB/s), reads peak near 2048.
  const int grids[4] = {1024, 2048, 4096, 16384};
  for (int k = 0; k < n && k < 3 && e == hipSuccess; ++k) {
    float best = 1e30f;
    for (int rep = 0; rep < 4 * 4 && e == hipSuccess; ++rep) {
      const int blocks = grids[rep / 4];
      e = hipEventRecord(a, s);
      if (e != hipSuccess) break;
      if (k == 0) {
        hipLaunchKernelGGL(calib_read, dim3(blocks), dim3(256), 0, s, (const uint4*)src, words, (uint32_t*)dst);
      } else if (k == 1) {
        hipLaunchKernelGGL(calib_copy, dim3(blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, words);
      } else {
        hipLaunchKernelGGL(calib_mix21, dim3(blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, words / 2);
      }
      e = hipGetLastError();
      if (e == hipSuccess) e = hipEventRecord(b, s);
      if (e == hipSuccess) e = hipEventSynchronize(b);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
      if (e == hipSuccess && rep % 4 > 0 && ms < best) best = ms;
    }
    // bytes moved per second: read probe reads `bytes`, copy reads and writes them, the 2:1
    // probe reads them and writes half as many
    if (e == hipSuccess) rates[k] = (k == 0 ? 1.0 : (k == 1 ? 2.0 : 1.5)) * (double)bytes / (best * 1e-3);
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return e;
}

hipError_t calibrate(hipStream_t s, double* rates, int n) {
  // grid: 8 blocks per CU of a 256-CU device, 256 threads each
  const int blocks = 2048, threads = 256, iters = 4096;
  void* buf = nullptr;
  hipError_t e = hipMalloc(&buf, (size_t)blocks * threads * 8);
  if (e != hipSuccess) return e;
  hipEvent_t a = nullptr, b = nullptr;
  e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  for (int k = 0; k < n && k < CALIB_HBM_READ && e == hipSuccess; ++k) {
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {  // rep 0 warms up
      e = hipEventRecord(a, s);
      if (e != hipSuccess) break;
      switch (k) {
        case CALIB_MAD_U64:
          hipLaunchKernelGGL(calib_mad_u64, dim3(blocks), dim3(threads), 0, s, 0x9E3779B9u, iters, (uint64_t*)buf);
          break;
        case CALIB_MUL_HI_U32:
          hipLaunchKernelGGL(calib_mul_hi_u32, dim3(blocks), dim3(threads), 0, s, 0x9E3779B9u, iters, (uint32_t*)buf);
          break;
        case CALIB_ADD_CO:
          hipLaunchKernelGGL(calib_add_co, dim3(blocks), dim3(threads), 0, s, 0x9E3779B9u, iters, (uint32_t*)buf);
          break;
        default:
          hipLaunchKernelGGL(calib_fma_f64, dim3(blocks), dim3(threads), 0, s, 1.0000001, iters, (double*)buf);
          break;
      }
      e = hipGetLastError();
      if (e == hipSuccess) e = hipEventRecord(b, s);
      if (e == hipSuccess) e = hipEventSynchronize(b);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
      // the 64-bit add kernel does two adds per chain step
      const double per_step = k == CALIB_ADD_CO ? 2.0 : 1.0;
      if (e == hipSuccess) rates[k] = per_step * blocks * threads * iters * CHAINS / (ms * 1e-3);
    }
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipFree(buf);
  if (e == hipSuccess && n > CALIB_HBM_READ) e = calibrate_bandwidth(s, rates + CALIB_HBM_READ, n - CALIB_HBM_READ);
  return e;
}

}  // namespace sc
