// Shared helpers of the yasps_amd HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstddef>

#include "../../include/yasps_amd.h"

#define YSP_CHECK(expr)                                                          \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      ysp_set_error(__FILE__, __LINE__, hipGetErrorString(_e));                  \
      return YSP_ERR_HIP;                                                        \
    }                                                                            \
  } while (0)

#define YSP_CHECK_LAUNCH() YSP_CHECK(hipGetLastError())

namespace ysp {

constexpr int kWave = 64;

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(size_t n, unsigned block) {
  size_t g = (n + block - 1) / block;
  return static_cast<unsigned>(g == 0 ? 1 : g);
}

// Queue a check of a device flag word written by the kernels just launched on s; the
// flag is examined (error `what` when nonzero) at the next ysp_check_deferred(s).
int defer_flag_check(const int* flag, hipStream_t s, const char* what);

// Device scratch that grows on demand and is released by ysp_release_scratch().
// Each slot is an independent buffer so nested users do not alias.
void* scratch(int slot, size_t bytes);

// FP64 atomic add through the hardware global_atomic_add_f64 path.
__device__ __forceinline__ void atomic_add_f64(double* p, double v) {
  unsafeAtomicAdd(p, v);
}
__device__ __forceinline__ void atomic_add_f32(float* p, float v) {
  unsafeAtomicAdd(p, v);
}

// Number of lanes below the calling lane whose bit is set in mask.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Reserve `need` (0..3) consecutive entries of a counter for each active lane of the wave
// with one atomic per wave, in lane order; returns the lane's first entry.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counter, uint32_t need) {
  const uint64_t b0 = __ballot(need & 1u), b1 = __ballot((need >> 1) & 1u);
  const uint64_t active = __ballot(1);
  const uint32_t prefix = lanes_below(b0) + 2u * lanes_below(b1);
  const uint32_t total = __popcll(b0) + 2u * __popcll(b1);
  const int leader = __ffsll(static_cast<unsigned long long>(active)) - 1;
  uint32_t base = 0;
  if (lanes_below(active) == 0 && total) base = atomicAdd(counter, total);
  base = __shfl(base, leader);
  return base + prefix;
}

// Sum over a 64-lane wavefront (all lanes active).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum for blocks of NT threads (multiple of 64); result valid in all threads.
template <int NT, typename T>
__device__ __forceinline__ T block_sum(T v, T* lds /* >= NT/64 entries */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) s += lds[i];
  __syncthreads();
  return s;
}

}  // namespace ysp

extern "C" void ysp_set_error(const char* file, int line, const char* msg);
