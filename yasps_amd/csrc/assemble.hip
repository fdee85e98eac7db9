// Atomic-free assembly: element kernels write their oriented 3x3 blocks (one per
// coordinate) and local gradients to staging buffers; these kernels sum every unique
// block and every gradient segment over its contributions in a fixed order.
//
// The contribution lists are the inverse of the dedupe map: a stable sort of
// (unique ordinal, coordinate) pairs groups the coordinates of each unique block in
// ascending coordinate order, so the sum order - and the assembled bits - do not depend
// on thread scheduling.
#include "common.hpp"

#include <hipcub/hipcub.hpp>

namespace ysp {
namespace {

constexpr unsigned kB = 256;

__global__ void iota_kernel(uint32_t* out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

// Run starts of the sorted keys (keys >= num_keys sort last and are dropped): entry k
// starts every key after the previous entry's key up to its own (absent keys included);
// the first dropped entry (or k = n) closes the remaining keys up to num_keys.
__global__ void run_starts_kernel(const uint32_t* __restrict__ sorted, uint32_t n,
                                  uint32_t num_keys, uint32_t* __restrict__ offsets) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > n) return;
  const uint32_t key = k < n ? sorted[k] : num_keys;
  const uint32_t rk = key < num_keys ? key : num_keys;
  uint32_t first = 0;
  if (k > 0) {
    const uint32_t prev = sorted[k - 1];
    first = prev < num_keys ? prev + 1 : num_keys + 1;
  }
  for (uint32_t r = first; r <= rk; ++r) offsets[r] = k;
}

// One slot per (element, slot): key = vertex of the gradient segment (or ~0u when the slot
// does not reach the gradient), value = its offset in the gradient staging buffer.
__global__ void gradient_contrib_kernel(const uint32_t* __restrict__ idx,
                                        const uint16_t* __restrict__ sizes, uint32_t n, uint32_t K,
                                        uint64_t staging_base, uint32_t grad_local,
                                        uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                        int* __restrict__ error_flag) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  uint32_t col = 0;
  for (uint32_t i = 0; i < K; ++i) {
    const size_t o = static_cast<size_t>(e) * K + i;
    const uint32_t ii = idx[o];
    const uint32_t sz = sizes[o];
    if (ii >= 2 && sz != 3) atomicOr(error_flag, 1);
    const uint64_t pos = staging_base + static_cast<uint64_t>(e) * grad_local + col;
    if (ii >= 2 && (ii - 2) % 3 == 0 && pos < 0xffffffffull) {
      keys[o] = (ii - 2) / 3;
      vals[o] = static_cast<uint32_t>(pos);
    } else {
      keys[o] = 0xffffffffu;
      vals[o] = 0xffffffffu;
      if (ii >= 2) atomicOr(error_flag, 1);
    }
    col += sz;
  }
}

// Lanes (unit, entry): each wavefront holds kUnitsPerWave whole units (global blocks of 9
// entries, or gradient vertices of 3) with one lane per entry, so that the entries of a
// contribution are read as one contiguous run and the result is written contiguously.
template <int W>
struct LaneUnit {
  static constexpr uint32_t kUnitsPerWave = 64 / W;
  uint32_t unit;
  int entry;
  bool valid;
  __device__ LaneUnit(uint32_t count) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = tid % 64, slot = lane / W;
    unit = (tid / 64) * kUnitsPerWave + slot;
    entry = static_cast<int>(lane % W);
    valid = slot < kUnitsPerWave && unit < count;
  }
  static unsigned grid(uint32_t count) {
    const size_t waves = (static_cast<size_t>(count) + kUnitsPerWave - 1) / kUnitsPerWave;
    return grid_for(waves * 64, kB);
  }
};

// Sum of the staged contributions of every global block (contributions in their sorted
// order, so the sums are reproducible), and its copy into the diagonal blocks.
__global__ void __launch_bounds__(kB)
    gather_blocks3_kernel(const double* __restrict__ staging, const uint32_t* __restrict__ offsets,
                          const uint32_t* __restrict__ items, uint32_t nu,
                          const uint32_t* __restrict__ positions, double* __restrict__ values,
                          double* __restrict__ diag_blocks, const uint32_t* __restrict__ diag_start,
                          const uint32_t* __restrict__ grad_start, uint32_t num_attr,
                          int accumulate_diag) {
  const LaneUnit<9> lu(nu);
  if (!lu.valid) return;
  const uint32_t u = lu.unit;
  const int j = lu.entry;
  double acc = 0.0;
  const uint32_t k1 = offsets[u + 1];
#pragma unroll 4
  for (uint32_t k = offsets[u]; k < k1; ++k) acc += staging[9 * static_cast<size_t>(items[k]) + j];
  values[9 * static_cast<size_t>(u) + j] = acc;
  const uint32_t r = positions[2 * u], c = positions[2 * u + 1];
  if (r == c && diag_blocks) {
    uint32_t a = 0;
    while (a < num_attr && r >= grad_start[a + 1]) ++a;
    if (a < num_attr) {
      double* d = diag_blocks + diag_start[a] + ((r - grad_start[a]) / 3) * 9 + j;
      *d = accumulate_diag ? *d + acc : acc;
    }
  }
}

__global__ void __launch_bounds__(kB)
    gather_gradient3_kernel(const double* __restrict__ staging, const uint32_t* __restrict__ offsets,
                            const uint32_t* __restrict__ items, uint32_t nv,
                            double* __restrict__ gradient) {
  const LaneUnit<3> lu(nv);
  if (!lu.valid) return;
  const uint32_t v = lu.unit;
  double g = 0.0;
  const uint32_t k1 = offsets[v + 1];
#pragma unroll 4
  for (uint32_t k = offsets[v]; k < k1; ++k) g += staging[items[k] + lu.entry];
  gradient[3 * static_cast<size_t>(v) + lu.entry] = g;
}

}  // namespace
}  // namespace ysp

using namespace ysp;

extern "C" int ysp_inverse_map(const uint32_t* keys, const uint32_t* values, uint32_t n,
                               uint32_t num_keys, uint32_t* offsets, uint32_t* items,
                               void* stream) {
  hipStream_t s = as_stream(stream);
  if (num_keys == 0) return YSP_OK;
  if (n == 0) {
    YSP_CHECK(hipMemsetAsync(offsets, 0, sizeof(uint32_t) * (num_keys + 1), s));
    return YSP_OK;
  }
  uint32_t* k_sorted = static_cast<uint32_t*>(scratch(50, sizeof(uint32_t) * n));
  uint32_t* v_in = static_cast<uint32_t*>(scratch(51, sizeof(uint32_t) * n));
  uint32_t* v_out = static_cast<uint32_t*>(scratch(52, sizeof(uint32_t) * n));
  if (!k_sorted || !v_in || !v_out) return YSP_ERR_HIP;
  if (values) {
    YSP_CHECK(hipMemcpyAsync(v_in, values, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  } else {
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n, kB)), dim3(kB), 0, s, v_in, n);
  }
  // keys equal to ~0u (no contribution) must sort after every real key: all 32 bits
  const int end_bit = 32;
  size_t temp = 0;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys, k_sorted, v_in, v_out, n, 0,
                                               end_bit, s));
  void* tmp = scratch(54, temp);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, temp, keys, k_sorted, v_in, v_out, n, 0,
                                               end_bit, s));
  YSP_CHECK(hipMemcpyAsync(items, v_out, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(run_starts_kernel, dim3(grid_for(static_cast<size_t>(n) + 1, kB)), dim3(kB), 0,
                     s, k_sorted, n, num_keys, offsets);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_gradient_contributions(const uint32_t* segment_indices,
                                          const uint16_t* segment_sizes, uint32_t num_instances,
                                          uint32_t K, uint64_t staging_base,
                                          uint32_t gradient_local_size, uint32_t* keys,
                                          uint32_t* values, void* stream) {
  if (num_instances == 0) return YSP_OK;
  hipStream_t s = as_stream(stream);
  int* flag = static_cast<int*>(scratch(55, sizeof(int)));
  if (!flag) return YSP_ERR_HIP;
  YSP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
  hipLaunchKernelGGL(gradient_contrib_kernel, dim3(grid_for(num_instances, kB)), dim3(kB), 0, s,
                     segment_indices, segment_sizes, num_instances, K, staging_base,
                     gradient_local_size, keys, values, flag);
  YSP_CHECK_LAUNCH();
  int h = 0;
  YSP_CHECK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  if (h) {
    ysp_set_error(__FILE__, __LINE__,
                  "ysp_gradient_contributions: staged gradients need 3-vector segments");
    return YSP_ERR_UNSUPPORTED;
  }
  return YSP_OK;
}

extern "C" int ysp_gather_blocks3(const double* block_staging, const uint32_t* offsets,
                                  const uint32_t* items, uint32_t num_blocks,
                                  const uint32_t* positions, double* values,
                                  double* diagonal_blocks, const uint32_t* diagonal_blocks_start,
                                  const uint32_t* gradient_segments_start,
                                  uint32_t num_attributes, int32_t accumulate_diagonal,
                                  void* stream) {
  if (num_blocks == 0) return YSP_OK;
  hipLaunchKernelGGL(gather_blocks3_kernel, dim3(LaneUnit<9>::grid(num_blocks)), dim3(kB), 0,
                     as_stream(stream), block_staging, offsets, items, num_blocks, positions,
                     values, diagonal_blocks, diagonal_blocks_start, gradient_segments_start,
                     num_attributes, accumulate_diagonal);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_gather_gradient3(const double* gradient_staging, const uint32_t* offsets,
                                    const uint32_t* items, uint32_t n, double* gradient,
                                    void* stream) {
  if (n % 3) return YSP_ERR_ARG;
  if (n == 0) return YSP_OK;
  hipLaunchKernelGGL(gather_gradient3_kernel, dim3(LaneUnit<3>::grid(n / 3)), dim3(kB), 0,
                     as_stream(stream), gradient_staging, offsets, items, n / 3, gradient);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}
