// Sparsity pipeline on the GPU: slot indices, local compression, coordinates and the
// global (h, w, row, col) dedupe. Integer work, bit-exact with the reference layout.
#include "common.hpp"

#include <hipcub/hipcub.hpp>
#include <vector>

namespace ysp {
namespace {

struct ProgramPtrs {
  const uint32_t* conns[YSP_MAX_PROGRAM_CONNS];
  const uint32_t* counts[YSP_MAX_PROGRAM_UNIONS];
};

enum : int32_t { OP_LEAF = 0, OP_JOIN = 1, OP_UNION = 2, OP_CONCAT = 3 };

// Emit the slots of the node at code[pc] for instance idx into oi/os. The depth
// template bounds the nesting without device recursion.
template <int DEPTH>
__device__ bool emit(const int32_t* code, int pc, uint32_t idx, const ProgramPtrs& P,
                     const uint32_t* wrt_start, uint32_t* oi, uint16_t* os) {
  const int32_t op = code[pc];
  if (op == OP_LEAF) {
    const int32_t size = code[pc + 2];
    const int32_t wrt_pos = code[pc + 4];
    const int32_t per_instance = code[pc + 5];
    if (wrt_pos >= 0) {
      oi[0] = wrt_start[wrt_pos] + idx * (per_instance ? static_cast<uint32_t>(size) : 0u) + 2u;
    } else {
      oi[0] = 1u;
    }
    os[0] = static_cast<uint16_t>(size);
    return true;
  }
  if constexpr (DEPTH > 0) {
    if (op == OP_JOIN) {
      const int32_t k = code[pc + 1];
      const uint32_t* conn = P.conns[code[pc + 4]];
      const int32_t dim = code[pc + 5];
      const int32_t nchild = code[pc + 6];
      const int32_t per = k / dim;
      for (int32_t i = 0; i < dim; ++i) {
        const uint32_t target = conn[static_cast<size_t>(idx) * dim + i];
        int cpc = pc + 7;
        int32_t acc = 0;
        for (int32_t c = 0; c < nchild; ++c) {
          if (!emit<DEPTH - 1>(code, cpc, target, P, wrt_start, oi + per * i + acc,
                               os + per * i + acc))
            return false;
          acc += code[cpc + 1];
          cpc += code[cpc + 3];
        }
      }
      return true;
    }
    if (op == OP_UNION) {
      const int32_t k_union = code[pc + 1];
      const int32_t g_union = code[pc + 2];
      const uint32_t* counts = P.counts[code[pc + 4]];
      const int32_t nuc = code[pc + 5];
      uint32_t total = 0;
      int32_t which = 0;
      for (int32_t c = 0; c < nuc; ++c) {
        if (total + counts[c] > idx) {
          which = c;
          break;
        }
        total += counts[c];
      }
      const uint32_t shifted = idx - total;
      int cpc = pc + 6;
      for (int32_t c = 0; c < which; ++c) {  // skip the other union children
        const int32_t nm = code[cpc];
        cpc += 1;
        for (int32_t m = 0; m < nm; ++m) cpc += code[cpc + 3];
      }
      const int32_t nm = code[cpc];
      cpc += 1;
      int32_t shift = 0;
      int32_t actual = 0;
      for (int32_t m = 0; m < nm; ++m) {
        if (!emit<DEPTH - 1>(code, cpc, shifted, P, wrt_start, oi + shift, os + shift))
          return false;
        shift += code[cpc + 1];
        actual += code[cpc + 2];
        cpc += code[cpc + 3];
      }
      os[k_union - 1] = static_cast<uint16_t>(g_union - actual);
      return true;
    }
    if (op == OP_CONCAT) {
      const int32_t nchild = code[pc + 4];
      int cpc = pc + 5;
      int32_t acc = 0;
      for (int32_t c = 0; c < nchild; ++c) {
        if (!emit<DEPTH - 1>(code, cpc, idx, P, wrt_start, oi + acc, os + acc)) return false;
        acc += code[cpc + 1];
        cpc += code[cpc + 3];
      }
      return true;
    }
  }
  return false;
}

__global__ void get_indices_kernel(const int32_t* __restrict__ code, ProgramPtrs P,
                                   const uint32_t* __restrict__ wrt_start,
                                   uint32_t* __restrict__ out_idx, uint16_t* __restrict__ out_sz,
                                   uint32_t n, uint32_t K, int* __restrict__ error_flag) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  if (!emit<6>(code, 0, e, P, wrt_start, out_idx + static_cast<size_t>(e) * K,
               out_sz + static_cast<size_t>(e) * K)) {
    atomicOr(error_flag, 1);
  }
}

// computePermutation (gradientIndicesKernel.pyx:62-123).
__global__ void permutation_kernel(const uint32_t* __restrict__ indices,
                                   const uint16_t* __restrict__ sizes,
                                   int16_t* __restrict__ perm, uint16_t* __restrict__ total,
                                   uint32_t* __restrict__ counts_plus1, uint32_t n, uint32_t K) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t* idx = indices + static_cast<size_t>(e) * K;
  const uint16_t* sz = sizes + static_cast<size_t>(e) * K;
  int16_t* pm = perm + static_cast<size_t>(e) * K;
  int offset = 0;
  unsigned total_size = 0;
  unsigned m = 0;
  for (uint32_t i = 0; i < K; ++i) {
    const uint32_t ii = idx[i];
    if (ii == 0) continue;
    bool found = false;
    if (ii >= 2) {
      for (uint32_t j = 0; j < i; ++j) {
        if (idx[j] == ii) {
          pm[i] = static_cast<int16_t>(-pm[j]);
          found = true;
          break;
        }
      }
    }
    if (!found) {
      pm[i] = static_cast<int16_t>(offset + 1);
      offset += sz[i];
      total_size += sz[i];
      if (ii >= 2) ++m;
    }
  }
  total[e] = static_cast<uint16_t>(total_size);
  counts_plus1[e] = m;
}

struct Triangular {
  __host__ __device__ uint32_t operator()(uint32_t x) const { return x * (x + 1) / 2; }
};

__global__ void iota_kernel(uint32_t* out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

__global__ void finish_groups_kernel(uint32_t* offsets, const int* num_runs,
                                     uint16_t* num_unique) {
  offsets[0] = 0;
  num_unique[0] = static_cast<uint16_t>(num_runs[0]);
}

__global__ void zero_u32_kernel(uint32_t* p) { p[0] = 0; }

// computeCoordinates (gradientIndicesKernel.pyx:234-296).
__global__ void coordinates_kernel(const uint32_t* __restrict__ indices,
                                   const int16_t* __restrict__ perms,
                                   const uint16_t* __restrict__ sizes,
                                   uint32_t* __restrict__ coords, uint16_t* __restrict__ dims,
                                   const uint32_t* __restrict__ outer, uint32_t K, uint32_t n) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint32_t* idx = indices + static_cast<size_t>(e) * K;
  const int16_t* pm = perms + static_cast<size_t>(e) * K;
  const uint16_t* sz = sizes + static_cast<size_t>(e) * K;
  size_t t = static_cast<size_t>(outer[e]) * 2;
  for (uint32_t i = 0; i < K; ++i) {
    const uint32_t ii = idx[i];
    if (!(pm[i] > 0 && ii >= 2)) continue;
    for (uint32_t j = i; j < K; ++j) {
      const uint32_t jj = idx[j];
      if (!(pm[j] > 0 && jj >= 2)) continue;
      if (ii < jj) {
        coords[t] = ii - 2;
        coords[t + 1] = jj - 2;
        dims[t] = sz[i];
        dims[t + 1] = sz[j];
      } else {
        coords[t] = jj - 2;
        coords[t + 1] = ii - 2;
        dims[t] = sz[j];
        dims[t + 1] = sz[i];
      }
      t += 2;
    }
  }
}

__device__ __forceinline__ uint32_t bit_length(uint32_t v) { return v ? 32u - __clz(v) : 0u; }

// Largest coordinate and dimension, in two passes without atomics: every workgroup
// writes its maxima to partial[2 * block], then one workgroup reduces the partials.
constexpr unsigned kRedBlock = 256;
__device__ __forceinline__ void block_max2(uint32_t& a, uint32_t& b) {
  __shared__ uint32_t sa[kRedBlock / 64], sb[kRedBlock / 64];
  for (int off = 32; off > 0; off >>= 1) {
    a = max(a, static_cast<uint32_t>(__shfl_xor(static_cast<int>(a), off, 64)));
    b = max(b, static_cast<uint32_t>(__shfl_xor(static_cast<int>(b), off, 64)));
  }
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x / 64] = a;
    sb[threadIdx.x / 64] = b;
  }
  __syncthreads();
  a = sa[0];
  b = sb[0];
  for (unsigned w = 1; w < kRedBlock / 64; ++w) {
    a = max(a, sa[w]);
    b = max(b, sb[w]);
  }
}

__global__ void __launch_bounds__(kRedBlock)
    max_fields_kernel(const uint32_t* __restrict__ coords, const uint16_t* __restrict__ dims,
                      uint32_t n, uint32_t* __restrict__ partial) {
  uint32_t mc = 0, md = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    mc = max(mc, max(coords[2 * i], coords[2 * i + 1]));
    md = max(md, static_cast<uint32_t>(max(dims[2 * i], dims[2 * i + 1])));
  }
  block_max2(mc, md);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = mc;
    partial[2 * blockIdx.x + 1] = md;
  }
}

__global__ void __launch_bounds__(kRedBlock)
    max_partials_kernel(const uint32_t* __restrict__ partial, uint32_t nblocks,
                        uint32_t* __restrict__ out /* [2]: max coord, max dim */) {
  uint32_t mc = 0, md = 0;
  for (uint32_t i = threadIdx.x; i < nblocks; i += blockDim.x) {
    mc = max(mc, partial[2 * i]);
    md = max(md, partial[2 * i + 1]);
  }
  block_max2(mc, md);
  if (threadIdx.x == 0) {
    out[0] = mc;
    out[1] = md;
  }
}

// One 64-bit key per occurrence, ordered as (h, w, row, col).
__global__ void pack_keys_kernel(const uint32_t* __restrict__ coords,
                                 const uint16_t* __restrict__ dims, uint32_t n, uint32_t bc,
                                 uint32_t bd, uint64_t* __restrict__ keys,
                                 uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t h = dims[2 * i], w = dims[2 * i + 1];
  const uint64_t r = coords[2 * i], c = coords[2 * i + 1];
  keys[i] = (h << (bd + 2 * bc)) | (w << (2 * bc)) | (r << bc) | c;
  vals[i] = i;
}

// Two-pass fallback keys: (row, col) first, then a stable pass on (h, w).
__global__ void pack_rc_keys_kernel(const uint32_t* __restrict__ coords, uint32_t n,
                                    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = (static_cast<uint64_t>(coords[2 * i]) << 32) | coords[2 * i + 1];
  vals[i] = i;
}

__global__ void gather_hw_keys_kernel(const uint16_t* __restrict__ dims,
                                      const uint32_t* __restrict__ vals, uint32_t n,
                                      uint32_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = vals[i];
  keys[i] = (static_cast<uint32_t>(dims[2 * v]) << 16) | dims[2 * v + 1];
}

__device__ __forceinline__ bool same_coord(const uint32_t* coords, const uint16_t* dims,
                                           uint32_t a, uint32_t b) {
  return coords[2 * a] == coords[2 * b] && coords[2 * a + 1] == coords[2 * b + 1] &&
         dims[2 * a] == dims[2 * b] && dims[2 * a + 1] == dims[2 * b + 1];
}

__global__ void run_flags_kernel(const uint32_t* __restrict__ sorted_vals,
                                 const uint32_t* __restrict__ coords,
                                 const uint16_t* __restrict__ dims, uint32_t n,
                                 uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = (i == 0 || !same_coord(coords, dims, sorted_vals[i], sorted_vals[i - 1])) ? 1u : 0u;
}

__global__ void scatter_ordinals_kernel(const uint32_t* __restrict__ sorted_vals,
                                        const uint32_t* __restrict__ ordinals,
                                        const uint32_t* __restrict__ flags,
                                        const uint32_t* __restrict__ coords,
                                        const uint16_t* __restrict__ dims, uint32_t n,
                                        uint32_t* __restrict__ lookup,
                                        ysp_coord_dim* __restrict__ uniq) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = sorted_vals[i];
  const uint32_t ord = ordinals[i] - 1u;
  lookup[v] = ord;
  if (flags[i]) {
    ysp_coord_dim r;
    r.row = coords[2 * v];
    r.col = coords[2 * v + 1];
    r.h = dims[2 * v];
    r.w = dims[2 * v + 1];
    r.placeholder = v;
    uniq[ord] = r;
  }
}

// prepare_unique_metadata (coordinateCompressionKernel.pyx:264-287).
__global__ void prepare_unique_kernel(const ysp_coord_dim* __restrict__ u, uint32_t nu,
                                      uint32_t* __restrict__ coords_out,
                                      uint32_t* __restrict__ block_sizes,
                                      uint32_t* __restrict__ dim_flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nu) return;
  const ysp_coord_dim c = u[i];
  coords_out[2 * i] = c.row;
  coords_out[2 * i + 1] = c.col;
  block_sizes[i] = static_cast<uint32_t>(c.h) * c.w;
  dim_flags[i] = (i == 0 || c.h != u[i - 1].h || c.w != u[i - 1].w) ? 1u : 0u;
}

// write_dimension_metadata (coordinateCompressionKernel.pyx:289-317).
__global__ void dim_metadata_kernel(const ysp_coord_dim* __restrict__ u, uint32_t nu,
                                    const uint32_t* __restrict__ offsets,
                                    const uint32_t* __restrict__ dim_ord,
                                    uint16_t* __restrict__ dims_out,
                                    uint32_t* __restrict__ outer, uint32_t* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nu) return;
  const ysp_coord_dim c = u[i];
  if (i == 0 || c.h != u[i - 1].h || c.w != u[i - 1].w) {
    const uint32_t d = dim_ord[i] - 1u;
    dims_out[2 * d] = c.h;
    dims_out[2 * d + 1] = c.w;
    outer[d] = offsets[i];
    cnt[d] = i;
  }
}

__global__ void map_lookup_kernel(uint32_t* __restrict__ lookup,
                                  const uint32_t* __restrict__ offsets, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lookup[i] = offsets[lookup[i]];
}

// finalize_dimension_metadata (coordinateCompressionKernel.pyx:329-348).
__global__ void finalize_dims_kernel(const ysp_coord_dim* __restrict__ u, uint32_t nu,
                                     const uint32_t* __restrict__ offsets, uint32_t nd,
                                     uint32_t* __restrict__ outer, uint32_t* __restrict__ cnt) {
  for (uint32_t d = 0; d < nd; ++d) {
    const uint32_t start = cnt[d];
    const uint32_t end = (d + 1 < nd) ? cnt[d + 1] : nu;
    cnt[d] = end - start;
  }
  const ysp_coord_dim last = u[nu - 1];
  outer[nd] = offsets[nu - 1] + static_cast<uint32_t>(last.h) * last.w;
}

constexpr unsigned kBlock = 256;

// Stable radix sort of (key, value) pairs through library scratch buffers.
template <typename K>
int sort_pairs(K* keys, uint32_t* vals, uint32_t n, int end_bit, int slot_base,
               hipStream_t s, K** keys_out, uint32_t** vals_out) {
  K* keys_alt = static_cast<K*>(scratch(slot_base, sizeof(K) * n));
  uint32_t* vals_alt = static_cast<uint32_t*>(scratch(slot_base + 1, sizeof(uint32_t) * n));
  if (!keys_alt || !vals_alt) {
    ysp_set_error(__FILE__, __LINE__, "scratch allocation failed");
    return YSP_ERR_HIP;
  }
  hipcub::DoubleBuffer<K> kb(keys, keys_alt);
  hipcub::DoubleBuffer<uint32_t> vb(vals, vals_alt);
  size_t temp = 0;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, kb, vb, n, 0, end_bit, s));
  void* tmp = scratch(slot_base + 2, temp);
  if (!tmp) {
    ysp_set_error(__FILE__, __LINE__, "scratch allocation failed");
    return YSP_ERR_HIP;
  }
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, temp, kb, vb, n, 0, end_bit, s));
  *keys_out = kb.Current();
  *vals_out = vb.Current();
  return YSP_OK;
}

}  // namespace
}  // namespace ysp

using namespace ysp;

extern "C" int ysp_get_indices(const int32_t* program, int32_t program_len,
                               const uint32_t* const* conns, int32_t num_conns,
                               const uint32_t* const* union_counts, int32_t num_unions,
                               const uint32_t* wrt_start_indices, uint32_t* out_indices,
                               uint16_t* out_sizes, uint32_t num_instances,
                               uint32_t num_indices_for_each_instance, void* stream) {
  if (num_instances == 0) return YSP_OK;
  if (!program || program_len < 6 || num_conns > YSP_MAX_PROGRAM_CONNS ||
      num_unions > YSP_MAX_PROGRAM_UNIONS || num_conns < 0 || num_unions < 0) {
    ysp_set_error(__FILE__, __LINE__, "ysp_get_indices: invalid program or pointer tables");
    return YSP_ERR_ARG;
  }
  ProgramPtrs P{};
  for (int i = 0; i < num_conns; ++i) P.conns[i] = conns[i];
  for (int i = 0; i < num_unions; ++i) P.counts[i] = union_counts[i];
  hipStream_t s = as_stream(stream);
  int* flag = static_cast<int*>(scratch(0, sizeof(int)));
  if (!flag) return YSP_ERR_HIP;
  YSP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
  hipLaunchKernelGGL(get_indices_kernel, dim3(grid_for(num_instances, kBlock)), dim3(kBlock), 0,
                     s, program, P, wrt_start_indices, out_indices, out_sizes, num_instances,
                     num_indices_for_each_instance, flag);
  YSP_CHECK_LAUNCH();
  int h = 0;
  YSP_CHECK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  if (h) {
    ysp_set_error(__FILE__, __LINE__, "ysp_get_indices: program nesting too deep or bad opcode");
    return YSP_ERR_UNSUPPORTED;
  }
  return YSP_OK;
}

extern "C" int ysp_compress_indices(const uint32_t* indices, const uint16_t* index_sizes,
                                    int16_t* permutations, uint16_t* total_gradient_sizes,
                                    uint32_t* coordinate_counts_outer, uint16_t* unique_sizes,
                                    uint32_t* grouped_indices, uint32_t* offsets,
                                    uint16_t* num_unique, uint32_t num_instances,
                                    uint32_t num_indices_for_each_instance,
                                    uint32_t compute_coordinate_metadata, void* stream) {
  if (num_instances == 0) return YSP_OK;
  hipStream_t s = as_stream(stream);
  const uint32_t n = num_instances;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, coordinate_counts_outer);
  hipLaunchKernelGGL(permutation_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, indices,
                     index_sizes, permutations, total_gradient_sizes, coordinate_counts_outer + 1,
                     n, num_indices_for_each_instance);
  YSP_CHECK_LAUNCH();
  if (!compute_coordinate_metadata) return YSP_OK;

  // Stable grouping of instances by compressed size (the sizes are sorted in place).
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, grouped_indices, n);
  uint16_t* keys_alt = static_cast<uint16_t*>(scratch(1, sizeof(uint16_t) * n));
  uint32_t* vals_alt = static_cast<uint32_t*>(scratch(2, sizeof(uint32_t) * n));
  if (!keys_alt || !vals_alt) return YSP_ERR_HIP;
  hipcub::DoubleBuffer<uint16_t> kb(total_gradient_sizes, keys_alt);
  hipcub::DoubleBuffer<uint32_t> vb(grouped_indices, vals_alt);
  size_t t1 = 0;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, kb, vb, n, 0, 16, s));
  void* tmp = scratch(3, t1);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, t1, kb, vb, n, 0, 16, s));
  if (kb.Current() != total_gradient_sizes)
    YSP_CHECK(hipMemcpyAsync(total_gradient_sizes, kb.Current(), sizeof(uint16_t) * n,
                             hipMemcpyDeviceToDevice, s));
  if (vb.Current() != grouped_indices)
    YSP_CHECK(hipMemcpyAsync(grouped_indices, vb.Current(), sizeof(uint32_t) * n,
                             hipMemcpyDeviceToDevice, s));
  int* num_runs = static_cast<int*>(scratch(4, sizeof(int)));
  if (!num_runs) return YSP_ERR_HIP;
  size_t t2 = 0;
  YSP_CHECK(hipcub::DeviceRunLengthEncode::Encode(nullptr, t2, total_gradient_sizes, unique_sizes,
                                                  offsets + 1, num_runs, n, s));
  tmp = scratch(3, t2);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceRunLengthEncode::Encode(tmp, t2, total_gradient_sizes, unique_sizes,
                                                  offsets + 1, num_runs, n, s));
  hipLaunchKernelGGL(finish_groups_kernel, dim3(1), dim3(1), 0, s, offsets, num_runs, num_unique);
  // offsets[1..U] inclusive scan: run over n entries is safe only for U entries; scan the
  // counts in place with the device run count read on the host.
  int h_runs = 0;
  YSP_CHECK(hipMemcpyAsync(&h_runs, num_runs, sizeof(int), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  size_t t3 = 0;
  YSP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, t3, offsets + 1, offsets + 1, h_runs, s));
  tmp = scratch(3, t3);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, t3, offsets + 1, offsets + 1, h_runs, s));
  // coordinate counts m -> m(m+1)/2, inclusive prefix
  hipcub::TransformInputIterator<uint32_t, Triangular, uint32_t*> tri(coordinate_counts_outer + 1,
                                                                      Triangular());
  size_t t4 = 0;
  YSP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, t4, tri, coordinate_counts_outer + 1, n, s));
  tmp = scratch(3, t4);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, t4, tri, coordinate_counts_outer + 1, n, s));
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_compute_coordinates(const uint32_t* indices, const int16_t* permutations,
                                       const uint16_t* index_sizes, uint32_t* coordinates,
                                       uint16_t* dimensions,
                                       const uint32_t* coordinate_counts_outer,
                                       uint32_t num_indices_for_each_instance,
                                       uint32_t num_instances, void* stream) {
  if (num_instances == 0) return YSP_OK;
  hipLaunchKernelGGL(coordinates_kernel, dim3(grid_for(num_instances, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), indices, permutations, index_sizes, coordinates,
                     dimensions, coordinate_counts_outer, num_indices_for_each_instance,
                     num_instances);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_get_unique_coords(const uint32_t* coords, const uint16_t* dims,
                                     uint32_t num_coordinates, ysp_coord_dim* coord_dim_tmp,
                                     uint32_t* lookup, uint32_t* num_unique_coords,
                                     void* stream) {
  if (!num_unique_coords) return YSP_ERR_ARG;
  *num_unique_coords = 0;
  const uint32_t n = num_coordinates;
  if (n == 0) return YSP_OK;
  hipStream_t s = as_stream(stream);
  const unsigned mblocks = std::min<unsigned>(grid_for(n, kRedBlock), 1024);
  uint32_t* maxes = static_cast<uint32_t*>(scratch(10, 2 * sizeof(uint32_t) * (mblocks + 1)));
  uint64_t* keys = static_cast<uint64_t*>(scratch(11, sizeof(uint64_t) * n));
  uint32_t* vals = static_cast<uint32_t*>(scratch(12, sizeof(uint32_t) * n));
  uint32_t* flags = static_cast<uint32_t*>(scratch(13, sizeof(uint32_t) * n));
  uint32_t* ords = static_cast<uint32_t*>(scratch(14, sizeof(uint32_t) * n));
  if (!maxes || !keys || !vals || !flags || !ords) return YSP_ERR_HIP;
  hipLaunchKernelGGL(max_fields_kernel, dim3(mblocks), dim3(kRedBlock), 0, s, coords, dims, n,
                     maxes + 2);
  hipLaunchKernelGGL(max_partials_kernel, dim3(1), dim3(kRedBlock), 0, s, maxes + 2, mblocks,
                     maxes);
  uint32_t hm[2];
  YSP_CHECK(hipMemcpyAsync(hm, maxes, sizeof(hm), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  auto bits = [](uint32_t v) { uint32_t b = 0; while (v) { ++b; v >>= 1; } return b; };
  const uint32_t bc = std::max<uint32_t>(1, bits(hm[0]));
  const uint32_t bd = std::max<uint32_t>(1, bits(hm[1]));
  uint32_t* sorted_vals = nullptr;
  if (2 * bc + 2 * bd <= 64) {
    hipLaunchKernelGGL(pack_keys_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, coords,
                       dims, n, bc, bd, keys, vals);
    uint64_t* ko = nullptr;
    int rc = sort_pairs<uint64_t>(keys, vals, n, static_cast<int>(2 * bc + 2 * bd), 15, s, &ko,
                                  &sorted_vals);
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL(pack_rc_keys_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, coords,
                       n, keys, vals);
    uint64_t* ko = nullptr;
    uint32_t* v1 = nullptr;
    int rc = sort_pairs<uint64_t>(keys, vals, n, 64, 15, s, &ko, &v1);
    if (rc) return rc;
    uint32_t* hw = static_cast<uint32_t*>(scratch(18, sizeof(uint32_t) * n));
    if (!hw) return YSP_ERR_HIP;
    hipLaunchKernelGGL(gather_hw_keys_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, dims,
                       v1, n, hw);
    // values must live outside the first pass's double buffer
    uint32_t* v2 = static_cast<uint32_t*>(scratch(19, sizeof(uint32_t) * n));
    if (!v2) return YSP_ERR_HIP;
    YSP_CHECK(hipMemcpyAsync(v2, v1, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    uint32_t* ko2 = nullptr;
    rc = sort_pairs<uint32_t>(hw, v2, n, 32, 20, s, &ko2, &sorted_vals);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(run_flags_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, sorted_vals,
                     coords, dims, n, flags);
  size_t t = 0;
  YSP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, t, flags, ords, n, s));
  void* tmp = scratch(23, t);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, t, flags, ords, n, s));
  hipLaunchKernelGGL(scatter_ordinals_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s,
                     sorted_vals, ords, flags, coords, dims, n, lookup, coord_dim_tmp);
  YSP_CHECK_LAUNCH();
  uint32_t nu = 0;
  YSP_CHECK(hipMemcpyAsync(&nu, ords + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  *num_unique_coords = nu;
  return YSP_OK;
}

extern "C" int ysp_compress_unique_coords(const ysp_coord_dim* coord_dim_tmp,
                                          uint32_t num_unique_coords, uint32_t num_coordinates,
                                          uint32_t* unique_coords_out, uint16_t* unique_dims_out,
                                          uint32_t* unique_dims_outer_indices,
                                          uint32_t* unique_dims_block_counts, uint32_t* lookup,
                                          uint32_t* num_unique_dims, void* stream) {
  if (!num_unique_dims) return YSP_ERR_ARG;
  *num_unique_dims = 0;
  const uint32_t nu = num_unique_coords;
  if (nu == 0 || num_coordinates == 0) return YSP_OK;
  hipStream_t s = as_stream(stream);
  uint32_t* offsets = static_cast<uint32_t*>(scratch(24, sizeof(uint32_t) * nu));
  uint32_t* dflags = static_cast<uint32_t*>(scratch(25, sizeof(uint32_t) * nu));
  uint32_t* dords = static_cast<uint32_t*>(scratch(26, sizeof(uint32_t) * nu));
  if (!offsets || !dflags || !dords) return YSP_ERR_HIP;
  hipLaunchKernelGGL(prepare_unique_kernel, dim3(grid_for(nu, kBlock)), dim3(kBlock), 0, s,
                     coord_dim_tmp, nu, unique_coords_out, offsets, dflags);
  size_t t = 0;
  YSP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t, offsets, offsets, nu, s));
  void* tmp = scratch(27, t);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, t, offsets, offsets, nu, s));
  t = 0;
  YSP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, t, dflags, dords, nu, s));
  tmp = scratch(27, t);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, t, dflags, dords, nu, s));
  uint32_t nd = 0;
  YSP_CHECK(hipMemcpyAsync(&nd, dords + (nu - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  hipLaunchKernelGGL(dim_metadata_kernel, dim3(grid_for(nu, kBlock)), dim3(kBlock), 0, s,
                     coord_dim_tmp, nu, offsets, dords, unique_dims_out, unique_dims_outer_indices,
                     unique_dims_block_counts);
  hipLaunchKernelGGL(map_lookup_kernel, dim3(grid_for(num_coordinates, kBlock)), dim3(kBlock), 0, s,
                     lookup, offsets, num_coordinates);
  hipLaunchKernelGGL(finalize_dims_kernel, dim3(1), dim3(1), 0, s, coord_dim_tmp, nu, offsets, nd,
                     unique_dims_outer_indices, unique_dims_block_counts);
  YSP_CHECK_LAUNCH();
  YSP_CHECK(hipStreamSynchronize(s));
  *num_unique_dims = nd;
  return YSP_OK;
}
