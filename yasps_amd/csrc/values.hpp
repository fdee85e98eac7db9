// Where the 3x3 block of entry k of the mirrored full block-CSR comes from.
//
// * Mirrored<VT>: the component-major copy (9 arrays of pad64(nnz) values, FP64 or FP32),
//   gathered once per assembly; coalesced loads, every off-diagonal block stored twice.
// * Upper32 / Upper64: the assembled upper block storage itself, read through each entry's
//   source word (the same words the gathers use): 32-bit words for 3x3 layouts (bit 31
//   transposed, bit 30 dynamic part, low bits the block index), 64-bit words for split
//   affine-body layouts (bits 0-39 value offset, 40-41 source row width / 3 - 1, 42
//   transposed, 43 dynamic part). Every stored block is read by both of its entries, so an
//   operator beyond the 256 MB Infinity Cache moves about half the value bytes from HBM.
#pragma once
#include "common.hpp"

namespace ysp {

// Entries per component array of the component-major copy.
__host__ __device__ __forceinline__ uint32_t pad64(uint32_t n) { return (n + 63u) & ~63u; }

constexpr uint32_t kTrans = 1u << 31;
constexpr uint32_t kDyn = 1u << 30;
constexpr uint64_t kSplitOffset = (1ull << 40) - 1;
constexpr uint64_t kSplitTrans = 1ull << 42;
constexpr uint64_t kSplitDyn = 1ull << 43;

template <typename VT>
struct Mirrored {
  const VT* v;
  uint32_t stride;
  __device__ __forceinline__ void load(uint32_t k, double (&b)[9]) const {
#pragma unroll
    for (int j = 0; j < 9; ++j) b[j] = static_cast<double>(v[j * static_cast<size_t>(stride) + k]);
  }
};

struct Upper32 {
  const uint32_t* src;
  const double* sv;
  const double* dv;
  __device__ __forceinline__ void load(uint32_t k, double (&b)[9]) const {
    const uint32_t s = src[k];
    const double* B = ((s & kDyn) ? dv : sv) + 9 * static_cast<size_t>(s & ~(kTrans | kDyn));
    double a[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) a[j] = B[j];
    const bool tr = (s & kTrans) != 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) b[3 * i + j] = tr ? a[3 * j + i] : a[3 * i + j];
  }
};

struct Upper64 {
  const uint64_t* src;
  const double* sv;
  const double* dv;
  __device__ __forceinline__ void load(uint32_t k, double (&b)[9]) const {
    const uint64_t word = src[k];
    const double* B = ((word & kSplitDyn) ? dv : sv) + (word & kSplitOffset);
    const uint32_t w = 3 * (static_cast<uint32_t>((word >> 40) & 3u) + 1);
    const bool tr = (word & kSplitTrans) != 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) b[3 * i + j] = tr ? B[j * w + i] : B[i * w + j];
  }
};

// y += B x for one 3x3 block as an explicit chain of fused multiply-adds, so that every
// value source (and every kernel instantiation) rounds the same way.
__device__ __forceinline__ void block_fma(const double (&b)[9], double x0, double x1, double x2,
                                          double& y0, double& y1, double& y2) {
  y0 = fma(b[0], x0, y0);
  y0 = fma(b[1], x1, y0);
  y0 = fma(b[2], x2, y0);
  y1 = fma(b[3], x0, y1);
  y1 = fma(b[4], x1, y1);
  y1 = fma(b[5], x2, y1);
  y2 = fma(b[6], x0, y2);
  y2 = fma(b[7], x1, y2);
  y2 = fma(b[8], x2, y2);
}

// Calls f(source) with the value source an entry point's (values, format) describe:
// YSP_VALUES_F64 / YSP_VALUES_F32 (values: the component-major copy) or YSP_VALUES_UPPER
// (values: a host ysp_upper_values). Returns YSP_ERR_ARG for anything else.
template <typename F>
int with_values(const void* values, int32_t format, uint32_t nnz, const char* who, F&& f) {
  const uint32_t stride = pad64(nnz);
  if (format == YSP_VALUES_F64) return f(Mirrored<double>{static_cast<const double*>(values), stride});
  if (format == YSP_VALUES_F32) return f(Mirrored<float>{static_cast<const float*>(values), stride});
  if (format == YSP_VALUES_UPPER && values) {
    const ysp_upper_values* u = static_cast<const ysp_upper_values*>(values);
    if (u->src && u->static_values) {
      const double* dv = u->dynamic_values ? u->dynamic_values : u->static_values;
      if (u->src_64)
        return f(Upper64{static_cast<const uint64_t*>(u->src), u->static_values, dv});
      return f(Upper32{static_cast<const uint32_t*>(u->src), u->static_values, dv});
    }
  }
  ysp_set_error(__FILE__, __LINE__, who);
  return YSP_ERR_ARG;
}

}  // namespace ysp
