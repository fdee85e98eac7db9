// Rectangular mixed second derivatives d g / d theta of the generated energies (the
// blocks of the reference's secondOrderJacobian, kernel/Jacobian/
// secondOrderJacobianKernelNoProject.pyx, for row targets = the energy's DoF and column
// targets = an attribute its parameters depend on).
//
// Per element e the generated code gives the derivatives of the element gradient w.r.t.
// the element's parameters, Gp = d g / d p (N x P) and Gs = d g / d s (N x S)
// (generated_mixed.hpp); the caller supplies the derivatives of those parameters w.r.t.
// theta, Cp_e = d p / d theta (P x w) and Cs_e = d s / d theta (S x w), and the element's
// block is B_e = Gp Cp_e + Gs Cs_e (N x w, row-major). No projection: the reference's
// rectangular kernels do not project either.
//
// A block's rows are the element's DoF in the order of its index program (vertex order
// of the connectivity), mapped once to global rows from the get_indices output
// (ysp_jacobian_rows, -1 for DoF outside the row targets); its columns are w consecutive
// columns starting at e w (theta per element) or at 0 (theta with one instance, shared by
// every element). Products: y += B x over rows (FP64 atomics: elements share vertices)
// and y += B^T x over columns (per-element columns are written by their element alone;
// shared columns are reduced per wavefront, then one atomic per wavefront and column).
#include "jacobian_kernels.hpp"
#include "generated_mixed.hpp"

namespace ysp {
namespace {

#define YSP_MIXED_KIND(ID, NAME, ARITY, WIDTH, NPAR, NSCENE)                                 \
  template <>                                                                              \
  struct MixedKind<ID> {                                                                   \
    static constexpr int A = ARITY, W = WIDTH, N = ARITY * WIDTH, P = NPAR, S = NSCENE;    \
    static constexpr int PM = gen::kMixed_##NAME##_P;                                      \
    __device__ static void mixed(const double* x, const double* p, const double* s,        \
                                 double* Gp, double* Gs) {                                 \
      gen::mixed_##NAME(x, p, s, Gp, Gs);                                                  \
    }                                                                                      \
  };

YSP_MIXED_KIND(YSP_ENERGY_POINT_POINT, point_point, 2, 3, 0, 2)
YSP_MIXED_KIND(YSP_ENERGY_POINT_EDGE, point_edge, 3, 3, 0, 2)
YSP_MIXED_KIND(YSP_ENERGY_POINT_TRIANGLE, point_triangle, 4, 3, 0, 2)
YSP_MIXED_KIND(YSP_ENERGY_EDGE_EDGE, edge_edge, 4, 3, 0, 2)
YSP_MIXED_KIND(YSP_ENERGY_BENDING, bending, 4, 3, 2, 1)
YSP_MIXED_KIND(YSP_ENERGY_BARAFF_WITKIN, baraff_witkin, 3, 3, 5, 5)
YSP_MIXED_KIND(YSP_ENERGY_STVK_MEMBRANE, stvk_membrane, 3, 3, 5, 3)
YSP_MIXED_KIND(YSP_ENERGY_AFFINE_ORTHOGONALITY, affine_orthogonality, 1, 9, 0, 1)
YSP_MIXED_KIND(YSP_ENERGY_SNH, snh, 4, 3, 10, 3)
YSP_MIXED_KIND(YSP_ENERGY_INERTIA, inertia, 1, 3, 7, 1)
#undef YSP_MIXED_KIND

// Global row of every DoF of every element: group k of element e (K groups, sizes from
// get_indices) covers the next size DoF; index 0 / 1 (union padding / not a row target)
// map to -1. flag: set when an element's group sizes do not add up to dofs.
__global__ void __launch_bounds__(kJB)
    jacobian_rows_kernel(const uint32_t* __restrict__ seg_idx, const uint16_t* __restrict__ seg_size,
                         uint32_t K, uint32_t n, uint32_t dofs, int32_t* __restrict__ rows,
                         int* __restrict__ flag) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  uint32_t off = 0;
  for (uint32_t k = 0; k < K; ++k) {
    const uint32_t idx = seg_idx[static_cast<size_t>(e) * K + k];
    const uint32_t sz = seg_size[static_cast<size_t>(e) * K + k];
    for (uint32_t i = 0; i < sz && off + i < dofs; ++i)
      rows[static_cast<size_t>(e) * dofs + off + i] =
          idx >= 2 ? static_cast<int32_t>(idx - 2 + i) : -1;
    off += sz;
  }
  if (off != dofs) {
    for (uint32_t i = off; i < dofs; ++i) rows[static_cast<size_t>(e) * dofs + i] = -1;
    atomicOr(flag, 1);
  }
}

// y[rows] += B x[cols]
__global__ void __launch_bounds__(kJB)
    jacobian_product_kernel(const double* __restrict__ blocks, uint32_t n, uint32_t dofs,
                            uint32_t w, const int32_t* __restrict__ rows, int32_t shared,
                            const double* __restrict__ x, double* __restrict__ y) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double xc[kMaxCols];
  const size_t c0 = shared ? 0 : static_cast<size_t>(e) * w;
  for (uint32_t c = 0; c < w; ++c) xc[c] = x[c0 + c];
  const double* B = blocks + static_cast<size_t>(e) * dofs * w;
  const int32_t* rr = rows + static_cast<size_t>(e) * dofs;
  for (uint32_t r = 0; r < dofs; ++r) {
    const int32_t row = rr[r];
    if (row < 0) continue;
    double acc = 0.0;
    for (uint32_t c = 0; c < w; ++c) acc += B[r * w + c] * xc[c];
    atomic_add_f64(y + row, acc);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// y[cols] += B^T x[rows]
__global__ void __launch_bounds__(kJB)
    jacobian_transpose_kernel(const double* __restrict__ blocks, uint32_t n, uint32_t dofs,
                              uint32_t w, const int32_t* __restrict__ rows, int32_t shared,
                              const double* __restrict__ x, double* __restrict__ y) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = e < n;
  double acc[kMaxCols];
  for (uint32_t c = 0; c < w; ++c) acc[c] = 0.0;
  if (live) {
    const double* B = blocks + static_cast<size_t>(e) * dofs * w;
    const int32_t* rr = rows + static_cast<size_t>(e) * dofs;
    for (uint32_t r = 0; r < dofs; ++r) {
      const int32_t row = rr[r];
      if (row < 0) continue;
      const double xv = x[row];
      for (uint32_t c = 0; c < w; ++c) acc[c] += B[r * w + c] * xv;
    }
  }
  if (!shared) {
    if (live)
      for (uint32_t c = 0; c < w; ++c) y[static_cast<size_t>(e) * w + c] += acc[c];
    return;
  }
  // one instance of theta: every element adds to the same w columns
  for (uint32_t c = 0; c < w; ++c) {
    const double s = wave_sum(acc[c]);
    if ((threadIdx.x & 63u) == 0) atomic_add_f64(y + c, s);
  }
}

}  // namespace
}  // namespace ysp

using namespace ysp;

extern "C" int ysp_element_param_jacobian(int32_t kind, const ysp_union_f64* positions,
                                          const uint32_t* element_vertices, uint32_t n,
                                          const double* element_params, const double* scene_params,
                                          uint32_t scene_stride, const double* dparams,
                                          uint32_t dparams_stride, const double* dscene,
                                          uint32_t dscene_stride, uint32_t w, double* blocks,
                                          void* stream) {
  const char* who = "ysp_element_param_jacobian: bad arguments (vertex width, element "
                    "parameters or column count do not match the energy kind)";
  switch (kind) {
#define YSP_CASE(ID)                                                                         \
  case ID:                                                                                   \
    return param_jacobian<ID>(positions, element_vertices, n, element_params, scene_params,  \
                              scene_stride, dparams, dparams_stride, dscene, dscene_stride, w, \
                              blocks, stream, who);
    YSP_CASE(YSP_ENERGY_POINT_POINT)
    YSP_CASE(YSP_ENERGY_POINT_EDGE)
    YSP_CASE(YSP_ENERGY_POINT_TRIANGLE)
    YSP_CASE(YSP_ENERGY_EDGE_EDGE)
    YSP_CASE(YSP_ENERGY_BENDING)
    YSP_CASE(YSP_ENERGY_BARAFF_WITKIN)
    YSP_CASE(YSP_ENERGY_STVK_MEMBRANE)
    YSP_CASE(YSP_ENERGY_AFFINE_ORTHOGONALITY)
    YSP_CASE(YSP_ENERGY_SNH)
    YSP_CASE(YSP_ENERGY_INERTIA)
#undef YSP_CASE
    default:
      ysp_set_error(__FILE__, __LINE__,
                    "ysp_element_param_jacobian: no mixed derivatives for this energy kind");
      return YSP_ERR_ARG;
  }
}

extern "C" int ysp_jacobian_rows(const uint32_t* segment_indices, const uint16_t* segment_sizes,
                                 uint32_t max_num_indices, uint32_t n, uint32_t dofs, int32_t* rows,
                                 void* stream) {
  if (n == 0) return YSP_OK;
  if (!segment_indices || !segment_sizes || !rows || max_num_indices == 0 || dofs == 0) {
    ysp_set_error(__FILE__, __LINE__, "ysp_jacobian_rows: bad arguments");
    return YSP_ERR_ARG;
  }
  hipStream_t s = as_stream(stream);
  int* flag = static_cast<int*>(scratch(120, sizeof(int)));
  if (!flag) return YSP_ERR_HIP;
  YSP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
  hipLaunchKernelGGL(jacobian_rows_kernel, dim3(grid_for(n, kJB)), dim3(kJB), 0, s, segment_indices,
                     segment_sizes, max_num_indices, n, dofs, rows, flag);
  YSP_CHECK_LAUNCH();
  int h = 0;
  YSP_CHECK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  if (h) {
    ysp_set_error(__FILE__, __LINE__,
                  "ysp_jacobian_rows: the index segments of an element do not cover its DoF "
                  "(elements on affine-body vertices have no mixed Jacobian here)");
    return YSP_ERR_ARG;
  }
  return YSP_OK;
}

extern "C" int ysp_jacobian_spmv(const double* blocks, uint32_t n, uint32_t dofs, uint32_t w,
                                 const int32_t* rows, int32_t shared_columns, const double* x,
                                 double* y, int32_t transpose, void* stream) {
  if (n == 0) return YSP_OK;
  if (!blocks || !rows || !x || !y || w == 0 || w > kMaxCols || dofs == 0) {
    ysp_set_error(__FILE__, __LINE__, "ysp_jacobian_spmv: bad arguments");
    return YSP_ERR_ARG;
  }
  hipStream_t s = as_stream(stream);
  const dim3 g(grid_for(n, kJB)), b(kJB);
  if (transpose)
    hipLaunchKernelGGL(jacobian_transpose_kernel, g, b, 0, s, blocks, n, dofs, w, rows,
                       shared_columns, x, y);
  else
    hipLaunchKernelGGL(jacobian_product_kernel, g, b, 0, s, blocks, n, dofs, w, rows,
                       shared_columns, x, y);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}
