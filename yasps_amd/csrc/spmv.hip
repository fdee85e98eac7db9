// Block SpMV over symmetric upper block storage (row-sorted within each (h, w) category).
//
// y_r += B x_c for every stored block and y_c += B^T x_r for r != c. The 3x3 kernel
// gives one block per lane: the row part is reduced across the lanes of a wavefront that
// share a block row (blocks are sorted by (row, col), so they are contiguous lanes) and
// flushed with one FP64 atomic per row segment; the transposed part is scattered with
// FP64 atomics. Optionally the kernel also accumulates the quadratic form x^T H x
// (sum over blocks of x_r^T B x_c, doubled off the diagonal) used by CG as p^T A p, one
// partial sum per workgroup into a set of 64 slots.
#include "spmv.hpp"

namespace ysp {

template <typename VT>
__global__ void __launch_bounds__(256)
    bsr3_spmv_kernel(const VT* __restrict__ values, const uint32_t* __restrict__ pos,
                     uint32_t nb, const double* __restrict__ x, double* __restrict__ y,
                     double* __restrict__ quad, const double* __restrict__ stop) {
  if (stop && *stop != 0.0) return;
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t row = 0xffffffffu, col = 0;
  double yr0 = 0, yr1 = 0, yr2 = 0;
  double q = 0.0;
  if (b < nb) {
    row = pos[2 * b];
    col = pos[2 * b + 1];
    const VT* B = values + 9 * static_cast<size_t>(b);
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = static_cast<double>(B[k]);
    const double xc0 = x[col], xc1 = x[col + 1], xc2 = x[col + 2];
    yr0 = v[0] * xc0 + v[1] * xc1 + v[2] * xc2;
    yr1 = v[3] * xc0 + v[4] * xc1 + v[5] * xc2;
    yr2 = v[6] * xc0 + v[7] * xc1 + v[8] * xc2;
    if (row != col || quad) {
      const double xr0 = x[row], xr1 = x[row + 1], xr2 = x[row + 2];
      if (quad) q = (xr0 * yr0 + xr1 * yr1 + xr2 * yr2) * (row != col ? 2.0 : 1.0);
      if (row != col) {
        atomic_add_f64(y + col, v[0] * xr0 + v[3] * xr1 + v[6] * xr2);
        atomic_add_f64(y + col + 1, v[1] * xr0 + v[4] * xr1 + v[7] * xr2);
        atomic_add_f64(y + col + 2, v[2] * xr0 + v[5] * xr1 + v[8] * xr2);
      }
    }
  }
  // segmented suffix sums over lanes with equal row
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t orow = __shfl_down(row, off, 64);
    const double o0 = __shfl_down(yr0, off, 64);
    const double o1 = __shfl_down(yr1, off, 64);
    const double o2 = __shfl_down(yr2, off, 64);
    if (lane + off < 64 && orow == row) {
      yr0 += o0;
      yr1 += o1;
      yr2 += o2;
    }
  }
  const uint32_t prow = __shfl_up(row, 1, 64);
  if (b < nb && (lane == 0 || prow != row)) {
    atomic_add_f64(y + row, yr0);
    atomic_add_f64(y + row + 1, yr1);
    atomic_add_f64(y + row + 2, yr2);
  }
  if (quad) {
    __shared__ double red[4];
    const double s = block_sum<256>(q, red);
    if (threadIdx.x == 0) atomic_add_f64(quad + (blockIdx.x % kQuadSlots) * kQuadSpace, s);
  }
}

template <typename VT>
__global__ void __launch_bounds__(256)
    generic_block_spmv_kernel(const VT* __restrict__ values, const uint32_t* __restrict__ pos,
                              uint32_t nb, uint32_t h, uint32_t w, const double* __restrict__ x,
                              double* __restrict__ y, int transpose, double* __restrict__ quad,
                              const double* __restrict__ stop) {
  if (stop && *stop != 0.0) return;
  // H holds B at (row, col) and B^T at (col, row) when row != col. For row != col,
  // H x and H^T x coincide; a diagonal block contributes B x (or B^T x when transposed).
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  double q = 0.0;
  if (b < nb) {
    const uint32_t row = pos[2 * b], col = pos[2 * b + 1];
    const VT* B = values + static_cast<size_t>(b) * h * w;
    const bool diag = row == col;
    if (!diag || !transpose) {
      for (uint32_t i = 0; i < h; ++i) {
        double acc = 0.0;
        for (uint32_t j = 0; j < w; ++j) acc += static_cast<double>(B[i * w + j]) * x[col + j];
        if (quad) q += x[row + i] * acc * (diag ? 1.0 : 2.0);
        atomic_add_f64(y + row + i, acc);
      }
    }
    if (!diag || transpose) {
      for (uint32_t j = 0; j < w; ++j) {
        double acc = 0.0;
        for (uint32_t i = 0; i < h; ++i) acc += static_cast<double>(B[i * w + j]) * x[row + i];
        if (quad && diag) q += x[col + j] * acc;
        atomic_add_f64(y + col + j, acc);
      }
    }
  }
  if (quad) {
    __shared__ double red[4];
    const double s = block_sum<256>(q, red);
    if (threadIdx.x == 0) atomic_add_f64(quad + (blockIdx.x % kQuadSlots) * kQuadSpace, s);
  }
}

int launch_part_spmv(const ysp_block_matrix* part, const double* x, double* y, int transpose,
                     double* quad, const double* stop, hipStream_t s) {
  if (!part || part->num_categories == 0) return YSP_OK;
  uint32_t pos_offset = 0;
  for (uint32_t c = 0; c < part->num_categories; ++c) {
    const uint32_t nb = part->counts[c];
    const uint32_t h = part->dims[2 * c], w = part->dims[2 * c + 1];
    if (nb == 0) continue;
    const double* vals = part->values + part->starts[c];
    const uint32_t* pos = part->positions + 2 * static_cast<size_t>(pos_offset);
    if (h == 3 && w == 3 && !transpose) {
      hipLaunchKernelGGL(bsr3_spmv_kernel<double>, dim3(grid_for(nb, 256)), dim3(256), 0, s, vals,
                         pos, nb, x, y, quad, stop);
    } else {
      hipLaunchKernelGGL(generic_block_spmv_kernel<double>, dim3(grid_for(nb, 256)), dim3(256), 0,
                         s, vals, pos, nb, h, w, x, y, transpose, quad, stop);
    }
    YSP_CHECK_LAUNCH();
    pos_offset += nb;
  }
  return YSP_OK;
}

}  // namespace ysp

using namespace ysp;

extern "C" int ysp_block_spmv(const ysp_block_matrix* part, const ysp_block_matrix* dynamic_part,
                              const double* x, double* y, uint32_t n, int transpose,
                              int accumulate, void* stream) {
  hipStream_t s = as_stream(stream);
  if (!accumulate && n) YSP_CHECK(hipMemsetAsync(y, 0, sizeof(double) * n, s));
  int rc = launch_part_spmv(part, x, y, transpose, nullptr, nullptr, s);
  if (rc) return rc;
  return launch_part_spmv(dynamic_part, x, y, transpose, nullptr, nullptr, s);
}
