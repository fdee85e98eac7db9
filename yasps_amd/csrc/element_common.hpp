// Element-kernel plumbing shared by the energy kernels: the union view of a vertex
// attribute (fixed/moving children concatenated), the scatter arguments of one term and
// the diagonal-block offset of a segment.
#pragma once
#include "common.hpp"

namespace ysp {

struct UnionView {
  uint32_t nc;
  uint32_t width;
  const uint32_t* counts;
  const double* data[YSP_MAX_UNION_CHILDREN];
};

inline UnionView make_view(const ysp_union_f64* u) {
  UnionView v{};
  v.nc = u->num_children;
  v.width = u->width;
  v.counts = u->counts;
  for (uint32_t i = 0; i < YSP_MAX_UNION_CHILDREN; ++i) v.data[i] = u->data[i];
  return v;
}

__device__ __forceinline__ const double* union_row(const UnionView& U, uint32_t v) {
  uint32_t total = 0;
  for (uint32_t c = 0; c < U.nc; ++c) {
    const uint32_t cnt = U.counts[c];
    if (v < total + cnt) return U.data[c] + static_cast<size_t>(v - total) * U.width;
    total += cnt;
  }
  return nullptr;
}

struct ScatterArgs {
  const uint32_t* seg_idx;
  const uint16_t* seg_size;
  const int16_t* perm;
  const uint32_t* lookups;
  const uint32_t* coord_outer;
  uint32_t K;
  uint32_t num_attr;
  const uint32_t* diag_start;
  const uint32_t* grad_start;
  double* gradient;
  double* blocks;
  double* diag_blocks;
  double* block_staging;     // staged mode: 9 doubles per coordinate of this term
  double* grad_staging;      // staged mode: grad_local doubles per element
  uint32_t grad_local;
};

inline ScatterArgs make_scatter(const ysp_scatter_args* a) {
  return ScatterArgs{a->segment_indices, a->segment_sizes, a->local_permutations, a->lookups,
                     a->coordinates_outer, a->max_num_indices, a->num_attributes,
                     a->diagonal_blocks_start, a->gradient_segments_start, a->gradient,
                     a->hessian_blocks, a->diagonal_blocks, a->block_staging,
                     a->gradient_staging, a->gradient_local_size};
}

// Baraff-Witkin stretch coefficients of one element from the current deformation
// gradient F = [x1 - x0, x2 - x0] Minv (helpers.py:252-253: 0.01 below unit stretch,
// else 1); x: the 3 vertices (9 doubles), p: Minv (4, row-major) first.
__device__ __forceinline__ void bw_stretch_coefficients(const double* x, const double* p,
                                                        double& cu, double& cv) {
  double fu2 = 0.0, fv2 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double d1 = x[3 + i] - x[i], d2 = x[6 + i] - x[i];
    const double fu = d1 * p[0] + d2 * p[2], fv = d1 * p[1] + d2 * p[3];
    fu2 += fu * fu;
    fv2 += fv * fv;
  }
  cu = sqrt(fu2) < 1.0 ? 0.01 : 1.0;
  cv = sqrt(fv2) < 1.0 ? 0.01 : 1.0;
}

__device__ __forceinline__ uint32_t diag_offset(const ScatterArgs& S, uint32_t seg, uint32_t size) {
  uint32_t attr = 0;
  while (attr < S.num_attr && seg >= S.grad_start[attr + 1]) ++attr;
  const uint32_t inst = (seg - S.grad_start[attr]) / size;
  return S.diag_start[attr] + inst * size * size;
}

}  // namespace ysp
