// Internal interface of the MAS apply (shared with the PCG driver).
#pragma once
#include "common.hpp"

#include <cstddef>
#include <vector>

namespace ysp {

constexpr int kMasMaxCoarseLevels = 8;

struct MasApply {
  const uint32_t* restrict_offsets;   // [num_coarse_nodes + 1]
  const uint32_t* restrict_items;     // level-0 descendants of the nodes of levels 1 to
                                      // 3; level-3 descendants (coarse ids) of deeper ones
  uint32_t num_coarse_nodes;          // all coarse levels, concatenated ids
  uint32_t num_coarse_levels;
  uint32_t level_offsets[kMasMaxCoarseLevels + 1];  // coarse id range of every level
  const uint32_t* prolong_maps;       // [num_coarse_levels x n/3] global coarse ids
  const float* inverses;              // [num_domains x 32 x 32] column-major
  const uint32_t* domain_nodes;
  const uint32_t* domain_node_offsets;  // [num_domains + 1]
  uint32_t num_domains;
  uint32_t level0_domains;            // domains [0, level0_domains) are level-0 domains
  const float* weights;               // [1 + num_coarse_levels]
  double* coarse_residual;            // [3 x num_coarse_nodes]
  float* fine_correction;             // [n]
  float* coarse_correction;           // [3 x num_coarse_nodes]
  const uint16_t* inverses_bf16;      // nullptr, or BF16 inverses of domains >= bf16_first
  uint32_t bf16_first_domain;
  const float* domain_scale;          // nullptr, or a factor per domain on its correction
  uint32_t domain_size;               // padded domain size: 32 (or 0) or 48 scalars
};
static_assert(sizeof(MasApply) == sizeof(ysp_mas_args) &&
                  offsetof(MasApply, bf16_first_domain) == offsetof(ysp_mas_args, bf16_first_domain) &&
                  offsetof(MasApply, domain_scale) == offsetof(ysp_mas_args, domain_scale) &&
                  offsetof(MasApply, domain_size) == offsetof(ysp_mas_args, domain_size),
              "MasApply mirrors ysp_mas_args");

// BF16 (round to nearest even) and back.
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// z = M r; if rz_slots != nullptr the partial sums of r.z are added to the 64 slots
// (spaced 16 doubles) starting there. When stop != nullptr and *stop != 0 at run time the
// kernels do nothing (a converged Krylov loop).
int launch_mas_apply(const MasApply& m, const double* r, uint32_t n, double* z,
                     double* rz_slots, const double* stop, hipStream_t s);

// One system of a batched apply: its argument block, the PCG vectors it reads (r) and
// writes (z), its scalar state (status word and the r.z slots of iteration k at
// st + rz_offset + (k % 3) * rz_stride), the iteration number of pass 0 and the last
// iteration; the launch geometry is filled by mas_batch_geometry.
struct MasBatchItem {
  MasApply m;
  const double* r;
  double* z;
  double* st;
  uint32_t n;
  int32_t k0, kmax;
  uint32_t status_offset, rz_offset, rz_stride;
  uint32_t nv, n1, n2, n3, b1, b2, first;
};
struct MasBatchWork {
  uint32_t item;   // index into the item table
  uint32_t blk;    // workgroup index within that system's launch
};
struct MasBatchTables {
  std::vector<MasBatchWork> restrict_, deep, solve, prolong, solve48;  // solve48: D = 48
};

// Fills the geometry of item `item` and appends its workgroups to the four work tables.
void mas_batch_geometry(MasBatchItem& it, uint32_t item, MasBatchTables& t);

// The launches of one batched apply (tables: restrict, deep restrict, local solve of the
// 32-scalar domains, prolong, local solve of the 48-scalar domains; device arrays of
// counts[i] entries; empty tables launch nothing).
constexpr int kMasBatchTables = 5;
int launch_mas_apply_batch(const MasBatchItem* items,
                           const MasBatchWork* const tables[kMasBatchTables],
                           const uint32_t counts[kMasBatchTables], int pass, hipStream_t s);

// YSP_OK when the argument block is consistent (sizes, level ranges, pointers).
int check_mas_args(const MasApply& m, uint32_t n);

}  // namespace ysp
