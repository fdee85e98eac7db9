// Multilevel additive Schwarz (MAS) preconditioner on the GPU.
//
// Per numeric update (each Newton solve):
//   * coarse operators, level by level: a level-l block is the sum of the level-(l-1)
//     blocks that map to it (identity aggregation, Galerkin R A R^T); level 0 is the
//     mirrored full 3x3 block-CSR of the solve loop (component-major values);
//   * domain matrices: every level block whose two nodes share a domain is copied into
//     that domain's dense, zero-padded D x D FP64 matrix (one contribution per entry,
//     no reduction; D = 32 or 48 scalars);
//   * inverses: in-place Gauss-Jordan in registers, two domains per wavefront (one for
//     D = 48), FP64, stored FP32 column-major (the reference's FP64-factor / FP32-bank split,
//     solver/mas/cuda/local_inverse.cu:244-308). A pivot below 1e-12 * max|diag| marks the
//     domain as failed.
// Per application z = sum_l w_l P_l B_l^-1 R_l r:
//   restrict (a node of levels 1 to 3 sums the residual of its level-0 descendants, a
//   node of a deeper level the level-3 residuals of its level-3 descendants: a second
//   launch), one 32-lane group per domain (64 lanes for D = 48) multiplies its inverse
//   with the local residual, prolong adds the weighted corrections of every level and accumulates
//   r.z.
#include "mas.hpp"

namespace ysp {
namespace {

constexpr unsigned kB = 256;
// Padded domain sizes (scalars): 32, the configurations' leaf size, and 48, the
// reference's default (solver/mas/solver.py:91); lanes per domain in the local solve and
// the inversion, and the domains one wavefront holds.
template <int D>
struct DomainShape {
  static_assert(D == 32 || D == 48, "domains of 32 or 48 scalars");
  static constexpr int lanes = D <= 32 ? 32 : 64;
  static constexpr int per_wave = 64 / lanes;
};

__host__ __device__ __forceinline__ uint32_t pad64(uint32_t n) { return (n + 63u) & ~63u; }

// level-l block values (AoS, 9 doubles) from the previous level's blocks
template <bool kPrevSoA>
__global__ void __launch_bounds__(kB)
    coarsen_kernel(const double* __restrict__ prev, uint32_t prev_nnz,
                   const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ items,
                   uint32_t nnz, double* __restrict__ out) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const size_t stride = pad64(prev_nnz);
  for (uint32_t k = offsets[e]; k < offsets[e + 1]; ++k) {
    const uint32_t s = items[k];
#pragma unroll
    for (int j = 0; j < 9; ++j)
      acc[j] += kPrevSoA ? prev[j * stride + s] : prev[9 * static_cast<size_t>(s) + j];
  }
#pragma unroll
  for (int j = 0; j < 9; ++j) out[9 * static_cast<size_t>(e) + j] = acc[j];
}

// copy the in-domain blocks of one level into the dense domain matrices
template <bool kSoA, int D>
__global__ void __launch_bounds__(kB)
    domain_scatter_kernel(const double* __restrict__ vals, uint32_t nnz,
                          const uint32_t* __restrict__ dest, double* __restrict__ dom) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const uint32_t d = dest[e];
  if (d == 0xffffffffu) return;
  const size_t stride = pad64(nnz);
  double* out = dom + d;  // d = domain * D*D + 3a * D + 3b
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      out[i * D + j] = kSoA ? vals[(3 * i + j) * stride + e] : vals[9 * static_cast<size_t>(e) + 3 * i + j];
}

__global__ void pad_identity_kernel(double* __restrict__ dom, const uint32_t* __restrict__ sizes,
                                    uint32_t ndom, uint32_t D) {
  const uint32_t d = blockIdx.x;
  if (d >= ndom) return;
  const uint32_t P = sizes[d];
  for (uint32_t i = P + threadIdx.x; i < D; i += blockDim.x)
    dom[static_cast<size_t>(d) * D * D + i * D + i] = 1.0;
}

// In-place Gauss-Jordan inversion (no pivoting: the domain matrices are principal
// submatrices of an SPD operator), two 32-scalar domains per wavefront (one 48-scalar
// domain, lanes 48-63 idle): in each domain's lane group, lane j holds column j of its
// domain in registers. At step k the half's lane k publishes its
// column c = A[:, k] through LDS (broadcast reads) and, with p = c_k,
//   lane k:      A_kk <- 1/p,  A_ik <- -c_i / p      (i != k)
//   lane j != k: A_kj <- A_kj / p,  A_ij <- A_ij - c_i A_kj   (i != k)
// which leaves A^-1 in place after D steps. The wavefront works alone: no workgroup
// barriers.
constexpr int kInvWaves = 4;                   // wavefronts per workgroup

// LDS written by some lanes of a wavefront and then read by others of the same wavefront.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int D>
__global__ void __launch_bounds__(64 * kInvWaves)
    invert_domains_kernel(const double* __restrict__ dom, float* __restrict__ inv,
                          uint16_t* __restrict__ inv_bf16, uint32_t bf16_first, uint32_t ndom,
                          int* __restrict__ status) {
  constexpr int G = DomainShape<D>::lanes, H = DomainShape<D>::per_wave;
  __shared__ double colk[kInvWaves][H][D];
  __shared__ double tr[kInvWaves][H][D][D + 1];
  const int w = threadIdx.x / 64, h = (threadIdx.x % 64) / G, j = threadIdx.x % G;
  const uint32_t d = (blockIdx.x * kInvWaves + w) * H + h;
  const bool live = d < ndom;                  // a group past the end idles along
  const bool col = j < D;                      // lanes past the domain's columns idle along
  const double* M = dom + static_cast<size_t>(live ? d : 0) * D * D;
  double x[D];
  double scale = 0.0;
  if (live && col) {                           // column j (the matrix is symmetric)
#pragma unroll
    for (int i = 0; i < D; ++i) x[i] = M[i * D + j];
    scale = fabs(M[j * D + j]);
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) x[i] = i == j ? 1.0 : 0.0;
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) scale = fmax(scale, __shfl_xor(scale, o, G));
  bool bad = false;
  // fully unrolled (x[k] must name a register, never a private array in scratch)
#pragma clang loop unroll(full)
  for (int k = 0; k < D; ++k) {
    if (j == k) {
#pragma unroll
      for (int i = 0; i < D; ++i) colk[w][h][i] = x[i];
    }
    wave_lds_sync();
    const double p = colk[w][h][k];
    if (!(fabs(p) > 1e-12 * scale)) bad = true;
    const double ip = 1.0 / p;
    if (j == k) {
#pragma unroll
      for (int i = 0; i < D; ++i) x[i] = i == k ? ip : -x[i] * ip;
    } else if (col) {
      const double xk = x[k] * ip;
#pragma unroll
      for (int i = 0; i < D; ++i)
        if (i != k) x[i] = fma(-colk[w][h][i], xk, x[i]);
      x[k] = xk;
    }
    wave_lds_sync();
  }
  // stored FP32 column-major, symmetrised: out[c * D + r] = (B[r][c] + B[c][r]) / 2
  if (col) {
#pragma unroll
    for (int i = 0; i < D; ++i) tr[w][h][i][j] = x[i];
  }
  wave_lds_sync();
  if (live && col) {
    if (inv_bf16 && d >= bf16_first) {
      uint16_t* out = inv_bf16 + static_cast<size_t>(d - bf16_first) * D * D;
#pragma unroll
      for (int c = 0; c < D; ++c)
        out[c * D + j] = to_bf16(static_cast<float>(0.5 * (tr[w][h][j][c] + tr[w][h][c][j])));
    } else {
      float* out = inv + static_cast<size_t>(d) * D * D;
#pragma unroll
      for (int c = 0; c < D; ++c)
        out[c * D + j] = static_cast<float>(0.5 * (tr[w][h][j][c] + tr[w][h][c][j]));
    }
    if (bad && j == 0) atomicOr(status, 1);
  }
}

__device__ __forceinline__ bool mas_stopped(const double* stop) {
  return stop && *stop != 0.0;
}

// Sum of r over the level-0 descendants items[offsets[J] .. offsets[J+1]) of coarse node J,
// by a group of G lanes (lane strides, then a fixed shuffle tree: deterministic).
template <int G>
__device__ __forceinline__ void descendant_sum(const double* __restrict__ r,
                                               const uint32_t* __restrict__ offsets,
                                               const uint32_t* __restrict__ items, uint32_t J,
                                               int lane, double s[3]) {
  s[0] = s[1] = s[2] = 0.0;
  for (uint32_t k = offsets[J] + lane; k < offsets[J + 1]; k += G) {
    const size_t i = 3 * static_cast<size_t>(items[k]);
    s[0] += r[i];
    s[1] += r[i + 1];
    s[2] += r[i + 2];
  }
  const int width = G < 64 ? G : 64;
#pragma unroll
  for (int o = width / 2; o > 0; o >>= 1) {
    s[0] += __shfl_xor(s[0], o, width);
    s[1] += __shfl_xor(s[1], o, width);
    s[2] += __shfl_xor(s[2], o, width);
  }
}

// Residuals of levels 1 to 3, straight from the level-0 residual: workgroups [0, b1)
// give 8 lanes to each level-1 node, [b1, b1 + b2) a wavefront to each level-2 node, the
// rest a whole workgroup to each level-3 node (few nodes with many descendants).
__device__ __forceinline__ void restrict_body(const double* __restrict__ r,
                                              const uint32_t* __restrict__ offsets,
                                              const uint32_t* __restrict__ items, uint32_t n1,
                                              uint32_t n2, uint32_t b1, uint32_t b2,
                                              double* __restrict__ rc, uint32_t bid) {
  double sum[3];
  uint32_t J;
  bool write;
  if (bid < b1) {
    J = bid * (kB / 8) + threadIdx.x / 8;
    if (J >= n1) return;                       // whole 8-lane groups leave together
    descendant_sum<8>(r, offsets, items, J, threadIdx.x % 8, sum);
    write = threadIdx.x % 8 == 0;
  } else if (bid < b1 + b2) {
    J = n1 + (bid - b1) * (kB / 64) + threadIdx.x / 64;
    if (J >= n2) return;                       // whole wavefronts leave together
    descendant_sum<64>(r, offsets, items, J, threadIdx.x % 64, sum);
    write = threadIdx.x % 64 == 0;
  } else {
    J = n2 + (bid - b1 - b2);                  // grid sized so that J < n3
    descendant_sum<kB>(r, offsets, items, J, threadIdx.x, sum);
    __shared__ double part[3][kB / 64];
    if (threadIdx.x % 64 == 0)
      for (int c = 0; c < 3; ++c) part[c][threadIdx.x / 64] = sum[c];
    __syncthreads();
    for (int c = 0; c < 3; ++c) {
      double t = 0.0;
      for (int w = 0; w < static_cast<int>(kB / 64); ++w) t += part[c][w];
      sum[c] = t;
    }
    write = threadIdx.x == 0;
  }
  if (write) {
    rc[3 * static_cast<size_t>(J)] = sum[0];
    rc[3 * static_cast<size_t>(J) + 1] = sum[1];
    rc[3 * static_cast<size_t>(J) + 2] = sum[2];
  }
}

__global__ void __launch_bounds__(kB)
    restrict_kernel(const double* __restrict__ r, const uint32_t* __restrict__ offsets,
                    const uint32_t* __restrict__ items, uint32_t n1, uint32_t n2, uint32_t b1,
                    uint32_t b2, double* __restrict__ rc, const double* __restrict__ stop) {
  if (mas_stopped(stop)) return;
  restrict_body(r, offsets, items, n1, n2, b1, b2, rc, blockIdx.x);
}

// Residuals of levels 4 and deeper, a wavefront per node, from the level-3 residuals of
// its level-3 descendants (coarse ids in [n2, n3), written by restrict_kernel; the nodes
// written here, [n3, nc), are never read in this launch).
__device__ __forceinline__ void restrict_deep_body(const uint32_t* __restrict__ offsets,
                                                   const uint32_t* __restrict__ items,
                                                   uint32_t n3, uint32_t nc, double* rc,
                                                   uint32_t bid) {
  const uint32_t J = n3 + bid * (kB / 64) + threadIdx.x / 64;
  if (J >= nc) return;                         // whole wavefronts leave together
  double sum[3];
  descendant_sum<64>(rc, offsets, items, J, threadIdx.x % 64, sum);
  if (threadIdx.x % 64 == 0) {
    rc[3 * static_cast<size_t>(J)] = sum[0];
    rc[3 * static_cast<size_t>(J) + 1] = sum[1];
    rc[3 * static_cast<size_t>(J) + 2] = sum[2];
  }
}

__global__ void __launch_bounds__(kB)
    restrict_deep_kernel(const uint32_t* __restrict__ offsets,
                         const uint32_t* __restrict__ items, uint32_t n3, uint32_t nc,
                         double* rc, const double* __restrict__ stop) {
  if (mas_stopped(stop)) return;
  restrict_deep_body(offsets, items, n3, nc, rc, blockIdx.x);
}

// One lane group per domain (32 lanes, 64 for D = 48 with lanes 48-63 idle):
// c_loc = B^-1 r_loc. Lane i reads row i of the inverse, which is its column i (the stored
// inverse is exactly symmetric): D/4 16-byte loads from the FP32 bank, or D/8 from the
// BF16 bank; the domain's residual reaches every lane through LDS (the group sits in one
// wavefront).
template <bool kBF16, int D>
__device__ __forceinline__ void local_solve_domain(
    const void* __restrict__ bank, const uint32_t* __restrict__ dom_nodes,
    const uint32_t* __restrict__ dom_node_off, uint32_t d, size_t row,
    const double* __restrict__ r0, const double* __restrict__ rc, uint32_t level0_domains,
    const float* __restrict__ dscale, float* __restrict__ c0, float* __restrict__ cc,
    float* __restrict__ rs, int lane) {
  const uint32_t n0 = dom_node_off[d], n1 = dom_node_off[d + 1];
  const uint32_t P = 3 * (n1 - n0);
  float rl = 0.0f;
  uint32_t node = 0;
  if (static_cast<uint32_t>(lane) < P) {
    node = dom_nodes[n0 + lane / 3];
    const uint32_t s = 3 * node + lane % 3;
    rl = static_cast<float>(d < level0_domains ? r0[s] : rc[s]);
  }
  if (lane < D) rs[lane] = rl;
  wave_lds_sync();
  if (lane >= D) return;                       // no row of the bank
  const float4* R = reinterpret_cast<const float4*>(rs);
  float acc = 0.0f;
  if (kBF16) {
    const uint4* Bi = reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(bank) + row * D);
#pragma unroll
    for (int q = 0; q < D / 8; ++q) {
      const uint4 b = Bi[q];
      const float4 ra = R[2 * q], rb = R[2 * q + 1];
      acc = fmaf(__uint_as_float(b.x << 16), ra.x, acc);
      acc = fmaf(__uint_as_float(b.x & 0xffff0000u), ra.y, acc);
      acc = fmaf(__uint_as_float(b.y << 16), ra.z, acc);
      acc = fmaf(__uint_as_float(b.y & 0xffff0000u), ra.w, acc);
      acc = fmaf(__uint_as_float(b.z << 16), rb.x, acc);
      acc = fmaf(__uint_as_float(b.z & 0xffff0000u), rb.y, acc);
      acc = fmaf(__uint_as_float(b.w << 16), rb.z, acc);
      acc = fmaf(__uint_as_float(b.w & 0xffff0000u), rb.w, acc);
    }
  } else {
    const float4* Bi = reinterpret_cast<const float4*>(static_cast<const float*>(bank) + row * D);
#pragma unroll
    for (int q = 0; q < D / 4; ++q) {
      const float4 b = Bi[q], rq = R[q];
      acc = fmaf(b.x, rq.x, acc);
      acc = fmaf(b.y, rq.y, acc);
      acc = fmaf(b.z, rq.z, acc);
      acc = fmaf(b.w, rq.w, acc);
    }
  }
  if (dscale) acc *= dscale[d];
  if (static_cast<uint32_t>(lane) < P) {
    const uint32_t s = 3 * node + lane % 3;
    if (d < level0_domains) c0[s] = acc;
    else cc[s] = acc;
  }
}

// Workgroups [0, nb32) hold the FP32-bank domains [0, bf16_first), the rest the BF16-bank
// domains [bf16_first, ndom): the bank type is uniform per workgroup.
template <int D>
__device__ __forceinline__ void local_solve_body(
    const float* __restrict__ inv, const uint16_t* __restrict__ inv_bf16, uint32_t bf16_first,
    const uint32_t* __restrict__ dom_nodes, const uint32_t* __restrict__ dom_node_off,
    uint32_t ndom, const double* __restrict__ r0, const double* __restrict__ rc,
    uint32_t level0_domains, const float* __restrict__ dscale, float* __restrict__ c0,
    float* __restrict__ cc, uint32_t bid) {
  constexpr int G = DomainShape<D>::lanes;
  constexpr uint32_t kPer = kB / G;
  __shared__ __align__(16) float rs[kPer][D];
  const int g = threadIdx.x / G, lane = threadIdx.x % G;
  const uint32_t nb32 = (bf16_first + kPer - 1) / kPer;
  if (bid < nb32) {
    const uint32_t d = bid * kPer + g;
    if (d >= bf16_first) return;
    local_solve_domain<false, D>(inv, dom_nodes, dom_node_off, d,
                                 static_cast<size_t>(d) * D + lane, r0, rc, level0_domains,
                                 dscale, c0, cc, rs[g], lane);
  } else {
    const uint32_t d = bf16_first + (bid - nb32) * kPer + g;
    if (d >= ndom) return;
    local_solve_domain<true, D>(inv_bf16, dom_nodes, dom_node_off, d,
                                static_cast<size_t>(d - bf16_first) * D + lane, r0, rc,
                                level0_domains, dscale, c0, cc, rs[g], lane);
  }
}

// Workgroups of the local solve for the FP32 bank domains [0, first) and the BF16 ones.
inline uint32_t local_solve_blocks(int D, uint32_t first, uint32_t ndom) {
  const uint32_t per = kB / (D <= 32 ? 32 : 64);
  return grid_for(first, per) + grid_for(ndom - first, per);
}

template <int D>
__global__ void __launch_bounds__(kB)
    local_solve_kernel(const float* __restrict__ inv, const uint16_t* __restrict__ inv_bf16,
                       uint32_t bf16_first, const uint32_t* __restrict__ dom_nodes,
                       const uint32_t* __restrict__ dom_node_off, uint32_t ndom,
                       const double* __restrict__ r0, const double* __restrict__ rc,
                       uint32_t level0_domains, const float* __restrict__ dscale,
                       float* __restrict__ c0, float* __restrict__ cc,
                       const double* __restrict__ stop) {
  if (mas_stopped(stop)) return;
  local_solve_body<D>(inv, inv_bf16, bf16_first, dom_nodes, dom_node_off, ndom, r0, rc,
                      level0_domains, dscale, c0, cc, blockIdx.x);
}

// z = w0 c0 + sum_l w_l c_l[map_l(node)], and r.z
__device__ __forceinline__ void prolong_body(const float* __restrict__ c0,
                                             const float* __restrict__ cc,
                                             const uint32_t* __restrict__ maps, uint32_t nv,
                                             uint32_t ncoarse_levels,
                                             const float* __restrict__ weights,
                                             const double* __restrict__ r,
                                             double* __restrict__ z, double* __restrict__ rz_slot,
                                             uint32_t bid) {
  const uint32_t v = bid * kB + threadIdx.x;
  double rz = 0.0;
  if (v < nv) {
    float s0 = weights[0] * c0[3 * v], s1 = weights[0] * c0[3 * v + 1],
          s2 = weights[0] * c0[3 * v + 2];
    for (uint32_t l = 0; l < ncoarse_levels; ++l) {
      const uint32_t J = maps[static_cast<size_t>(l) * nv + v];
      const float w = weights[l + 1];
      s0 = fmaf(w, cc[3 * J], s0);
      s1 = fmaf(w, cc[3 * J + 1], s1);
      s2 = fmaf(w, cc[3 * J + 2], s2);
    }
    const double z0 = s0, z1 = s1, z2 = s2;
    z[3 * v] = z0;
    z[3 * v + 1] = z1;
    z[3 * v + 2] = z2;
    rz = r[3 * v] * z0 + r[3 * v + 1] * z1 + r[3 * v + 2] * z2;
  }
  if (rz_slot) {
    __shared__ double lds[kB / 64];
    rz = block_sum<kB>(rz, lds);
    if (threadIdx.x == 0) atomic_add_f64(rz_slot + (bid % 64) * 16, rz);
  }
}

__global__ void __launch_bounds__(kB)
    prolong_kernel(const float* __restrict__ c0, const float* __restrict__ cc,
                   const uint32_t* __restrict__ maps, uint32_t nv, uint32_t ncoarse_levels,
                   const float* __restrict__ weights, const double* __restrict__ r,
                   double* __restrict__ z, double* __restrict__ rz_slot,
                   const double* __restrict__ stop) {
  if (mas_stopped(stop)) return;
  prolong_body(c0, cc, maps, nv, ncoarse_levels, weights, r, z, rz_slot, blockIdx.x);
}

// ------------------------------------------------------------------ batched apply
// The same four steps for several systems per launch: a workgroup reads from a work table
// which system's item it serves and which of that system's workgroups it is, so each
// system computes exactly what its own launch would. Finished systems (status set by the
// product pass) and passes beyond a system's last iteration do nothing.
__device__ __forceinline__ bool batch_idle(const MasBatchItem& it, int pass) {
  return it.k0 + pass > it.kmax || it.st[it.status_offset] != 0.0;
}

__global__ void __launch_bounds__(kB)
    batch_restrict_kernel(const MasBatchItem* __restrict__ I, const MasBatchWork* __restrict__ W,
                          int pass) {
  const MasBatchWork w = W[blockIdx.x];
  const MasBatchItem& it = I[w.item];
  if (batch_idle(it, pass)) return;
  restrict_body(it.r, it.m.restrict_offsets, it.m.restrict_items, it.n1, it.n2, it.b1, it.b2,
                it.m.coarse_residual, w.blk);
}

__global__ void __launch_bounds__(kB)
    batch_restrict_deep_kernel(const MasBatchItem* __restrict__ I,
                               const MasBatchWork* __restrict__ W, int pass) {
  const MasBatchWork w = W[blockIdx.x];
  const MasBatchItem& it = I[w.item];
  if (batch_idle(it, pass)) return;
  restrict_deep_body(it.m.restrict_offsets, it.m.restrict_items, it.n3, it.m.num_coarse_nodes,
                     it.m.coarse_residual, w.blk);
}

template <int D>
__global__ void __launch_bounds__(kB)
    batch_local_solve_kernel(const MasBatchItem* __restrict__ I,
                             const MasBatchWork* __restrict__ W, int pass) {
  const MasBatchWork w = W[blockIdx.x];
  const MasBatchItem& it = I[w.item];
  if (batch_idle(it, pass)) return;
  local_solve_body<D>(it.m.inverses, it.m.inverses_bf16, it.first, it.m.domain_nodes,
                   it.m.domain_node_offsets, it.m.num_domains, it.r, it.m.coarse_residual,
                   it.m.level0_domains, it.m.domain_scale, it.m.fine_correction,
                   it.m.coarse_correction, w.blk);
}

__global__ void __launch_bounds__(kB)
    batch_prolong_kernel(const MasBatchItem* __restrict__ I, const MasBatchWork* __restrict__ W,
                         int pass) {
  const MasBatchWork w = W[blockIdx.x];
  const MasBatchItem& it = I[w.item];
  if (batch_idle(it, pass)) return;
  const int k = it.k0 + pass;
  prolong_body(it.m.fine_correction, it.m.coarse_correction, it.m.prolong_maps, it.nv,
               it.m.num_coarse_levels, it.m.weights, it.r, it.z,
               it.st + it.rz_offset + static_cast<size_t>(k % 3) * it.rz_stride, w.blk);
}

}  // namespace
}  // namespace ysp

using namespace ysp;

extern "C" int ysp_mas_coarsen(const double* prev_values, int32_t prev_is_full_bsr,
                               uint32_t prev_nnz, const uint32_t* offsets, const uint32_t* items,
                               uint32_t nnz, double* values, void* stream) {
  if (nnz == 0) return YSP_OK;
  const dim3 g(grid_for(nnz, kB)), b(kB);
  if (prev_is_full_bsr)
    hipLaunchKernelGGL(coarsen_kernel<true>, g, b, 0, as_stream(stream), prev_values, prev_nnz,
                       offsets, items, nnz, values);
  else
    hipLaunchKernelGGL(coarsen_kernel<false>, g, b, 0, as_stream(stream), prev_values, prev_nnz,
                       offsets, items, nnz, values);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

namespace {
// 0 stands for 32; anything but 32 or 48 is refused.
int domain_size_of(uint32_t domain_size) {
  const uint32_t D = domain_size ? domain_size : 32;
  return D == 32 || D == 48 ? static_cast<int>(D) : 0;
}
}  // namespace

extern "C" int ysp_mas_domain_scatter(const double* values, int32_t is_full_bsr, uint32_t nnz,
                                      const uint32_t* dest, double* domains, uint32_t domain_size,
                                      void* stream) {
  const int D = domain_size_of(domain_size);
  if (!D) {
    ysp_set_error(__FILE__, __LINE__, "ysp_mas_domain_scatter: domain_size must be 32 or 48");
    return YSP_ERR_ARG;
  }
  if (nnz == 0) return YSP_OK;
  const dim3 g(grid_for(nnz, kB)), b(kB);
  hipStream_t s = as_stream(stream);
  if (is_full_bsr && D == 32)
    hipLaunchKernelGGL((domain_scatter_kernel<true, 32>), g, b, 0, s, values, nnz, dest, domains);
  else if (is_full_bsr)
    hipLaunchKernelGGL((domain_scatter_kernel<true, 48>), g, b, 0, s, values, nnz, dest, domains);
  else if (D == 32)
    hipLaunchKernelGGL((domain_scatter_kernel<false, 32>), g, b, 0, s, values, nnz, dest, domains);
  else
    hipLaunchKernelGGL((domain_scatter_kernel<false, 48>), g, b, 0, s, values, nnz, dest, domains);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_mas_invert_domains(double* domains, const uint32_t* domain_sizes,
                                      uint32_t num_domains, float* inverses,
                                      uint16_t* inverses_bf16, uint32_t bf16_first_domain,
                                      int32_t* status, uint32_t domain_size, void* stream) {
  const int D = domain_size_of(domain_size);
  const uint32_t fp32_domains = inverses_bf16 ? bf16_first_domain : num_domains;
  if (!D || (num_domains && (!domains || !domain_sizes || !status ||
                             (inverses_bf16 && bf16_first_domain > num_domains) ||
                             (fp32_domains && !inverses)))) {
    ysp_set_error(__FILE__, __LINE__, "ysp_mas_invert_domains: bad arguments");
    return YSP_ERR_ARG;
  }
  if (num_domains == 0) return YSP_OK;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(pad_identity_kernel, dim3(num_domains), dim3(64), 0, s, domains, domain_sizes,
                     num_domains, static_cast<uint32_t>(D));
  if (D == 32)
    hipLaunchKernelGGL(invert_domains_kernel<32>,
                       dim3(grid_for(num_domains, DomainShape<32>::per_wave * kInvWaves)),
                       dim3(64 * kInvWaves), 0, s, domains, inverses, inverses_bf16,
                       bf16_first_domain, num_domains, status);
  else
    hipLaunchKernelGGL(invert_domains_kernel<48>,
                       dim3(grid_for(num_domains, DomainShape<48>::per_wave * kInvWaves)),
                       dim3(64 * kInvWaves), 0, s, domains, inverses, inverses_bf16,
                       bf16_first_domain, num_domains, status);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

namespace ysp {
int launch_mas_apply(const MasApply& m, const double* r, uint32_t n, double* z,
                     double* rz_slots, const double* stop, hipStream_t s) {
  const uint32_t nv = n / 3;
  if (m.num_coarse_levels) {
    const uint32_t nc = m.num_coarse_nodes;
    const uint32_t n1 = m.level_offsets[1];
    const uint32_t n2 = m.level_offsets[m.num_coarse_levels >= 2 ? 2 : 1];
    const uint32_t n3 = m.level_offsets[m.num_coarse_levels >= 3 ? 3 : m.num_coarse_levels];
    const uint32_t b1 = grid_for(n1, kB / 8), b2 = grid_for(n2 - n1, kB / 64);
    const uint32_t blocks = b1 + b2 + (n3 - n2);
    if (blocks)
      hipLaunchKernelGGL(restrict_kernel, dim3(blocks), dim3(kB), 0, s, r, m.restrict_offsets,
                         m.restrict_items, n1, n2, b1, b2, m.coarse_residual, stop);
    if (nc > n3)
      hipLaunchKernelGGL(restrict_deep_kernel, dim3(grid_for(nc - n3, kB / 64)), dim3(kB), 0, s,
                         m.restrict_offsets, m.restrict_items, n3, nc, m.coarse_residual, stop);
  }
  if (m.num_domains) {
    const uint32_t first = m.inverses_bf16 ? m.bf16_first_domain : m.num_domains;
    const int D = m.domain_size == 48 ? 48 : 32;
    const dim3 grid(local_solve_blocks(D, first, m.num_domains));
    if (D == 32)
      hipLaunchKernelGGL(local_solve_kernel<32>, grid, dim3(kB), 0, s, m.inverses,
                         m.inverses_bf16, first, m.domain_nodes, m.domain_node_offsets,
                         m.num_domains, r, m.coarse_residual, m.level0_domains, m.domain_scale,
                         m.fine_correction, m.coarse_correction, stop);
    else
      hipLaunchKernelGGL(local_solve_kernel<48>, grid, dim3(kB), 0, s, m.inverses,
                         m.inverses_bf16, first, m.domain_nodes, m.domain_node_offsets,
                         m.num_domains, r, m.coarse_residual, m.level0_domains, m.domain_scale,
                         m.fine_correction, m.coarse_correction, stop);
  }
  hipLaunchKernelGGL(prolong_kernel, dim3(grid_for(nv, kB)), dim3(kB), 0, s, m.fine_correction,
                     m.coarse_correction, m.prolong_maps, nv, m.num_coarse_levels, m.weights, r, z,
                     rz_slots, stop);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

void mas_batch_geometry(MasBatchItem& it, uint32_t item, MasBatchTables& t) {
  const MasApply& m = it.m;
  it.nv = it.n / 3;
  it.n1 = it.n2 = it.n3 = it.b1 = it.b2 = 0;
  if (m.num_coarse_levels) {
    it.n1 = m.level_offsets[1];
    it.n2 = m.level_offsets[m.num_coarse_levels >= 2 ? 2 : 1];
    it.n3 = m.level_offsets[m.num_coarse_levels >= 3 ? 3 : m.num_coarse_levels];
    it.b1 = grid_for(it.n1, kB / 8);
    it.b2 = grid_for(it.n2 - it.n1, kB / 64);
    for (uint32_t b = 0, nb = it.b1 + it.b2 + (it.n3 - it.n2); b < nb; ++b)
      t.restrict_.push_back({item, b});
    if (m.num_coarse_nodes > it.n3)
      for (uint32_t b = 0, nb = grid_for(m.num_coarse_nodes - it.n3, kB / 64); b < nb; ++b)
        t.deep.push_back({item, b});
  }
  it.first = m.inverses_bf16 ? m.bf16_first_domain : m.num_domains;
  if (m.num_domains)
    for (uint32_t b = 0, nb = local_solve_blocks(m.domain_size == 48 ? 48 : 32, it.first,
                                                 m.num_domains);
         b < nb; ++b)
      (m.domain_size == 48 ? t.solve48 : t.solve).push_back({item, b});
  for (uint32_t b = 0, nb = grid_for(it.nv, kB); b < nb; ++b) t.prolong.push_back({item, b});
}

int launch_mas_apply_batch(const MasBatchItem* items,
                           const MasBatchWork* const tables[kMasBatchTables],
                           const uint32_t counts[kMasBatchTables], int pass, hipStream_t s) {
  if (counts[0])
    hipLaunchKernelGGL(batch_restrict_kernel, dim3(counts[0]), dim3(kB), 0, s, items, tables[0],
                       pass);
  if (counts[1])
    hipLaunchKernelGGL(batch_restrict_deep_kernel, dim3(counts[1]), dim3(kB), 0, s, items,
                       tables[1], pass);
  if (counts[2])
    hipLaunchKernelGGL(batch_local_solve_kernel<32>, dim3(counts[2]), dim3(kB), 0, s, items,
                       tables[2], pass);
  if (counts[4])
    hipLaunchKernelGGL(batch_local_solve_kernel<48>, dim3(counts[4]), dim3(kB), 0, s, items,
                       tables[4], pass);
  if (counts[3])
    hipLaunchKernelGGL(batch_prolong_kernel, dim3(counts[3]), dim3(kB), 0, s, items, tables[3],
                       pass);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

// Host-side check of the argument block against the kernels' assumptions.
int check_mas_args(const MasApply& m, uint32_t n) {
  if (n % 3 || m.num_coarse_levels > kMasMaxCoarseLevels || !m.weights || !m.fine_correction ||
      (m.num_domains && (!m.domain_nodes || !m.domain_node_offsets)) ||
      (m.inverses_bf16 && m.bf16_first_domain > m.num_domains) ||
      ((m.inverses_bf16 ? m.bf16_first_domain : m.num_domains) && !m.inverses) ||
      m.level0_domains > m.num_domains || (m.domain_size && m.domain_size != 32 &&
                                           m.domain_size != 48) ||
      (m.num_coarse_levels &&
       (!m.prolong_maps || !m.restrict_offsets || !m.restrict_items || !m.coarse_residual ||
        !m.coarse_correction || m.level_offsets[0] != 0 ||
        m.level_offsets[m.num_coarse_levels] != m.num_coarse_nodes)))
    return YSP_ERR_ARG;
  for (uint32_t l = 0; l < m.num_coarse_levels; ++l)
    if (m.level_offsets[l + 1] < m.level_offsets[l]) return YSP_ERR_ARG;
  return YSP_OK;
}
}  // namespace ysp

extern "C" int ysp_mas_apply(const ysp_mas_args* args, const double* r, uint32_t n, double* z,
                             void* stream) {
  if (!args) return YSP_ERR_ARG;
  const MasApply& m = *reinterpret_cast<const MasApply*>(args);
  if (check_mas_args(m, n) != YSP_OK) {
    ysp_set_error(__FILE__, __LINE__, "ysp_mas_apply: inconsistent arguments");
    return YSP_ERR_ARG;
  }
  if (n == 0) return YSP_OK;
  return launch_mas_apply(m, r, n, z, nullptr, nullptr, as_stream(stream));
}
