// Full (mirrored) 3x3 block-CSR used by the solve loop, its SpMV and the 3x3 PCG.
//
// The assembled Hessian keeps the reference's symmetric upper storage. For the Krylov
// loop it is mirrored once per assembly into a full block-CSR whose rows own every block
// they touch, so the SpMV is a pure row gather: no FP64 atomics (device-scope atomics
// cross the eight XCDs' L2 caches on MI355X) and no zero-fill of the output.
//
// Layout: row_ptr[nbr + 1], col[nnz] (block column), src[nnz] (bit 31: stored
// transposed, bit 30: dynamic part, low bits: upper block index) and component-major
// values v[j * stride + k] (j = 0..8 row-major entry of block k, stride = nnz rounded up
// to 64) so that 16 lanes reading 16 consecutive blocks issue fully coalesced loads.
#include "common.hpp"
#include "dense.hpp"
#include "mas.hpp"
#include "values.hpp"

#include <cmath>

#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <mutex>
#include <vector>

namespace ysp {
namespace {

constexpr unsigned kB = 256;

// Blocks per XCD-contiguous chunk: consecutive blockIdx values land on different XCDs
// (round robin over 8), so remap them so that each XCD walks one contiguous range of
// block rows and keeps its share of the gathered vector in its own L2.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t bid, uint32_t nblocks) {
  constexpr uint32_t X = 8;
  if (nblocks < X) return bid;
  const uint32_t per = nblocks / X;
  const uint32_t xcd = bid % X, slot = bid / X;
  if (slot >= per) return bid;  // tail of a grid that is not a multiple of 8
  return xcd * per + slot;
}

// Block b of a part writes the entries 2b and 2b + 1 of the part's range: (r, c) and
// its mirror (c, r), or, for a diagonal block, the sentinel key nbr^2 that sorts past
// every real entry. Fixed slots keep the order of equal keys (a block present in both
// parts) independent of scheduling, so the sums of the SpMV are reproducible.
__global__ void emit_entries_kernel(const uint32_t* __restrict__ pos, uint32_t nb, uint32_t tag,
                                    uint32_t nbr, uint64_t* __restrict__ keys,
                                    uint32_t* __restrict__ vals) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const uint32_t r = pos[2 * b] / 3, c = pos[2 * b + 1] / 3;
  const uint64_t sentinel = static_cast<uint64_t>(nbr) * nbr;
  keys[2 * static_cast<size_t>(b)] = static_cast<uint64_t>(r) * nbr + c;
  vals[2 * static_cast<size_t>(b)] = b | tag;
  keys[2 * static_cast<size_t>(b) + 1] = r == c ? sentinel : static_cast<uint64_t>(c) * nbr + r;
  vals[2 * static_cast<size_t>(b) + 1] = b | tag | kTrans;
}

// key / nbr for keys below 2^52: the double-precision quotient, corrected by one.
__device__ __forceinline__ uint32_t key_row(uint64_t key, uint32_t nbr) {
  uint64_t q = static_cast<uint64_t>(static_cast<double>(key) / static_cast<double>(nbr));
  if (q * nbr > key) --q;
  else if ((q + 1) * nbr <= key) ++q;
  return static_cast<uint32_t>(q);
}

// Number of real entries: the first sentinel in the sorted keys (one thread, binary
// search).
__global__ void count_entries_kernel(const uint64_t* __restrict__ keys, uint32_t n,
                                     uint64_t sentinel, uint32_t* __restrict__ nnz) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < sentinel) lo = mid + 1;
    else hi = mid;
  }
  *nnz = lo;
}

// Columns and row starts from the sorted keys, without atomics: entry k starts every
// row after the previous entry's row up to its own (empty rows included); k = nnz closes
// the remaining rows.
__global__ void split_keys_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ nnz_p,
                                  uint32_t nbr, uint32_t* __restrict__ col,
                                  uint32_t* __restrict__ row_ptr) {
  const uint32_t nnz = *nnz_p;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nnz) return;
  const uint32_t rk = k < nnz ? key_row(keys[k], nbr) : nbr;
  const uint32_t first = k == 0 ? 0 : key_row(keys[k - 1], nbr) + 1;
  for (uint32_t rr = first; rr <= rk; ++rr) row_ptr[rr] = k;
  if (k < nnz) col[k] = static_cast<uint32_t>(keys[k] - static_cast<uint64_t>(rk) * nbr);
}

template <typename VT>
__global__ void gather_values_kernel(const double* __restrict__ sv, const double* __restrict__ dv,
                                     const uint32_t* __restrict__ src, uint32_t nnz,
                                     VT* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint32_t s = src[k];
  const double* B = ((s & kDyn) ? dv : sv) + 9 * static_cast<size_t>(s & ~(kTrans | kDyn));
  const uint32_t stride = pad64(nnz);
  double v[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) v[j] = B[j];
  if (s & kTrans) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) out[(3 * i + j) * static_cast<size_t>(stride) + k] = static_cast<VT>(v[3 * j + i]);
  } else {
#pragma unroll
    for (int j = 0; j < 9; ++j) out[j * static_cast<size_t>(stride) + k] = static_cast<VT>(v[j]);
  }
}

constexpr int kG = 16;  // lanes per block row

// Block rows per workgroup of the fused PCG product: about 2048 workgroups, in multiples
// of the 16 rows one pass of a workgroup covers, at most kMaxRowsPerWg.
constexpr uint32_t kMaxRowsPerWg = 64;
// Block rows with more entries than this get a workgroup of their own in the PCG product.
constexpr uint32_t kLongRow3 = 128;
constexpr uint32_t kNoLongSlot = 0xffffffffu;
static_assert(3 * kMaxRowsPerWg <= kB, "one thread per scalar of a workgroup's rows");
inline uint32_t spmv_rows_per_wg(uint32_t nv) {
  constexpr uint32_t kSpmvWgs = 2048, kPass = kB / kG;
  const uint32_t rows = (grid_for(nv, kSpmvWgs) + kPass - 1) / kPass * kPass;
  return std::min<uint32_t>(kMaxRowsPerWg, std::max<uint32_t>(kPass, rows));
}

// One 16-lane group per block row: (y0, y1, y2) = sum over the row's blocks of B x_c,
// the blocks read from the value source (values.hpp).
template <typename Src>
__device__ __forceinline__ void row_product(const uint32_t* __restrict__ row_ptr,
                                            const uint32_t* __restrict__ col, const Src& vals,
                                            const double* __restrict__ x, uint32_t r, int l,
                                            double& y0, double& y1, double& y2) {
  y0 = y1 = y2 = 0.0;
  const uint32_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  for (uint32_t k = k0 + l; k < k1; k += kG) {
    const uint32_t c = 3 * col[k];
    const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
    double b[9];
    vals.load(k, b);
    block_fma(b, x0, x1, x2, y0, y1, y2);
  }
#pragma unroll
  for (int off = kG / 2; off > 0; off >>= 1) {
    y0 += __shfl_xor(y0, off, kG);
    y1 += __shfl_xor(y1, off, kG);
    y2 += __shfl_xor(y2, off, kG);
  }
}

template <typename Src>
__global__ void __launch_bounds__(kB)
    bsr3_full_spmv_kernel(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                          Src vals, uint32_t nbr, const double* __restrict__ x,
                          double* __restrict__ y) {
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint32_t r = blk * (kB / kG) + threadIdx.x / kG;
  const int l = threadIdx.x % kG;
  if (r >= nbr) return;
  double y0, y1, y2;
  row_product(row_ptr, col, vals, x, r, l, y0, y1, y2);
  if (l == 0) {
    y[3 * r] = y0;
    y[3 * r + 1] = y1;
    y[3 * r + 2] = y2;
  }
}

// ------------------------------------------------------------------ PCG (3x3 blocks)
// Scalar state: a header of named doubles, then the partial-sum slots. Each workgroup
// adds its partial dot into one of kSlots slots (blockIdx mod kSlots) spaced one cache
// line apart, so the device-scope FP64 atomics of thousands of workgroups spread over
// 64 addresses instead of serialising on one. Every set has its own 64 lines: a line
// that takes atomics in a kernel is never one that the same kernel's workgroups read
// (sharing lines across sets measured 25% slower). Consumers reduce the 64 slots with
// one wavefront load and a fixed shuffle tree, so every workgroup sees the same sum.
enum : int {
  S_RZ0 = 0, S_BB = 1, S_STATUS = 2, S_ITER = 3, S_THR = 4, S_MODE = 5,
  S_FINAL_RZ = 17, S_FINAL_RR = 18, S_CUR_ITER = 19, S_HEADER = 64
};
constexpr int kSlots = 64;
constexpr int kSpace = 16;                       // doubles between slots (128 B)
constexpr int kSet = kSlots * kSpace;            // one partial-sum array
enum : int { SET_PAP = 0, SET_RZ = 3, SET_RR = 6, SET_RZ0 = 9, SET_BB = 10, SET_COUNT = 11 };
constexpr int kStateDoubles = S_HEADER + SET_COUNT * kSet;

__host__ __device__ constexpr int slots_base(int set) { return S_HEADER + set * kSet; }

__device__ __forceinline__ double* slot_ptr(double* st, int set) {
  return st + slots_base(set) + (blockIdx.x % kSlots) * kSpace;
}

__device__ __forceinline__ double* slot_ptr_at(double* st, int set, uint32_t block) {
  return st + slots_base(set) + (block % kSlots) * kSpace;
}

// Sum of one slot array, identical in every thread of the workgroup.
__device__ __forceinline__ double read_slots(const double* st, int set, double* lds) {
  if (threadIdx.x < 64) {
    double v = st[slots_base(set) + threadIdx.x * kSpace];
    v = wave_sum(v);
    if (threadIdx.x == 0) lds[0] = v;
  }
  __syncthreads();
  const double r = lds[0];
  __syncthreads();
  return r;
}

// Three slot sums read by three wavefronts at once (one barrier instead of six).
__device__ __forceinline__ void read_slots3(const double* st, int set_a, int set_b, int set_c,
                                            double* lds, double& a, double& b, double& c) {
  const int w = threadIdx.x / 64;
  if (w < 3) {
    const int set = w == 0 ? set_a : (w == 1 ? set_b : set_c);
    double v = st[slots_base(set) + (threadIdx.x % 64) * kSpace];
    v = wave_sum(v);
    if (threadIdx.x % 64 == 0) lds[w] = v;
  }
  __syncthreads();
  a = lds[0];
  b = lds[1];
  c = lds[2];
  __syncthreads();
}

// Status read once per workgroup so that every thread takes the same early exit.
__device__ __forceinline__ bool stopped(const double* st) {
  __shared__ double s_status;
  if (threadIdx.x == 0) s_status = st[S_STATUS];
  __syncthreads();
  return s_status != 0.0;
}

__device__ __forceinline__ void wg_add(double v, double* dst, double* lds) {
  v = block_sum<kB>(v, lds);
  if (threadIdx.x == 0) atomic_add_f64(dst, v);
}

__device__ __forceinline__ void zero_set(double* st, int set) {
  if (threadIdx.x < kSlots) st[slots_base(set) + threadIdx.x * kSpace] = 0.0;
}

template <typename MT>
__device__ __forceinline__ void apply3(const MT* __restrict__ minv, uint32_t v, const double r[3],
                                       double z[3]) {
  const MT* M = minv + 9 * static_cast<size_t>(v);
#pragma unroll
  for (int i = 0; i < 3; ++i)
    z[i] = static_cast<double>(M[3 * i]) * r[0] + static_cast<double>(M[3 * i + 1]) * r[1] +
           static_cast<double>(M[3 * i + 2]) * r[2];
}

// r = b (or b - A x0 held in ax0), x = x0 or 0 and, with the block-Jacobi
// preconditioner (minv != nullptr), z = M^-1 r and the dots r.z, b.M^-1 b.
template <typename MT>
__global__ void __launch_bounds__(kB)
    pcg3_init_kernel(const MT* __restrict__ minv, const double* __restrict__ b,
                     const double* __restrict__ x0, const double* __restrict__ ax0,
                     double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                     uint32_t nv, double* __restrict__ st) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0, rr = 0, bmb = 0, bb = 0;
  if (v < nv) {
    double rl[3], bl[3], zl[3], tmp[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const uint32_t d = 3 * v + i;
      bl[i] = b[d];
      rl[i] = ax0 ? bl[i] - ax0[d] : bl[i];
      x[d] = x0 ? x0[d] : 0.0;
      r[d] = rl[i];
      rr += rl[i] * rl[i];
      bb += bl[i] * bl[i];
    }
    if (minv) {
      apply3(minv, v, rl, zl);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        z[3 * v + i] = zl[i];
        rz += rl[i] * zl[i];
      }
      if (x0) {
        apply3(minv, v, bl, tmp);
#pragma unroll
        for (int i = 0; i < 3; ++i) bmb += bl[i] * tmp[i];
      } else {
        bmb = rz;
      }
    }
  }
  __shared__ double lds[kB / 64];
  wg_add(rz, slot_ptr(st, SET_RZ + 0), lds);
  wg_add(rr, slot_ptr(st, SET_RR + 0), lds);
  wg_add(bmb, slot_ptr(st, SET_RZ0), lds);
  wg_add(bb, slot_ptr(st, SET_BB), lds);
}

__global__ void __launch_bounds__(64) pcg3_init_finish_kernel(double* st, double tol, int mode,
                                                             int rz0_is_rz) {
  __shared__ double lds[1];
  const double rz0 = read_slots(st, rz0_is_rz ? SET_RZ + 0 : SET_RZ0, lds);
  const double rz = read_slots(st, SET_RZ + 0, lds);
  const double bb = read_slots(st, SET_BB, lds);
  const double rr = read_slots(st, SET_RR + 0, lds);
  if (threadIdx.x != 0) return;
  st[S_RZ0] = rz0;
  st[S_BB] = bb;
  st[S_MODE] = mode;
  st[S_FINAL_RZ] = rz;
  st[S_FINAL_RR] = rr;
  if (!isfinite(rz0) || rz0 < 0.0 || !isfinite(rz) || rz < 0.0) {
    st[S_STATUS] = -2.0;
    return;
  }
  const double thr = (mode == 0) ? tol * rz0 : tol * tol * bb;
  st[S_THR] = thr;
  if (!isfinite(thr) || thr < 0.0) {
    st[S_STATUS] = -2.0;
    return;
  }
  if ((mode == 0 ? rz : rr) <= thr) st[S_STATUS] = 1.0;
}

// Iteration k, kernel 1: convergence test of iteration k-1, beta, then
//   p_k = z + beta p_{k-1},  q_k = A p_k = A z + beta q_{k-1},  p_k^T q_k.
// Only z is gathered (one neighbour gather per block); p and q are updated in place per
// row, so the product costs one pass over the blocks like a plain SpMV.
// Body of one workgroup of the product kernel: blk is the workgroup's chunk of block rows
// of this system, sb its partial-sum slot, lead the workgroup that writes the status.
template <typename Src>
__device__ __forceinline__ void pcg3_spmv_body(const uint32_t* __restrict__ row_ptr,
                                               const uint32_t* __restrict__ col, const Src& vals,
                                               uint32_t nbr, const double* __restrict__ z,
                                               double* __restrict__ p, double* __restrict__ q,
                                               double* __restrict__ st, int k,
                                               uint32_t rows_per_wg, uint32_t blk, uint32_t sb,
                                               bool lead_wg,
                                               const uint32_t* __restrict__ long_slot = nullptr,
                                               const double* __restrict__ ylong = nullptr,
                                               const uint32_t* __restrict__ chunk_start = nullptr) {
  if (stopped(st)) return;
  const int cur = k % 3, prev = (k + 2) % 3, prev2 = (k + 1) % 3;
  __shared__ double lds[kB / 64];
  double rz1, rr1, rz2;
  read_slots3(st, SET_RZ + prev, SET_RR + prev, SET_RZ + prev2, lds, rz1, rr1, rz2);
  const bool lead = lead_wg && threadIdx.x == 0;
  double beta = 0.0;
  if (k > 1) {
    if (!isfinite(rz1) || rz1 < 0.0 || !isfinite(rz2) || !(rz2 > 0.0)) {
      if (lead) {
        st[S_STATUS] = -1.0;
        st[S_ITER] = k - 1;
      }
      return;
    }
    if ((st[S_MODE] == 0.0 ? rz1 : rr1) <= st[S_THR]) {
      if (lead) {
        st[S_STATUS] = 1.0;
        st[S_ITER] = k - 1;
        st[S_FINAL_RZ] = rz1;
        st[S_FINAL_RR] = rr1;
      }
      return;
    }
    beta = rz1 / rz2;
  }
  if (lead) {
    st[S_FINAL_RZ] = rz1;
    st[S_FINAL_RR] = rr1;
    st[S_CUR_ITER] = k - 1;
  }
  // Each workgroup walks a contiguous chunk of block rows (rows_per_wg of them, or the
  // entry-balanced chunk_start[blk] .. chunk_start[blk + 1]; at most kMaxRowsPerWg), 16
  // at a time, so the grid stays near a few thousand workgroups: fewer partial-sum
  // atomics and fewer slot reads by the consumers than one workgroup per 16 rows. The
  // row products meet in LDS; the chunk's 3 * rows scalars of p, q and z are then
  // updated one per thread, in coalesced order.
  __shared__ double ys[3 * kMaxRowsPerWg];
  const uint32_t rbase = chunk_start ? chunk_start[blk] : blk * rows_per_wg;
  const uint32_t r_end = chunk_start ? chunk_start[blk + 1] : min(nbr, rbase + rows_per_wg);
  const int l = threadIdx.x % kG;
  for (uint32_t r = rbase + threadIdx.x / kG; r < r_end; r += kB / kG) {
    double y0, y1, y2;
    const uint32_t ls = long_slot ? long_slot[r] : kNoLongSlot;
    if (ls != kNoLongSlot) {           // a long row: its product came from pcg3_long_rows_kernel
      y0 = ylong[3 * static_cast<size_t>(ls)];
      y1 = ylong[3 * static_cast<size_t>(ls) + 1];
      y2 = ylong[3 * static_cast<size_t>(ls) + 2];
    } else {
      row_product(row_ptr, col, vals, z, r, l, y0, y1, y2);
    }
    if (l < 3) ys[3 * (r - rbase) + l] = l == 0 ? y0 : (l == 1 ? y1 : y2);
  }
  __syncthreads();
  double pap = 0.0;
  if (threadIdx.x < 3 * (r_end - rbase)) {
    const size_t d = 3 * static_cast<size_t>(rbase) + threadIdx.x;
    double pi = z[d], qi = ys[threadIdx.x];
    if (k > 1) {                               // first direction: p = z, q = A z
      pi = fma(p[d], beta, pi);
      qi = fma(q[d], beta, qi);
    }
    p[d] = pi;
    q[d] = qi;
    pap = pi * qi;
  }
  wg_add(pap, slot_ptr_at(st, SET_PAP + cur, sb), lds);
}

template <typename Src>
__global__ void __launch_bounds__(kB)
    pcg3_spmv_kernel(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                     Src vals, uint32_t nbr, const double* __restrict__ z, double* __restrict__ p,
                     double* __restrict__ q, double* __restrict__ st, int k,
                     uint32_t rows_per_wg, const uint32_t* __restrict__ long_slot,
                     const double* __restrict__ ylong, const uint32_t* __restrict__ chunk_start,
                     int reverse) {
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  pcg3_spmv_body(row_ptr, col, vals, nbr, z, p, q, st, k, rows_per_wg,
                     reverse ? gridDim.x - 1 - blk : blk, blockIdx.x, blockIdx.x == 0, long_slot,
                     ylong, chunk_start);
}

// The block rows longer than kLongRow3 entries (the coupling rows of affine bodies, the
// cloth rows under a ball): one workgroup per row computes (A z)_r into ylong ahead of
// the product kernel, which reads it instead of walking the row with 16 lanes.
// `reverse` walks the rows (and the product kernel its chunks) in the opposite order
// every other iteration: an operator larger than the Infinity Cache then starts each
// pass on the blocks the previous pass read last, which are still cached.
template <typename Src>
__global__ void __launch_bounds__(kB)
    pcg3_long_rows_kernel(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                          Src vals, const uint32_t* __restrict__ long_rows,
                          const double* __restrict__ z,
                          double* __restrict__ ylong, const double* __restrict__ st, int reverse) {
  if (stopped(st)) return;
  const uint32_t i = reverse ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
  const uint32_t r = long_rows[i];
  const uint32_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  double y0 = 0.0, y1 = 0.0, y2 = 0.0;
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += kB) {
    const uint32_t c = 3 * col[k];
    const double x0 = z[c], x1 = z[c + 1], x2 = z[c + 2];
    double b[9];
    vals.load(k, b);
    block_fma(b, x0, x1, x2, y0, y1, y2);
  }
  __shared__ double red[kB / 64];
  y0 = block_sum<kB>(y0, red);
  y1 = block_sum<kB>(y1, red);
  y2 = block_sum<kB>(y2, red);
  if (threadIdx.x == 0) {
    ylong[3 * static_cast<size_t>(i)] = y0;
    ylong[3 * static_cast<size_t>(i) + 1] = y1;
    ylong[3 * static_cast<size_t>(i) + 2] = y2;
  }
}

// Iteration k, kernel 2: alpha, x += alpha p, r -= alpha q, z = M^-1 r, r.z and r.r.
// One thread per scalar (coalesced vector traffic); a workgroup of 192 threads holds 64
// whole vertices, whose residual 3-vectors meet in LDS for the block-Jacobi product.
constexpr unsigned kUB = 192;
template <typename MT>
__device__ __forceinline__ void pcg3_update_body(const MT* __restrict__ minv,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 const double* __restrict__ p,
                                                 const double* __restrict__ q,
                                                 double* __restrict__ z, uint32_t n,
                                                 double* __restrict__ st, int k, uint32_t blk) {
  if (stopped(st)) return;
  const int cur = k % 3, prev = (k + 2) % 3, next = (k + 1) % 3;
  __shared__ double lds[4];
  double pap, rz1, unused;
  read_slots3(st, SET_PAP + cur, SET_RZ + prev, SET_RZ + prev, lds, pap, rz1, unused);
  const uint32_t t = blk * kUB + threadIdx.x;
  if (!isfinite(pap) || pap <= 0.0) {
    if (t == 0) {
      st[S_STATUS] = -1.0;
      st[S_ITER] = k;
    }
    return;
  }
  const double alpha = rz1 / pap;
  if (blk == 0) {
    zero_set(st, SET_PAP + next);
    zero_set(st, SET_RZ + next);
    zero_set(st, SET_RR + next);
  }
  __shared__ double rs[kUB];
  double rl = 0.0, zl = 0.0;
  if (t < n) {
    x[t] = fma(p[t], alpha, x[t]);
    rl = fma(-q[t], alpha, r[t]);
    r[t] = rl;
  }
  if (minv) {
    rs[threadIdx.x] = rl;
    __syncthreads();
    if (t < n) {
      const uint32_t c = threadIdx.x % 3, base = threadIdx.x - c;
      const MT* M = minv + 3 * static_cast<size_t>(t);    // row c of vertex t / 3
      zl = static_cast<double>(M[0]) * rs[base] + static_cast<double>(M[1]) * rs[base + 1] +
           static_cast<double>(M[2]) * rs[base + 2];
      z[t] = zl;
    }
  }
  // both partial sums in one block reduction
  double rz = wave_sum(rl * zl), rr = wave_sum(rl * rl);
  __shared__ double part[2][kUB / 64];
  if (threadIdx.x % 64 == 0) {
    part[0][threadIdx.x / 64] = rz;
    part[1][threadIdx.x / 64] = rr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    rz = part[0][0] + part[0][1] + part[0][2];
    rr = part[1][0] + part[1][1] + part[1][2];
    if (minv) atomic_add_f64(slot_ptr_at(st, SET_RZ + cur, blk), rz);
    atomic_add_f64(slot_ptr_at(st, SET_RR + cur, blk), rr);
  }
}

template <typename MT>
__global__ void __launch_bounds__(kUB)
    pcg3_update_kernel(const MT* __restrict__ minv, double* __restrict__ x, double* __restrict__ r,
                       const double* __restrict__ p, const double* __restrict__ q,
                       double* __restrict__ z, uint32_t n, double* __restrict__ st, int k) {
  pcg3_update_body<MT>(minv, x, r, p, q, z, n, st, k, blockIdx.x);
}

// Pinned ring of header copies, one per chunk of iterations in flight.
constexpr int kRing = 4;
constexpr int kBoxDoubles = 32;                  // header entries copied back per chunk
double* status_box() {
  static double* box = nullptr;
  if (!box && hipHostMalloc(reinterpret_cast<void**>(&box), kRing * kBoxDoubles * sizeof(double),
                            hipHostMallocDefault) != hipSuccess)
    box = nullptr;
  return box;
}

// Iterations still to enqueue so that the convergence test of the predicted last
// iteration runs, from two observations of the tested quantity (geometric rate);
// -1 when no prediction is possible.
long predict_more(double it_a, double val_a, double it_b, double val_b, double thr,
                  uint32_t enqueued) {
  if (!(it_b > it_a) || !(val_b > 0.0) || !(val_a > val_b) || !(thr > 0.0)) return -1;
  const double rate = (std::log(val_b) - std::log(val_a)) / (it_b - it_a);
  if (!(rate < 0.0)) return -1;
  const double left = std::max(0.0, (std::log(thr) - std::log(val_b)) / rate);
  const double detect = it_b + std::ceil(left) + 1.0;   // spmv pass that sees convergence
  return static_cast<long>(detect - static_cast<double>(enqueued));
}

template <typename Src>
void launch_full_spmv(const uint32_t* row_ptr, const uint32_t* col, const Src& vals, uint32_t nbr,
                      const double* x, double* y, hipStream_t s) {
  const unsigned rows_per_block = kB / kG;
  hipLaunchKernelGGL(bsr3_full_spmv_kernel<Src>, dim3(grid_for(nbr, rows_per_block)), dim3(kB), 0,
                     s, row_ptr, col, vals, nbr, x, y);
}

}  // namespace
}  // namespace ysp

using namespace ysp;

extern "C" int ysp_bsr3_build(const uint32_t* pos_static, uint32_t nb_static,
                              const uint32_t* pos_dynamic, uint32_t nb_dynamic, uint32_t n,
                              uint32_t* row_ptr, uint32_t* col, uint32_t* src, uint32_t capacity,
                              uint32_t* nnz_out, void* stream) {
  if (!nnz_out || n % 3) {
    ysp_set_error(__FILE__, __LINE__, "ysp_bsr3_build: bad arguments");
    return YSP_ERR_ARG;
  }
  *nnz_out = 0;
  hipStream_t s = as_stream(stream);
  const uint32_t nbr = n / 3;
  const size_t cap = 2 * (static_cast<size_t>(nb_static) + nb_dynamic);
  if (cap == 0) {
    YSP_CHECK(hipMemsetAsync(row_ptr, 0, sizeof(uint32_t) * (nbr + 1), s));
    return YSP_OK;
  }
  if (cap > 0xffffffffull) {
    ysp_set_error(__FILE__, __LINE__, "ysp_bsr3_build: too many blocks");
    return YSP_ERR_ARG;
  }
  uint64_t* keys = static_cast<uint64_t*>(ysp::scratch(40, sizeof(uint64_t) * cap));
  uint32_t* vals = static_cast<uint32_t*>(ysp::scratch(41, sizeof(uint32_t) * cap));
  uint64_t* keys2 = static_cast<uint64_t*>(ysp::scratch(42, sizeof(uint64_t) * cap));
  uint32_t* count = static_cast<uint32_t*>(ysp::scratch(43, sizeof(uint32_t)));
  uint32_t* vals2 = static_cast<uint32_t*>(ysp::scratch(45, sizeof(uint32_t) * cap));
  if (!keys || !vals || !keys2 || !count || !vals2) return YSP_ERR_HIP;
  if (nb_static)
    hipLaunchKernelGGL(emit_entries_kernel, dim3(grid_for(nb_static, kB)), dim3(kB), 0, s,
                       pos_static, nb_static, 0u, nbr, keys, vals);
  if (nb_dynamic)
    hipLaunchKernelGGL(emit_entries_kernel, dim3(grid_for(nb_dynamic, kB)), dim3(kB), 0, s,
                       pos_dynamic, nb_dynamic, kDyn, nbr, keys + 2 * static_cast<size_t>(nb_static),
                       vals + 2 * static_cast<size_t>(nb_static));
  YSP_CHECK_LAUNCH();
  // sort (row, col) keys (stable); the value keeps the source block and its orientation
  const uint64_t sentinel = static_cast<uint64_t>(nbr) * nbr;
  int end_bit = 1;
  while (end_bit < 64 && (1ull << end_bit) <= sentinel) ++end_bit;
  hipcub::DoubleBuffer<uint64_t> kb(keys, keys2);
  hipcub::DoubleBuffer<uint32_t> vb(vals, vals2);
  const uint32_t total = static_cast<uint32_t>(cap);
  size_t temp = 0;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, kb, vb, total, 0, end_bit, s));
  void* tmp = ysp::scratch(44, temp);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, temp, kb, vb, total, 0, end_bit, s));
  hipLaunchKernelGGL(count_entries_kernel, dim3(1), dim3(1), 0, s, kb.Current(), total, sentinel,
                     count);
  uint32_t nnz = 0;
  YSP_CHECK(hipMemcpyAsync(&nnz, count, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  if (nnz > capacity) {
    ysp_set_error(__FILE__, __LINE__, "ysp_bsr3_build: capacity too small");
    return YSP_ERR_ARG;
  }
  if (nnz)
    YSP_CHECK(hipMemcpyAsync(src, vb.Current(), sizeof(uint32_t) * nnz, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(split_keys_kernel, dim3(grid_for(static_cast<size_t>(nnz) + 1, kB)), dim3(kB),
                     0, s, kb.Current(), count, nbr, col, row_ptr);
  YSP_CHECK_LAUNCH();
  YSP_CHECK(hipStreamSynchronize(s));
  *nnz_out = nnz;
  return YSP_OK;
}

extern "C" int ysp_bsr3_gather_values(const double* static_values, const double* dynamic_values,
                                      const uint32_t* src, uint32_t nnz, double* out64,
                                      float* out32, void* stream) {
  if (nnz == 0) return YSP_OK;
  hipStream_t s = as_stream(stream);
  if (out64)
    hipLaunchKernelGGL(gather_values_kernel<double>, dim3(grid_for(nnz, kB)), dim3(kB), 0, s,
                       static_values, dynamic_values, src, nnz, out64);
  if (out32)
    hipLaunchKernelGGL(gather_values_kernel<float>, dim3(grid_for(nnz, kB)), dim3(kB), 0, s,
                       static_values, dynamic_values, src, nnz, out32);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

// ------------------------------------------------------------------ split build
// Upper block storage whose block dimensions are multiples of 3 (affine-body segments:
// 9x9, 9x3, 3x9 next to 3x3) as a mirrored full 3x3 block-CSR: every stored h x w block
// becomes (h/3)(w/3) sub-blocks, each entered at (r, c) and mirrored at (c, r); the
// sub-blocks of a diagonal block already hold both triangles and are entered once. The
// source of entry k is a 64-bit word: bits 0-39 the value offset of the sub-block's first
// element in its part, 40-41 w/3 - 1 (the row stride of the source block), 42 transposed,
// 43 dynamic part.
namespace {

// Entries 2i and 2i + 1 of this category's range for sub-block i = (block, a, b).
__global__ void emit_split_kernel(const uint32_t* __restrict__ pos, uint32_t nb, uint32_t h,
                                  uint32_t w, uint64_t value_start, uint64_t tag, uint32_t nbr,
                                  uint64_t* __restrict__ keys, uint64_t* __restrict__ vals) {
  const uint32_t sh = h / 3, sw = w / 3;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(nb) * sh * sw) return;
  const uint32_t b = static_cast<uint32_t>(i / (sh * sw));
  const uint32_t sub = static_cast<uint32_t>(i % (sh * sw));
  const uint32_t a = sub / sw, bb = sub % sw;
  const uint32_t pr = pos[2 * b], pc = pos[2 * b + 1];
  const uint32_t r = pr / 3 + a, c = pc / 3 + bb;
  const uint64_t off = value_start + static_cast<uint64_t>(b) * h * w + 3ull * a * w + 3ull * bb;
  const uint64_t word = off | (static_cast<uint64_t>(sw - 1) << 40) | tag;
  keys[2 * i] = static_cast<uint64_t>(r) * nbr + c;
  vals[2 * i] = word;
  keys[2 * i + 1] = pr == pc ? static_cast<uint64_t>(nbr) * nbr : static_cast<uint64_t>(c) * nbr + r;
  vals[2 * i + 1] = word | kSplitTrans;
}

__global__ void gather_split_kernel(const double* __restrict__ sv, const double* __restrict__ dv,
                                    const uint64_t* __restrict__ src, uint32_t nnz,
                                    double* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint64_t word = src[k];
  const double* B = ((word & kSplitDyn) ? dv : sv) + (word & kSplitOffset);
  const uint32_t w = 3 * (static_cast<uint32_t>((word >> 40) & 3u) + 1);
  const bool tr = (word & kSplitTrans) != 0;
  const size_t stride = pad64(nnz);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      out[(3 * i + j) * stride + k] = tr ? B[j * w + i] : B[i * w + j];
}

// Sub-block count of a part, 0 with *ok = false when a block dimension is not a
// multiple of 3 or wider than 12.
size_t split_entries(const ysp_block_matrix* part, bool* ok) {
  size_t total = 0;
  if (!part) return 0;
  for (uint32_t c = 0; c < part->num_categories; ++c) {
    const uint32_t h = part->dims[2 * c], w = part->dims[2 * c + 1];
    if (h % 3 || w % 3 || h == 0 || w == 0 || w > 12) *ok = false;
    total += 2ull * part->counts[c] * (h / 3) * (w / 3);
  }
  return total;
}
}  // namespace

extern "C" int ysp_bsr3_build_split(const ysp_block_matrix* part, const ysp_block_matrix* dynamic_part,
                                    uint32_t n, uint32_t* row_ptr, uint32_t* col, uint64_t* src,
                                    uint32_t capacity, uint32_t* nnz_out, void* stream) {
  if (!nnz_out || n % 3) {
    ysp_set_error(__FILE__, __LINE__, "ysp_bsr3_build_split: bad arguments");
    return YSP_ERR_ARG;
  }
  *nnz_out = 0;
  bool ok = true;
  const size_t cap = split_entries(part, &ok) + split_entries(dynamic_part, &ok);
  if (!ok) {
    ysp_set_error(__FILE__, __LINE__,
                  "ysp_bsr3_build_split: block dimensions must be multiples of 3, at most 12");
    return YSP_ERR_UNSUPPORTED;
  }
  hipStream_t s = as_stream(stream);
  const uint32_t nbr = n / 3;
  if (cap == 0) {
    YSP_CHECK(hipMemsetAsync(row_ptr, 0, sizeof(uint32_t) * (nbr + 1), s));
    return YSP_OK;
  }
  if (cap > 0xffffffffull) {
    ysp_set_error(__FILE__, __LINE__, "ysp_bsr3_build_split: too many blocks");
    return YSP_ERR_ARG;
  }
  uint64_t* keys = static_cast<uint64_t*>(ysp::scratch(60, sizeof(uint64_t) * cap));
  uint64_t* vals = static_cast<uint64_t*>(ysp::scratch(61, sizeof(uint64_t) * cap));
  uint64_t* keys2 = static_cast<uint64_t*>(ysp::scratch(62, sizeof(uint64_t) * cap));
  uint64_t* vals2 = static_cast<uint64_t*>(ysp::scratch(63, sizeof(uint64_t) * cap));
  uint32_t* count = static_cast<uint32_t*>(ysp::scratch(64, sizeof(uint32_t)));
  if (!keys || !vals || !keys2 || !vals2 || !count) return YSP_ERR_HIP;
  size_t base = 0;
  for (int which = 0; which < 2; ++which) {
    const ysp_block_matrix* m = which ? dynamic_part : part;
    if (!m) continue;
    uint32_t pos_offset = 0;
    for (uint32_t c = 0; c < m->num_categories; ++c) {
      const uint32_t nb = m->counts[c], h = m->dims[2 * c], w = m->dims[2 * c + 1];
      const size_t subs = static_cast<size_t>(nb) * (h / 3) * (w / 3);
      if (subs)
        hipLaunchKernelGGL(emit_split_kernel, dim3(grid_for(subs, kB)), dim3(kB), 0, s,
                           m->positions + 2 * static_cast<size_t>(pos_offset), nb, h, w,
                           static_cast<uint64_t>(m->starts[c]), which ? kSplitDyn : 0ull, nbr,
                           keys + base, vals + base);
      base += 2 * subs;
      pos_offset += nb;
    }
  }
  YSP_CHECK_LAUNCH();
  const uint64_t sentinel = static_cast<uint64_t>(nbr) * nbr;
  int end_bit = 1;
  while (end_bit < 64 && (1ull << end_bit) <= sentinel) ++end_bit;
  hipcub::DoubleBuffer<uint64_t> kb(keys, keys2);
  hipcub::DoubleBuffer<uint64_t> vb(vals, vals2);
  const uint32_t total = static_cast<uint32_t>(cap);
  size_t temp = 0;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, kb, vb, total, 0, end_bit, s));
  void* tmp = ysp::scratch(65, temp);
  if (!tmp) return YSP_ERR_HIP;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, temp, kb, vb, total, 0, end_bit, s));
  hipLaunchKernelGGL(count_entries_kernel, dim3(1), dim3(1), 0, s, kb.Current(), total, sentinel,
                     count);
  uint32_t nnz = 0;
  YSP_CHECK(hipMemcpyAsync(&nnz, count, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  if (nnz > capacity) {
    ysp_set_error(__FILE__, __LINE__, "ysp_bsr3_build_split: capacity too small");
    return YSP_ERR_ARG;
  }
  if (nnz)
    YSP_CHECK(hipMemcpyAsync(src, vb.Current(), sizeof(uint64_t) * nnz, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(split_keys_kernel, dim3(grid_for(static_cast<size_t>(nnz) + 1, kB)), dim3(kB),
                     0, s, kb.Current(), count, nbr, col, row_ptr);
  YSP_CHECK_LAUNCH();
  YSP_CHECK(hipStreamSynchronize(s));
  *nnz_out = nnz;
  return YSP_OK;
}

extern "C" int ysp_bsr3_gather_split(const double* static_values, const double* dynamic_values,
                                     const uint64_t* src, uint32_t nnz, double* out64,
                                     void* stream) {
  if (nnz == 0) return YSP_OK;
  if (!src || !out64) {
    ysp_set_error(__FILE__, __LINE__, "ysp_bsr3_gather_split: bad arguments");
    return YSP_ERR_ARG;
  }
  hipLaunchKernelGGL(gather_split_kernel, dim3(grid_for(nnz, kB)), dim3(kB), 0, as_stream(stream),
                     static_values, dynamic_values, src, nnz, out64);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_bsr3_spmv(const uint32_t* row_ptr, const uint32_t* col, const void* values,
                             int32_t values_f32, uint32_t nnz, uint32_t n, const double* x,
                             double* y, void* stream) {
  if (n % 3) return YSP_ERR_ARG;
  if (n == 0) return YSP_OK;
  hipStream_t s = as_stream(stream);
  return with_values(values, values_f32, nnz, "ysp_bsr3_spmv: bad value source", [&](auto vals) {
    launch_full_spmv(row_ptr, col, vals, n / 3, x, y, s);
    YSP_CHECK_LAUNCH();
    return YSP_OK;
  });
}

namespace {
__global__ void cast_f32_kernel(const double* __restrict__ in, float* __restrict__ out, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = static_cast<float>(in[i]);
}
}  // namespace

extern "C" int ysp_cast_f32(const double* in, float* out, uint64_t n, void* stream) {
  if (n == 0) return YSP_OK;
  hipLaunchKernelGGL(cast_f32_kernel, dim3(grid_for(n, kB)), dim3(kB), 0, as_stream(stream), in,
                     out, static_cast<size_t>(n));
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

namespace {
__global__ void __launch_bounds__(kB)
    axpy_kernel(const double* __restrict__ d, double* __restrict__ x, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) x[i] += d[i];
}

// r = b - A x over the full block-CSR, with ||r||^2 and ||b||^2 partial sums (sums[0..1]).
template <typename Src>
__global__ void __launch_bounds__(kB)
    residual_kernel(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                    Src vals, uint32_t nbr, const double* __restrict__ x,
                    const double* __restrict__ b,
                    double* __restrict__ r, double* __restrict__ sums) {
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint32_t row = blk * (kB / kG) + threadIdx.x / kG;
  const int l = threadIdx.x % kG;
  double rr = 0.0, bb = 0.0;
  if (row < nbr) {
    double y0, y1, y2;
    row_product(row_ptr, col, vals, x, row, l, y0, y1, y2);
    if (l < 3) {
      const size_t d = 3 * static_cast<size_t>(row) + l;
      const double bl = b[d];
      const double rl = bl - (l == 0 ? y0 : (l == 1 ? y1 : y2));
      r[d] = rl;
      rr = rl * rl;
      bb = bl * bl;
    }
  }
  __shared__ double lds[kB / 64];
  rr = block_sum<kB>(rr, lds);
  bb = block_sum<kB>(bb, lds);
  if (threadIdx.x == 0) {
    atomic_add_f64(sums + 2 * (blockIdx.x % kSlots) * kSpace, rr);
    atomic_add_f64(sums + (2 * (blockIdx.x % kSlots) + 1) * kSpace, bb);
  }
}
}  // namespace

extern "C" int ysp_bsr3_refine(const uint32_t* row_ptr, const uint32_t* col, const void* values,
                               int32_t values_f32, uint32_t nnz, uint32_t n, const double* rhs,
                               double* solution, const double* correction, double* residual,
                               double* residual_sq, double* rhs_sq, void* stream) {
  if (n == 0) {
    if (residual_sq) *residual_sq = 0.0;
    if (rhs_sq) *rhs_sq = 0.0;
    return YSP_OK;
  }
  if (n % 3 || !row_ptr || !col || !values || !rhs || !solution || !residual) {
    ysp_set_error(__FILE__, __LINE__, "ysp_bsr3_refine: bad arguments");
    return YSP_ERR_ARG;
  }
  hipStream_t s = as_stream(stream);
  const size_t nsum = 2 * kSlots * kSpace;
  double* sums = static_cast<double*>(ysp::scratch(33, nsum * sizeof(double)));
  if (!sums) return YSP_ERR_HIP;
  YSP_CHECK(hipMemsetAsync(sums, 0, nsum * sizeof(double), s));
  if (correction)
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n, kB)), dim3(kB), 0, s, correction, solution,
                       static_cast<size_t>(n));
  const uint32_t nbr = n / 3;
  const dim3 g(grid_for(nbr, kB / kG)), b(kB);
  if (int rc = with_values(values, values_f32, nnz, "ysp_bsr3_refine: bad value source",
                           [&](auto vals) {
                             hipLaunchKernelGGL(residual_kernel<decltype(vals)>, g, b, 0, s,
                                                row_ptr, col, vals, nbr, solution, rhs, residual,
                                                sums);
                             return YSP_OK;
                           }))
    return rc;
  YSP_CHECK_LAUNCH();
  double h[nsum];
  YSP_CHECK(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  double rr = 0.0, bb = 0.0;
  for (int i = 0; i < kSlots; ++i) {
    rr += h[2 * i * kSpace];
    bb += h[(2 * i + 1) * kSpace];
  }
  if (residual_sq) *residual_sq = rr;
  if (rhs_sq) *rhs_sq = bb;
  return YSP_OK;
}

namespace {
__global__ void __launch_bounds__(kB) fill_zero_kernel(double* __restrict__ p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// The long-row list cache, scratch slots 67-69 and the pinned status ring below are
// process-wide: single-system solves issued from several host threads run one at a time.
std::mutex g_pcg3_solve_mutex;

// Entry-balanced chunks of block rows for the PCG product, or none when the uniform
// chunks of spmv_rows_per_wg rows are balanced already (no short row longer than twice the
// mean). Cost of a row: its entries (one for a long row, whose product comes from the
// long-row kernel) plus 2 for its share of the p / q update. A chunk closes at
// kMaxRowsPerWg rows or when it holds the work of kChunkMeanRows mean rows; the result is
// the chunks' first rows and nv. On the contact scene's operator (121,576 block rows of 1
// to 1,005 entries) chunks of 8 / 16 / 32 / 64 mean rows measured 77.8 / 66.7 / 62.9 /
// 67.0 us per product, uniform 64-row chunks 77.6 us.
constexpr uint32_t kChunkMeanRows = 32;
std::vector<uint32_t> balanced_chunks(const std::vector<uint32_t>& rp, uint32_t nv) {
  std::vector<uint32_t> out;
  if (nv == 0) return out;
  uint64_t total = 0;
  uint32_t longest = 0;
  for (uint32_t r = 0; r < nv; ++r) {
    const uint32_t len = rp[r + 1] - rp[r];
    const uint32_t cost = (len > kLongRow3 ? 1u : len) + 2u;
    total += cost;
    if (len <= kLongRow3) longest = std::max(longest, len);
  }
  const double mean = static_cast<double>(total) / nv;
  if (longest + 2 <= 2.0 * mean) return out;
  const double target = kChunkMeanRows * mean;
  out.push_back(0);
  double acc = 0.0;
  uint32_t rows = 0;
  for (uint32_t r = 0; r < nv; ++r) {
    const uint32_t len = rp[r + 1] - rp[r];
    acc += (len > kLongRow3 ? 1u : len) + 2u;
    ++rows;
    if (rows == kMaxRowsPerWg || acc >= target) {
      if (r + 1 < nv) out.push_back(r + 1);
      acc = 0.0;
      rows = 0;
    }
  }
  out.push_back(nv);
  return out;
}

// no_alloc: the variant behind ysp_pcg3_solve_noalloc. It allocates nothing (no pinned
// status ring, no events, no long-row tables: every row goes through the 16-lane product)
// and clears the scalar state with a kernel; the host waits for each chunk's status with
// a stream synchronisation.
template <typename Src, typename MT>
int pcg3_run(uint32_t max_iterations, double threshold, int32_t mode, const uint32_t* row_ptr,
             const uint32_t* col, const Src& vals, uint32_t nnz, const MT* minv, const MasApply* mas,
             const double* rhs, uint32_t n, double* x, const double* x0, double* work,
             double* st, ysp_pcg_stats* stats, hipStream_t s, bool no_alloc = false) {
  std::unique_lock<std::mutex> solve_lock(g_pcg3_solve_mutex, std::defer_lock);
  if (!no_alloc) solve_lock.lock();   // the no-alloc variant touches none of that state
  const uint32_t nv = n / 3;
  double* r = work;
  double* z = work + n;
  double* pa = work + 2 * static_cast<size_t>(n);
  double* q = work + 4 * static_cast<size_t>(n);
  if (no_alloc)
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(kStateDoubles, kB)), dim3(kB), 0, s, st,
                       static_cast<uint32_t>(kStateDoubles));
  else
    YSP_CHECK(hipMemsetAsync(st, 0, kStateDoubles * sizeof(double), s));
  const double* ax0 = nullptr;
  if (x0) {
    launch_full_spmv(row_ptr, col, vals, nv, x0, q, s);
    ax0 = q;
  }
  // product grid: contiguous chunks of block rows, a multiple of 16 rows per workgroup
  const uint32_t rows_per_wg = spmv_rows_per_wg(nv);
  // rows longer than kLongRow3 (none in vertex meshes), listed when the operator's
  // (row_ptr, rows, entries) differ from the previous solve's. Any list of rows gives the
  // same product (a listed row is computed whole by its workgroup), so a list kept for an
  // operator rebuilt in place costs speed at most, never correctness.
  // With row lengths that vary widely (contact couplings), the product's chunks are cut
  // to about equal work instead (entry-balanced, listed with the long rows): a uniform
  // chunk of long rows would set the kernel's time while chunks of short rows idle.
  static struct {
    const uint32_t* row_ptr;
    uint32_t nv, nnz, num_long, num_chunks;
  } cache{nullptr, 0, 0, 0, 0};
  uint32_t num_long = 0, num_chunks = 0;
  uint32_t* long_rows = nullptr;
  uint32_t* long_slot = nullptr;
  double* ylong = nullptr;
  uint32_t* chunk_start = nullptr;
  if (no_alloc) {
    // every row through the product kernel, uniform chunks
  } else if (cache.row_ptr == row_ptr && cache.nv == nv && cache.nnz == nnz) {
    num_long = cache.num_long;
    num_chunks = cache.num_chunks;
    if (num_long) {
      long_rows = static_cast<uint32_t*>(ysp::scratch(67, sizeof(uint32_t) * num_long));
      long_slot = static_cast<uint32_t*>(ysp::scratch(68, sizeof(uint32_t) * nv));
      ylong = static_cast<double*>(ysp::scratch(69, sizeof(double) * 3 * num_long));
      if (!long_rows || !long_slot || !ylong) return YSP_ERR_HIP;
    }
    if (num_chunks) {
      chunk_start = static_cast<uint32_t*>(ysp::scratch(76, sizeof(uint32_t) * (num_chunks + 1)));
      if (!chunk_start) return YSP_ERR_HIP;
    }
  } else {
    cache = {nullptr, 0, 0, 0, 0};
    std::vector<uint32_t> rp(static_cast<size_t>(nv) + 1);
    YSP_CHECK(hipMemcpyAsync(rp.data(), row_ptr, sizeof(uint32_t) * rp.size(),
                             hipMemcpyDeviceToHost, s));
    YSP_CHECK(hipStreamSynchronize(s));
    std::vector<uint32_t> longs;
    for (uint32_t r = 0; r < nv; ++r)
      if (rp[r + 1] - rp[r] > kLongRow3) longs.push_back(r);
    const std::vector<uint32_t> chunks = balanced_chunks(rp, nv);
    if (!chunks.empty()) {
      num_chunks = static_cast<uint32_t>(chunks.size() - 1);
      chunk_start = static_cast<uint32_t*>(ysp::scratch(76, sizeof(uint32_t) * chunks.size()));
      if (!chunk_start) return YSP_ERR_HIP;
      YSP_CHECK(hipMemcpyAsync(chunk_start, chunks.data(), sizeof(uint32_t) * chunks.size(),
                               hipMemcpyHostToDevice, s));
      YSP_CHECK(hipStreamSynchronize(s));
    }
    if (!longs.empty()) {
      num_long = static_cast<uint32_t>(longs.size());
      std::vector<uint32_t> slot(nv, kNoLongSlot);
      for (uint32_t i = 0; i < num_long; ++i) slot[longs[i]] = i;
      long_rows = static_cast<uint32_t*>(ysp::scratch(67, sizeof(uint32_t) * num_long));
      long_slot = static_cast<uint32_t*>(ysp::scratch(68, sizeof(uint32_t) * nv));
      ylong = static_cast<double*>(ysp::scratch(69, sizeof(double) * 3 * num_long));
      if (!long_rows || !long_slot || !ylong) return YSP_ERR_HIP;
      YSP_CHECK(hipMemcpyAsync(long_rows, longs.data(), sizeof(uint32_t) * num_long,
                               hipMemcpyHostToDevice, s));
      YSP_CHECK(hipMemcpyAsync(long_slot, slot.data(), sizeof(uint32_t) * nv,
                               hipMemcpyHostToDevice, s));
      YSP_CHECK(hipStreamSynchronize(s));
    }
    cache = {row_ptr, nv, nnz, num_long, num_chunks};
  }
  static_assert(kStateDoubles <= YSP_PCG3_SCALARS, "cg_scalars too small");
  const dim3 vg(grid_for(nv, kB)), sg(num_chunks ? num_chunks : grid_for(nv, rows_per_wg)),
      ug(grid_for(n, kUB)), bl(kB);
  hipLaunchKernelGGL(pcg3_init_kernel<MT>, vg, bl, 0, s, minv, rhs, x0, ax0, x, r, z, nv, st);
  if (mas) {
    if (x0) {
      if (int rc = launch_mas_apply(*mas, rhs, n, z, st + slots_base(SET_RZ0), nullptr, s))
        return rc;
    }
    if (int rc = launch_mas_apply(*mas, r, n, z, st + slots_base(SET_RZ + 0), nullptr, s))
      return rc;
  }
  hipLaunchKernelGGL(pcg3_init_finish_kernel, dim3(1), dim3(64), 0, s, st, threshold, mode,
                     (mas && !x0) ? 1 : 0);
  YSP_CHECK_LAUNCH();
  const uint32_t end = max_iterations + 1;
  auto launch_pass = [&](uint32_t it) -> int {
    const int reverse = static_cast<int>(it & 1u);
    if (num_long)
      hipLaunchKernelGGL(pcg3_long_rows_kernel<Src>, dim3(num_long), bl, 0, s, row_ptr, col, vals,
                         long_rows, z, ylong, st, reverse);
    hipLaunchKernelGGL(pcg3_spmv_kernel<Src>, sg, bl, 0, s, row_ptr, col, vals, nv, z, pa, q,
                       st, static_cast<int>(it), rows_per_wg, long_slot, ylong, chunk_start,
                       reverse);
    if (it <= max_iterations) {
      hipLaunchKernelGGL(pcg3_update_kernel<MT>, ug, dim3(kUB), 0, s, minv, x, r, pa, q, z, n, st,
                         static_cast<int>(it));
      if (mas) {
        if (int rc = launch_mas_apply(*mas, r, n, z, st + slots_base(SET_RZ + (it % 3)),
                                      st + S_STATUS, s))
          return rc;
      }
    }
    return YSP_OK;
  };
  auto finish = [&]() -> int {
    double fin[S_HEADER];
    YSP_CHECK(hipMemcpyAsync(fin, st, sizeof(fin), hipMemcpyDeviceToHost, s));
    YSP_CHECK(hipStreamSynchronize(s));
    int code;
    const int iters = static_cast<int>(fin[S_ITER]);
    if (fin[S_STATUS] == 1.0) code = iters;
    else if (fin[S_STATUS] == -1.0) code = -iters - 4;
    else if (fin[S_STATUS] == -2.0) code = -5;
    else code = -4;
    if (stats) {
      stats->iterations = fin[S_STATUS] == 0.0 ? static_cast<int>(max_iterations) : iters;
      stats->status = code;
      stats->delta0 = fin[S_RZ0];
      stats->delta = fin[S_FINAL_RZ];
      stats->rr0 = fin[S_BB];
      stats->rr = fin[S_FINAL_RR];
    }
    return code;
  };
  if (no_alloc) {
    // chunks of passes, each followed by a synchronous read of the header; the chunk
    // size follows the observed convergence rate as in the pipelined loop below
    double hb[kBoxDoubles];
    uint32_t next = 1;
    double it_a = -1.0, val_a = 0.0;
    uint32_t chunk = 8;
    while (next <= end) {
      const uint32_t last = std::min<uint32_t>(end, next + chunk - 1);
      for (uint32_t it = next; it <= last; ++it)
        if (int rc = launch_pass(it)) return rc;
      YSP_CHECK_LAUNCH();
      next = last + 1;
      YSP_CHECK(hipMemcpyAsync(hb, st, sizeof(hb), hipMemcpyDeviceToHost, s));
      YSP_CHECK(hipStreamSynchronize(s));
      if (hb[S_STATUS] != 0.0) break;
      const double it_b = hb[S_CUR_ITER];
      const double val_b = hb[S_MODE] == 0.0 ? hb[S_FINAL_RZ] : hb[S_FINAL_RR];
      const long more = predict_more(it_a, val_a, it_b, val_b, hb[S_THR], next - 1);
      if (it_b > it_a) {
        it_a = it_b;
        val_a = val_b;
      }
      chunk = more < 0 ? 8u : static_cast<uint32_t>(std::min<long>(32, std::max<long>(1, more)));
    }
    return finish();
  }
  double* box = status_box();
  if (!box) {
    ysp_set_error(__FILE__, __LINE__, "pinned status allocation failed");
    return YSP_ERR_HIP;
  }
  hipEvent_t ev[kRing];
  for (int i = 0; i < kRing; ++i)
    YSP_CHECK(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
  struct EventGuard {
    hipEvent_t* e;
    ~EventGuard() {
      for (int i = 0; i < kRing; ++i) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  // Chunks of iterations are enqueued up to two ahead of the one the host waits for, so
  // the device never waits for the host; chunk sizes shrink as the observed convergence
  // rate predicts the last iteration, which bounds the no-op passes after convergence.
  // The extra spmv pass max+1 records convergence of iteration max.
  uint32_t next = 1;
  int head = 0, pending = 0;
  auto enqueue = [&](uint32_t size) -> int {
    const uint32_t last = std::min<uint32_t>(end, next + size - 1);
    for (uint32_t it = next; it <= last; ++it)
      if (int rc = launch_pass(it)) return rc;
    YSP_CHECK_LAUNCH();
    const int slot = (head + pending) % kRing;
    YSP_CHECK(hipMemcpyAsync(box + slot * kBoxDoubles, st, kBoxDoubles * sizeof(double),
                             hipMemcpyDeviceToHost, s));
    YSP_CHECK(hipEventRecord(ev[slot], s));
    ++pending;
    next = last + 1;
    return YSP_OK;
  };
  for (int i = 0; i < 2 && next <= end; ++i)
    if (int rc = enqueue(8)) return rc;
  double it_a = -1.0, val_a = 0.0;
  while (pending > 0) {
    const int slot = head;
    YSP_CHECK(hipEventSynchronize(ev[slot]));
    head = (head + 1) % kRing;
    --pending;
    const double* hb = box + slot * kBoxDoubles;
    if (hb[S_STATUS] != 0.0) break;
    if (next > end) continue;
    const double it_b = hb[S_CUR_ITER];
    const double val_b = hb[S_MODE] == 0.0 ? hb[S_FINAL_RZ] : hb[S_FINAL_RR];
    const long more = predict_more(it_a, val_a, it_b, val_b, hb[S_THR], next - 1);
    if (it_b > it_a) {
      it_a = it_b;
      val_a = val_b;
    }
    if (more < 0) {                        // no prediction yet: keep two chunks of 8 queued
      if (int rc = enqueue(8)) return rc;
      if (pending < 2 && next <= end)
        if (int rc = enqueue(8)) return rc;
      continue;
    }
    if (more == 0 && pending > 0) continue;  // the queued passes should reach convergence
    const uint32_t first = static_cast<uint32_t>(std::min<long>(16, std::max<long>(1, more)));
    if (int rc = enqueue(first)) return rc;
    if (more > 16 && pending < 2 && next <= end)
      if (int rc = enqueue(static_cast<uint32_t>(std::min<long>(16, more - 16)))) return rc;
  }
  return finish();
}
}  // namespace

namespace {
// Dispatch on the value source (values.hpp) and the inverse type (FP64 / FP32).
int pcg3_dispatch(uint32_t max_iterations, double threshold, int32_t threshold_mode,
                  const uint32_t* row_ptr, const uint32_t* col, const void* values,
                  int32_t value_format, uint32_t nnz, const void* diagonal_inverse,
                  int32_t inverse_f32, const double* rhs, uint32_t n, double* solution,
                  const double* x0, double* work, double* cg_scalars, ysp_pcg_stats* stats,
                  hipStream_t s, bool no_alloc) {
  return with_values(values, value_format, nnz, "PCG: bad value source", [&](auto vals) {
    if (inverse_f32)
      return pcg3_run(max_iterations, threshold, threshold_mode, row_ptr, col, vals, nnz,
                      static_cast<const float*>(diagonal_inverse), nullptr, rhs, n, solution, x0,
                      work, cg_scalars, stats, s, no_alloc);
    return pcg3_run(max_iterations, threshold, threshold_mode, row_ptr, col, vals, nnz,
                    static_cast<const double*>(diagonal_inverse), nullptr, rhs, n, solution, x0,
                    work, cg_scalars, stats, s, no_alloc);
  });
}
}  // namespace

extern "C" int ysp_pcg3_solve(uint32_t max_iterations, double threshold, int32_t threshold_mode,
                              const uint32_t* row_ptr, const uint32_t* col, const void* values,
                              int32_t values_f32, uint32_t nnz, const void* diagonal_inverse,
                              int32_t inverse_f32, const double* rhs, uint32_t n,
                              double* solution, const double* initial_guess,
                              int32_t zero_initial_guess, double* work, double* cg_scalars,
                              ysp_pcg_stats* stats, void* stream) {
  if (stats) *stats = ysp_pcg_stats{0, 0, 0, 0, 0, 0};
  if (n == 0) return 0;
  if (n % 3 || !work || !cg_scalars || (threshold_mode != 0 && threshold_mode != 1)) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg3_solve: bad arguments");
    return YSP_ERR_ARG;
  }
  const double* x0 = (!zero_initial_guess && initial_guess) ? initial_guess : nullptr;
  return pcg3_dispatch(max_iterations, threshold, threshold_mode, row_ptr, col, values,
                       values_f32, nnz, diagonal_inverse, inverse_f32, rhs, n, solution, x0, work,
                       cg_scalars, stats, as_stream(stream), false);
}

// ------------------------------------------------------------ one product pass
namespace {
// A running iteration k = 2 of the loop: r.z of the two previous iterations and r.r equal
// to 1 (beta = 1), the reference stopping rule with threshold 0 (never met); z, p, q set
// to nonzero data.
__global__ void __launch_bounds__(kB)
    product_state_kernel(double* __restrict__ st, double* __restrict__ z,
                         double* __restrict__ p, double* __restrict__ q, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    st[slots_base(SET_RZ + 1)] = 1.0;
    st[slots_base(SET_RR + 1)] = 1.0;
    st[slots_base(SET_RZ + 0)] = 1.0;
  }
  if (i < n) {
    z[i] = 1.0 + 0.01 * static_cast<double>(i % 7);
    p[i] = 0.5;
    q[i] = 0.25;
  }
}
}  // namespace

extern "C" int ysp_pcg3_product_pass(const uint32_t* row_ptr, const uint32_t* col,
                                     const void* values, int32_t values_f32, uint32_t nnz,
                                     uint32_t n, double* work, double* cg_scalars,
                                     int32_t prepare, void* stream) {
  if (n == 0) return YSP_OK;
  if (n % 3 || !row_ptr || !col || !values || !work || !cg_scalars) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg3_product_pass: bad arguments");
    return YSP_ERR_ARG;
  }
  hipStream_t s = as_stream(stream);
  const uint32_t nv = n / 3;
  double* z = work + n;
  double* pa = work + 2 * static_cast<size_t>(n);
  double* q = work + 4 * static_cast<size_t>(n);
  if (prepare) {
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(kStateDoubles, kB)), dim3(kB), 0, s,
                       cg_scalars, static_cast<uint32_t>(kStateDoubles));
    hipLaunchKernelGGL(product_state_kernel, dim3(grid_for(n, kB)), dim3(kB), 0, s, cg_scalars,
                       z, pa, q, n);
  }
  const uint32_t rows_per_wg = spmv_rows_per_wg(nv);
  const dim3 sg(grid_for(nv, rows_per_wg)), bl(kB);
  return with_values(values, values_f32, nnz, "ysp_pcg3_product_pass: bad value source",
                     [&](auto vals) {
                       hipLaunchKernelGGL(pcg3_spmv_kernel<decltype(vals)>, sg, bl, 0, s, row_ptr,
                                          col, vals, nv, z, pa, q, cg_scalars, 2, rows_per_wg,
                                          nullptr, nullptr, nullptr, 0);
                       YSP_CHECK_LAUNCH();
                       return YSP_OK;
                     });
}

// ---------------------------------------------------------- PCG without allocations
// Scalar state (first YSP_PCG3_SCALARS doubles) and work vectors (5 n doubles) of
// ysp_pcg3_solve_noalloc. An array of the code object: every GPU the library runs on
// holds its own copy, at its own address, until the process ends.
static __device__ __align__(256) double
    scratch[YSP_PCG3_SCALARS + 5 * static_cast<size_t>(YSP_PCG3_NOALLOC_MAX_N)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
double* g_noalloc_base[kMaxDevices] = {};

// Device address of the current GPU's copy of the array, looked up once per device. In
// host code the array's name stands for a host-side shadow, not for device memory: only
// the looked-up address may reach a kernel or a copy.
int noalloc_base(double** out) {
  int dev = -1;
  YSP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg3_solve_noalloc: device ordinal out of range");
    return YSP_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_base[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !p) {
      ysp_set_error(__FILE__, __LINE__, e != hipSuccess ? hipGetErrorString(e)
                                                        : "hipGetSymbolAddress returned NULL");
      return YSP_ERR_HIP;
    }
    g_noalloc_base[dev] = static_cast<double*>(p);
  }
  *out = g_noalloc_base[dev];
  return YSP_OK;
}
}  // namespace

extern "C" int ysp_pcg3_solve_noalloc(uint32_t max_iterations, double threshold,
                                      int32_t threshold_mode, const uint32_t* row_ptr,
                                      const uint32_t* col, const void* values, int32_t values_f32,
                                      uint32_t nnz, const void* diagonal_inverse,
                                      int32_t inverse_f32, const double* rhs, uint32_t n,
                                      double* solution, const double* initial_guess,
                                      int32_t zero_initial_guess, ysp_pcg_stats* stats,
                                      void* stream) {
  if (stats) *stats = ysp_pcg_stats{0, 0, 0, 0, 0, 0};
  if (n == 0) return 0;
  if (n % 3 || (threshold_mode != 0 && threshold_mode != 1)) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg3_solve_noalloc: bad arguments");
    return YSP_ERR_ARG;
  }
  if (n > YSP_PCG3_NOALLOC_MAX_N) {
    ysp_set_error(__FILE__, __LINE__,
                  "ysp_pcg3_solve_noalloc: n exceeds YSP_PCG3_NOALLOC_MAX_N (393216 unknowns); "
                  "use ysp_pcg3_solve with caller-provided work buffers");
    return YSP_ERR_ARG;
  }
  double* base = nullptr;
  if (int rc = noalloc_base(&base)) return rc;
  const double* x0 = (!zero_initial_guess && initial_guess) ? initial_guess : nullptr;
  return pcg3_dispatch(max_iterations, threshold, threshold_mode, row_ptr, col, values,
                       values_f32, nnz, diagonal_inverse, inverse_f32, rhs, n, solution, x0,
                       base + YSP_PCG3_SCALARS, base, stats, as_stream(stream), true);
}

extern "C" int ysp_pcg3_mas_solve(uint32_t max_iterations, double threshold,
                                  int32_t threshold_mode, const uint32_t* row_ptr,
                                  const uint32_t* col, const void* values, int32_t values_f32,
                                  uint32_t nnz, const ysp_mas_args* mas, const double* rhs,
                                  uint32_t n, double* solution, const double* initial_guess,
                                  int32_t zero_initial_guess, double* work, double* cg_scalars,
                                  ysp_pcg_stats* stats, void* stream) {
  if (stats) *stats = ysp_pcg_stats{0, 0, 0, 0, 0, 0};
  if (n == 0) return 0;
  if (n % 3 || !work || !cg_scalars || !mas || (threshold_mode != 0 && threshold_mode != 1)) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg3_mas_solve: bad arguments");
    return YSP_ERR_ARG;
  }
  hipStream_t s = as_stream(stream);
  const double* x0 = (!zero_initial_guess && initial_guess) ? initial_guess : nullptr;
  const MasApply* m = reinterpret_cast<const MasApply*>(mas);
  if (check_mas_args(*m, n) != YSP_OK) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg3_mas_solve: inconsistent MAS arguments");
    return YSP_ERR_ARG;
  }
  const float* no_minv = nullptr;
  return with_values(values, values_f32, nnz, "ysp_pcg3_mas_solve: bad value source",
                     [&](auto vals) {
                       return pcg3_run(max_iterations, threshold, threshold_mode, row_ptr, col,
                                       vals, nnz, no_minv, m, rhs, n, solution, x0, work,
                                       cg_scalars, stats, s);
                     });
}

// ------------------------------------------------------------------ batched PCG
// Several independent 3x3 systems advance together: every product / update launch
// covers all of them, a workgroup reads which system and which chunk of it it owns from
// a work table, and each system keeps its own scalar state and partial-sum slots
// (segmented reductions). The arithmetic of a system does not depend on the others in
// the batch (same row chunks and slots as the single-system loop).
namespace {
struct BatchSystem {
  const uint32_t* row_ptr;
  const uint32_t* col;
  const void* values;
  const void* minv;
  double* x;
  double* r;
  double* z;
  double* p;
  double* q;
  double* st;
  uint32_t stride;
  uint32_t nbr;
  uint32_t n;
  uint32_t rows_per_wg;
  int32_t k0;     // iteration number of the first pass of this run
  int32_t kmax;   // max_iterations
};

struct BatchWork {
  uint32_t sys;
  uint32_t blk;
};


template <typename VT>
__global__ void __launch_bounds__(kB)
    batch_spmv_kernel(const BatchSystem* __restrict__ S, const BatchWork* __restrict__ W,
                      int pass) {
  const BatchWork w = W[blockIdx.x];
  const BatchSystem& sy = S[w.sys];
  const int k = sy.k0 + pass;
  if (k > sy.kmax + 1) return;
  pcg3_spmv_body(sy.row_ptr, sy.col, Mirrored<VT>{static_cast<const VT*>(sy.values), sy.stride},
                 sy.nbr,
                     sy.z, sy.p, sy.q, sy.st, k, sy.rows_per_wg, w.blk, w.blk, w.blk == 0);
}

template <typename MT>
__global__ void __launch_bounds__(kUB)
    batch_update_kernel(const BatchSystem* __restrict__ S, const BatchWork* __restrict__ W,
                        int pass) {
  const BatchWork w = W[blockIdx.x];
  const BatchSystem& sy = S[w.sys];
  const int k = sy.k0 + pass;
  if (k > sy.kmax) return;
  pcg3_update_body<MT>(static_cast<const MT*>(sy.minv), sy.x, sy.r, sy.p, sy.q, sy.z, sy.n, sy.st,
                       k, w.blk);
}

constexpr int kBatchOut = 5;
__global__ void batch_status_kernel(const BatchSystem* __restrict__ S, uint32_t count,
                                    double* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double* st = S[i].st;
  out[kBatchOut * i + 0] = st[S_STATUS];
  out[kBatchOut * i + 1] = st[S_ITER];
  out[kBatchOut * i + 2] = st[S_FINAL_RZ];
  out[kBatchOut * i + 3] = st[S_FINAL_RR];
  out[kBatchOut * i + 4] = st[S_RZ0];
}

template <typename VT, typename MT>
int batch_run(ysp_pcg3_system* systems, uint32_t count, uint32_t passes, hipStream_t s) {
  std::vector<BatchSystem> host(count);
  std::vector<BatchWork> wsp, wup;
  for (uint32_t i = 0; i < count; ++i) {
    ysp_pcg3_system& y = systems[i];
    const uint32_t n = y.n, nv = n / 3;
    BatchSystem& b = host[i];
    b.row_ptr = y.row_ptr;
    b.col = y.col;
    b.values = y.values;
    b.minv = y.mas ? nullptr : y.diagonal_inverse;   // MAS: the update leaves z to the apply
    b.x = y.solution;
    b.r = y.work;
    b.z = y.work + n;
    b.p = y.work + 2 * static_cast<size_t>(n);
    b.q = y.work + 4 * static_cast<size_t>(n);
    b.st = y.cg_scalars;
    b.stride = pad64(y.nnz);
    b.nbr = nv;
    b.n = n;
    b.rows_per_wg = spmv_rows_per_wg(nv);
    b.k0 = static_cast<int32_t>(y.passes_done) + 1;
    b.kmax = static_cast<int32_t>(y.max_iterations);
    for (uint32_t k = 0, nb = grid_for(nv, b.rows_per_wg); k < nb; ++k) wsp.push_back({i, k});
    for (uint32_t k = 0, nb = grid_for(n, kUB); k < nb; ++k) wup.push_back({i, k});
  }
  BatchSystem* dsys = static_cast<BatchSystem*>(ysp::scratch(35, sizeof(BatchSystem) * count));
  BatchWork* dsp = static_cast<BatchWork*>(ysp::scratch(36, sizeof(BatchWork) * wsp.size()));
  BatchWork* dup = static_cast<BatchWork*>(ysp::scratch(37, sizeof(BatchWork) * wup.size()));
  double* dout = static_cast<double*>(ysp::scratch(38, sizeof(double) * kBatchOut * count));
  if (!dsys || !dsp || !dup || !dout) return YSP_ERR_HIP;
  YSP_CHECK(hipMemcpyAsync(dsys, host.data(), sizeof(BatchSystem) * count, hipMemcpyHostToDevice, s));
  YSP_CHECK(hipMemcpyAsync(dsp, wsp.data(), sizeof(BatchWork) * wsp.size(), hipMemcpyHostToDevice, s));
  YSP_CHECK(hipMemcpyAsync(dup, wup.data(), sizeof(BatchWork) * wup.size(), hipMemcpyHostToDevice, s));
  // systems preconditioned by MAS: one item each and the work tables of the apply
  std::vector<MasBatchItem> items;
  MasBatchTables mt;
  for (uint32_t i = 0; i < count; ++i) {
    if (!systems[i].mas) continue;
    const BatchSystem& b = host[i];
    MasBatchItem it{};
    it.m = *reinterpret_cast<const MasApply*>(systems[i].mas);
    it.r = b.r;
    it.z = b.z;
    it.st = b.st;
    it.n = b.n;
    it.k0 = b.k0;
    it.kmax = b.kmax;
    it.status_offset = S_STATUS;
    it.rz_offset = slots_base(SET_RZ);
    it.rz_stride = kSet;
    mas_batch_geometry(it, static_cast<uint32_t>(items.size()), mt);
    items.push_back(it);
  }
  const MasBatchWork* dtab[kMasBatchTables] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  uint32_t ntab[kMasBatchTables] = {0, 0, 0, 0, 0};
  MasBatchItem* ditems = nullptr;
  if (!items.empty()) {
    ditems = static_cast<MasBatchItem*>(ysp::scratch(70, sizeof(MasBatchItem) * items.size()));
    if (!ditems) return YSP_ERR_HIP;
    YSP_CHECK(hipMemcpyAsync(ditems, items.data(), sizeof(MasBatchItem) * items.size(),
                             hipMemcpyHostToDevice, s));
    const std::vector<MasBatchWork>* src[kMasBatchTables] = {&mt.restrict_, &mt.deep, &mt.solve,
                                                             &mt.prolong, &mt.solve48};
    for (int t = 0; t < kMasBatchTables; ++t) {
      ntab[t] = static_cast<uint32_t>(src[t]->size());
      if (!ntab[t]) continue;
      MasBatchWork* d = static_cast<MasBatchWork*>(ysp::scratch(71 + t, sizeof(MasBatchWork) * ntab[t]));
      if (!d) return YSP_ERR_HIP;
      YSP_CHECK(hipMemcpyAsync(d, src[t]->data(), sizeof(MasBatchWork) * ntab[t],
                               hipMemcpyHostToDevice, s));
      dtab[t] = d;
    }
  }
  const uint32_t nsp = static_cast<uint32_t>(wsp.size()), nup = static_cast<uint32_t>(wup.size());
  for (uint32_t pass = 0; pass < passes; ++pass) {
    hipLaunchKernelGGL(batch_spmv_kernel<VT>, dim3(nsp), dim3(kB), 0, s, dsys, dsp,
                       static_cast<int>(pass));
    hipLaunchKernelGGL(batch_update_kernel<MT>, dim3(nup), dim3(kUB), 0, s, dsys, dup,
                       static_cast<int>(pass));
    if (ditems)
      if (int rc = launch_mas_apply_batch(ditems, dtab, ntab, static_cast<int>(pass), s)) return rc;
  }
  hipLaunchKernelGGL(batch_status_kernel, dim3(grid_for(count, 64)), dim3(64), 0, s, dsys, count,
                     dout);
  YSP_CHECK_LAUNCH();
  std::vector<double> out(kBatchOut * count);
  YSP_CHECK(hipMemcpyAsync(out.data(), dout, sizeof(double) * out.size(), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  for (uint32_t i = 0; i < count; ++i) {
    ysp_pcg3_system& y = systems[i];
    y.passes_done += passes;
    const double st = out[kBatchOut * i], it = out[kBatchOut * i + 1];
    const int iters = static_cast<int>(it);
    if (st == 1.0) y.status = iters;
    else if (st == -1.0) y.status = -iters - 4;
    else if (st == -2.0) y.status = -5;
    else y.status = (y.passes_done > y.max_iterations) ? -4 : YSP_PCG3_RUNNING;
    y.iterations = st == 0.0 ? static_cast<int32_t>(std::min(y.passes_done, y.max_iterations)) : iters;
    y.delta = out[kBatchOut * i + 2];
    y.rr = out[kBatchOut * i + 3];
    y.delta0 = out[kBatchOut * i + 4];
  }
  return YSP_OK;
}
}  // namespace

extern "C" int ysp_pcg3_batch_start(ysp_pcg3_system* y, double threshold, int32_t threshold_mode,
                                    int32_t inverse_f32, void* stream) {
  if (!y || y->n % 3 || !y->work || !y->cg_scalars || !y->rhs || !y->solution ||
      (threshold_mode != 0 && threshold_mode != 1)) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg3_batch_start: bad arguments");
    return YSP_ERR_ARG;
  }
  y->passes_done = 0;
  y->status = YSP_PCG3_RUNNING;
  y->iterations = 0;
  if (y->n == 0) {
    y->status = 0;
    return YSP_OK;
  }
  hipStream_t s = as_stream(stream);
  const uint32_t n = y->n, nv = n / 3;
  YSP_CHECK(hipMemsetAsync(y->cg_scalars, 0, kStateDoubles * sizeof(double), s));
  const dim3 vg(grid_for(nv, kB)), bl(kB);
  if (y->mas) {
    // r = b, x = 0, then z = M r and r.z from the MAS apply (the single-system start)
    const MasApply& m = *reinterpret_cast<const MasApply*>(y->mas);
    if (check_mas_args(m, n) != YSP_OK) {
      ysp_set_error(__FILE__, __LINE__, "ysp_pcg3_batch_start: inconsistent MAS arguments");
      return YSP_ERR_ARG;
    }
    const float* no_minv = nullptr;
    hipLaunchKernelGGL(pcg3_init_kernel<float>, vg, bl, 0, s, no_minv, y->rhs, nullptr, nullptr,
                       y->solution, y->work, y->work + n, nv, y->cg_scalars);
    if (int rc = launch_mas_apply(m, y->work, n, y->work + n, y->cg_scalars + slots_base(SET_RZ + 0),
                                  nullptr, s))
      return rc;
    hipLaunchKernelGGL(pcg3_init_finish_kernel, dim3(1), dim3(64), 0, s, y->cg_scalars, threshold,
                       threshold_mode, 1);
    YSP_CHECK_LAUNCH();
    return YSP_OK;
  }
  if (inverse_f32)
    hipLaunchKernelGGL(pcg3_init_kernel<float>, vg, bl, 0, s,
                       static_cast<const float*>(y->diagonal_inverse), y->rhs, nullptr, nullptr,
                       y->solution, y->work, y->work + n, nv, y->cg_scalars);
  else
    hipLaunchKernelGGL(pcg3_init_kernel<double>, vg, bl, 0, s,
                       static_cast<const double*>(y->diagonal_inverse), y->rhs, nullptr, nullptr,
                       y->solution, y->work, y->work + n, nv, y->cg_scalars);
  hipLaunchKernelGGL(pcg3_init_finish_kernel, dim3(1), dim3(64), 0, s, y->cg_scalars, threshold,
                     threshold_mode, 0);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_pcg3_batch_run(ysp_pcg3_system* systems, uint32_t count, uint32_t passes,
                                  int32_t values_f32, int32_t inverse_f32, void* stream) {
  if (count == 0 || passes == 0) return YSP_OK;
  if (!systems) return YSP_ERR_ARG;
  for (uint32_t i = 0; i < count; ++i)
    if (systems[i].n == 0 || systems[i].n % 3 ||
        (!systems[i].diagonal_inverse && !systems[i].mas)) {
      ysp_set_error(__FILE__, __LINE__, "ysp_pcg3_batch_run: bad system");
      return YSP_ERR_ARG;
    }
  hipStream_t s = as_stream(stream);
  if (values_f32) {
    return inverse_f32 ? batch_run<float, float>(systems, count, passes, s)
                       : batch_run<float, double>(systems, count, passes, s);
  }
  return inverse_f32 ? batch_run<double, float>(systems, count, passes, s)
                     : batch_run<double, double>(systems, count, passes, s);
}

