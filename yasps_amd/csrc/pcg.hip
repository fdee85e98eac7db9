// Block-Jacobi PCG with device-resident recurrence scalars.
//
// One iteration k is three kernels on one stream:
//   (1) q = A p, plus p^T A p accumulated from the blocks' quadratic forms;
//   (2) alpha = rz / pAp, x += alpha p, r -= alpha q, q = 0, z = M^-1 r and the partial
//       dots r.z, r.r (one pass per diagonal block);
//   (3) beta = rz_new / rz_old, p = z + beta p and the convergence test.
// Scalars live in rotating sets (k mod 3) so that every kernel of iteration k reads
// complete sums and no kernel overwrites a set another kernel of the same launch reads.
// Each dot is a set of 64 partial-sum slots (one per 128-byte line, workgroup w adds into
// slot w mod 64); consumers reduce a set with one wavefront load and a fixed shuffle
// tree, so every workgroup sees the same sum.
// Every kernel returns immediately once a status was recorded, so the host can enqueue
// several iterations ahead and read the status once per chunk (the reference copies two
// doubles to the host after every iteration, solverKernel.pyx:667-668).
#include "common.hpp"
#include "dense.hpp"
#include "spmv.hpp"
#include "values.hpp"

#include <algorithm>
#include <vector>

namespace ysp {
namespace {

enum : int {
  S_RZ0 = 0,      // delta0 (reference: r0^T M^-1 r0, or b^T M^-1 b with an initial guess)
  S_BB = 1,       // ||b||^2
  S_STATUS = 2,   // 0 running, 1 converged, -1 non-positive curvature, -2 invalid start
  S_ITER = 3,     // iteration at which the status was set
  S_THR = 4,      // absolute threshold on rz (mode 0) or rr (mode 1)
  S_MODE = 5,
  S_FINAL_RZ = 17,
  S_FINAL_RR = 18,
  S_COUNT = 64    // header; the partial-sum sets follow
};
enum : int { SET_PAP = 0, SET_RZ = 3, SET_RR = 6, SET_RZ0 = 9, SET_BB = 10, SET_COUNT = 11 };
constexpr int kSet = kQuadSlots * kQuadSpace;
constexpr int kStateDoubles = S_COUNT + SET_COUNT * kSet;
static_assert(kStateDoubles <= YSP_PCG3_SCALARS, "cg_scalars too small");

__host__ __device__ constexpr int set_base(int set) { return S_COUNT + set * kSet; }

// This workgroup's slot of a set.
__device__ __forceinline__ double* my_slot(double* st, int set) {
  return st + set_base(set) + (blockIdx.x % kQuadSlots) * kQuadSpace;
}

// Sums of up to four sets, one wavefront each (blockDim >= 64 * count), identical in
// every thread of the workgroup.
template <int C>
__device__ __forceinline__ void set_sums(const double* st, const int (&sets)[C], double* lds,
                                         double (&out)[C]) {
  const int w = threadIdx.x / 64;
  if (w < C) {
    double v = st[set_base(sets[w]) + (threadIdx.x % 64) * kQuadSpace];
    v = wave_sum(v);
    if (threadIdx.x % 64 == 0) lds[w] = v;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < C; ++i) out[i] = lds[i];
  __syncthreads();
}

constexpr int kMaxAttr = 8;

struct DiagLayout {
  int num_attr;
  uint32_t cum[kMaxAttr + 1];     // prefix of block counts
  uint32_t grad_start[kMaxAttr];  // first DoF
  uint32_t diag_start[kMaxAttr];  // first value of the inverse blocks
  uint32_t size[kMaxAttr];
};

__device__ __forceinline__ bool locate(const DiagLayout& L, uint32_t t, uint32_t& dof0,
                                       uint32_t& inv0, uint32_t& s) {
  if (t >= L.cum[L.num_attr]) return false;
  int a = 0;
  while (t >= L.cum[a + 1]) ++a;
  const uint32_t i = t - L.cum[a];
  s = L.size[a];
  dof0 = L.grad_start[a] + i * s;
  inv0 = L.diag_start[a] + i * s * s;
  return true;
}

constexpr int kMaxBlock = 12;

// z = Minv r for one diagonal block; returns (r.z, r.r).
__device__ __forceinline__ void apply_block(const double* __restrict__ minv, const double* r,
                                            double* z, uint32_t s, double& rz, double& rr) {
  for (uint32_t i = 0; i < s; ++i) {
    double acc = 0.0;
    for (uint32_t j = 0; j < s; ++j) acc += minv[i * s + j] * r[j];
    z[i] = acc;
    rz += r[i] * acc;
    rr += r[i] * r[i];
  }
}

constexpr int kVecBlock = 256;

// One diagonal block of the update: x += alpha p, r -= alpha q, q = 0, z = Minv r, and
// its r.z, r.r terms (same arithmetic order as apply_block).
template <int S>
__device__ __forceinline__ void update_block(const double* __restrict__ minv, uint32_t dof0,
                                             double alpha, double* __restrict__ x,
                                             double* __restrict__ r, const double* __restrict__ p,
                                             double* __restrict__ q, double* __restrict__ z,
                                             double& rz, double& rr) {
  double rl[S];
#pragma unroll
  for (int i = 0; i < S; ++i) {
    const uint32_t d = dof0 + i;
    x[d] = x[d] + p[d] * alpha;
    rl[i] = r[d] - q[d] * alpha;
    r[d] = rl[i];
    q[d] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < S; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < S; ++j) acc += minv[i * S + j] * rl[j];
    z[dof0 + i] = acc;
    rz += rl[i] * acc;
    rr += rl[i] * rl[i];
  }
}

__global__ void __launch_bounds__(kVecBlock)
    init_kernel(DiagLayout L, const double* __restrict__ minv, const double* __restrict__ b,
                double* __restrict__ x, const double* __restrict__ x0, double* __restrict__ r,
                double* __restrict__ p, double* __restrict__ z, double* __restrict__ q,
                double* __restrict__ st, int have_guess) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0, bmb = 0.0, bb = 0.0;
  uint32_t dof0, inv0, s;
  if (locate(L, t, dof0, inv0, s)) {
    double rl[kMaxBlock], zl[kMaxBlock];
    for (uint32_t i = 0; i < s; ++i) {
      const double bi = b[dof0 + i];
      // with an initial guess q holds A x0 on entry
      rl[i] = have_guess ? bi - q[dof0 + i] : bi;
      x[dof0 + i] = have_guess ? x0[dof0 + i] : 0.0;
      q[dof0 + i] = 0.0;
    }
    apply_block(minv + inv0, rl, zl, s, rz, rr);
    if (have_guess) {
      double bl[kMaxBlock], tmp[kMaxBlock];
      for (uint32_t i = 0; i < s; ++i) bl[i] = b[dof0 + i];
      apply_block(minv + inv0, bl, tmp, s, bmb, bb);
    } else {
      bmb = rz;
      bb = rr;
    }
    for (uint32_t i = 0; i < s; ++i) {
      r[dof0 + i] = rl[i];
      z[dof0 + i] = zl[i];
      p[dof0 + i] = zl[i];
    }
  }
  __shared__ double red[kVecBlock / 64];
  rz = block_sum<kVecBlock>(rz, red);
  rr = block_sum<kVecBlock>(rr, red);
  bmb = block_sum<kVecBlock>(bmb, red);
  bb = block_sum<kVecBlock>(bb, red);
  if (threadIdx.x == 0) {
    atomic_add_f64(my_slot(st, SET_RZ + 0), rz);
    atomic_add_f64(my_slot(st, SET_RR + 0), rr);
    atomic_add_f64(my_slot(st, SET_RZ0), bmb);
    atomic_add_f64(my_slot(st, SET_BB), bb);
  }
}

// One workgroup of 256 threads.
__global__ void __launch_bounds__(256) init_finish_kernel(double* st, double tol, int mode) {
  __shared__ double lds[4];
  double v[4];
  const int sets[4] = {SET_RZ0, SET_RZ + 0, SET_BB, SET_RR + 0};
  set_sums<4>(st, sets, lds, v);
  if (threadIdx.x != 0) return;
  const double rz0 = v[0], rz = v[1], bb = v[2], rr = v[3];
  st[S_RZ0] = rz0;
  st[S_BB] = bb;
  st[S_MODE] = mode;
  if (!isfinite(rz0) || rz0 < 0.0 || !isfinite(rz) || rz < 0.0) {
    st[S_STATUS] = -2.0;
    st[S_ITER] = 0.0;
    return;
  }
  const double thr = (mode == 0) ? tol * rz0 : tol * tol * bb;
  st[S_THR] = thr;
  st[S_FINAL_RZ] = rz;
  st[S_FINAL_RR] = rr;
  if (!isfinite(thr) || thr < 0.0) {
    st[S_STATUS] = -2.0;
    st[S_ITER] = 0.0;
    return;
  }
  if ((mode == 0 ? rz : rr) <= thr) {
    st[S_STATUS] = 1.0;
    st[S_ITER] = 0.0;
  }
}

__global__ void __launch_bounds__(kVecBlock)
    update_kernel(DiagLayout L, const double* __restrict__ minv, double* __restrict__ x,
                  double* __restrict__ r, const double* __restrict__ p, double* __restrict__ q,
                  double* __restrict__ z, double* __restrict__ st, int k) {
  if (st[S_STATUS] != 0.0) return;
  const int cur = k % 3, prev = (k + 2) % 3, next = (k + 1) % 3;
  __shared__ double lds[2];
  double v[2];
  const int sets[2] = {SET_PAP + cur, SET_RZ + prev};
  set_sums<2>(st, sets, lds, v);
  const double pap = v[0], rzo = v[1];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (!isfinite(pap) || pap <= 0.0) {
    if (t == 0) {
      st[S_STATUS] = -1.0;
      st[S_ITER] = k;
    }
    return;
  }
  const double alpha = rzo / pap;
  if (blockIdx.x == 0 && threadIdx.x < kQuadSlots) {
    st[set_base(SET_PAP + next) + threadIdx.x * kQuadSpace] = 0.0;
    st[set_base(SET_RZ + next) + threadIdx.x * kQuadSpace] = 0.0;
    st[set_base(SET_RR + next) + threadIdx.x * kQuadSpace] = 0.0;
  }
  double rz = 0.0, rr = 0.0;
  uint32_t dof0, inv0, s;
  if (locate(L, t, dof0, inv0, s)) {
    // compile-time block sizes keep the per-block arrays in registers
    switch (s) {
      case 3: update_block<3>(minv + inv0, dof0, alpha, x, r, p, q, z, rz, rr); break;
      case 9: update_block<9>(minv + inv0, dof0, alpha, x, r, p, q, z, rz, rr); break;
      case 12: update_block<12>(minv + inv0, dof0, alpha, x, r, p, q, z, rz, rr); break;
      case 1: update_block<1>(minv + inv0, dof0, alpha, x, r, p, q, z, rz, rr); break;
      case 2: update_block<2>(minv + inv0, dof0, alpha, x, r, p, q, z, rz, rr); break;
      case 4: update_block<4>(minv + inv0, dof0, alpha, x, r, p, q, z, rz, rr); break;
      case 6: update_block<6>(minv + inv0, dof0, alpha, x, r, p, q, z, rz, rr); break;
      default: {
        double rl[kMaxBlock], zl[kMaxBlock];
        for (uint32_t i = 0; i < s; ++i) {
          const uint32_t d = dof0 + i;
          x[d] = x[d] + p[d] * alpha;
          rl[i] = r[d] - q[d] * alpha;
          r[d] = rl[i];
          q[d] = 0.0;
        }
        apply_block(minv + inv0, rl, zl, s, rz, rr);
        for (uint32_t i = 0; i < s; ++i) z[dof0 + i] = zl[i];
      }
    }
  }
  __shared__ double red[kVecBlock / 64];
  rz = block_sum<kVecBlock>(rz, red);
  rr = block_sum<kVecBlock>(rr, red);
  if (threadIdx.x == 0) {
    atomic_add_f64(my_slot(st, SET_RZ + cur), rz);
    atomic_add_f64(my_slot(st, SET_RR + cur), rr);
  }
}

__global__ void __launch_bounds__(kVecBlock)
    direction_kernel(const double* __restrict__ z, double* __restrict__ p, double* __restrict__ st,
                     uint32_t n, int k) {
  if (st[S_STATUS] != 0.0) return;
  const int cur = k % 3, prev = (k + 2) % 3;
  __shared__ double lds[3];
  double v[3];
  const int sets[3] = {SET_RZ + cur, SET_RZ + prev, SET_RR + cur};
  set_sums<3>(st, sets, lds, v);
  const double rzn = v[0], rzo = v[1], rr = v[2];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (!isfinite(rzn) || rzn < 0.0 || !isfinite(rzo) || rzo <= 0.0) {
    if (t == 0) {
      st[S_STATUS] = -1.0;
      st[S_ITER] = k;
    }
    return;
  }
  const double beta = rzn / rzo;
  const bool conv = (st[S_MODE] == 0.0 ? rzn : rr) <= st[S_THR];
  for (uint32_t i = t; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + p[i] * beta;
  if (t == 0) {
    st[S_FINAL_RZ] = rzn;
    st[S_FINAL_RR] = rr;
    if (conv) {
      st[S_STATUS] = 1.0;
      st[S_ITER] = k;
    }
  }
}

// Pseudo-inverse of diagonal blocks (diagonalBlockInverseKernel.pyx:51-104).
template <int N>
__global__ void __launch_bounds__(64)
    invert_blocks_kernel(const double* __restrict__ in, double* __restrict__ out, uint32_t count) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= count) return;
  double A[N][N], V[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) A[i][j] = in[static_cast<size_t>(b) * N * N + i * N + j];
  jacobi_eigen<N>(A, V);
  double inv[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double l = A[i][i];
    inv[i] = fabs(l) < 1e-6 ? fabs(l) : fabs(1.0 / l);
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) acc += V[i][k] * inv[k] * V[j][k];
      out[static_cast<size_t>(b) * N * N + i * N + j] = acc;
    }
}

int launch_invert(const double* in, double* out, uint32_t count, uint32_t size, hipStream_t s) {
  const dim3 g(grid_for(count, 64)), bl(64);
  switch (size) {
    case 1: hipLaunchKernelGGL(invert_blocks_kernel<1>, g, bl, 0, s, in, out, count); break;
    case 2: hipLaunchKernelGGL(invert_blocks_kernel<2>, g, bl, 0, s, in, out, count); break;
    case 3: hipLaunchKernelGGL(invert_blocks_kernel<3>, g, bl, 0, s, in, out, count); break;
    case 4: hipLaunchKernelGGL(invert_blocks_kernel<4>, g, bl, 0, s, in, out, count); break;
    case 6: hipLaunchKernelGGL(invert_blocks_kernel<6>, g, bl, 0, s, in, out, count); break;
    case 9: hipLaunchKernelGGL(invert_blocks_kernel<9>, g, bl, 0, s, in, out, count); break;
    case 12: hipLaunchKernelGGL(invert_blocks_kernel<12>, g, bl, 0, s, in, out, count); break;
    default:
      ysp_set_error(__FILE__, __LINE__, "ysp_invert_diagonal_blocks: unsupported block size");
      return YSP_ERR_UNSUPPORTED;
  }
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

// y = A x over a mirrored full 3x3 block-CSR (component-major values, `stride` entries
// per component): one 16-lane group per block row, every row written; with quad, the
// workgroup's sum of x_r^T y_r goes into its partial-sum slot. `reverse` walks the rows
// from the end: alternating directions from one product to the next, an operator larger
// than the 256 MB Infinity Cache finds the part it streamed last still resident instead
// of evicting it just before use (LRU over a cyclic stream never hits).
constexpr unsigned kFullThreads = 256;
constexpr unsigned kFullLanes = 16;
constexpr unsigned kFullRows = kFullThreads / kFullLanes;
// Rows with more entries than this (by default) go to full3_long_row_kernel (a
// workgroup per row) so that a few long rows (the coupling rows of affine bodies) do not
// set the time.
constexpr uint32_t kLongRow = 128;
template <typename Src>
__global__ void __launch_bounds__(kFullThreads)
    full3_product_kernel(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                         Src vals, uint32_t nbr,
                         uint32_t long_row, int reverse, const double* __restrict__ x,
                         double* __restrict__ y, double* __restrict__ quad,
                         const double* __restrict__ stop) {
  if (stop && *stop != 0.0) return;
  const uint32_t wg = reverse ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
  const uint32_t r = wg * kFullRows + threadIdx.x / kFullLanes;
  const int l = threadIdx.x % kFullLanes;
  double y0 = 0.0, y1 = 0.0, y2 = 0.0, xq = 0.0;
  if (r < nbr && row_ptr[r + 1] - row_ptr[r] <= long_row) {
    const uint32_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
    for (uint32_t k = k0 + l; k < k1; k += kFullLanes) {
      const uint32_t c = 3 * col[k];
      const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
      double b[9];
      vals.load(k, b);
      block_fma(b, x0, x1, x2, y0, y1, y2);
    }
#pragma unroll
    for (int off = kFullLanes / 2; off > 0; off >>= 1) {
      y0 += __shfl_xor(y0, off, kFullLanes);
      y1 += __shfl_xor(y1, off, kFullLanes);
      y2 += __shfl_xor(y2, off, kFullLanes);
    }
    if (l == 0) {
      y[3 * r] = y0;
      y[3 * r + 1] = y1;
      y[3 * r + 2] = y2;
      xq = x[3 * r] * y0 + x[3 * r + 1] * y1 + x[3 * r + 2] * y2;
    }
  }
  if (quad) {
    __shared__ double red[kFullThreads / 64];
    const double sum = block_sum<kFullThreads>(xq, red);
    if (threadIdx.x == 0) atomic_add_f64(quad + (blockIdx.x % kQuadSlots) * kQuadSpace, sum);
  }
}

// The rows longer than kLongRow (listed in long_rows): one workgroup per row.
template <typename Src>
__global__ void __launch_bounds__(kFullThreads)
    full3_long_row_kernel(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                          Src vals,
                          const uint32_t* __restrict__ long_rows, int reverse,
                          const double* __restrict__ x, double* __restrict__ y,
                          double* __restrict__ quad, const double* __restrict__ stop) {
  if (stop && *stop != 0.0) return;
  const uint32_t r = long_rows[reverse ? gridDim.x - 1 - blockIdx.x : blockIdx.x];
  const uint32_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  double y0 = 0.0, y1 = 0.0, y2 = 0.0;
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += kFullThreads) {
    const uint32_t c = 3 * col[k];
    const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
    double b[9];
    vals.load(k, b);
    block_fma(b, x0, x1, x2, y0, y1, y2);
  }
  __shared__ double red[kFullThreads / 64];
  y0 = block_sum<kFullThreads>(y0, red);
  y1 = block_sum<kFullThreads>(y1, red);
  y2 = block_sum<kFullThreads>(y2, red);
  if (threadIdx.x == 0) {
    y[3 * r] = y0;
    y[3 * r + 1] = y1;
    y[3 * r + 2] = y2;
    if (quad)
      atomic_add_f64(quad + (blockIdx.x % kQuadSlots) * kQuadSpace,
                     x[3 * r] * y0 + x[3 * r + 1] * y1 + x[3 * r + 2] * y2);
  }
}

// Pinned status mailbox shared by the solves issued from this process.
double* status_mailbox() {
  static double* box = nullptr;
  if (!box) {
    if (hipHostMalloc(reinterpret_cast<void**>(&box), 8 * sizeof(double), hipHostMallocDefault) !=
        hipSuccess)
      box = nullptr;
  }
  return box;
}

}  // namespace
}  // namespace ysp

using namespace ysp;

namespace ysp {
namespace {
// diagonal[g0 + i * sz + k] = blocks[d0 + i * sz * sz + k * (sz + 1)] for one target.
__global__ void __launch_bounds__(256)
    block_diagonal_kernel(const double* __restrict__ blocks, double* __restrict__ diagonal,
                          uint32_t count, uint32_t size) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * size) return;
  const uint32_t i = t / size, k = t % size;
  diagonal[t] = blocks[static_cast<size_t>(i) * size * size + k * (size + 1)];
}
}  // namespace
}  // namespace ysp

extern "C" int ysp_diagonal_from_blocks(const double* diagonal_blocks, double* diagonal,
                                        const uint32_t* starts, const uint32_t* counts,
                                        const uint32_t* sizes,
                                        const uint32_t* gradient_segments_start,
                                        int32_t num_attributes, void* stream) {
  hipStream_t s = as_stream(stream);
  for (int a = 0; a < num_attributes; ++a) {
    const uint32_t total = counts[a] * sizes[a];
    if (total == 0) continue;
    hipLaunchKernelGGL(ysp::block_diagonal_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s,
                       diagonal_blocks + starts[a], diagonal + gradient_segments_start[a],
                       counts[a], sizes[a]);
  }
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_invert_diagonal_blocks(const double* diagonal_blocks, double* inverse,
                                          const uint32_t* starts, const uint32_t* counts,
                                          const uint32_t* sizes, int32_t num_attributes,
                                          void* stream) {
  hipStream_t s = as_stream(stream);
  for (int a = 0; a < num_attributes; ++a) {
    if (counts[a] == 0) continue;
    int rc = launch_invert(diagonal_blocks + starts[a], inverse + starts[a], counts[a], sizes[a], s);
    if (rc) return rc;
  }
  return YSP_OK;
}

namespace ysp {
namespace {
// The PCG loop of ysp_pcg_solve over any operator: product(x, y, quad, stop) launches
// y = A x (accumulate == 0: y is overwritten for every row; otherwise y is zero on entry
// and receives sums), adding x^T A x into the partial-sum set quad when it is not null.
template <typename Product>
int generic_run(Product&& product, uint32_t max_iterations, double threshold,
                int32_t threshold_mode, const double* diagonal_blocks_inverse,
                const uint32_t* diagonal_blocks_start, const uint32_t* diagonal_blocks_count,
                const uint32_t* diagonal_blocks_size, const uint32_t* gradient_segments_start,
                int32_t num_attributes, const double* rhs, uint32_t n, double* solution,
                const double* initial_guess, int32_t zero_initial_guess, double* work,
                double* cg_scalars, ysp_pcg_stats* stats, hipStream_t s) {
  if (stats) *stats = ysp_pcg_stats{0, 0, 0, 0, 0, 0};
  if (n == 0) return 0;
  if (num_attributes <= 0 || num_attributes > kMaxAttr || !work || !cg_scalars ||
      (threshold_mode != 0 && threshold_mode != 1)) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg_solve: bad arguments");
    return YSP_ERR_ARG;
  }
  DiagLayout L{};
  L.num_attr = num_attributes;
  L.cum[0] = 0;
  uint32_t total_dof = 0;
  for (int a = 0; a < num_attributes; ++a) {
    if (diagonal_blocks_size[a] > static_cast<uint32_t>(kMaxBlock)) {
      ysp_set_error(__FILE__, __LINE__, "ysp_pcg_solve: diagonal block larger than 12");
      return YSP_ERR_UNSUPPORTED;
    }
    L.cum[a + 1] = L.cum[a] + diagonal_blocks_count[a];
    L.grad_start[a] = gradient_segments_start[a];
    L.diag_start[a] = diagonal_blocks_start[a];
    L.size[a] = diagonal_blocks_size[a];
    total_dof += diagonal_blocks_count[a] * diagonal_blocks_size[a];
  }
  if (total_dof != n) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg_solve: diagonal layout does not cover n");
    return YSP_ERR_ARG;
  }
  double* r = work;
  double* p = work + n;
  double* z = work + 2 * static_cast<size_t>(n);
  double* q = work + 3 * static_cast<size_t>(n);
  double* x = solution;
  double* st = cg_scalars;
  const uint32_t nblocks = L.cum[num_attributes];
  const dim3 vgrid(grid_for(nblocks, kVecBlock)), vblock(kVecBlock);
  const dim3 dgrid(std::min<unsigned>(grid_for(n, kVecBlock), 2048u));

  YSP_CHECK(hipMemsetAsync(st, 0, kStateDoubles * sizeof(double), s));
  const int have_guess = (!zero_initial_guess && initial_guess) ? 1 : 0;
  if (have_guess) {
    YSP_CHECK(hipMemsetAsync(q, 0, sizeof(double) * n, s));
    if (int rc = product(initial_guess, q, nullptr, nullptr)) return rc;
  }
  hipLaunchKernelGGL(init_kernel, vgrid, vblock, 0, s, L, diagonal_blocks_inverse, rhs, x,
                     initial_guess, r, p, z, q, st, have_guess);
  hipLaunchKernelGGL(init_finish_kernel, dim3(1), dim3(256), 0, s, st, threshold, threshold_mode);
  YSP_CHECK_LAUNCH();

  double* box = status_mailbox();
  if (!box) {
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg_solve: pinned status allocation failed");
    return YSP_ERR_HIP;
  }
  auto read_status = [&](double* dst) -> int {
    YSP_CHECK(hipMemcpyAsync(dst, st, 8 * sizeof(double), hipMemcpyDeviceToHost, s));
    YSP_CHECK(hipStreamSynchronize(s));
    return YSP_OK;
  };
  double h[8];
  if (int rc = read_status(box)) return rc;
  for (int i = 0; i < 8; ++i) h[i] = box[i];

  uint32_t k = 1;
  uint32_t chunk = 4;
  while (h[S_STATUS] == 0.0 && k <= max_iterations) {
    const uint32_t last = std::min<uint32_t>(max_iterations, k + chunk - 1);
    for (uint32_t it = k; it <= last; ++it) {
      const int cur = static_cast<int>(it % 3);
      if (int rc = product(p, q, st + set_base(SET_PAP + cur), st + S_STATUS)) return rc;
      hipLaunchKernelGGL(update_kernel, vgrid, vblock, 0, s, L, diagonal_blocks_inverse, x, r, p,
                         q, z, st, static_cast<int>(it));
      hipLaunchKernelGGL(direction_kernel, dgrid, vblock, 0, s, z, p, st, n,
                         static_cast<int>(it));
    }
    YSP_CHECK_LAUNCH();
    if (int rc = read_status(box)) return rc;
    for (int i = 0; i < 8; ++i) h[i] = box[i];
    k = last + 1;
    chunk = std::min<uint32_t>(chunk * 2, 64);
  }
  double fin[64];
  YSP_CHECK(hipMemcpyAsync(fin, st, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  int code;
  const int iters = static_cast<int>(fin[S_ITER]);
  if (fin[S_STATUS] == 1.0) {
    code = iters;
  } else if (fin[S_STATUS] == -1.0) {
    code = -iters - 4;
  } else if (fin[S_STATUS] == -2.0) {
    code = -5;
  } else {
    code = -4;
  }
  if (stats) {
    stats->iterations = fin[S_STATUS] == 0.0 ? static_cast<int>(max_iterations) : iters;
    stats->status = code;
    stats->delta0 = fin[S_RZ0];
    stats->delta = fin[S_FINAL_RZ];
    stats->rr0 = fin[S_BB];
    stats->rr = fin[S_FINAL_RR];
  }
  return code;
}
}  // namespace
}  // namespace ysp

extern "C" int ysp_pcg_solve(uint32_t max_iterations, double threshold, int32_t threshold_mode,
                             const ysp_block_matrix* part, const ysp_block_matrix* dynamic_part,
                             const double* diagonal_blocks_inverse,
                             const uint32_t* diagonal_blocks_start,
                             const uint32_t* diagonal_blocks_count,
                             const uint32_t* diagonal_blocks_size,
                             const uint32_t* gradient_segments_start, int32_t num_attributes,
                             const double* rhs, uint32_t n, double* solution,
                             const double* initial_guess, int32_t zero_initial_guess,
                             double* work, double* cg_scalars, ysp_pcg_stats* stats,
                             void* stream) {
  hipStream_t s = as_stream(stream);
  auto product = [&](const double* x, double* y, double* quad, const double* stop) -> int {
    if (int rc = launch_part_spmv(part, x, y, 0, quad, stop, s)) return rc;
    return launch_part_spmv(dynamic_part, x, y, 0, quad, stop, s);
  };
  return generic_run(product, max_iterations, threshold, threshold_mode, diagonal_blocks_inverse,
                     diagonal_blocks_start, diagonal_blocks_count, diagonal_blocks_size,
                     gradient_segments_start, num_attributes, rhs, n, solution, initial_guess,
                     zero_initial_guess, work, cg_scalars, stats, s);
}

extern "C" int ysp_pcg_solve_full(uint32_t max_iterations, double threshold,
                                  int32_t threshold_mode, const uint32_t* row_ptr,
                                  const uint32_t* col, const void* values, int32_t values_f32,
                                  uint32_t nnz,
                                  uint32_t long_row_threshold,
                                  const double* diagonal_blocks_inverse,
                                  const uint32_t* diagonal_blocks_start,
                                  const uint32_t* diagonal_blocks_count,
                                  const uint32_t* diagonal_blocks_size,
                                  const uint32_t* gradient_segments_start,
                                  int32_t num_attributes, const double* rhs, uint32_t n,
                                  double* solution, const double* initial_guess,
                                  int32_t zero_initial_guess, double* work, double* cg_scalars,
                                  ysp_pcg_stats* stats, void* stream) {
  if (n % 3 || (nnz && (!row_ptr || !col || !values)) ||
      (values_f32 != YSP_VALUES_F64 && values_f32 != YSP_VALUES_F32 &&
       values_f32 != YSP_VALUES_UPPER)) {
    if (stats) *stats = ysp_pcg_stats{0, 0, 0, 0, 0, 0};
    ysp_set_error(__FILE__, __LINE__, "ysp_pcg_solve_full: bad arguments");
    return YSP_ERR_ARG;
  }
  hipStream_t s = as_stream(stream);
  const uint32_t nbr = n / 3;
  // rows longer than the threshold, listed once per solve from the row starts
  const uint32_t long_row = long_row_threshold ? long_row_threshold : kLongRow;
  uint32_t num_long_rows = 0;
  uint32_t* long_rows = nullptr;
  if (nbr && nnz) {
    std::vector<uint32_t> rp(static_cast<size_t>(nbr) + 1);
    YSP_CHECK(hipMemcpyAsync(rp.data(), row_ptr, sizeof(uint32_t) * rp.size(),
                             hipMemcpyDeviceToHost, s));
    YSP_CHECK(hipStreamSynchronize(s));
    std::vector<uint32_t> longs;
    for (uint32_t r = 0; r < nbr; ++r)
      if (rp[r + 1] - rp[r] > long_row) longs.push_back(r);
    if (!longs.empty()) {
      long_rows = static_cast<uint32_t*>(scratch(66, sizeof(uint32_t) * longs.size()));
      if (!long_rows) return YSP_ERR_HIP;
      YSP_CHECK(hipMemcpyAsync(long_rows, longs.data(), sizeof(uint32_t) * longs.size(),
                               hipMemcpyHostToDevice, s));
      YSP_CHECK(hipStreamSynchronize(s));
      num_long_rows = static_cast<uint32_t>(longs.size());
    }
  }
  int reverse = 0;                 // direction of the next product's row walk
  // value source (values.hpp): FP64 or FP32 copy (FP32 values, the mixed-precision inner
  // solve, are widened to FP64 as they are read) or the upper storage read in place
  auto product = [&](const double* x, double* y, double* quad, const double* stop) -> int {
    if (nbr == 0) return YSP_OK;
    if (int rc = with_values(values, values_f32, nnz, "ysp_pcg_solve_full: bad value source",
                             [&](auto vals) {
                               using Src = decltype(vals);
                               hipLaunchKernelGGL(full3_product_kernel<Src>,
                                                  dim3(grid_for(nbr, kFullRows)),
                                                  dim3(kFullThreads), 0, s, row_ptr, col, vals,
                                                  nbr, long_row, reverse, x, y, quad, stop);
                               if (num_long_rows)
                                 hipLaunchKernelGGL(full3_long_row_kernel<Src>,
                                                    dim3(num_long_rows), dim3(kFullThreads), 0, s,
                                                    row_ptr, col, vals, long_rows, reverse, x, y,
                                                    quad, stop);
                               return YSP_OK;
                             }))
      return rc;
    reverse ^= 1;
    YSP_CHECK_LAUNCH();
    return YSP_OK;
  };
  return generic_run(product, max_iterations, threshold, threshold_mode, diagonal_blocks_inverse,
                     diagonal_blocks_start, diagonal_blocks_count, diagonal_blocks_size,
                     gradient_segments_start, num_attributes, rhs, n, solution, initial_guess,
                     zero_initial_guess, work, cg_scalars, stats, s);
}
