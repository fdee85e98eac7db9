// Element kernel machinery shared by the library's energies (energies.hip) and run-time
// generated user energies (yasps_amd/usercode.py): gather of an element's vertex data
// through the union view (vertices driven by affine bodies included), energy / gradient /
// packed Hessian from Kind<ID>::eval, PSD projection with the reference's rules, and the
// scatter into block storage for segments of any size (hessianKernelNoProject.pyx:37-264
// semantics). A translation unit specialises Kind<ID> for its energies (A vertices of W
// values, N = A W, P element parameters, S scene parameters, and
//   __device__ static void eval(x, p, s, E, g, H)
// with H the packed upper triangle) and instantiates eval_energy<ID> / energy_values<ID>.
//
// Vertices of an affine union child follow x = t + A r (JOIN of the body's A (9, row-major)
// and t (3) through a vertex -> body connectivity, one_bunny_partial_abd.py:127-144).
// The map is linear, so the projection acts on the Hessian in vertex positions and the
// chain rule maps it to the body's segments (differentiator.pyx:795-801): with slot
// entries s -> (position entry i(s), coefficient c(s)), g_s = c(s) g_i(s) and
// H_st = c(s) c(t) H_i(s)i(t): t entries map to x_j with c = 1, A entries (j, l) to x_j
// with c = r_l.
#pragma once
#include "common.hpp"
#include "dense.hpp"
#include "element_common.hpp"

#include <algorithm>
#include <type_traits>

namespace ysp {
namespace {

constexpr unsigned kEB = 64;
constexpr int kSlotsMax = 16;

template <int ID>
struct Kind;


// Quadratic form of a linear map of the element's vertex data, y = J x + c with J 6 x 12:
// E = y^T Q y / 2 + l^T y, g = J^T (Q y + l), H = J^T spd(Q) J (the projection acts on the
// inner Hessian: the map is linear). Scene parameters are read from memory in place
// (element_eval), not copied per element.
constexpr int kQfRows = 6;
template <int N>
__device__ __forceinline__ void project_packed(double (&H)[N * (N + 1) / 2], uint32_t method);

template <>
struct Kind<YSP_ENERGY_QUADRATIC_FORM> {
  static constexpr int A = 4, W = 3, N = 12, P = 0, S = kQfRows * 12 + kQfRows * kQfRows + 2 * kQfRows;
};

__device__ __forceinline__ void quadratic_form_eval(const double (&x)[12], const double* __restrict__ sp,
                                                    uint32_t method, double& E, double (&g)[12],
                                                    double (&H)[78]) {
  const double* J = sp;
  const double* Q = sp + kQfRows * 12;
  const double* lin = Q + kQfRows * kQfRows;
  const double* off = lin + kQfRows;
  double y[kQfRows], qy[kQfRows];
#pragma unroll
  for (int i = 0; i < kQfRows; ++i) {
    double acc = off[i];
#pragma unroll
    for (int j = 0; j < 12; ++j) acc += J[12 * i + j] * x[j];
    y[i] = acc;
  }
  E = 0.0;
#pragma unroll
  for (int i = 0; i < kQfRows; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kQfRows; ++k) acc += Q[kQfRows * i + k] * y[k];
    qy[i] = acc + lin[i];
    E += y[i] * (0.5 * acc + lin[i]);
  }
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < kQfRows; ++i) acc += J[12 * i + j] * qy[i];
    g[j] = acc;
  }
  double Qp[kQfRows * (kQfRows + 1) / 2];
#pragma unroll
  for (int i = 0; i < kQfRows; ++i)
#pragma unroll
    for (int k = i; k < kQfRows; ++k) Qp[upper_index<kQfRows>(i, k)] = Q[kQfRows * i + k];
  if (method != 0) project_packed<kQfRows>(Qp, method);
  double JQ[kQfRows][12];                       // spd(Q) J
#pragma unroll
  for (int i = 0; i < kQfRows; ++i)
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < kQfRows; ++k)
        acc += Qp[i <= k ? upper_index<kQfRows>(i, k) : upper_index<kQfRows>(k, i)] * J[12 * k + j];
      JQ[i][j] = acc;
    }
#pragma unroll
  for (int a = 0; a < 12; ++a)
#pragma unroll
    for (int b = a; b < 12; ++b) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < kQfRows; ++i) acc += J[12 * i + a] * JQ[i][b];
      H[upper_index<12>(a, b)] = acc;
    }
}

struct AffineView {
  uint32_t mask;                                  // bit c: union child c is affine
  const uint32_t* body[YSP_MAX_UNION_CHILDREN];   // vertex -> body
  const double* rest[YSP_MAX_UNION_CHILDREN];     // rest position r per vertex
};

// Union child of instance v and its index inside the child.
__device__ __forceinline__ uint32_t union_child(const UnionView& U, uint32_t v, uint32_t& local) {
  uint32_t total = 0;
  for (uint32_t c = 0; c < U.nc; ++c) {
    const uint32_t cnt = U.counts[c];
    if (v < total + cnt) {
      local = v - total;
      return c;
    }
    total += cnt;
  }
  local = 0;
  return 0;
}

// ----------------------------------------------------------------- projection
// hessianKernelHeader.pyx:79-252 for N >= 4: the packed-Schur PSD test on the element
// Hessian; a failing matrix is replaced by V max(L, 0) V^T (method 2) or V |L| V^T
// (method 1) from a cyclic Jacobi eigen-decomposition.
template <int N>
__device__ __forceinline__ void project_packed(double (&H)[N * (N + 1) / 2], uint32_t method) {
  double L[N * (N + 1) / 2];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c <= r; ++c) L[r * (r + 1) / 2 + c] = H[upper_index<N>(c, r)];
  if (psd_test_packed<N>(L)) return;
  double V[N][N];
  jacobi_eigen_packed<N>(H, V);
  double lam[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double l = H[upper_index<N>(i, i)];
    lam[i] = (l < 0.0) ? (method == 1 ? -l : 0.0) : l;
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i; j < N; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) acc += V[i][k] * lam[k] * V[j][k];
      H[upper_index<N>(i, j)] = acc;
    }
}

// 4-vertex kinds whose energy depends on vertex differences only (barriers, dihedral
// bending, SNH): their 12x12 Hessians annihilate the three translations, so a failing
// Hessian is projected in the 9-dimensional complement, H = (U x I3) M (U x I3)^T with U
// the orthonormal 4x3 basis orthogonal to (1,1,1,1): the same result as the 12x12
// decomposition up to the rounding-level translation eigenvalues, on a 9x9 Jacobi
// (0.42 of the rotations' work and 81 instead of 144 eigenvector registers). The PSD
// test stays on the full 12x12 matrix, as the reference runs it.
template <int ID>
__host__ __device__ constexpr bool translation_free() {
  return ID == YSP_ENERGY_POINT_TRIANGLE || ID == YSP_ENERGY_EDGE_EDGE ||
         ID == YSP_ENERGY_BENDING || ID == YSP_ENERGY_SNH;
}

__device__ __forceinline__ double basis4(int a, int k) {
  // columns of U: (1,-1,0,0)/sqrt2, (1,1,-2,0)/sqrt6, (1,1,1,-3)/sqrt12
  constexpr double c0 = 0.70710678118654752440, c1 = 0.40824829046386301637,
                   c2 = 0.28867513459481288225;
  return k == 0 ? (a == 0 ? c0 : (a == 1 ? -c0 : 0.0))
       : k == 1 ? (a <= 1 ? c1 : (a == 2 ? -2.0 * c1 : 0.0))
                : (a <= 2 ? c2 : -3.0 * c2);
}

__device__ __forceinline__ void project_translation_free(double (&H)[78], uint32_t method) {
  double L[78];
#pragma unroll
  for (int r = 0; r < 12; ++r)
#pragma unroll
    for (int c = 0; c <= r; ++c) L[r * (r + 1) / 2 + c] = H[upper_index<12>(c, r)];
  if (psd_test_packed<12>(L)) return;
  auto h12 = [&](int r, int c) { return r <= c ? H[upper_index<12>(r, c)] : H[upper_index<12>(c, r)]; };
  double M[45];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int l = k; l < 3; ++l)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          if (k == l && j < i) continue;
          double acc = 0.0;
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
              acc += basis4(a, k) * basis4(b, l) * h12(3 * a + i, 3 * b + j);
          M[upper_index<9>(3 * k + i, 3 * l + j)] = acc;
        }
  double V[9][9];
  jacobi_eigen_packed<9>(M, V);
  double lam[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const double l = M[upper_index<9>(i, i)];
    lam[i] = (l < 0.0) ? (method == 1 ? -l : 0.0) : l;
  }
  double P[45];
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = i; j < 9; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) acc += V[i][k] * lam[k] * V[j][k];
      P[upper_index<9>(i, j)] = acc;
    }
  auto p9 = [&](int r, int c) { return r <= c ? P[upper_index<9>(r, c)] : P[upper_index<9>(c, r)]; };
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = a; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          if (a == b && j < i) continue;
          double acc = 0.0;
#pragma unroll
          for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int l = 0; l < 3; ++l)
              acc += basis4(a, k) * basis4(b, l) * p9(3 * k + i, 3 * l + j);
          H[upper_index<12>(3 * a + i, 3 * b + j)] = acc;
        }
}

template <int ID>
__device__ __forceinline__ void project_element(double (&H)[Kind<ID>::N * (Kind<ID>::N + 1) / 2],
                                                uint32_t method) {
  if constexpr (translation_free<ID>() && Kind<ID>::A == 4 && Kind<ID>::W == 3)
    project_translation_free(H, method);
  else
    project_packed<Kind<ID>::N>(H, method);
}

// ----------------------------------------------------------------- slots (any size)
// Slot groups by |perm| in first-occurrence order (hessianKernelNoProject.pyx:93-163),
// with every slot's size and local column offset (sum of the previous slot sizes).
template <int KT>
struct GSlots {
  uint32_t idx[KT];
  uint8_t size[KT];
  uint8_t col[KT];
  uint8_t xbase[KT];   // first position entry of the slot's vertex
  uint8_t amode[KT];   // 1: affine-matrix slot of an affine vertex (x_j, coefficient r_l)
  uint8_t aff[KT];     // 1: a slot of an affine vertex (its A or its t)
  int8_t group_first[KT];
  uint32_t group_mask[KT];
  int ng;
  bool ok;
};

template <int KT>
__device__ __forceinline__ void load_gslots(const ScatterArgs& S, uint32_t e, int N,
                                            GSlots<KT>& si) {
  const uint32_t K = S.K;
  si.ng = 0;
  si.ok = K <= static_cast<uint32_t>(KT);
  if (!si.ok) return;
  int placement[KT];
  uint32_t col = 0;
#pragma unroll
  for (int i = 0; i < KT; ++i) {
    si.idx[i] = 0;
    si.size[i] = 0;
    si.col[i] = 0;
    si.xbase[i] = static_cast<uint8_t>(col);
    si.amode[i] = 0;
    si.aff[i] = 0;
    if (static_cast<uint32_t>(i) >= K) continue;
    const size_t o = static_cast<size_t>(e) * K + i;
    const uint32_t ii = S.seg_idx[o];
    const uint32_t sz = S.seg_size[o];
    const int pm = S.perm[o];
    si.idx[i] = ii;
    si.size[i] = static_cast<uint8_t>(sz);
    si.col[i] = static_cast<uint8_t>(col);
    col += sz;
    if (pm == 0) continue;
    const int p = (pm < 0 ? -pm : pm) - 1;
    int g = -1;
#pragma unroll
    for (int k = 0; k < KT; ++k)
      if (k < si.ng && placement[k] == p) g = k;
    if (g < 0) {
      g = si.ng++;
#pragma unroll
      for (int k = 0; k < KT; ++k)
        if (k == g) {
          placement[k] = p;
          si.group_first[k] = static_cast<int8_t>(i);
          si.group_mask[k] = 0;
        }
    }
#pragma unroll
    for (int k = 0; k < KT; ++k)
      if (k == g) si.group_mask[k] |= 1u << i;
  }
  if (N >= 0 && col != static_cast<uint32_t>(N)) si.ok = false;   // slots tile the DoFs
}

template <int N>
__device__ __forceinline__ double hget(const double (&H)[N * (N + 1) / 2], int i, int j) {
  return i <= j ? H[upper_index<N>(i, j)] : H[upper_index<N>(j, i)];
}

// Position entry and coefficient of entry k of slot i (identity unless the slot is the
// affine matrix of an affine vertex).
template <int KT>
__device__ __forceinline__ int xmap(const GSlots<KT>& si, const double (*r)[3], int i, int k,
                                    double& c) {
  if (si.amode[i]) {
    c = r[si.xbase[i] / 3][k % 3];
    return si.xbase[i] + k / 3;
  }
  c = 1.0;
  return si.xbase[i] + k;
}

// Per-workgroup sums of the contended destinations of the general kernel (atomic mode):
// every element of an affine body adds into the same (A, A), (A, t), (t, t) blocks and the
// same 12 gradient entries of that body, so a device-scope FP64 atomic per element and
// entry serialises on a few addresses. Blocks coupling two affine groups and gradient
// segments of affine groups are summed in LDS instead (LDS atomics, a few slots keyed by
// the destination's offset) and leave with one global atomic per entry and workgroup; a
// destination that finds no free slot takes the global atomic directly.
constexpr uint32_t kHotEmpty = 0xffffffffu;
constexpr int kHotBlocks = 8, kHotBlockMax = 81, kHotGrad = 4, kHotGradMax = 9;
struct HotTable {
  uint32_t bkey[kHotBlocks];    // value offset of the block in S.blocks
  uint32_t bsize[kHotBlocks];   // h * w
  uint32_t bdiag[kHotBlocks];   // offset in S.diag_blocks, kHotEmpty without one
  uint32_t gkey[kHotGrad];      // gradient offset
  uint32_t gsize[kHotGrad];
  double bval[kHotBlocks][kHotBlockMax];
  double gval[kHotGrad][kHotGradMax];
};

__device__ __forceinline__ int hot_slot(uint32_t* keys, int cap, uint32_t key) {
  for (int j = 0; j < cap; ++j) {
    const uint32_t cur = __hip_atomic_load(keys + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == key) return j;
    if (cur == kHotEmpty) {
      const uint32_t old = atomicCAS(keys + j, kHotEmpty, key);
      if (old == kHotEmpty || old == key) return j;
    }
  }
  return -1;
}

__device__ __forceinline__ void hot_init(HotTable& t) {
  for (uint32_t i = threadIdx.x; i < kHotBlocks * kHotBlockMax; i += blockDim.x)
    t.bval[i / kHotBlockMax][i % kHotBlockMax] = 0.0;
  for (uint32_t i = threadIdx.x; i < kHotGrad * kHotGradMax; i += blockDim.x)
    t.gval[i / kHotGradMax][i % kHotGradMax] = 0.0;
  if (threadIdx.x < kHotBlocks) t.bkey[threadIdx.x] = kHotEmpty;
  if (threadIdx.x < kHotGrad) t.gkey[threadIdx.x] = kHotEmpty;
  __syncthreads();
}

__device__ __forceinline__ void hot_flush(HotTable& t, const ScatterArgs& S) {
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kHotBlocks * kHotBlockMax; i += blockDim.x) {
    const uint32_t j = i / kHotBlockMax, o = i % kHotBlockMax;
    if (t.bkey[j] == kHotEmpty || o >= t.bsize[j]) continue;
    const double v = t.bval[j][o];
    atomic_add_f64(S.blocks + t.bkey[j] + o, v);
    if (t.bdiag[j] != kHotEmpty) atomic_add_f64(S.diag_blocks + t.bdiag[j] + o, v);
  }
  for (uint32_t i = threadIdx.x; i < kHotGrad * kHotGradMax; i += blockDim.x) {
    const uint32_t j = i / kHotGradMax, o = i % kHotGradMax;
    if (t.gkey[j] == kHotEmpty || o >= t.gsize[j]) continue;
    atomic_add_f64(S.gradient + t.gkey[j] + o, t.gval[j][o]);
  }
}

template <int KT, int N>
__device__ __forceinline__ void scatter_generic(const ScatterArgs& S, uint32_t e,
                                                const GSlots<KT>& si, const double (*r)[3],
                                                const double (&g)[N],
                                                const double (&H)[N * (N + 1) / 2],
                                                HotTable* ht) {
  // gradient
  if (S.grad_staging) {
    double* out = S.grad_staging + static_cast<size_t>(e) * S.grad_local;
#pragma unroll
    for (int i = 0; i < KT; ++i)
      for (int k = 0; k < si.size[i]; ++k) {
        double c;
        const int x = xmap(si, r, i, k, c);
        out[si.col[i] + k] = c * g[x];
      }
  } else {
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      if (si.idx[i] < 2) continue;
      const uint32_t p = si.idx[i] - 2;
      int gs = -1;
      if (ht && si.aff[i] && si.size[i] <= kHotGradMax) {
        gs = hot_slot(ht->gkey, kHotGrad, p);
        if (gs >= 0) ht->gsize[gs] = si.size[i];
      }
      for (int k = 0; k < si.size[i]; ++k) {
        double c;
        const int x = xmap(si, r, i, k, c);
        if (gs >= 0) atomicAdd(&ht->gval[gs][k], c * g[x]);
        else atomic_add_f64(S.gradient + p + k, c * g[x]);
      }
    }
  }
  // Hessian blocks of every valid group pair, in coordinate order
  uint32_t count = 0;
  const uint32_t cstart = S.coord_outer[e];
#pragma unroll
  for (int gi = 0; gi < KT; ++gi) {
    if (gi >= si.ng) continue;
    const int fi = si.group_first[gi];
    const uint32_t ii = si.idx[fi];
    if (ii < 2) continue;
    const int h = si.size[fi];
#pragma unroll
    for (int gj = gi; gj < KT; ++gj) {
      if (gj >= si.ng) continue;
      const int fj = si.group_first[gj];
      const uint32_t jj = si.idx[fj];
      if (jj < 2) continue;
      const int w = si.size[fj];
      const uint32_t coord = cstart + count;
      double* dst = S.block_staging ? S.block_staging + 9 * static_cast<size_t>(coord)
                                    : S.blocks + S.lookups[coord];
      double* dd = nullptr;
      if (!S.block_staging && gi == gj && S.diag_blocks) dd = S.diag_blocks + diag_offset(S, ii - 2, h);
      int hs = -1;
      if (ht && !S.block_staging && si.aff[fi] && si.aff[fj] && h * w <= kHotBlockMax) {
        hs = hot_slot(ht->bkey, kHotBlocks, S.lookups[coord]);
        if (hs >= 0) {
          ht->bsize[hs] = h * w;
          ht->bdiag[hs] = dd ? static_cast<uint32_t>(dd - S.diag_blocks) : kHotEmpty;
        }
      }
      const bool tr = !(ii < jj);
      for (int k = 0; k < h; ++k)
        for (int l = 0; l < w; ++l) {
          double v = 0.0;
#pragma unroll
          for (int a = 0; a < KT; ++a) {
            if (!((si.group_mask[gi] >> a) & 1u) || si.size[a] == 0) continue;
            double ca;
            const int xa = xmap(si, r, a, k, ca);
#pragma unroll
            for (int b = 0; b < KT; ++b) {
              if (!((si.group_mask[gj] >> b) & 1u) || si.size[b] == 0) continue;
              double cb;
              const int xb = xmap(si, r, b, l, cb);
              v += ca * cb * hget<N>(H, xa, xb);
            }
          }
          // stored block: rows of the lower global index; (l, k) of a transposed block
          const int o = tr ? l * h + k : k * w + l;
          if (S.block_staging) {
            dst[o] = v;
          } else if (hs >= 0) {
            atomicAdd(&ht->bval[hs][o], v);
          } else {
            atomic_add_f64(dst + o, v);
            if (dd) atomic_add_f64(dd + o, v);
          }
        }
      ++count;
    }
  }
}

template <int ID>
__device__ __forceinline__ void gather(const UnionView& P, const uint32_t* e2v, uint32_t e,
                                       double (&x)[Kind<ID>::N]) {
  using T = Kind<ID>;
#pragma unroll
  for (int a = 0; a < T::A; ++a) {
    const uint32_t v = e2v[static_cast<size_t>(e) * T::A + a];
    const double* row = union_row(P, v);
#pragma unroll
    for (int i = 0; i < T::W; ++i) x[T::W * a + i] = row[i];
  }
}

// Scene parameters of one element; Baraff-Witkin picks its stretch coefficients from
// the current F (helpers.py:252-253: 0.01 below unit stretch, else 1).
template <int ID>
__device__ __forceinline__ void scene_params(const double* sp, const double* x, const double* p,
                                             double (&s)[Kind<ID>::S]) {
#pragma unroll
  for (int i = 0; i < Kind<ID>::S; ++i) s[i] = sp[i];
  if constexpr (ID == YSP_ENERGY_BARAFF_WITKIN) bw_stretch_coefficients(x, p, s[3], s[4]);
}

// Energy, gradient and (projected) packed Hessian of one element.
template <int ID>
__device__ __forceinline__ void element_eval(const double (&x)[Kind<ID>::N], const double* p,
                                             const double* sp, uint32_t method, double& E,
                                             double (&g)[Kind<ID>::N],
                                             double (&H)[Kind<ID>::N * (Kind<ID>::N + 1) / 2]) {
  if constexpr (ID == YSP_ENERGY_QUADRATIC_FORM) {
    quadratic_form_eval(x, sp, method, E, g, H);
  } else {
    double s[Kind<ID>::S];
    scene_params<ID>(sp, x, p, s);
    Kind<ID>::eval(x, p, s, E, g, H);
    if (method != 0) project_element<ID>(H, method);
  }
}

// One element through the general slot bookkeeping (any slot sizes, affine vertices).
template <int ID, int KT>
__device__ __forceinline__ void eval_element_general(const UnionView& P, const AffineView& AF,
                                                     const uint32_t* __restrict__ e2v, uint32_t e,
                                                     const double* __restrict__ ep,
                                                     const double* __restrict__ sp, uint32_t method,
                                                     const ScatterArgs& SA,
                                                     int* __restrict__ error_flag,
                                                     HotTable* ht) {
  using T = Kind<ID>;
  GSlots<KT> si;
  load_gslots<KT>(SA, e, AF.mask ? -1 : T::N, si);
  double r[T::A][3];
  if (AF.mask) {
    // slots of vertex a: [a Kv, (a + 1) Kv); an affine vertex owns A then t
    const int Kv = static_cast<int>(SA.K) / T::A;
    if (Kv * T::A != static_cast<int>(SA.K) || T::W != 3) si.ok = false;
#pragma unroll
    for (int a = 0; a < T::A; ++a) {
      uint32_t local;
      const uint32_t c = union_child(P, e2v[static_cast<size_t>(e) * T::A + a], local);
      const bool aff = (AF.mask >> c) & 1u;
      r[a][0] = r[a][1] = r[a][2] = 0.0;
      if (aff) {
        const double* rr = AF.rest[c] + 3 * static_cast<size_t>(local);
        r[a][0] = rr[0];
        r[a][1] = rr[1];
        r[a][2] = rr[2];
      }
#pragma unroll
      for (int q = 0; q < KT; ++q) {
        if (q >= Kv) continue;
        const int i = a * Kv + q;
        if (i >= KT) continue;
        si.xbase[i] = static_cast<uint8_t>(3 * a);
        si.amode[i] = aff && q == 0;
        si.aff[i] = aff;
        const int want = aff ? (q == 0 ? 9 : (q == 1 ? 3 : -1)) : (q == 0 ? 3 : -1);
        if (want > 0 && si.size[i] != want) si.ok = false;
        if (want < 0 && si.idx[i] >= 2) si.ok = false;
      }
    }
  }
  if (!si.ok || (SA.block_staging && (T::W != 3 || AF.mask))) {
    atomicOr(error_flag, 1);
    return;
  }
  double x[T::N];
  gather<ID>(P, e2v, e, x);
  const double* p = ep + static_cast<size_t>(e) * T::P;
  double E, g[T::N], H[T::N * (T::N + 1) / 2];
  element_eval<ID>(x, p, sp, method, E, g, H);
  scatter_generic<KT, T::N>(SA, e, si, r, g, H, ht);
}

// General kernel: every element (queue == nullptr) or the elements queued by the vertex
// kernel (count in flags[1]), in a grid-stride loop. With affine children and atomic
// assembly, the affine destinations of the workgroup's elements are summed in LDS first.
constexpr unsigned kGB = 256;
template <int ID, int KT>
__global__ void __launch_bounds__(kGB)
    energy_eval_kernel(UnionView P, AffineView AF, const uint32_t* __restrict__ e2v, uint32_t n,
                       const double* __restrict__ ep, const double* __restrict__ sp,
                       uint32_t method, ScatterArgs SA, int* __restrict__ flags,
                       const uint32_t* __restrict__ queue) {
  __shared__ HotTable table;
  const bool hot = AF.mask && !SA.block_staging;   // uniform over the grid
  if (hot) hot_init(table);
  const uint32_t count = queue ? static_cast<uint32_t>(flags[1]) : n;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x)
    eval_element_general<ID, KT>(P, AF, e2v, queue ? queue[t] : t, ep, sp, method, SA, flags,
                                 hot ? &table : nullptr);
  if (hot) hot_flush(table, SA);
}

// ----------------------------------------------------------------- vertex fast path
// Elements without affine vertices whose slots are one W-wide slot per local vertex (or
// padding): slot bookkeeping and scatter at compile-time vertex indices, so that no
// per-element array is indexed at run time (such an array would be placed in scratch
// memory). An element whose layout is anything else (a target vertex repeated inside
// the element) goes to a queue for the general kernel.
template <int NV>
struct VSlots {
  uint32_t idx[NV];   // seg_idx of the vertex's slot, 0 without one
  bool has[NV];       // the vertex owns a W-wide slot
  bool target[NV];    // ... with a target index (>= 2) in a group
  int rank[NV];       // rank of its group among the element's target groups
  int ntarget;
  bool ok;
};

template <int KT, int NV, int W>
__device__ __forceinline__ void load_vslots(const ScatterArgs& S, uint32_t e, VSlots<NV>& vs) {
  const uint32_t K = S.K;
  vs.ok = K <= static_cast<uint32_t>(KT);
  int grp[NV];
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    vs.idx[a] = 0;
    vs.has[a] = false;
    grp[a] = -1;
  }
  int placement[KT];
  int ng = 0;
  uint32_t col = 0;
#pragma unroll
  for (int i = 0; i < KT; ++i) {
    if (!vs.ok || static_cast<uint32_t>(i) >= K) continue;
    const size_t o = static_cast<size_t>(e) * K + i;
    const uint32_t ii = S.seg_idx[o];
    const uint32_t sz = S.seg_size[o];
    const int pm = S.perm[o];
    // a union's padding slot (index 0) holds no DoFs, whatever its size (the size makes
    // a vertex of a small child as wide as the union's widest child, e.g. next to
    // affine bodies)
    if (ii == 0) continue;
    int g = -1;
    if (pm != 0) {
      const int p = (pm < 0 ? -pm : pm) - 1;
#pragma unroll
      for (int k = 0; k < KT; ++k)
        if (k < ng && placement[k] == p) g = k;
      if (g < 0) {
        g = ng++;
#pragma unroll
        for (int k = 0; k < KT; ++k)
          if (k == g) placement[k] = p;
      }
    }
    if (sz == W) {
      const uint32_t v = col / W;
      if (v >= static_cast<uint32_t>(NV)) vs.ok = false;
#pragma unroll
      for (int a = 0; a < NV; ++a)
        if (v == static_cast<uint32_t>(a)) {
          vs.idx[a] = ii;
          vs.has[a] = true;
          grp[a] = g;
        }
    } else if (sz != 0) {
      vs.ok = false;
    }
    col += sz;
  }
  if (col != static_cast<uint32_t>(NV * W)) vs.ok = false;      // slots tile the DoFs
  vs.ntarget = 0;
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    vs.target[a] = vs.has[a] && vs.idx[a] >= 2 && grp[a] >= 0;
    vs.ntarget += vs.target[a];
  }
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    int r = 0;
#pragma unroll
    for (int b = 0; b < NV; ++b) {
      r += vs.target[b] && grp[b] < grp[a];
      if (b > a && vs.target[a] && vs.target[b] && grp[a] == grp[b]) vs.ok = false;
    }
    vs.rank[a] = r;
  }
}

// Gradient segments and Hessian blocks of a vertex-layout element (the semantics of
// scatter_generic: pair of target-group ranks lo <= hi at coordinate coordStart +
// lo*n - lo(lo-1)/2 + (hi-lo), block oriented with the rows of the smaller target index).
template <int NV, int W, int N>
__device__ __forceinline__ void scatter_vertex(const ScatterArgs& S, uint32_t e,
                                               const VSlots<NV>& vs, const double (&g)[N],
                                               const double (&H)[N * (N + 1) / 2]) {
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    if (S.grad_staging ? !vs.has[a] : (!vs.has[a] || vs.idx[a] < 2)) continue;
    if (S.grad_staging) {
      double* out = S.grad_staging + static_cast<size_t>(e) * S.grad_local + W * a;
#pragma unroll
      for (int k = 0; k < W; ++k) out[k] = g[W * a + k];
    } else {
      const uint32_t p = vs.idx[a] - 2;
#pragma unroll
      for (int k = 0; k < W; ++k) atomic_add_f64(S.gradient + p + k, g[W * a + k]);
    }
  }
  const uint32_t cstart = S.coord_outer[e];
  const int n = vs.ntarget;
#pragma unroll
  for (int a = 0; a < NV; ++a)
#pragma unroll
    for (int b = a; b < NV; ++b) {
      if (!vs.target[a] || !vs.target[b]) continue;
      const int lo = min(vs.rank[a], vs.rank[b]), hi = max(vs.rank[a], vs.rank[b]);
      const uint32_t coord = cstart + static_cast<uint32_t>(lo * n - lo * (lo - 1) / 2 + (hi - lo));
      const bool tr = !(vs.idx[a] < vs.idx[b]);   // includes the diagonal block a == b
      if (S.block_staging) {
        double* dst = S.block_staging + 9 * static_cast<size_t>(coord);
#pragma unroll
        for (int k = 0; k < W; ++k)
#pragma unroll
          for (int l = 0; l < W; ++l)
            dst[tr ? l * W + k : k * W + l] = hget<N>(H, W * a + k, W * b + l);
      } else {
        double* dst = S.blocks + S.lookups[coord];
        double* dd = nullptr;
        if (a == b && S.diag_blocks) dd = S.diag_blocks + diag_offset(S, vs.idx[a] - 2, W);
#pragma unroll
        for (int k = 0; k < W; ++k)
#pragma unroll
          for (int l = 0; l < W; ++l) {
            const double v = hget<N>(H, W * a + k, W * b + l);
            const int o = tr ? l * W + k : k * W + l;
            atomic_add_f64(dst + o, v);
            if (dd) atomic_add_f64(dd + o, v);
          }
      }
    }
}

template <int ID, int KT>
__global__ void __launch_bounds__(kEB)
    energy_eval_vertex_kernel(UnionView P, const uint32_t* __restrict__ e2v, uint32_t n,
                              const double* __restrict__ ep, const double* __restrict__ sp,
                              uint32_t method, ScatterArgs SA, int* __restrict__ flags,
                              uint32_t* __restrict__ queue) {
  using T = Kind<ID>;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  VSlots<T::A> vs;
  load_vslots<KT, T::A, T::W>(SA, e, vs);
  if (!vs.ok) {
    queue[wave_reserve(reinterpret_cast<uint32_t*>(flags) + 1, 1u)] = e;
    return;
  }
  if (SA.block_staging && T::W != 3) {
    atomicOr(flags, 1);
    return;
  }
  double x[T::N];
  gather<ID>(P, e2v, e, x);
  const double* p = ep + static_cast<size_t>(e) * T::P;
  double E, g[T::N], H[T::N * (T::N + 1) / 2];
  element_eval<ID>(x, p, sp, method, E, g, H);
  scatter_vertex<T::A, T::W, T::N>(SA, e, vs, g, H);
}

template <int ID>
__global__ void __launch_bounds__(kEB)
    energy_value_kernel(UnionView P, const uint32_t* __restrict__ e2v, uint32_t n,
                        const double* __restrict__ ep, const double* __restrict__ sp,
                        double* __restrict__ energy) {
  using T = Kind<ID>;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double x[T::N];
  gather<ID>(P, e2v, e, x);
  const double* p = ep + static_cast<size_t>(e) * T::P;
  double E, g[T::N], H[T::N * (T::N + 1) / 2];
  element_eval<ID>(x, p, sp, 0, E, g, H);
  energy[e] = E;
}

template <int ID>
int launch_eval(const UnionView& U, const AffineView& AF, const uint32_t* e2v, uint32_t n,
                const double* ep, const double* sp, uint32_t method, const ScatterArgs& SA,
                int* flags, uint32_t* queue, hipStream_t s) {
  const dim3 g(grid_for(n, kEB)), b(kEB);
  auto run = [&](auto kt) {
    constexpr int KT = decltype(kt)::value;
    // vertex-layout elements in the fast kernel; the others (affine-driven vertices, other
    // slot layouts) are queued for the general kernel: a small grid-stride launch when
    // the queue is usually empty, a full grid when the union has affine children
    hipLaunchKernelGGL((energy_eval_vertex_kernel<ID, KT>), g, b, 0, s, U, e2v, n, ep, sp, method,
                       SA, flags, queue);
    const unsigned gg = grid_for(n, kGB);
    const unsigned qg = AF.mask ? gg : std::min<unsigned>(gg, 64);
    hipLaunchKernelGGL((energy_eval_kernel<ID, KT>), dim3(qg), dim3(kGB), 0, s, U, AF, e2v, n, ep,
                       sp, method, SA, flags, queue);
  };
  if (SA.K <= 8) run(std::integral_constant<int, 8>());
  else run(std::integral_constant<int, kSlotsMax>());
  return YSP_OK;
}


template <int ID>
void launch_values(const UnionView& U, const uint32_t* e2v, uint32_t n, const double* ep,
                   const double* sp, double* energy, hipStream_t s) {
  hipLaunchKernelGGL(energy_value_kernel<ID>, dim3(grid_for(n, kEB)), dim3(kEB), 0, s, U, e2v, n,
                     ep, sp, energy);
}


int check_view(const ysp_union_f64* u, uint32_t width, const char* who) {
  if (!u || u->num_children == 0 || u->num_children > YSP_MAX_UNION_CHILDREN || u->width != width) {
    ysp_set_error(__FILE__, __LINE__, who);
    return YSP_ERR_ARG;
  }
  return YSP_OK;
}

// Body of ysp_eval_element_energy for energy ID (flags: an unsupported slot layout, the
// elements queued for the general kernel; checked at the next ysp_check_deferred).
template <int ID>
int eval_energy(const ysp_union_f64* positions, const ysp_affine_children* affine,
                const uint32_t* elem2v, uint32_t n, const double* element_params,
                const double* scene_params, uint32_t projection_method,
                const ysp_scatter_args* out, void* stream, const char* who) {
  if (!out || projection_method > 2 || out->max_num_indices > static_cast<uint32_t>(kSlotsMax)) {
    ysp_set_error(__FILE__, __LINE__, who);
    return YSP_ERR_ARG;
  }
  if (n == 0) return YSP_OK;
  if (int rc = check_view(positions, Kind<ID>::W, who)) return rc;
  hipStream_t s = as_stream(stream);
  int* flag = static_cast<int*>(scratch(34, 2 * sizeof(int)));
  uint32_t* queue = static_cast<uint32_t*>(scratch(46, sizeof(uint32_t) * n));
  if (!flag || !queue) return YSP_ERR_HIP;
  YSP_CHECK(hipMemsetAsync(flag, 0, 2 * sizeof(int), s));
  const UnionView U = make_view(positions);
  const ScatterArgs SA = make_scatter(out);
  AffineView AF{};
  if (affine) {
    AF.mask = affine->mask & ((1u << positions->num_children) - 1u);
    for (int c = 0; c < YSP_MAX_UNION_CHILDREN; ++c) {
      AF.body[c] = affine->vertex_body[c];
      AF.rest[c] = affine->rest_position[c];
      if (((AF.mask >> c) & 1u) && !AF.rest[c]) {
        ysp_set_error(__FILE__, __LINE__, "element energy: affine child without rest data");
        return YSP_ERR_ARG;
      }
    }
  }
  launch_eval<ID>(U, AF, elem2v, n, element_params, scene_params, projection_method, SA, flag,
                  queue, s);
  YSP_CHECK_LAUNCH();
  return defer_flag_check(flag, s, "element energy: slot layout does not match the element's vertices");
}

// Body of ysp_element_energy_values for energy ID: energy[e] of every element.
template <int ID>
int energy_values(const ysp_union_f64* positions, const uint32_t* elem2v, uint32_t n,
                  const double* element_params, const double* scene_params, double* energy,
                  void* stream, const char* who) {
  if (n == 0) return YSP_OK;
  if (int rc = check_view(positions, Kind<ID>::W, who)) return rc;
  launch_values<ID>(make_view(positions), elem2v, n, element_params, scene_params, energy,
                    as_stream(stream));
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

}  // namespace
}  // namespace ysp
