// Element gradient/Hessian kernels: Stable Neo-Hookean tet and inertia, PSD projection
// and the scatter into upper block storage.
//
// SNH closed form (derived by hand from examples/one_bunny/helpers.py:42-67):
//   F = Ds Dm^-1, I_C = |F|^2, J = det F, cof F = [f1 x f2, f2 x f0, f0 x f1] (columns),
//   a1 = mu (1 - 1/(I_C+1)), a2 = 2 mu/(I_C+1)^2, a3 = lam (J - alpha), alpha = 1 + 3mu/(4lam),
//   shape vectors g_a (a = 0..3): g_0 = -sum_k Dm^-1[k,:], g_{k+1} = Dm^-1[k,:],
//   u_a = F g_a, w_a = cof(F) g_a, s = dt^2 vol,
//   grad_a = s (a1 u_a + a3 w_a),
//   H_ab   = s (a1 (g_a.g_b) I + a2 u_a u_b^T + lam w_a w_b^T - a3 [F (g_a x g_b)]x).
// The Hessian annihilates translations, so with an orthonormal basis U (4x3) of the
// complement of (1,1,1,1), H = (U x I3) M (U x I3)^T where M is the same formula on the
// three "virtual vertices" c_k = sum_a U[a,k] g_a. The PSD projection of H is therefore
// (U x I3) proj(M) (U x I3)^T: a 9x9 eigenproblem instead of 12x12.
//
// Projection decision: the packed Schur PSD test of hessianKernelHeader.pyx:79-141 runs on
// the 12x12 H; only elements that fail it are projected (second kernel, work queue).
#include "common.hpp"
#include "dense.hpp"
#include "element_common.hpp"

#include <type_traits>

namespace ysp {
namespace {

constexpr unsigned kBlock = 128;

constexpr int kMaxSlots = 16;

// Slot bookkeeping of one element (hessianKernelNoProject.pyx:93-163), per local vertex.
// Slots come in first-occurrence groups of equal |perm|; every 3-wide slot belongs to the
// next local vertex (its column / 3). Here each local vertex owns at most one slot, so the
// element is described by, per vertex: its target index, whether the slot is a target
// (index >= 2, in a group) and the rank of its group among the element's target groups.
// Everything is indexed by compile-time vertex numbers (a run-time index into a private
// array would place it in scratch memory). KT is the compile-time slot capacity, S.K the
// slot count; NV the element's local vertices. An element that repeats a target vertex
// (two local vertices in one group) is rejected (ok = false).
template <int NV>
struct VertexSlots {
  uint32_t idx[NV];   // seg_idx of the vertex's slot, 0 without one
  bool has[NV];       // the vertex owns a 3-wide slot
  bool target[NV];    // ... with a target index (>= 2) in a group
  int rank[NV];       // rank of the vertex's group among the target groups
  int ntarget;
  bool ok;
};

template <int KT, int NV>
__device__ __forceinline__ void load_vertex_slots(const ScatterArgs& S, uint32_t e,
                                                  VertexSlots<NV>& vs) {
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
    if (sz == 3) {
      const uint32_t v = col / 3;
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

// Gradient segments (hessianKernelNoProject.pyx:63-88): FP64 atomics into the gradient,
// or, staged, the element's whole local gradient written in place. grad(a, g3) gives the
// 3-vector of local vertex a.
template <int NV, typename GradFn>
__device__ __forceinline__ void scatter_gradient(const ScatterArgs& S, uint32_t e,
                                                 const VertexSlots<NV>& vs, GradFn grad) {
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    if (S.grad_staging ? !vs.has[a] : (!vs.has[a] || vs.idx[a] < 2)) continue;
    double g3[3];
    grad(a, g3);
    if (S.grad_staging) {
      double* out = S.grad_staging + static_cast<size_t>(e) * S.grad_local + 3 * a;
      out[0] = g3[0];
      out[1] = g3[1];
      out[2] = g3[2];
    } else {
      const uint32_t p = vs.idx[a] - 2;
      atomic_add_f64(S.gradient + p, g3[0]);
      atomic_add_f64(S.gradient + p + 1, g3[1]);
      atomic_add_f64(S.gradient + p + 2, g3[2]);
    }
  }
}

// Hessian blocks (hessianKernelNoProject.pyx:165-264). blockfn(a, b, out[3][3]) adds the
// local 3x3 block between local vertices a and b into out. The target-group pair of ranks
// lo <= hi is coordinate coordStart + lo*n - lo(lo-1)/2 + (hi-lo) (pairs in first-
// occurrence order); its block is oriented with the rows of the smaller target index:
// FP64 atomics at lookups[coordinate], or, staged, a plain store at staging[coordinate].
template <int NV, typename BlockFn>
__device__ __forceinline__ void scatter_hessian(const ScatterArgs& S, uint32_t e,
                                                const VertexSlots<NV>& vs, BlockFn blockfn) {
  const uint32_t cstart = S.coord_outer[e];
  const int n = vs.ntarget;
#pragma unroll
  for (int a = 0; a < NV; ++a)
#pragma unroll
    for (int b = a; b < NV; ++b) {
      if (!vs.target[a] || !vs.target[b]) continue;
      double blk[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      blockfn(a, b, blk);
      const int lo = min(vs.rank[a], vs.rank[b]), hi = max(vs.rank[a], vs.rank[b]);
      const uint32_t coord = cstart + static_cast<uint32_t>(lo * n - lo * (lo - 1) / 2 + (hi - lo));
      const bool tr = !(vs.idx[a] < vs.idx[b]);   // includes the diagonal block a == b
      if (S.block_staging) {
        double* dst = S.block_staging + 9 * static_cast<size_t>(coord);
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) dst[3 * k + l] = tr ? blk[l][k] : blk[k][l];
      } else {
        double* dst = S.blocks + S.lookups[coord];
        double* dd = nullptr;
        if (a == b && S.diag_blocks) dd = S.diag_blocks + diag_offset(S, vs.idx[a] - 2, 3);
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            const double v = tr ? blk[l][k] : blk[k][l];
            atomic_add_f64(dst + 3 * k + l, v);
            if (dd) atomic_add_f64(dd + 3 * k + l, v);
          }
      }
    }
}

// ------------------------------------------------------------------ SNH element
struct Snh {
  double F[3][3];    // F[i][j]
  double cof[3][3];  // cof[i][j] = column j of cof F
  double g[4][3];    // shape vectors
  double u[4][3];
  double w[4][3];
  double s, a1, a2, a3, lam, ic, J;
};

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ void snh_setup(const double x[4][3], const double* Di, double vol,
                                          double mu, double lam, double dt, Snh& S) {
  double d[3][3];  // d[k] = x_{k+1} - x_0
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i) d[k][i] = x[k + 1][i] - x[0][i];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      S.F[i][j] = d[0][i] * Di[0 * 3 + j] + d[1][i] * Di[1 * 3 + j] + d[2][i] * Di[2 * 3 + j];
  double f[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) f[j][i] = S.F[i][j];
  double c0[3], c1[3], c2[3];
  cross3(f[1], f[2], c0);
  cross3(f[2], f[0], c1);
  cross3(f[0], f[1], c2);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    S.cof[i][0] = c0[i];
    S.cof[i][1] = c1[i];
    S.cof[i][2] = c2[i];
  }
  double ic = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) ic += S.F[i][j] * S.F[i][j];
  const double J = f[0][0] * c0[0] + f[0][1] * c0[1] + f[0][2] * c0[2];
  const double alpha = 1.0 + 0.75 * mu / lam;
  const double ip1 = 1.0 / (ic + 1.0);
  S.ic = ic;
  S.J = J;
  S.a1 = mu * (1.0 - ip1);
  S.a2 = 2.0 * mu * ip1 * ip1;
  S.a3 = lam * (J - alpha);
  S.lam = lam;
  S.s = dt * dt * vol;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    S.g[0][j] = -(Di[0 * 3 + j] + Di[1 * 3 + j] + Di[2 * 3 + j]);
    S.g[1][j] = Di[0 * 3 + j];
    S.g[2][j] = Di[1 * 3 + j];
    S.g[3][j] = Di[2 * 3 + j];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      S.u[a][i] = S.F[i][0] * S.g[a][0] + S.F[i][1] * S.g[a][1] + S.F[i][2] * S.g[a][2];
      S.w[a][i] = S.cof[i][0] * S.g[a][0] + S.cof[i][1] * S.g[a][1] + S.cof[i][2] * S.g[a][2];
    }
}

// 3x3 block of the SNH Hessian for shape vectors ga, gb with u/w images.
__device__ __forceinline__ void snh_block(const Snh& S, const double* ga, const double* gb,
                                          const double* ua, const double* ub, const double* wa,
                                          const double* wb, double out[3][3]) {
  const double gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
  double c[3];
  cross3(ga, gb, c);
  double t[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = S.F[i][0] * c[0] + S.F[i][1] * c[1] + S.F[i][2] * c[2];
  // [t]x = [[0,-t2,t1],[t2,0,-t0],[-t1,t0,0]]
  const double tx[3][3] = {{0.0, -t[2], t[1]}, {t[2], 0.0, -t[0]}, {-t[1], t[0], 0.0}};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double v = S.a2 * ua[i] * ub[k] + S.lam * wa[i] * wb[k] - S.a3 * tx[i][k];
      if (i == k) v += S.a1 * gg;
      out[i][k] = S.s * v;
    }
}

// Gradient of local vertex a.
__device__ __forceinline__ void snh_gradient_at(const Snh& S, int a, double g3[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) g3[i] = S.s * (S.a1 * S.u[a][i] + S.a3 * S.w[a][i]);
}

__device__ __forceinline__ double snh_energy_value(const Snh& S, double mu) {
  const double alpha_gap = S.a3 / S.lam;  // J - alpha
  return S.s * (0.5 * mu * (S.ic - 3.0) - 0.5 * mu * log(S.ic + 1.0) +
                0.5 * S.lam * alpha_gap * alpha_gap);
}

__device__ __forceinline__ void load_tet(const UnionView& P, const uint32_t* tet2v, uint32_t e,
                                         double x[4][3], uint32_t vid[4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    vid[a] = tet2v[4 * static_cast<size_t>(e) + a];
    const double* p = union_row(P, vid[a]);
#pragma unroll
    for (int i = 0; i < 3; ++i) x[a][i] = p[i];
  }
}

__global__ void __launch_bounds__(kBlock) snh_rest_kernel(UnionView R, const uint32_t* tet2v,
                                                          uint32_t n, double* dm_inv, double* vol) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double x[4][3];
  uint32_t vid[4];
  load_tet(R, tet2v, e, x, vid);
  double m[3][3];  // columns = rest edges
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i) m[i][k] = x[k + 1][i] - x[0][i];
  const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
  const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
  const double id = 1.0 / det;
  double* o = dm_inv + 9 * static_cast<size_t>(e);
  o[0] = c00 * id;
  o[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
  o[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
  o[3] = c01 * id;
  o[4] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
  o[5] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
  o[6] = c02 * id;
  o[7] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
  o[8] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
  vol[e] = det / 6.0;
}

struct SnhParams {
  const double* dm_inv;
  const double* vol;
  const double* mu;
  uint32_t mu_stride;
  const double* lam;
  uint32_t lam_stride;
  double dt;
  uint32_t method;
};

__device__ __forceinline__ void snh_load(const UnionView& P, const uint32_t* tet2v,
                                         const SnhParams& prm, uint32_t e, Snh& S, double& mu) {
  double x[4][3];
  uint32_t vid[4];
  load_tet(P, tet2v, e, x, vid);
  double Di[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Di[k] = prm.dm_inv[9 * static_cast<size_t>(e) + k];
  mu = prm.mu[static_cast<size_t>(e) * prm.mu_stride];
  const double lam = prm.lam[static_cast<size_t>(e) * prm.lam_stride];
  snh_setup(x, Di, prm.vol[e], mu, lam, prm.dt, S);
}

// Pass 1: gradient, PSD test, scatter of PSD (or unprojected) Hessians, queue the rest.
template <int KT>
__global__ void __launch_bounds__(kBlock)
    snh_eval_kernel(UnionView P, const uint32_t* __restrict__ tet2v, uint32_t n, SnhParams prm,
                    ScatterArgs SA, uint32_t* __restrict__ queue, uint32_t* __restrict__ qcount,
                    int* __restrict__ error_flag) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  VertexSlots<4> vs;
  load_vertex_slots<KT>(SA, e, vs);
  if (!vs.ok) {
    atomicOr(error_flag, 1);
    return;
  }
  auto scatter_blocks = [&](const Snh& S) {
    scatter_hessian(SA, e, vs, [&](int a, int b, double out[3][3]) {
      double blk[3][3];
      snh_block(S, S.g[a], S.g[b], S.u[a], S.u[b], S.w[a], S.w[b], blk);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) out[i][k] += blk[i][k];
    });
  };
  double L[78];   // lower packed element Hessian for the PSD test
  {
    Snh S;
    double mu;
    snh_load(P, tet2v, prm, e, S, mu);
    scatter_gradient(SA, e, vs, [&](int a, double* g3) { snh_gradient_at(S, a, g3); });
    if (prm.method == 0) {
      scatter_blocks(S);
      return;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b <= a; ++b) {
        double blk[3][3];
        snh_block(S, S.g[a], S.g[b], S.u[a], S.u[b], S.w[a], S.w[b], blk);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const int r = 3 * a + i, c = 3 * b + k;
            if (c <= r) L[r * (r + 1) / 2 + c] = blk[i][k];
          }
      }
  }
  // The test destroys L; a PSD element re-derives its blocks from the tet (the element
  // state is not kept alive across the test, which keeps the kernel within 256 registers).
  if (psd_test_packed<12>(L)) {
    Snh S;
    double mu;
    snh_load(P, tet2v, prm, e, S, mu);
    scatter_blocks(S);
  } else {
    queue[wave_reserve(qcount, 1u)] = e;
  }
}

// Orthonormal basis of the complement of (1,1,1,1): columns of U (4x3).
__constant__ double kU[4][3] = {
    {0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225},
    {-0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225},
    {0.0, -0.81649658092772603273, 0.28867513459481288225},
    {0.0, 0.0, -0.86602540378443864676}};

// Pass 2: project the queued (non-PSD) SNH Hessians through the 9x9 reduction.
template <int KT>
__global__ void __launch_bounds__(64)
    snh_project_kernel(UnionView P, const uint32_t* __restrict__ tet2v, SnhParams prm,
                       ScatterArgs SA, const uint32_t* __restrict__ queue,
                       const uint32_t* __restrict__ qcount) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= *qcount) return;
  const uint32_t e = queue[q];
  Snh S;
  double mu;
  snh_load(P, tet2v, prm, e, S, mu);
  // virtual vertices
  double cg[3][3], cu[3][3], cw[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cg[k][j] = kU[0][k] * S.g[0][j] + kU[1][k] * S.g[1][j] + kU[2][k] * S.g[2][j] +
                 kU[3][k] * S.g[3][j];
      cu[k][j] = kU[0][k] * S.u[0][j] + kU[1][k] * S.u[1][j] + kU[2][k] * S.u[2][j] +
                 kU[3][k] * S.u[3][j];
      cw[k][j] = kU[0][k] * S.w[0][j] + kU[1][k] * S.w[1][j] + kU[2][k] * S.w[2][j] +
                 kU[3][k] * S.w[3][j];
    }
  constexpr int kP = upper_index<9>(8, 8) + 1;   // 45 packed entries
  double M[kP];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int l = k; l < 3; ++l) {
      double blk[3][3];
      snh_block(S, cg[k], cg[l], cu[k], cu[l], cw[k], cw[l], blk);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          if (k < l || i <= j) M[upper_index<9>(3 * k + i, 3 * l + j)] = blk[i][j];
    }
  double V[9][9];
  jacobi_eigen_packed<9>(M, V);
  double lam_c[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const double l = M[upper_index<9>(i, i)];
    lam_c[i] = (l < 0.0) ? (prm.method == 1 ? -l : 0.0) : l;
  }
  // projected M' = V diag(lam_c) V^T (packed upper)
  double Mp[kP];
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = i; j < 9; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) acc += V[i][k] * lam_c[k] * V[j][k];
      Mp[upper_index<9>(i, j)] = acc;
    }
  VertexSlots<4> vs;
  load_vertex_slots<KT>(SA, e, vs);
  scatter_hessian(SA, e, vs, [&](int a, int b, double out[3][3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        const double w = kU[a][k] * kU[b][l];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const int r = 3 * k + i, c = 3 * l + j;
            out[i][j] += w * Mp[r <= c ? upper_index<9>(r, c) : upper_index<9>(c, r)];
          }
      }
  });
}

__global__ void __launch_bounds__(kBlock)
    snh_energy_kernel(UnionView P, const uint32_t* __restrict__ tet2v, uint32_t n, SnhParams prm,
                      double* __restrict__ energy) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  Snh S;
  double mu;
  snh_load(P, tet2v, prm, e, S, mu);
  energy[e] = snh_energy_value(S, mu);
}

// Adjoint of the material parameter theta_e = log E_e of every tet: with mu and lam both
// proportional to E, the SNH energy (and so its gradient) is homogeneous of degree one
// in (mu, lam), hence dg_e/dtheta_e = mu dg_e/dmu + lam dg_e/dlam = g_e, and
// dL/dtheta_e = -lambda_e . g_e for the adjoint solution lambda of H lambda = dL/dx
// (gathered at the element's target segments).
template <int KT>
__global__ void __launch_bounds__(kBlock)
    snh_theta_adjoint_kernel(UnionView P, const uint32_t* __restrict__ tet2v, uint32_t n,
                             SnhParams prm, ScatterArgs SA, const double* __restrict__ adjoint,
                             double* __restrict__ out, int* __restrict__ error_flag) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  Snh S;
  double mu;
  snh_load(P, tet2v, prm, e, S, mu);
  VertexSlots<4> vs;
  load_vertex_slots<KT>(SA, e, vs);
  if (!vs.ok) {
    atomicOr(error_flag, 1);
    return;
  }
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (!vs.has[a] || vs.idx[a] < 2) continue;
    const double* l = adjoint + (vs.idx[a] - 2);
    double g3[3];
    snh_gradient_at(S, a, g3);
#pragma unroll
    for (int k = 0; k < 3; ++k) acc += l[k] * g3[k];
  }
  out[e] = -acc;
}

// ------------------------------------------------------------------ inertia
struct InertiaArgs {
  const double* x;
  const double* xp;
  const double* v;
  const double* m;
  double dt;
};

__device__ __forceinline__ void inertia_diff(const InertiaArgs& I, uint32_t e, double d[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double target = I.xp[3 * static_cast<size_t>(e) + i] + I.v[3 * static_cast<size_t>(e) + i] * I.dt;
    if (i == 1) target -= 9.8 * I.dt * I.dt;
    d[i] = I.x[3 * static_cast<size_t>(e) + i] - target;
  }
}

template <int KT>
__global__ void __launch_bounds__(kBlock)
    inertia_eval_kernel(InertiaArgs I, uint32_t n, uint32_t method, ScatterArgs SA,
                        int* __restrict__ error_flag) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double d[3];
  inertia_diff(I, e, d);
  double m = I.m[e];
  // Hessian m I: PSD for m >= -tol; otherwise the clamp rule on the eigenvalue m.
  double mh = m;
  if (method != 0 && m < -1.0e-10 * fmax(1.0, fabs(m))) mh = (method == 1) ? -m : 0.0;
  VertexSlots<1> vs;
  load_vertex_slots<KT>(SA, e, vs);
  if (!vs.ok) {
    atomicOr(error_flag, 1);
    return;
  }
  scatter_gradient(SA, e, vs, [&](int, double* g3) {
#pragma unroll
    for (int i = 0; i < 3; ++i) g3[i] = m * d[i];
  });
  scatter_hessian(SA, e, vs, [&](int, int, double out[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i][i] += mh;
  });
}

__global__ void __launch_bounds__(kBlock)
    inertia_energy_kernel(InertiaArgs I, uint32_t n, double* __restrict__ energy) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double d[3];
  inertia_diff(I, e, d);
  energy[e] = 0.5 * I.m[e] * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

int check_flag(int* flag, hipStream_t s, const char* what) {
  return defer_flag_check(flag, s, what);
}

}  // namespace
}  // namespace ysp

using namespace ysp;

extern "C" int ysp_snh_rest_data(const ysp_union_f64* rest_positions, const uint32_t* tet2v,
                                  uint32_t num_tets, double* dm_inv, double* vol, void* stream) {
  if (num_tets == 0) return YSP_OK;
  if (!rest_positions || rest_positions->num_children == 0 ||
      rest_positions->num_children > YSP_MAX_UNION_CHILDREN || rest_positions->width != 3) {
    ysp_set_error(__FILE__, __LINE__, "ysp_snh_rest_data: bad union view");
    return YSP_ERR_ARG;
  }
  hipLaunchKernelGGL(snh_rest_kernel, dim3(grid_for(num_tets, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), make_view(rest_positions), tet2v, num_tets, dm_inv, vol);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_eval_snh_tet(const ysp_union_f64* positions, const uint32_t* tet2v,
                                uint32_t num_tets, const double* dm_inv, const double* vol,
                                const double* mu, uint32_t mu_stride, const double* lam,
                                uint32_t lam_stride, double dt, uint32_t projection_method,
                                const ysp_scatter_args* out, void* stream) {
  if (num_tets == 0) return YSP_OK;
  if (!positions || positions->num_children == 0 ||
      positions->num_children > YSP_MAX_UNION_CHILDREN || positions->width != 3 || !out ||
      out->max_num_indices > kMaxSlots || projection_method > 2) {
    ysp_set_error(__FILE__, __LINE__, "ysp_eval_snh_tet: bad arguments");
    return YSP_ERR_ARG;
  }
  hipStream_t s = as_stream(stream);
  uint32_t* queue = static_cast<uint32_t*>(scratch(30, sizeof(uint32_t) * num_tets));
  uint32_t* qcount = static_cast<uint32_t*>(scratch(31, 2 * sizeof(uint32_t) + sizeof(int)));
  if (!queue || !qcount) return YSP_ERR_HIP;
  int* flag = reinterpret_cast<int*>(qcount + 1);
  YSP_CHECK(hipMemsetAsync(qcount, 0, 2 * sizeof(uint32_t), s));
  SnhParams prm{dm_inv, vol, mu, mu_stride, lam, lam_stride, dt, projection_method};
  const UnionView P = make_view(positions);
  const ScatterArgs SA = make_scatter(out);
  auto run = [&](auto kt) {
    constexpr int KT = decltype(kt)::value;
    hipLaunchKernelGGL(snh_eval_kernel<KT>, dim3(grid_for(num_tets, kBlock)), dim3(kBlock), 0, s, P,
                       tet2v, num_tets, prm, SA, queue, qcount, flag);
    if (projection_method != 0)
      hipLaunchKernelGGL(snh_project_kernel<KT>, dim3(grid_for(num_tets, 64)), dim3(64), 0, s, P,
                         tet2v, prm, SA, queue, qcount);
  };
  if (out->max_num_indices == 8) run(std::integral_constant<int, 8>());
  else run(std::integral_constant<int, kMaxSlots>());
  YSP_CHECK_LAUNCH();
  return check_flag(flag, s, "ysp_eval_snh_tet: slot layout not supported (K>16, size!=3 or a repeated vertex)");
}

extern "C" int ysp_energy_snh_tet(const ysp_union_f64* positions, const uint32_t* tet2v,
                                  uint32_t num_tets, const double* dm_inv, const double* vol,
                                  const double* mu, uint32_t mu_stride, const double* lam,
                                  uint32_t lam_stride, double dt, double* energy, void* stream) {
  if (num_tets == 0) return YSP_OK;
  if (!positions || positions->num_children == 0 ||
      positions->num_children > YSP_MAX_UNION_CHILDREN || positions->width != 3) {
    ysp_set_error(__FILE__, __LINE__, "ysp_energy_snh_tet: bad union view");
    return YSP_ERR_ARG;
  }
  SnhParams prm{dm_inv, vol, mu, mu_stride, lam, lam_stride, dt, 0};
  hipLaunchKernelGGL(snh_energy_kernel, dim3(grid_for(num_tets, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), make_view(positions), tet2v, num_tets, prm, energy);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_snh_theta_adjoint(const ysp_union_f64* positions, const uint32_t* tet2v,
                                     uint32_t num_tets, const double* dm_inv, const double* vol,
                                     const double* mu, uint32_t mu_stride, const double* lam,
                                     uint32_t lam_stride, double dt,
                                     const ysp_scatter_args* slots, const double* adjoint,
                                     double* dtheta, void* stream) {
  if (num_tets == 0) return YSP_OK;
  if (!positions || positions->num_children == 0 ||
      positions->num_children > YSP_MAX_UNION_CHILDREN || positions->width != 3 || !slots ||
      slots->max_num_indices > kMaxSlots || !adjoint || !dtheta) {
    ysp_set_error(__FILE__, __LINE__, "ysp_snh_theta_adjoint: bad arguments");
    return YSP_ERR_ARG;
  }
  hipStream_t s = as_stream(stream);
  int* flag = static_cast<int*>(scratch(39, sizeof(int)));
  if (!flag) return YSP_ERR_HIP;
  YSP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
  SnhParams prm{dm_inv, vol, mu, mu_stride, lam, lam_stride, dt, 0};
  const ScatterArgs SA = make_scatter(slots);
  const dim3 g(grid_for(num_tets, kBlock)), b(kBlock);
  if (slots->max_num_indices == 8)
    hipLaunchKernelGGL(snh_theta_adjoint_kernel<8>, g, b, 0, s, make_view(positions), tet2v,
                       num_tets, prm, SA, adjoint, dtheta, flag);
  else
    hipLaunchKernelGGL(snh_theta_adjoint_kernel<kMaxSlots>, g, b, 0, s, make_view(positions),
                       tet2v, num_tets, prm, SA, adjoint, dtheta, flag);
  YSP_CHECK_LAUNCH();
  return check_flag(flag, s, "ysp_snh_theta_adjoint: slot layout not supported");
}

extern "C" int ysp_eval_inertia(const double* position, const double* last_position,
                                const double* velocity, const double* mass, double dt,
                                uint32_t num_vertices, uint32_t projection_method,
                                const ysp_scatter_args* out, void* stream) {
  if (num_vertices == 0) return YSP_OK;
  if (!out || out->max_num_indices > kMaxSlots || projection_method > 2) {
    ysp_set_error(__FILE__, __LINE__, "ysp_eval_inertia: bad arguments");
    return YSP_ERR_ARG;
  }
  hipStream_t s = as_stream(stream);
  int* flag = static_cast<int*>(scratch(32, sizeof(int)));
  if (!flag) return YSP_ERR_HIP;
  YSP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
  InertiaArgs I{position, last_position, velocity, mass, dt};
  if (out->max_num_indices == 1)
    hipLaunchKernelGGL(inertia_eval_kernel<1>, dim3(grid_for(num_vertices, kBlock)), dim3(kBlock), 0,
                       s, I, num_vertices, projection_method, make_scatter(out), flag);
  else
    hipLaunchKernelGGL(inertia_eval_kernel<kMaxSlots>, dim3(grid_for(num_vertices, kBlock)),
                       dim3(kBlock), 0, s, I, num_vertices, projection_method, make_scatter(out), flag);
  YSP_CHECK_LAUNCH();
  return check_flag(flag, s, "ysp_eval_inertia: slot layout not supported");
}

extern "C" int ysp_energy_inertia(const double* position, const double* last_position,
                                  const double* velocity, const double* mass, double dt,
                                  uint32_t num_vertices, double* energy, void* stream) {
  if (num_vertices == 0) return YSP_OK;
  InertiaArgs I{position, last_position, velocity, mass, dt};
  hipLaunchKernelGGL(inertia_energy_kernel, dim3(grid_for(num_vertices, kBlock)), dim3(kBlock), 0,
                     as_stream(stream), I, num_vertices, energy);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}
