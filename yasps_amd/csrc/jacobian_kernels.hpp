// The element kernel of the rectangular mixed second derivatives d g / d theta, shared by
// the library's energies (jacobian.hip) and run-time generated user energies
// (yasps_amd/usercode.py). A translation unit specialises MixedKind<ID> (A, W, N, P, S, PM:
// the element parameters whose d g / d p is generated, and
//   __device__ static void mixed(x, p, s, Gp, Gs)
// with Gp = d g / d p (N x PM) and Gs = d g / d s (N x S), row-major) and instantiates
// param_jacobian<ID>. Per element the block is B_e = Gp Cp_e + Gs Cs_e (N x w), Cp_e and
// Cs_e the caller's derivatives of the parameters w.r.t. theta.
#pragma once
#include "common.hpp"
#include "element_common.hpp"

namespace ysp {
namespace {

constexpr unsigned kJB = 64;
constexpr uint32_t kMaxCols = 9;

template <int ID>
struct MixedKind;

struct JacArgs {
  UnionView U;
  const uint32_t* elem2v;
  uint32_t n;
  const double* ep;             // [n, P] element parameters (P > 0)
  const double* se;             // scene parameters, se_stride apart per element
  uint32_t se_stride;           // 0: shared by every element
  const double* cp;             // [n, PM, w] d p / d theta, cp_stride apart (null: zero)
  uint32_t cp_stride;
  const double* cs;             // [n, S, w] d s / d theta, cs_stride apart (null: zero)
  uint32_t cs_stride;
  uint32_t w;
  double* blocks;               // [n, N, w]
};

template <int ID>
__global__ void __launch_bounds__(kJB) param_jacobian_kernel(JacArgs a) {
  using K = MixedKind<ID>;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n) return;
  double x[K::N];
#pragma unroll
  for (int v = 0; v < K::A; ++v) {
    const double* row = union_row(a.U, a.elem2v[static_cast<size_t>(e) * K::A + v]);
#pragma unroll
    for (int i = 0; i < K::W; ++i) x[K::W * v + i] = row ? row[i] : 0.0;
  }
  double p[K::P > 0 ? K::P : 1];
#pragma unroll
  for (int j = 0; j < K::P; ++j) p[j] = a.ep[static_cast<size_t>(e) * K::P + j];
  if constexpr (K::P == 0) p[0] = 0.0;
  double s[K::S];
  const double* sr = a.se + static_cast<size_t>(e) * a.se_stride;
#pragma unroll
  for (int j = 0; j < K::S; ++j) s[j] = sr[j];
  if constexpr (ID == YSP_ENERGY_BARAFF_WITKIN) bw_stretch_coefficients(x, p, s[3], s[4]);
  double Gp[K::N * (K::PM > 0 ? K::PM : 1)];
  double Gs[K::N * K::S];
  K::mixed(x, p, s, Gp, Gs);
  const double* cp = a.cp ? a.cp + static_cast<size_t>(e) * a.cp_stride : nullptr;
  const double* cs = a.cs ? a.cs + static_cast<size_t>(e) * a.cs_stride : nullptr;
  double* out = a.blocks + static_cast<size_t>(e) * K::N * a.w;
  for (uint32_t c = 0; c < a.w; ++c) {
    double dp[K::PM > 0 ? K::PM : 1], ds[K::S];
#pragma unroll
    for (int j = 0; j < K::PM; ++j) dp[j] = cp ? cp[j * a.w + c] : 0.0;
#pragma unroll
    for (int j = 0; j < K::S; ++j) ds[j] = cs ? cs[j * a.w + c] : 0.0;
#pragma unroll
    for (int r = 0; r < K::N; ++r) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < K::PM; ++j) acc += Gp[r * K::PM + j] * dp[j];
#pragma unroll
      for (int j = 0; j < K::S; ++j) acc += Gs[r * K::S + j] * ds[j];
      out[r * a.w + c] = acc;
    }
  }
}

template <int ID>
void launch_kind(const JacArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(param_jacobian_kernel<ID>, dim3(grid_for(a.n, kJB)), dim3(kJB), 0, s, a);
}


// Body of ysp_element_param_jacobian for energy ID.
template <int ID>
int param_jacobian(const ysp_union_f64* positions, const uint32_t* element_vertices, uint32_t n,
                   const double* element_params, const double* scene_params,
                   uint32_t scene_stride, const double* dparams, uint32_t dparams_stride,
                   const double* dscene, uint32_t dscene_stride, uint32_t w, double* blocks,
                   void* stream, const char* who) {
  if (n == 0) return YSP_OK;
  if (!positions || !element_vertices || !scene_params || !blocks || w == 0 || w > kMaxCols ||
      positions->num_children == 0 || positions->num_children > YSP_MAX_UNION_CHILDREN ||
      static_cast<int>(positions->width) != MixedKind<ID>::W ||
      (MixedKind<ID>::P > 0 && !element_params)) {
    ysp_set_error(__FILE__, __LINE__, who);
    return YSP_ERR_ARG;
  }
  JacArgs a{make_view(positions), element_vertices, n, element_params, scene_params, scene_stride,
            dparams, dparams_stride, dscene, dscene_stride, w, blocks};
  launch_kind<ID>(a, as_stream(stream));
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

}  // namespace
}  // namespace ysp
