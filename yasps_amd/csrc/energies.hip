// The library's element energies (IPC barriers, bending, Baraff-Witkin, StVK membrane,
// affine orthogonality, and SNH / inertia for vertices driven by affine bodies): the
// Kind specialisations over the generated functions (generated_energies.hpp), the rest
// data of the cloth energies, and the C ABI. The kernels are energy_kernels.hpp's.
#include "energy_kernels.hpp"
#include "generated_energies.hpp"

namespace ysp {
namespace {

#define YSP_ENERGY_KIND(ID, NAME, ARITY, WIDTH, NPAR, NSCENE)                               \
  template <>                                                                             \
  struct Kind<ID> {                                                                       \
    static constexpr int A = ARITY, W = WIDTH, N = ARITY * WIDTH, P = NPAR, S = NSCENE;   \
    static_assert(N == gen::kGen_##NAME##_N, "generated size mismatch");                 \
    __device__ static void eval(const double* x, const double* p, const double* s, double& E, \
                                double* g, double* H) {                                  \
      gen::gen_##NAME(x, p, s, E, g, H);                                                 \
    }                                                                                     \
  };

YSP_ENERGY_KIND(YSP_ENERGY_POINT_POINT, point_point, 2, 3, 0, 2)
YSP_ENERGY_KIND(YSP_ENERGY_POINT_EDGE, point_edge, 3, 3, 0, 2)
YSP_ENERGY_KIND(YSP_ENERGY_POINT_TRIANGLE, point_triangle, 4, 3, 0, 2)
YSP_ENERGY_KIND(YSP_ENERGY_EDGE_EDGE, edge_edge, 4, 3, 0, 2)
YSP_ENERGY_KIND(YSP_ENERGY_BENDING, bending, 4, 3, 2, 1)
YSP_ENERGY_KIND(YSP_ENERGY_BARAFF_WITKIN, baraff_witkin, 3, 3, 5, 5)
YSP_ENERGY_KIND(YSP_ENERGY_STVK_MEMBRANE, stvk_membrane, 3, 3, 5, 3)
YSP_ENERGY_KIND(YSP_ENERGY_AFFINE_ORTHOGONALITY, affine_orthogonality, 1, 9, 0, 1)
YSP_ENERGY_KIND(YSP_ENERGY_SNH, snh, 4, 3, 10, 3)
YSP_ENERGY_KIND(YSP_ENERGY_INERTIA, inertia, 1, 3, 7, 1)
#undef YSP_ENERGY_KIND

// ----------------------------------------------------------------- rest data
__device__ __forceinline__ void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Dihedral angle of examples/many_bunnies_one_cloth/helpers.py:193-203.
__device__ __forceinline__ double edge_theta(const double* q0, const double* q1,
                                             const double* q2, const double* q3) {
  double a[3], b[3], c[3], d[3], n0[3], n1[3], axis[3], m[3];
  for (int i = 0; i < 3; ++i) {
    a[i] = q0[i] - q2[i];
    b[i] = q1[i] - q2[i];
    c[i] = q1[i] - q3[i];
    d[i] = q0[i] - q3[i];
    axis[i] = q1[i] - q0[i];
  }
  cross(a, b, n0);
  cross(c, d, n1);
  cross(n0, n1, m);
  const double num = dot(m, axis) / sqrt(dot(axis, axis));
  const double den = dot(n0, n1) + sqrt(dot(n0, n0)) * sqrt(dot(n1, n1));
  return 2.0 * atan2(num, den);
}

// 2x2 rest shape of helpers.py:214-237 (triangle rotated so that its normal is +y, then
// (x, z) coordinates), inverted; area = thickness |v01 x v02| / 2.
__device__ __forceinline__ void triangle_rest(const double* v0, const double* v1,
                                              const double* v2, double thickness, double* out) {
  double e1[3], e2[3], nrm[3];
  for (int i = 0; i < 3; ++i) {
    e1[i] = v1[i] - v0[i];
    e2[i] = v2[i] - v0[i];
  }
  cross(e1, e2, nrm);
  const double len = sqrt(dot(nrm, nrm));
  for (int i = 0; i < 3; ++i) nrm[i] /= len;
  const double tgt[3] = {0.0, 1.0, 0.0};
  double vv[3];
  cross(nrm, tgt, vv);
  const double cs = dot(nrm, tgt);
  const double K[3][3] = {{0.0, -vv[2], vv[1]}, {vv[2], 0.0, -vv[0]}, {-vv[1], vv[0], 0.0}};
  double R[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double k2 = 0.0;
      for (int k = 0; k < 3; ++k) k2 += K[i][k] * K[k][j];
      R[i][j] = (i == j ? 1.0 : 0.0) + K[i][j] + k2 / (1.0 + cs);
    }
  double uv[3][2];
  const double* vs[3] = {v0, v1, v2};
  for (int t = 0; t < 3; ++t) {
    uv[t][0] = R[0][0] * vs[t][0] + R[0][1] * vs[t][1] + R[0][2] * vs[t][2];
    uv[t][1] = R[2][0] * vs[t][0] + R[2][1] * vs[t][1] + R[2][2] * vs[t][2];
  }
  const double m00 = uv[1][0] - uv[0][0], m01 = uv[2][0] - uv[0][0];
  const double m10 = uv[1][1] - uv[0][1], m11 = uv[2][1] - uv[0][1];
  const double det = m00 * m11 - m01 * m10;
  out[0] = m11 / det;
  out[1] = -m01 / det;
  out[2] = -m10 / det;
  out[3] = m00 / det;
  out[4] = thickness * len / 2.0;
}

__global__ void __launch_bounds__(kEB)
    rest_kernel(int kind, UnionView R, const uint32_t* __restrict__ e2v, uint32_t n,
                double thickness, double* __restrict__ ep) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  if (kind == YSP_ENERGY_BENDING) {
    const double* q[4];
    for (int a = 0; a < 4; ++a) q[a] = union_row(R, e2v[4 * static_cast<size_t>(e) + a]);
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = q[1][i] - q[0][i];
    ep[2 * static_cast<size_t>(e)] = edge_theta(q[0], q[1], q[2], q[3]);
    ep[2 * static_cast<size_t>(e) + 1] = sqrt(dot(d, d));
  } else {
    const double* v[3];
    for (int a = 0; a < 3; ++a) v[a] = union_row(R, e2v[3 * static_cast<size_t>(e) + a]);
    triangle_rest(v[0], v[1], v[2], thickness, ep + 5 * static_cast<size_t>(e));
  }
}

__global__ void __launch_bounds__(256)
    affine_positions_kernel(const double* __restrict__ A, const double* __restrict__ t,
                            const uint32_t* __restrict__ body, const double* __restrict__ rest,
                            uint32_t n, double* __restrict__ x) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint32_t b = body[v];
  const double* a = A + 9 * static_cast<size_t>(b);
  const double* r = rest + 3 * static_cast<size_t>(v);
#pragma unroll
  for (int j = 0; j < 3; ++j)
    x[3 * static_cast<size_t>(v) + j] =
        t[3 * static_cast<size_t>(b) + j] + a[3 * j] * r[0] + a[3 * j + 1] * r[1] + a[3 * j + 2] * r[2];
}


constexpr int kUnknownKind = -1000;   // dispatch_kind's result for an unknown kind
template <typename Fn>
int dispatch_kind(int32_t kind, Fn&& fn) {
  switch (kind) {
    case YSP_ENERGY_POINT_POINT: return fn(std::integral_constant<int, YSP_ENERGY_POINT_POINT>());
    case YSP_ENERGY_POINT_EDGE: return fn(std::integral_constant<int, YSP_ENERGY_POINT_EDGE>());
    case YSP_ENERGY_POINT_TRIANGLE: return fn(std::integral_constant<int, YSP_ENERGY_POINT_TRIANGLE>());
    case YSP_ENERGY_EDGE_EDGE: return fn(std::integral_constant<int, YSP_ENERGY_EDGE_EDGE>());
    case YSP_ENERGY_BENDING: return fn(std::integral_constant<int, YSP_ENERGY_BENDING>());
    case YSP_ENERGY_BARAFF_WITKIN: return fn(std::integral_constant<int, YSP_ENERGY_BARAFF_WITKIN>());
    case YSP_ENERGY_STVK_MEMBRANE: return fn(std::integral_constant<int, YSP_ENERGY_STVK_MEMBRANE>());
    case YSP_ENERGY_AFFINE_ORTHOGONALITY:
      return fn(std::integral_constant<int, YSP_ENERGY_AFFINE_ORTHOGONALITY>());
    case YSP_ENERGY_SNH: return fn(std::integral_constant<int, YSP_ENERGY_SNH>());
    case YSP_ENERGY_INERTIA: return fn(std::integral_constant<int, YSP_ENERGY_INERTIA>());
    case YSP_ENERGY_QUADRATIC_FORM:
      return fn(std::integral_constant<int, YSP_ENERGY_QUADRATIC_FORM>());
    default: return kUnknownKind;
  }
}


}  // namespace
}  // namespace ysp

using namespace ysp;

extern "C" int ysp_energy_arity(int32_t kind, uint32_t* arity, uint32_t* width,
                                uint32_t* element_params, uint32_t* scene_params) {
  return dispatch_kind(kind, [&](auto id) {
    using T = Kind<decltype(id)::value>;
    if (arity) *arity = T::A;
    if (width) *width = T::W;
    if (element_params) *element_params = T::P;
    if (scene_params) *scene_params = T::S;
    return YSP_OK;
  }) == YSP_OK ? YSP_OK : YSP_ERR_ARG;
}

extern "C" int ysp_element_rest_data(int32_t kind, const ysp_union_f64* rest_positions,
                                     const uint32_t* elem2v, uint32_t n, double thickness,
                                     double* element_params, void* stream) {
  if (kind != YSP_ENERGY_BENDING && kind != YSP_ENERGY_BARAFF_WITKIN &&
      kind != YSP_ENERGY_STVK_MEMBRANE) {
    ysp_set_error(__FILE__, __LINE__, "ysp_element_rest_data: energy has no rest data");
    return YSP_ERR_ARG;
  }
  if (n == 0) return YSP_OK;
  if (int rc = check_view(rest_positions, 3, "ysp_element_rest_data: bad union view")) return rc;
  hipLaunchKernelGGL(rest_kernel, dim3(grid_for(n, kEB)), dim3(kEB), 0, as_stream(stream), kind,
                     make_view(rest_positions), elem2v, n, thickness, element_params);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_affine_positions(const double* affine_matrix, const double* translation,
                                    const uint32_t* vertex_body, const double* rest_position,
                                    uint32_t n, double* position, void* stream) {
  if (n == 0) return YSP_OK;
  if (!affine_matrix || !translation || !vertex_body || !rest_position || !position) {
    ysp_set_error(__FILE__, __LINE__, "ysp_affine_positions: bad arguments");
    return YSP_ERR_ARG;
  }
  hipLaunchKernelGGL(affine_positions_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                     as_stream(stream), affine_matrix, translation, vertex_body, rest_position, n,
                     position);
  YSP_CHECK_LAUNCH();
  return YSP_OK;
}

extern "C" int ysp_eval_element_energy(int32_t kind, const ysp_union_f64* positions,
                                       const ysp_affine_children* affine,
                                       const uint32_t* elem2v, uint32_t n,
                                       const double* element_params, const double* scene_params,
                                       uint32_t projection_method,
                                       const ysp_scatter_args* out, void* stream) {
  const int rc = dispatch_kind(kind, [&](auto id) {
    return eval_energy<decltype(id)::value>(positions, affine, elem2v, n, element_params,
                                            scene_params, projection_method, out, stream,
                                            "ysp_eval_element_energy: bad arguments");
  });
  if (rc == kUnknownKind) {
    ysp_set_error(__FILE__, __LINE__, "ysp_eval_element_energy: unknown energy");
    return YSP_ERR_ARG;
  }
  return rc;
}

extern "C" int ysp_element_energy_values(int32_t kind, const ysp_union_f64* positions,
                                         const uint32_t* elem2v, uint32_t n,
                                         const double* element_params,
                                         const double* scene_params, double* energy,
                                         void* stream) {
  const int rc = dispatch_kind(kind, [&](auto id) {
    return energy_values<decltype(id)::value>(positions, elem2v, n, element_params, scene_params,
                                              energy, stream,
                                              "ysp_element_energy_values: bad union view");
  });
  if (rc == kUnknownKind) {
    ysp_set_error(__FILE__, __LINE__, "ysp_element_energy_values: unknown energy");
    return YSP_ERR_ARG;
  }
  return rc;
}
