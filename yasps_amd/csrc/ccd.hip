// Contact sets on the GPU: the collision module the reference's examples drive
// (examples/ccd/ccd.py over examples/ccd/mlbvh.cu and examples/ccd/ACCD.cu).
//
// * A linear BVH per primitive set (faces or edges): leaf boxes of the primitives'
//   vertices (at x, and at x - alpha d when the sweep is asked for), 63-bit Morton codes of
//   the box centres in the scene box, a radix sort, the binary radix tree of the sorted
//   codes (each internal node on its own: direction, range and split from common-prefix
//   lengths, equal codes ordered by position), boxes refitted bottom-up (a node's second
//   arriving child, seen through an agent-scope acquire-release counter, writes the union)
//   and escape links (the node after a skipped subtree), so a query walks the tree with no
//   stack.
// * Queries: every surface vertex against the face tree, every edge against the edge tree
//   (partners with a larger id); boxes overlap unless an axis separates them by the gap
//   sqrt(dhat); pairs sharing a vertex, or whose four vertices carry one nonzero mesh id,
//   are skipped. Candidates are compact int pairs: (-(p + 1), face) or (edge, edge).
// * Proximity: the distance type of a candidate (a vertex, an edge or the face of the
//   triangle; nine cases for two segments), its squared distance and, below dhat, the pair
//   of that type appended to the pp / pe / pt / ee arrays.
// * Step bound: additive CCD of each candidate along x - t d (eta = 1 - slackness), the
//   minimum of the times, 1 without candidates.
// Arithmetic is IEEE double operation by operation as written (no fused multiply-adds):
// the CPU restatement computes the same values.
#pragma clang fp contract(off)

#include "common.hpp"

#include <hipcub/hipcub.hpp>
#include <algorithm>

namespace ysp {
namespace {

constexpr unsigned kCB = 256;
constexpr uint32_t kNone = 0xffffffffu;
constexpr int kMaxAdvances = 1 << 20;

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 ld3(const double* a, uint32_t i) {
  const double* p = a + 3 * static_cast<size_t>(i);
  return {p[0], p[1], p[2]};
}
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ------------------------------------------------------------------ distances
__device__ __forceinline__ double d_pp(V3 a, V3 b) {
  const V3 d = sub(a, b);
  return dot(d, d);
}
__device__ __forceinline__ double d_pe(V3 p, V3 a, V3 b) {
  const V3 c = cross(sub(a, p), sub(b, p));
  const V3 e = sub(b, a);
  return dot(c, c) / dot(e, e);
}
__device__ __forceinline__ double d_pt(V3 p, V3 a, V3 b, V3 c) {
  const V3 n = cross(sub(b, a), sub(c, a));
  const double t = dot(sub(p, a), n);
  return t * t / dot(n, n);
}
__device__ __forceinline__ double d_ee(V3 a0, V3 a1, V3 b0, V3 b1) {
  const V3 n = cross(sub(a1, a0), sub(b1, b0));
  const double t = dot(sub(b0, a0), n);
  return t * t / dot(n, n);
}

// Coordinate of p - a along the edge a -> b and its offset along (b - a) x n.
__device__ __forceinline__ void edge_coords(V3 p, V3 a, V3 b, V3 n, double& s, double& side) {
  const V3 e = sub(b, a);
  const V3 w = sub(p, a);
  s = dot(w, e) / dot(e, e);
  side = dot(w, cross(e, n));
}

// Closest feature of triangle (t0, t1, t2) to p: 0-2 a vertex, 3-5 the edges (t0 t1),
// (t1 t2), (t2 t0), 6 the face (mlbvh.cu:453-522).
__device__ __forceinline__ int pt_type(V3 p, V3 t0, V3 t1, V3 t2) {
  const V3 n = cross(sub(t1, t0), sub(t2, t0));
  double s0, o0, s1, o1, s2, o2;
  edge_coords(p, t0, t1, n, s0, o0);
  if (s0 > 0.0 && s0 < 1.0 && o0 >= 0.0) return 3;
  edge_coords(p, t1, t2, n, s1, o1);
  if (s1 > 0.0 && s1 < 1.0 && o1 >= 0.0) return 4;
  edge_coords(p, t2, t0, n, s2, o2);
  if (s2 > 0.0 && s2 < 1.0 && o2 >= 0.0) return 5;
  if (s0 <= 0.0 && s2 >= 1.0) return 0;
  if (s1 <= 0.0 && s0 >= 1.0) return 1;
  if (s2 <= 0.0 && s1 >= 1.0) return 2;
  return 6;
}

// Closest features of segments (v0 v1) and (v2 v3), types 0-8 (mlbvh.cu:524-591).
__device__ __forceinline__ int ee_type(V3 v0, V3 v1, V3 v2, V3 v3) {
  const V3 u = sub(v1, v0), v = sub(v3, v2), w = sub(v0, v2);
  const double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
  const double D = a * c - b * b;
  double tD = D, tN;
  int kind = 8;
  const double sN = b * e - c * d;
  if (sN <= 0.0) {
    tN = e;
    tD = c;
    kind = 2;
  } else if (sN >= D) {
    tN = e + b;
    tD = c;
    kind = 5;
  } else {
    tN = a * e - b * d;
    const V3 uv = cross(u, v);
    if (tN > 0.0 && tN < tD && (dot(w, uv) == 0.0 || dot(uv, uv) < 1.0e-20 * a * c)) {
      if (sN < D / 2) {
        tN = e;
        tD = c;
        kind = 2;
      } else {
        tN = e + b;
        tD = c;
        kind = 5;
      }
    }
  }
  if (tN <= 0.0) {
    if (-d <= 0.0) return 0;
    if (-d >= a) return 3;
    return 6;
  }
  if (tN >= tD) {
    if (-d + b <= 0.0) return 1;
    if (-d + b >= a) return 4;
    return 7;
  }
  return kind;
}

__device__ __forceinline__ double pt_distance(V3 p, V3 t0, V3 t1, V3 t2) {
  switch (pt_type(p, t0, t1, t2)) {
    case 0: return d_pp(p, t0);
    case 1: return d_pp(p, t1);
    case 2: return d_pp(p, t2);
    case 3: return d_pe(p, t0, t1);
    case 4: return d_pe(p, t1, t2);
    case 5: return d_pe(p, t2, t0);
    default: return d_pt(p, t0, t1, t2);
  }
}

__device__ __forceinline__ double ee_distance(V3 a0, V3 a1, V3 b0, V3 b1) {
  switch (ee_type(a0, a1, b0, b1)) {
    case 0: return d_pp(a0, b0);
    case 1: return d_pp(a0, b1);
    case 2: return d_pe(a0, b0, b1);
    case 3: return d_pp(a1, b0);
    case 4: return d_pp(a1, b1);
    case 5: return d_pe(a1, b0, b1);
    case 6: return d_pe(b0, a0, a1);
    case 7: return d_pe(b1, a0, a1);
    default: return d_ee(a0, a1, b0, b1);
  }
}

// ---------------------------------------------------------------------- boxes
struct Box {
  double lo[3], hi[3];
};

__device__ __forceinline__ void grow(Box& b, V3 p) {
  b.lo[0] = fmin(b.lo[0], p.x);
  b.lo[1] = fmin(b.lo[1], p.y);
  b.lo[2] = fmin(b.lo[2], p.z);
  b.hi[0] = fmax(b.hi[0], p.x);
  b.hi[1] = fmax(b.hi[1], p.y);
  b.hi[2] = fmax(b.hi[2], p.z);
}

__device__ __forceinline__ Box empty_box() {
  Box b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = HUGE_VAL;
    b.hi[k] = -HUGE_VAL;
  }
  return b;
}

// A vertex's box contribution: x, and x - alpha d when sweeping.
__device__ __forceinline__ void grow_vertex(Box& b, const double* x, const double* d,
                                            double alpha, uint32_t v) {
  const V3 p = ld3(x, v);
  grow(b, p);
  if (d) grow(b, sub(p, scale(ld3(d, v), alpha)));
}

__device__ __forceinline__ Box ld_box(const double* boxes, uint32_t i) {
  const double* s = boxes + 6 * static_cast<size_t>(i);
  Box b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = s[k];
    b.hi[k] = s[3 + k];
  }
  return b;
}

__device__ __forceinline__ void st_box(double* boxes, uint32_t i, const Box& b) {
  double* s = boxes + 6 * static_cast<size_t>(i);
  for (int k = 0; k < 3; ++k) {
    s[k] = b.lo[k];
    s[3 + k] = b.hi[k];
  }
}

// Separated by at least the gap along an axis (the reference's `overlap`, mlbvh.cu:225).
__device__ __forceinline__ bool apart(const Box& q, const Box& b, double gap) {
  for (int k = 0; k < 3; ++k)
    if (b.lo[k] - q.hi[k] >= gap || q.lo[k] - b.hi[k] >= gap) return true;
  return false;
}

__global__ void __launch_bounds__(kCB)
    leaf_boxes_kernel(const double* __restrict__ x, const double* __restrict__ d, double alpha,
                      const uint32_t* __restrict__ prims, uint32_t arity, uint32_t n,
                      double* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Box b = empty_box();
  for (uint32_t k = 0; k < arity; ++k) grow_vertex(b, x, d, alpha, prims[arity * static_cast<size_t>(i) + k]);
  st_box(out, i, b);
}

// Scene box: per-block reductions of the leaf boxes, then one block over those.
__global__ void __launch_bounds__(kCB)
    box_reduce_kernel(const double* __restrict__ boxes, uint32_t n, double* __restrict__ out) {
  __shared__ double s[6][kCB];
  Box b = empty_box();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Box c = ld_box(boxes, i);
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = fmin(b.lo[k], c.lo[k]);
      b.hi[k] = fmax(b.hi[k], c.hi[k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    s[k][threadIdx.x] = b.lo[k];
    s[3 + k][threadIdx.x] = b.hi[k];
  }
  __syncthreads();
  for (unsigned h = kCB / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h)
      for (int k = 0; k < 3; ++k) {
        s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + h]);
        s[3 + k][threadIdx.x] = fmax(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + h]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[6 * static_cast<size_t>(blockIdx.x) + threadIdx.x] = s[threadIdx.x][0];
}

// 21 bits of a coordinate spread to every third bit.
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

__global__ void __launch_bounds__(kCB)
    morton_kernel(const double* __restrict__ boxes, uint32_t n, const double* __restrict__ scene,
                  uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Box b = ld_box(boxes, i);
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const double ext = scene[3 + k] - scene[k];
    const double c = (b.lo[k] + b.hi[k]) * 0.5;
    double u = ext > 0.0 ? (c - scene[k]) / ext : 0.0;
    u = fmin(fmax(u, 0.0), 1.0);
    q[k] = static_cast<uint32_t>(fmin(u * 2097152.0, 2097151.0));
  }
  keys[i] = spread3(q[0]) << 2 | spread3(q[1]) << 1 | spread3(q[2]);
  vals[i] = i;
}

// Common-prefix length of sorted keys i and j (-1 out of range; equal keys compare their
// positions).
__device__ __forceinline__ int prefix(const uint64_t* keys, uint32_t n, int i, int j) {
  if (j < 0 || j >= static_cast<int>(n)) return -1;
  const uint64_t a = keys[i], b = keys[j];
  if (a != b) return __clzll(static_cast<long long>(a ^ b));
  return 64 + __clz(static_cast<int>(static_cast<uint32_t>(i) ^ static_cast<uint32_t>(j)));
}

// Internal node i of the binary radix tree over n sorted leaves: internal nodes are
// [0, n - 1), leaf k is node n - 1 + k.
__global__ void __launch_bounds__(kCB)
    radix_tree_kernel(const uint64_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ child,
                      uint32_t* __restrict__ parent) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= static_cast<int>(n) - 1) return;
  const int dir = prefix(keys, n, i, i + 1) - prefix(keys, n, i, i - 1) >= 0 ? 1 : -1;
  const int dmin = prefix(keys, n, i, i - dir);
  int lmax = 2;
  while (prefix(keys, n, i, i + lmax * dir) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (prefix(keys, n, i, i + (l + t) * dir) > dmin) l += t;
  const int j = i + l * dir;
  const int dnode = prefix(keys, n, i, j);
  int s = 0;
  for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
    if (prefix(keys, n, i, i + (s + t) * dir) > dnode) s += t;
    if (t == 1) break;
  }
  const int split = i + s * dir + (dir < 0 ? -1 : 0);
  const int lo = min(i, j), hi = max(i, j);
  const uint32_t left = lo == split ? n - 1 + split : split;
  const uint32_t right = hi == split + 1 ? n - 1 + split + 1 : split + 1;
  child[2 * i] = left;
  child[2 * i + 1] = right;
  parent[left] = i;
  parent[right] = i;
}

__global__ void __launch_bounds__(kCB)
    leaf_setup_kernel(const uint32_t* __restrict__ sorted_vals, uint32_t n,
                      const double* __restrict__ leaf_boxes, double* __restrict__ boxes,
                      uint32_t* __restrict__ prim, uint32_t* __restrict__ flags,
                      uint32_t* __restrict__ parent) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t p = sorted_vals[k];
  prim[k] = p;
  st_box(boxes, n - 1 + k, ld_box(leaf_boxes, p));
  if (k + 1 < n) flags[k] = 0;
  if (k == 0) parent[0] = kNone;
}

// Bottom-up boxes: a leaf climbs until it is the first child to arrive at a node.
__global__ void __launch_bounds__(kCB)
    refit_kernel(uint32_t n, const uint32_t* __restrict__ child, const uint32_t* __restrict__ parent,
                 uint32_t* flags, double* boxes, int* bad) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || n < 2) return;
  uint32_t node = parent[n - 1 + k];
  for (uint32_t steps = 0; node != kNone; ++steps) {
    if (steps >= n || node >= n - 1) {        // deeper than n - 1 or not an internal node
      atomicOr(bad, 1);
      return;
    }
    const uint32_t seen = __hip_atomic_fetch_add(flags + node, 1u, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    if (seen == 0) return;
    Box a = ld_box(boxes, child[2 * node]);
    const Box b = ld_box(boxes, child[2 * node + 1]);
    for (int c = 0; c < 3; ++c) {
      a.lo[c] = fmin(a.lo[c], b.lo[c]);
      a.hi[c] = fmax(a.hi[c], b.hi[c]);
    }
    st_box(boxes, node, a);
    node = node == 0 ? kNone : parent[node];
  }
}

// Escape link of every node: the right sibling of the nearest ancestor-or-self that is a
// left child, -1 at the end of the walk.
__global__ void __launch_bounds__(kCB)
    escape_kernel(uint32_t n, const uint32_t* __restrict__ child, const uint32_t* __restrict__ parent,
                  int32_t* __restrict__ escape, int* bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n - 1) return;
  int32_t next = -1;
  uint32_t node = i;
  for (uint32_t steps = 0; node != 0; ++steps) {
    const uint32_t p = parent[node];
    if (steps >= n || p == kNone || p + 1 >= n) {   // no root within n - 1 levels
      atomicOr(bad, 1);
      break;
    }
    if (child[2 * p] == node) {
      next = static_cast<int32_t>(child[2 * p + 1]);
      break;
    }
    node = p;
  }
  escape[i] = next;
}

struct Tree {
  const double* boxes;
  const uint32_t* child;
  const int32_t* escape;
  const uint32_t* prim;
  uint32_t n;
};

__device__ __forceinline__ bool same_mesh(const uint32_t* mesh, uint32_t a, uint32_t b, uint32_t c,
                                          uint32_t d) {
  if (!mesh) return false;
  const uint32_t m = mesh[a];
  return m != 0u && m == mesh[b] && m == mesh[c] && m == mesh[d];
}

__device__ __forceinline__ void emit_candidate(int32_t a, int32_t b, int32_t* out, uint32_t cap,
                                               uint32_t* count) {
  const uint32_t slot = wave_reserve(count, 1u);
  if (slot < cap) {
    out[2 * static_cast<size_t>(slot)] = a;
    out[2 * static_cast<size_t>(slot) + 1] = b;
  }
}

__global__ void __launch_bounds__(kCB)
    face_query_kernel(const double* __restrict__ x, const double* __restrict__ d, double alpha,
                      const uint32_t* __restrict__ surface, uint32_t ns,
                      const uint32_t* __restrict__ faces, const uint32_t* __restrict__ mesh,
                      Tree T, double gap, int32_t* __restrict__ out, uint32_t cap,
                      uint32_t* count, int* bad) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const uint32_t p = surface[k];
  Box q = empty_box();
  grow_vertex(q, x, d, alpha, p);
  int32_t node = 0;
  // a walk visits every node at most once: more steps mean a malformed tree
  for (uint32_t steps = 0; node != -1; ++steps) {
    if (steps >= 2 * T.n) {
      atomicOr(bad, 1);
      return;
    }
    if (apart(q, ld_box(T.boxes, static_cast<uint32_t>(node)), gap)) {
      node = T.escape[node];
      continue;
    }
    if (static_cast<uint32_t>(node) + 1 < T.n) {        // internal
      node = static_cast<int32_t>(T.child[2 * node]);
      continue;
    }
    const uint32_t f = T.prim[node - (T.n - 1)];
    const uint32_t a = faces[3 * static_cast<size_t>(f)], b = faces[3 * static_cast<size_t>(f) + 1],
                   c = faces[3 * static_cast<size_t>(f) + 2];
    if (p != a && p != b && p != c && !same_mesh(mesh, p, a, b, c))
      emit_candidate(-static_cast<int32_t>(p) - 1, static_cast<int32_t>(f), out, cap, count);
    node = T.escape[node];
  }
}

__global__ void __launch_bounds__(kCB)
    edge_query_kernel(const uint32_t* __restrict__ edges, const uint32_t* __restrict__ mesh, Tree T,
                      double gap, int32_t* __restrict__ out, uint32_t cap, uint32_t* count,
                      int* bad) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T.n) return;
  const uint32_t self = T.prim[k];
  const uint32_t s0 = edges[2 * static_cast<size_t>(self)], s1 = edges[2 * static_cast<size_t>(self) + 1];
  const Box q = ld_box(T.boxes, T.n - 1 + k);
  int32_t node = 0;
  for (uint32_t steps = 0; node != -1; ++steps) {
    if (steps >= 2 * T.n) {
      atomicOr(bad, 1);
      return;
    }
    if (apart(q, ld_box(T.boxes, static_cast<uint32_t>(node)), gap)) {
      node = T.escape[node];
      continue;
    }
    if (static_cast<uint32_t>(node) + 1 < T.n) {
      node = static_cast<int32_t>(T.child[2 * node]);
      continue;
    }
    const uint32_t other = T.prim[node - (T.n - 1)];
    const uint32_t o0 = edges[2 * static_cast<size_t>(other)],
                   o1 = edges[2 * static_cast<size_t>(other) + 1];
    if (other > self && s0 != o0 && s0 != o1 && s1 != o0 && s1 != o1 &&
        !same_mesh(mesh, s0, s1, o0, o1))
      emit_candidate(static_cast<int32_t>(self), static_cast<int32_t>(other), out, cap, count);
    node = T.escape[node];
  }
}

// ---------------------------------------------------------------- narrow phase
// Case 0 pp (2 ids), 1 pe (3), 2 pt (4), 3 ee (4); -1 when not active.
__device__ __forceinline__ int classify(const double* x, const uint32_t* faces,
                                        const uint32_t* edges, int32_t ca, int32_t cb, double dhat,
                                        uint32_t (&ids)[4]) {
  double dist;
  int kase;
  if (ca < 0) {
    const uint32_t p = static_cast<uint32_t>(-ca - 1);
    const uint32_t a = faces[3 * static_cast<size_t>(cb)], b = faces[3 * static_cast<size_t>(cb) + 1],
                   c = faces[3 * static_cast<size_t>(cb) + 2];
    const V3 vp = ld3(x, p), va = ld3(x, a), vb = ld3(x, b), vc = ld3(x, c);
    const int t = pt_type(vp, va, vb, vc);
    ids[0] = p;
    if (t <= 2) {
      const uint32_t q = t == 0 ? a : (t == 1 ? b : c);
      ids[1] = q;
      dist = d_pp(vp, ld3(x, q));
      kase = 0;
    } else if (t <= 5) {
      const uint32_t e0 = t == 3 ? a : (t == 4 ? b : c), e1 = t == 3 ? b : (t == 4 ? c : a);
      ids[1] = e0;
      ids[2] = e1;
      dist = d_pe(vp, ld3(x, e0), ld3(x, e1));
      kase = 1;
    } else {
      ids[1] = a;
      ids[2] = b;
      ids[3] = c;
      dist = d_pt(vp, va, vb, vc);
      kase = 2;
    }
  } else {
    const uint32_t i0 = edges[2 * static_cast<size_t>(ca)], i1 = edges[2 * static_cast<size_t>(ca) + 1],
                   j0 = edges[2 * static_cast<size_t>(cb)], j1 = edges[2 * static_cast<size_t>(cb) + 1];
    const V3 a0 = ld3(x, i0), a1 = ld3(x, i1), b0 = ld3(x, j0), b1 = ld3(x, j1);
    switch (ee_type(a0, a1, b0, b1)) {
      case 0: ids[0] = i0; ids[1] = j0; dist = d_pp(a0, b0); kase = 0; break;
      case 1: ids[0] = i0; ids[1] = j1; dist = d_pp(a0, b1); kase = 0; break;
      case 2: ids[0] = i0; ids[1] = j0; ids[2] = j1; dist = d_pe(a0, b0, b1); kase = 1; break;
      case 3: ids[0] = i1; ids[1] = j0; dist = d_pp(a1, b0); kase = 0; break;
      case 4: ids[0] = i1; ids[1] = j1; dist = d_pp(a1, b1); kase = 0; break;
      case 5: ids[0] = i1; ids[1] = j0; ids[2] = j1; dist = d_pe(a1, b0, b1); kase = 1; break;
      case 6: ids[0] = j0; ids[1] = i0; ids[2] = i1; dist = d_pe(b0, a0, a1); kase = 1; break;
      case 7: ids[0] = j1; ids[1] = i0; ids[2] = i1; dist = d_pe(b1, a0, a1); kase = 1; break;
      default:
        ids[0] = i0; ids[1] = i1; ids[2] = j0; ids[3] = j1;
        dist = d_ee(a0, a1, b0, b1);
        kase = 3;
    }
  }
  return dist < dhat ? kase : -1;
}

__global__ void __launch_bounds__(kCB)
    active_pairs_kernel(const double* __restrict__ x, const uint32_t* __restrict__ faces,
                        const uint32_t* __restrict__ edges, const int32_t* __restrict__ cand,
                        uint32_t nc, double dhat, uint32_t* __restrict__ pp,
                        uint32_t* __restrict__ pe, uint32_t* __restrict__ pt,
                        uint32_t* __restrict__ ee, uint32_t cap, uint32_t* counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t ids[4] = {0, 0, 0, 0};
  int kase = -1;
  if (i < nc) kase = classify(x, faces, edges, cand[2 * static_cast<size_t>(i)],
                              cand[2 * static_cast<size_t>(i) + 1], dhat, ids);
  uint32_t* outs[4] = {pp, pe, pt, ee};
  const uint32_t width[4] = {2, 3, 4, 4};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t slot = wave_reserve(counts + c, kase == c ? 1u : 0u);
    if (kase == c && slot < cap)
      for (uint32_t k = 0; k < width[c]; ++k) outs[c][width[c] * static_cast<size_t>(slot) + k] = ids[k];
  }
}

// ------------------------------------------------------------------ additive CCD
// Time bound of one candidate moving by ds (ACCD.cu:258-420 with thickness 0): the
// displacements centred by their mean, conservative advancement by
// (1 - eta) d^2 / (d l_max) until the distance falls below eta d0 or the time passes 1. A
// pair at zero distance gets 0; the advancement stops (at the time reached) if the
// distance reaches zero or after kMaxAdvances steps.
__device__ double accd(V3 (&xs)[4], V3 (&ds)[4], double eta, bool edge_edge) {
  const V3 mov = scale(add(add(add(ds[0], ds[1]), ds[2]), ds[3]), -0.25);
  for (int k = 0; k < 4; ++k) ds[k] = add(ds[k], mov);
  double lmax;
  if (edge_edge)
    lmax = sqrt(fmax(dot(ds[0], ds[0]), dot(ds[1], ds[1]))) +
           sqrt(fmax(dot(ds[2], ds[2]), dot(ds[3], ds[3])));
  else
    lmax = sqrt(dot(ds[0], ds[0])) +
           sqrt(fmax(fmax(dot(ds[1], ds[1]), dot(ds[2], ds[2])), dot(ds[3], ds[3])));
  if (lmax == 0.0) return 1.0;
  auto dist2 = [&]() {
    if (!edge_edge) return pt_distance(xs[0], xs[1], xs[2], xs[3]);
    const double r = ee_distance(xs[0], xs[1], xs[2], xs[3]);
    if (r > 0.0) return r;
    return fmin(fmin(d_pp(xs[0], xs[2]), d_pp(xs[0], xs[3])), fmin(d_pp(xs[1], xs[2]), d_pp(xs[1], xs[3])));
  };
  double d2 = dist2();
  if (!(d2 > 0.0)) return 0.0;
  double d = sqrt(d2);
  const double gap = eta * d2 / d;
  double toc = 0.0;
  for (int it = 0; it < kMaxAdvances; ++it) {
    const double lb = (1.0 - eta) * d2 / (d * lmax);
    for (int k = 0; k < 4; ++k) xs[k] = add(xs[k], scale(ds[k], lb));
    d2 = dist2();
    if (!(d2 > 0.0)) break;
    d = sqrt(d2);
    if (toc != 0.0 && d2 / d < gap) break;
    toc += lb;
    if (toc > 1.0) return 1.0;
  }
  return toc;
}

__global__ void __launch_bounds__(kCB)
    step_bound_kernel(const double* __restrict__ x, const double* __restrict__ dir,
                      const uint32_t* __restrict__ faces, const uint32_t* __restrict__ edges,
                      const int32_t* __restrict__ cand, uint32_t nc, double eta,
                      double* __restrict__ times, double* __restrict__ partial) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double t = 1.0;
  if (i < nc) {
    const int32_t a = cand[2 * static_cast<size_t>(i)], b = cand[2 * static_cast<size_t>(i) + 1];
    uint32_t v[4];
    bool ee = a >= 0;
    if (!ee) {
      v[0] = static_cast<uint32_t>(-a - 1);
      v[1] = faces[3 * static_cast<size_t>(b)];
      v[2] = faces[3 * static_cast<size_t>(b) + 1];
      v[3] = faces[3 * static_cast<size_t>(b) + 2];
    } else {
      v[0] = edges[2 * static_cast<size_t>(a)];
      v[1] = edges[2 * static_cast<size_t>(a) + 1];
      v[2] = edges[2 * static_cast<size_t>(b)];
      v[3] = edges[2 * static_cast<size_t>(b) + 1];
    }
    V3 xs[4], ds[4];
    for (int k = 0; k < 4; ++k) {
      xs[k] = ld3(x, v[k]);
      const V3 dk = ld3(dir, v[k]);
      ds[k] = {-dk.x, -dk.y, -dk.z};           // the step moves x to x - t d
    }
    t = accd(xs, ds, eta, ee);
    if (times) times[i] = t;
  }
  __shared__ double s[kCB];
  s[threadIdx.x] = t;
  __syncthreads();
  for (unsigned h = kCB / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kCB)
    min_reduce_kernel(const double* __restrict__ in, uint32_t n, double* __restrict__ out) {
  __shared__ double s[kCB];
  double t = 1.0;
  for (uint32_t i = threadIdx.x; i < n; i += kCB) t = fmin(t, in[i]);
  s[threadIdx.x] = t;
  __syncthreads();
  for (unsigned h = kCB / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// ---------------------------------------------------------------- tree building
// Scratch slots: 80 + 16 * tree + k.
struct TreeBuffers {
  double* leaf_boxes;
  double* boxes;
  uint32_t* child;
  uint32_t* parent;
  int32_t* escape;
  uint32_t* prim;
  uint32_t* flags;
  uint64_t* keys[2];
  uint32_t* vals[2];
  double* partial;
};

int tree_buffers(int tree, uint32_t n, TreeBuffers& b) {
  const int s = 80 + 16 * tree;
  const size_t nn = 2 * static_cast<size_t>(n) - 1;
  b.leaf_boxes = static_cast<double*>(scratch(s, 6 * sizeof(double) * n));
  b.boxes = static_cast<double*>(scratch(s + 1, 6 * sizeof(double) * nn));
  b.child = static_cast<uint32_t*>(scratch(s + 2, 2 * sizeof(uint32_t) * std::max<size_t>(1, n - 1)));
  b.parent = static_cast<uint32_t*>(scratch(s + 3, sizeof(uint32_t) * nn));
  b.escape = static_cast<int32_t*>(scratch(s + 4, sizeof(int32_t) * nn));
  b.prim = static_cast<uint32_t*>(scratch(s + 5, sizeof(uint32_t) * n));
  b.flags = static_cast<uint32_t*>(scratch(s + 6, sizeof(uint32_t) * std::max<size_t>(1, n - 1)));
  b.keys[0] = static_cast<uint64_t*>(scratch(s + 7, sizeof(uint64_t) * n));
  b.keys[1] = static_cast<uint64_t*>(scratch(s + 8, sizeof(uint64_t) * n));
  b.vals[0] = static_cast<uint32_t*>(scratch(s + 9, sizeof(uint32_t) * n));
  b.vals[1] = static_cast<uint32_t*>(scratch(s + 10, sizeof(uint32_t) * n));
  b.partial = static_cast<double*>(scratch(s + 11, 6 * sizeof(double) * (kCB + 1)));
  if (!b.leaf_boxes || !b.boxes || !b.child || !b.parent || !b.escape || !b.prim || !b.flags ||
      !b.keys[0] || !b.keys[1] || !b.vals[0] || !b.vals[1] || !b.partial) {
    ysp_set_error(__FILE__, __LINE__, "contact BVH: scratch allocation failed");
    return YSP_ERR_HIP;
  }
  return YSP_OK;
}

// A linear BVH over n primitives of `arity` vertices (n >= 1).
int build_tree(int tree, const double* x, const double* d, double alpha, const uint32_t* prims,
               uint32_t arity, uint32_t n, Tree& T, int* bad, hipStream_t s) {
  TreeBuffers b;
  if (int rc = tree_buffers(tree, n, b)) return rc;
  const dim3 g(grid_for(n, kCB)), bl(kCB);
  hipLaunchKernelGGL(leaf_boxes_kernel, g, bl, 0, s, x, d, alpha, prims, arity, n, b.leaf_boxes);
  const unsigned nb = std::min<unsigned>(grid_for(n, kCB), kCB);
  hipLaunchKernelGGL(box_reduce_kernel, dim3(nb), bl, 0, s, b.leaf_boxes, n, b.partial + 6);
  hipLaunchKernelGGL(box_reduce_kernel, dim3(1), bl, 0, s, b.partial + 6, nb, b.partial);
  hipLaunchKernelGGL(morton_kernel, g, bl, 0, s, b.leaf_boxes, n, b.partial, b.keys[0], b.vals[0]);
  size_t temp = 0;
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, b.keys[0], b.keys[1], b.vals[0],
                                               b.vals[1], static_cast<int>(n), 0, 63, s));
  void* tmp = scratch(80 + 16 * tree + 12, temp);
  if (!tmp) {
    ysp_set_error(__FILE__, __LINE__, "contact BVH: scratch allocation failed");
    return YSP_ERR_HIP;
  }
  YSP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, temp, b.keys[0], b.keys[1], b.vals[0], b.vals[1],
                                               static_cast<int>(n), 0, 63, s));
  hipLaunchKernelGGL(leaf_setup_kernel, g, bl, 0, s, b.vals[1], n, b.leaf_boxes, b.boxes, b.prim,
                     b.flags, b.parent);
  if (n > 1) {
    hipLaunchKernelGGL(radix_tree_kernel, dim3(grid_for(n - 1, kCB)), bl, 0, s, b.keys[1], n,
                       b.child, b.parent);
    hipLaunchKernelGGL(refit_kernel, g, bl, 0, s, n, b.child, b.parent, b.flags, b.boxes, bad);
  }
  hipLaunchKernelGGL(escape_kernel, dim3(grid_for(2 * static_cast<size_t>(n) - 1, kCB)), bl, 0, s,
                     n, b.child, b.parent, b.escape, bad);
  YSP_CHECK_LAUNCH();
  T = Tree{b.boxes, b.child, b.escape, b.prim, n};
  return YSP_OK;
}

}  // namespace
}  // namespace ysp

using namespace ysp;

extern "C" int ysp_ccd_candidates(const double* x, const double* d, double alpha, uint32_t nv,
                                  const uint32_t* faces, uint32_t nf, const uint32_t* surface,
                                  uint32_t ns, const uint32_t* edges, uint32_t ne,
                                  const uint32_t* mesh, double dhat, int32_t* candidates,
                                  uint32_t capacity, uint32_t* count, void* stream) {
  if (!count || !(dhat >= 0.0) || !std::isfinite(dhat) || (d && !(alpha >= 0.0)) ||
      (nv && !x) || (nf && (!faces || !surface)) || (ne && !edges) ||
      (capacity && !candidates)) {
    ysp_set_error(__FILE__, __LINE__, "ysp_ccd_candidates: bad arguments");
    return YSP_ERR_ARG;
  }
  *count = 0;
  hipStream_t s = as_stream(stream);
  // [0] candidate count, [1] malformed-tree flag
  uint32_t* counter = static_cast<uint32_t*>(scratch(79, 2 * sizeof(uint32_t)));
  if (!counter) return YSP_ERR_HIP;
  int* bad = reinterpret_cast<int*>(counter + 1);
  YSP_CHECK(hipMemsetAsync(counter, 0, 2 * sizeof(uint32_t), s));
  const double gap = sqrt(dhat);
  if (nf && ns) {
    Tree T;
    if (int rc = build_tree(0, x, d, alpha, faces, 3, nf, T, bad, s)) return rc;
    hipLaunchKernelGGL(face_query_kernel, dim3(grid_for(ns, kCB)), dim3(kCB), 0, s, x, d, alpha,
                       surface, ns, faces, mesh, T, gap, candidates, capacity, counter, bad);
  }
  if (ne > 1) {
    Tree T;
    if (int rc = build_tree(1, x, d, alpha, edges, 2, ne, T, bad, s)) return rc;
    hipLaunchKernelGGL(edge_query_kernel, dim3(grid_for(ne, kCB)), dim3(kCB), 0, s, edges, mesh, T,
                       gap, candidates, capacity, counter, bad);
  }
  YSP_CHECK_LAUNCH();
  uint32_t host[2] = {0, 0};
  YSP_CHECK(hipMemcpyAsync(host, counter, sizeof(host), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  if (host[1]) {
    ysp_set_error(__FILE__, __LINE__, "ysp_ccd_candidates: malformed BVH (a walk exceeded its bound)");
    return YSP_ERR_HIP;
  }
  *count = host[0];
  return YSP_OK;
}

extern "C" int ysp_ccd_active_pairs(const double* x, const uint32_t* faces, const uint32_t* edges,
                                    const int32_t* candidates, uint32_t num_candidates,
                                    double dhat, uint32_t* pp, uint32_t* pe, uint32_t* pt,
                                    uint32_t* ee, uint32_t capacity, uint32_t* counts,
                                    void* stream) {
  if (!counts || (num_candidates && (!x || !candidates)) || (capacity && (!pp || !pe || !pt || !ee))) {
    ysp_set_error(__FILE__, __LINE__, "ysp_ccd_active_pairs: bad arguments");
    return YSP_ERR_ARG;
  }
  for (int c = 0; c < 4; ++c) counts[c] = 0;
  if (num_candidates == 0) return YSP_OK;
  hipStream_t s = as_stream(stream);
  uint32_t* dc = static_cast<uint32_t*>(scratch(78, 4 * sizeof(uint32_t)));
  if (!dc) return YSP_ERR_HIP;
  YSP_CHECK(hipMemsetAsync(dc, 0, 4 * sizeof(uint32_t), s));
  hipLaunchKernelGGL(active_pairs_kernel, dim3(grid_for(num_candidates, kCB)), dim3(kCB), 0, s, x,
                     faces, edges, candidates, num_candidates, dhat, pp, pe, pt, ee, capacity, dc);
  YSP_CHECK_LAUNCH();
  YSP_CHECK(hipMemcpyAsync(counts, dc, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  return YSP_OK;
}

extern "C" int ysp_ccd_step_bound(const double* x, const double* d, const uint32_t* faces,
                                  const uint32_t* edges, const int32_t* candidates,
                                  uint32_t num_candidates, double slackness, double* times,
                                  double* step, void* stream) {
  if (!step || !(slackness >= 0.0 && slackness < 1.0) ||
      (num_candidates && (!x || !d || !candidates))) {
    ysp_set_error(__FILE__, __LINE__, "ysp_ccd_step_bound: bad arguments");
    return YSP_ERR_ARG;
  }
  *step = 1.0;
  if (num_candidates == 0) return YSP_OK;
  hipStream_t s = as_stream(stream);
  const unsigned nb = grid_for(num_candidates, kCB);
  double* partial = static_cast<double*>(scratch(77, sizeof(double) * (nb + 1)));
  if (!partial) return YSP_ERR_HIP;
  hipLaunchKernelGGL(step_bound_kernel, dim3(nb), dim3(kCB), 0, s, x, d, faces, edges, candidates,
                     num_candidates, 1.0 - slackness, times, partial + 1);
  hipLaunchKernelGGL(min_reduce_kernel, dim3(1), dim3(kCB), 0, s, partial + 1, nb, partial);
  YSP_CHECK_LAUNCH();
  YSP_CHECK(hipMemcpyAsync(step, partial, sizeof(double), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  return YSP_OK;
}
