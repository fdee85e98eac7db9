// Small dense FP64 linear algebra held in registers: the packed-Schur PSD test and
// cyclic Jacobi symmetric eigensolvers (Golub & Van Loan, "Matrix Computations",
// section 8.5: symmetric Schur 2x2 rotations swept row by row).
#pragma once
#include "common.hpp"

namespace ysp {

// Packed-Schur PSD test (hessianKernelHeader.pyx:79-141) on a packed lower triangle
// L[r*(r+1)/2 + c] (c <= r) that it overwrites.
template <int N>
__device__ __forceinline__ bool psd_test_packed(double* L) {
  double scale = 1.0;
#pragma unroll
  for (int r = 0; r < N; ++r) scale = fmax(scale, fabs(L[r * (r + 1) / 2 + r]));
  const double tol = scale * 1.0e-10;
  bool ok = true;
#pragma unroll
  for (int c = 0; c < N; ++c) {
    const double pivot = L[c * (c + 1) / 2 + c];
    if (pivot < -tol) ok = false;
    const bool small = fabs(pivot) <= tol;
    if (small) {
#pragma unroll
      for (int r = c + 1; r < N; ++r)
        if (fabs(L[r * (r + 1) / 2 + c]) > tol) ok = false;
    } else {
      const double inv = 1.0 / pivot;
#pragma unroll
      for (int r = c + 1; r < N; ++r) {
        const double factor = L[r * (r + 1) / 2 + c] * inv;
#pragma unroll
        for (int t = c + 1; t <= r; ++t)
          L[r * (r + 1) / 2 + t] = fma(-factor, L[t * (t + 1) / 2 + c], L[r * (r + 1) / 2 + t]);
      }
    }
  }
  return ok;
}

// Cyclic Jacobi eigen-decomposition of a symmetric N x N matrix held in registers.
// On return A is diagonal (eigenvalues) and V holds the eigenvectors as columns.
template <int N>
__device__ __forceinline__ void jacobi_eigen(double (&A)[N][N], double (&V)[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
  double fro = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) fro += A[i][j] * A[i][j];
  const double stop = fro * 1.0e-32;
  for (int sweep = 0; sweep < 24; ++sweep) {
    double off = 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += A[p][q] * A[p][q];
    if (off <= stop) break;
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const double apq = A[p][q];
        if (apq == 0.0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
        const double c = 1.0 / sqrt(1.0 + t * t);
        const double s = t * c;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
}

// Index of (p, q), p <= q, in the packed upper triangle of an N x N symmetric matrix.
template <int N>
__host__ __device__ constexpr int upper_index(int p, int q) {
  return p * N - p * (p - 1) / 2 + (q - p);
}

// 1/x to full double precision for finite normal x: the hardware reciprocal estimate and
// two Newton steps (each at least doubles the correct bits).
__device__ __forceinline__ double recip_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}

// 1/sqrt(x) for x in [1, 2]: the hardware estimate and two Newton steps.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = y * fma(-hx * y, y, 1.5);
  return y * fma(-hx * y, y, 1.5);
}

// Round-robin (circle method) ordering of the N(N-1)/2 index pairs: with M = N rounded up
// to even, index M-1 stays fixed and round r pairs r with M-1 and (r+i) mod (M-1) with
// (r-i) mod (M-1), i = 1..M/2-1. The pairs of a round are disjoint (a pair with index N,
// the stand-in for odd N, is skipped).
template <int N>
struct RoundRobin {
  static constexpr int M = N + (N & 1);
  static constexpr int kRounds = M - 1, kPerRound = M / 2;
  __host__ __device__ static constexpr int a(int r, int i) { return i == 0 ? r : (r + i) % (M - 1); }
  __host__ __device__ static constexpr int b(int r, int i) {
    return i == 0 ? M - 1 : (r - i + M - 1) % (M - 1);
  }
  __host__ __device__ static constexpr int p(int r, int i) { return a(r, i) < b(r, i) ? a(r, i) : b(r, i); }
  __host__ __device__ static constexpr int q(int r, int i) { return a(r, i) < b(r, i) ? b(r, i) : a(r, i); }
};

// Jacobi on the packed upper triangle a (N(N+1)/2 entries) of a symmetric matrix, with the
// symmetric two-sided update of each rotation (Golub & Van Loan, section 8.5): only the
// 2(N-2) entries of rows/columns p and q and the three entries of the 2x2 pivot block
// change. The rotation t = tan(phi) = sgn(d) y / (|d| + sqrt(d^2 + y^2)) with
// d = a_qq - a_pp, y = 2 a_pq (the smaller root of t^2 + 2 (d/y) t - 1 = 0), written with
// one square root and one reciprocal; c = 1/sqrt(1 + t^2), s = t c; a_pq = 0 gives the
// identity. Sweeps follow the round-robin ordering: the rotations of a round touch
// disjoint pivot blocks, so their parameters are computed together from the state at the
// start of the round (the same values the one-at-a-time order computes) and the
// rotations are then applied in turn. Branch-free rounds let the scheduler overlap the
// independent parameter chains. On return the diagonal holds the eigenvalues and V's
// columns the eigenvectors.
template <int N>
__device__ __forceinline__ void jacobi_eigen_packed(double (&a)[N * (N + 1) / 2],
                                                    double (&V)[N][N]) {
  using RR = RoundRobin<N>;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
  double diag = 0.0, off = 0.0;
#pragma unroll
  for (int p = 0; p < N; ++p) {
    diag += a[upper_index<N>(p, p)] * a[upper_index<N>(p, p)];
#pragma unroll
    for (int q = p + 1; q < N; ++q) off += a[upper_index<N>(p, q)] * a[upper_index<N>(p, q)];
  }
  const double stop = (diag + 2.0 * off) * 1.0e-32;
  for (int sweep = 0; sweep < 24; ++sweep) {
    off = 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += a[upper_index<N>(p, q)] * a[upper_index<N>(p, q)];
    if (off <= stop) break;
#pragma unroll
    for (int rd = 0; rd < RR::kRounds; ++rd) {
      double cr[RR::kPerRound], sr[RR::kPerRound], tr[RR::kPerRound];
#pragma unroll
      for (int i = 0; i < RR::kPerRound; ++i) {
        const int p = RR::p(rd, i), q = RR::q(rd, i);
        if (q >= N) continue;
        const double apq = a[upper_index<N>(p, q)];
        const double d = a[upper_index<N>(q, q)] - a[upper_index<N>(p, p)], y = 2.0 * apq;
        const double r = sqrt(fma(d, d, y * y));
        double t = (d >= 0.0 ? y : -y) * recip_nr(fabs(d) + r);
        t = apq == 0.0 ? 0.0 : t;
        const double c = rsqrt_nr(fma(t, t, 1.0));
        cr[i] = c;
        sr[i] = t * c;
        tr[i] = t;
      }
#pragma unroll
      for (int i = 0; i < RR::kPerRound; ++i) {
        const int p = RR::p(rd, i), q = RR::q(rd, i);
        if (q >= N) continue;
        const double c = cr[i], s = sr[i], t = tr[i];
        const double apq = a[upper_index<N>(p, q)];
        a[upper_index<N>(p, p)] = fma(-t, apq, a[upper_index<N>(p, p)]);
        a[upper_index<N>(q, q)] = fma(t, apq, a[upper_index<N>(q, q)]);
        a[upper_index<N>(p, q)] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          if (k == p || k == q) continue;
          const int ip = k < p ? upper_index<N>(k, p) : upper_index<N>(p, k);
          const int iq = k < q ? upper_index<N>(k, q) : upper_index<N>(q, k);
          const double akp = a[ip], akq = a[iq];
          a[ip] = fma(c, akp, -s * akq);
          a[iq] = fma(s, akp, c * akq);
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = fma(c, vkp, -s * vkq);
          V[k][q] = fma(s, vkp, c * vkq);
        }
      }
    }
  }
}

}  // namespace ysp
