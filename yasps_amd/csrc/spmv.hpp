// Internal interface of the block SpMV kernels (shared with the PCG driver).
#pragma once
#include "common.hpp"

namespace ysp {

// Partial-sum sets: kQuadSlots doubles kQuadSpace apart (one 128-byte line each); a
// workgroup adds its partial sum into slot blockIdx mod kQuadSlots, so thousands of
// workgroups spread their FP64 atomics over 64 addresses instead of serialising on one.
constexpr int kQuadSlots = 64;
constexpr int kQuadSpace = 16;

// y += H x for one symmetric upper-storage part; if quad != nullptr the kernel also
// adds x^T H x (this part's contribution) into the partial-sum set at quad. A non-null
// stop flag that reads non-zero on the device turns the launch into a no-op.
int launch_part_spmv(const ysp_block_matrix* part, const double* x, double* y, int transpose,
                     double* quad, const double* stop, hipStream_t s);

}  // namespace ysp
