// Internal interface of the block SpMV kernels (shared with the PCG driver).
#pragma once
#include "common.hpp"

namespace ysp {

// y += H x for one symmetric upper-storage part; if quad != nullptr the kernel also
// adds x^T H x (this part's contribution) into *quad. A non-null stop flag that reads
// non-zero on the device turns the launch into a no-op.
int launch_part_spmv(const ysp_block_matrix* part, const double* x, double* y, int transpose,
                     double* quad, const double* stop, hipStream_t s);

}  // namespace ysp
