// Host entry points of the device Krylov loops (solver.hip).
#pragma once

#include "dfm_internal.hpp"

namespace dfm {

int solve_pcg(Sys& s, int precision, double* conv_ratio, unsigned* iteration);
int solve_pbicgstab(Sys& s, double* conv_ratio, unsigned* iteration);
// z = M^-1 src with the FP64 factor (src may equal z); kernels skip when *done != 0 (done may be null)
int bench_kernel(Sys& s, int which, int reps, double* avg_ms, double* bytes);
hipError_t precond_apply64(const Sys& s, const double* src, double* z, const int* done, hipStream_t st);

}  // namespace dfm
