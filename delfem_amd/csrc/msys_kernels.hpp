// Launchers of the multi-segment kernels (k_msys.hip). All run on the given stream.
#pragma once

#include <hip/hip_runtime.h>

namespace dfm {

// A coupling matrix on the device: nrow block rows of bi x bj blocks, CRS pattern (columns ascending).
struct MsBlkDev {
  int nrow, bi, bj;
  const unsigned* rowptr;
  const unsigned* colind;
};
// The ILU(0) factor of a diagonal segment on the device (natural numbering): pattern with the diagonal apart,
// first upper entry of each row, off-diagonal and (inverted) diagonal blocks of b x b.
struct MsDiaDev {
  int nrow, b;
  const unsigned* rowptr;
  const unsigned* colind;
  const unsigned* diaind;
  double* crs;
  double* dia;
};

hipError_t launch_ms_blk_matvec(const MsBlkDev& a, double alpha, const double* val, const double* x, double beta,
                                double* y, hipStream_t st);
hipError_t launch_ms_blk_bc(const MsBlkDev& a, double* val, const int* bc_i, const int* bc_j, hipStream_t st);
hipError_t launch_ms_dia_lowup(const MsDiaDev& d, const MsBlkDev& low, const double* lval, const MsBlkDev& up,
                               const double* uval, hipStream_t st);
hipError_t launch_ms_od_lowup(const MsBlkDev& f, double* fval, const MsBlkDev& low, const double* lval,
                              const MsBlkDev& up, const double* uval, hipStream_t st);
hipError_t launch_ms_od_low(const MsBlkDev& f, double* fval, const MsDiaDev& d, hipStream_t st);
// lptr[nlev+1] / lrows: the rows of each dependency level of the diagonal segment's lower factor (device arrays)
hipError_t launch_ms_od_up(int nlev, const int* lptr, const int* lrows, const MsBlkDev& f, double* fval,
                           const MsDiaDev& d, hipStream_t st);

}  // namespace dfm
