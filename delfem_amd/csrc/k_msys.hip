// Kernels of the multi-segment systems (msys.cpp): the coupling matrices of LsSol::CLinearSystem with several
// segments and the coupling half of CPreconditioner_ILU (src/ls/preconditioner.cpp:215-335), ILU(0).
//
// A coupling matrix is a block-CRS matrix of bi x bj blocks (segment i rows, segment j columns). Every kernel keeps
// one thread per block row and does that row's work in the reference's order, so the results are the reference's
// bit for bit (built with -ffp-contract=off): the elimination kernels update only their own row, reading rows that
// are final (other matrices, or, for the upper couplings, rows of earlier dependency levels).
#include <hip/hip_runtime.h>

#include "msys_kernels.hpp"

namespace dfm {
namespace {

constexpr int kT = 256;

// out -= a b, one product at a time (CalcSubMatPr, src/matvec/matfrac_blkcrs.cpp:43-54)
__device__ inline void sub_matpr(double* out, const double* a, const double* b, int ni, int nk, int nj) {
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < nj; j++)
      for (int k = 0; k < nk; k++) out[i * nj + j] -= a[i * nk + k] * b[k * nj + j];
}

// out := d out with each entry summed from 0.0 (CalcMatPr, :56-72); d is ni x ni, out ni x nj (at most 3 x 3)
__device__ inline void matpr(double* out, const double* d, int ni, int nj) {
  double tmp[9];
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < nj; j++) {
      tmp[i * nj + j] = 0.0;
      for (int k = 0; k < ni; k++) tmp[i * nj + j] += d[i * ni + k] * out[k * nj + j];
    }
  for (int i = 0; i < ni * nj; i++) out[i] = tmp[i];
}

// The entry of column c in the sorted row [b, e) of colind, or -1.
__device__ inline long long find_col(const unsigned* colind, unsigned b, unsigned e, unsigned c) {
  while (b < e) {
    const unsigned m = (b + e) >> 1;
    const unsigned v = colind[m];
    if (v == c) return m;
    if (v < c) b = m + 1;
    else e = m;
  }
  return -1;
}

// y_i = beta y_i + alpha sum_j A_ij x_j in the shape branches of CMat_BlkCrs::MatVec (src/matvec/mat_blkcrs.cpp:
// 955-1175): 1 x 1 with beta == 0 starts from 0.0; 1 x 2 and 1 x 3 add alpha (row of products); 1 x n adds one
// alpha a x at a time; n x 1 multiplies by x alpha; otherwise one alpha a x term at a time, row-major.
__global__ void k_ms_blk_matvec(int nrow, int bi, int bj, const unsigned* __restrict__ rowptr,
                                const unsigned* __restrict__ colind, const double* __restrict__ val, double alpha,
                                const double* __restrict__ x, double beta, double* __restrict__ y) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= nrow) return;
  double t[3];
  for (int d = 0; d < bi; d++) t[d] = y[(size_t)i * bi + d];
  if (bi == 1 && bj == 1 && beta == 0.0) {
    t[0] = 0.0;
  } else {
    for (int d = 0; d < bi; d++) t[d] *= beta;
  }
  const int bs = bi * bj;
  for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
    const double* a = val + (size_t)q * bs;
    const double* jx = x + (size_t)colind[q] * bj;
    if (bi == 1 && bj == 1) {
      t[0] += alpha * a[0] * jx[0];
    } else if (bi == 1 && bj == 2) {
      t[0] += alpha * (a[0] * jx[0] + a[1] * jx[1]);
    } else if (bi == 1 && bj == 3) {
      t[0] += alpha * (a[0] * jx[0] + a[1] * jx[1] + a[2] * jx[2]);
    } else if (bj == 1) {
      const double ax = jx[0] * alpha;
      for (int r = 0; r < bi; r++) t[r] += a[r] * ax;
    } else {
      for (int r = 0; r < bi; r++)
        for (int c = 0; c < bj; c++) t[r] += alpha * a[r * bj + c] * jx[c];
    }
  }
  for (int d = 0; d < bi; d++) y[(size_t)i * bi + d] = t[d];
}

// SetBoundaryCondition_Row by the row segment's flags, then _Colum by the column segment's (mat_blkcrs.cpp:829-874,
// 910-953): zero the fixed rows and the fixed columns of a coupling matrix.
__global__ void k_ms_blk_bc(int nrow, int bi, int bj, const unsigned* __restrict__ rowptr,
                            const unsigned* __restrict__ colind, double* __restrict__ val, const int* __restrict__ bc_i,
                            const int* __restrict__ bc_j) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= nrow) return;
  const int bs = bi * bj;
  for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
    double* a = val + (size_t)q * bs;
    const unsigned c = colind[q];
    for (int r = 0; r < bi; r++)
      for (int k = 0; k < bj; k++)
        if (bc_i[(size_t)i * bi + r] != 0 || bc_j[(size_t)c * bj + k] != 0) a[r * bj + k] = 0.0;
  }
}

// CMatDiaFrac_BlkCrs::DoILUDecompLowUp (src/matvec/matdiafrac_blkcrs.cpp:671-772): the diagonal segment's factor
// row i (pattern frowptr/fcolind, sorted, diagonal apart) less L_ik U_kj over the couplings low (i x k) and
// up (k x i), the row's low entries ascending; entries outside the factor pattern dropped.
__global__ void k_ms_dia_lowup(int nrow, int bi, int bk, const unsigned* __restrict__ frowptr,
                               const unsigned* __restrict__ fcolind, double* __restrict__ fcrs,
                               double* __restrict__ fdia, const unsigned* __restrict__ lrow,
                               const unsigned* __restrict__ lcol, const double* __restrict__ lval,
                               const unsigned* __restrict__ urow, const unsigned* __restrict__ ucol,
                               const double* __restrict__ uval) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= nrow) return;
  const int bii = bi * bi, bik = bi * bk;
  const unsigned fb = frowptr[i], fe = frowptr[i + 1];
  for (unsigned ik = lrow[i]; ik < lrow[i + 1]; ik++) {
    const unsigned k = lcol[ik];
    const double* pik = lval + (size_t)ik * bik;
    for (unsigned kj = urow[k]; kj < urow[k + 1]; kj++) {
      const unsigned j = ucol[kj];
      const double* pkj = uval + (size_t)kj * bik;
      if (j != (unsigned)i) {
        const long long ij = find_col(fcolind, fb, fe, j);
        if (ij < 0) continue;
        sub_matpr(fcrs + (size_t)ij * bii, pik, pkj, bi, bk, bi);
      } else {
        sub_matpr(fdia + (size_t)i * bii, pik, pkj, bi, bk, bi);
      }
    }
  }
}

// CMatFrac_BlkCrs::DoILUDecompLowUp (src/matvec/matfrac_blkcrs.cpp:540-599): F_ij -= L_ik U_kj
__global__ void k_ms_od_lowup(int nrow, int bi, int bk, int bj, const unsigned* __restrict__ frow,
                              const unsigned* __restrict__ fcol, double* __restrict__ fval,
                              const unsigned* __restrict__ lrow, const unsigned* __restrict__ lcol,
                              const double* __restrict__ lval, const unsigned* __restrict__ urow,
                              const unsigned* __restrict__ ucol, const double* __restrict__ uval) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= nrow) return;
  const unsigned fb = frow[i], fe = frow[i + 1];
  for (unsigned ik = lrow[i]; ik < lrow[i + 1]; ik++) {
    const unsigned k = lcol[ik];
    const double* pik = lval + (size_t)ik * bi * bk;
    for (unsigned kj = urow[k]; kj < urow[k + 1]; kj++) {
      const long long ij = find_col(fcol, fb, fe, ucol[kj]);
      if (ij < 0) continue;
      sub_matpr(fval + (size_t)ij * bi * bj, pik, uval + (size_t)kj * bk * bj, bi, bk, bj);
    }
  }
}

// CMatFrac_BlkCrs::DoILUDecompLow (:703-783): the lower coupling F (i x j) against the upper part of the factored
// diagonal segment j: F_ij -= F_ik U_kj, the row's entries ascending (each F_ik final when it is used).
__global__ void k_ms_od_low(int nrow, int bi, int bj, const unsigned* __restrict__ frow,
                            const unsigned* __restrict__ fcol, double* __restrict__ fval,
                            const unsigned* __restrict__ drow, const unsigned* __restrict__ dcol,
                            const unsigned* __restrict__ ddia, const double* __restrict__ dcrs) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= nrow) return;
  const unsigned fb = frow[i], fe = frow[i + 1];
  const int bs = bi * bj, bjj = bj * bj;
  for (unsigned ik = fb; ik < fe; ik++) {
    const unsigned k = fcol[ik];
    const double* pik = fval + (size_t)ik * bs;
    for (unsigned kj = ddia[k]; kj < drow[k + 1]; kj++) {
      const long long ij = find_col(fcol, fb, fe, dcol[kj]);
      if (ij < 0) continue;
      sub_matpr(fval + (size_t)ij * bs, pik, dcrs + (size_t)kj * bjj, bi, bj, bj);
    }
  }
}

// CMatFrac_BlkCrs::DoILUDecompUp (:601-701): the upper coupling F (i x j) eliminated with the lower part of the
// factored diagonal segment i, then each row scaled by its inverted diagonal block. Row i reads the scaled rows
// k < i of F, so rows go by dependency level: one workgroup walks the levels (lptr/lrows), a barrier between them.
__global__ void __launch_bounds__(1024) k_ms_od_up(int nlev, const int* __restrict__ lptr,
                                                   const int* __restrict__ lrows, int bi, int bj,
                                                   const unsigned* __restrict__ frow, const unsigned* __restrict__ fcol,
                                                   double* fval, const unsigned* __restrict__ drow,
                                                   const unsigned* __restrict__ dcol,
                                                   const unsigned* __restrict__ ddia,
                                                   const double* __restrict__ dcrs, const double* __restrict__ ddiav) {
  const int bs = bi * bj, bii = bi * bi;
  for (int l = 0; l < nlev; l++) {
    for (int p = lptr[l] + (int)threadIdx.x; p < lptr[l + 1]; p += (int)blockDim.x) {
      const int i = lrows[p];
      const unsigned fb = frow[i], fe = frow[i + 1];
      for (unsigned ik = drow[i]; ik < ddia[i]; ik++) {
        const unsigned k = dcol[ik];
        const double* pik = dcrs + (size_t)ik * bii;
        for (unsigned kj = frow[k]; kj < frow[k + 1]; kj++) {
          const long long ij = find_col(fcol, fb, fe, fcol[kj]);
          if (ij < 0) continue;
          sub_matpr(fval + (size_t)ij * bs, pik, fval + (size_t)kj * bs, bi, bi, bj);
        }
      }
      const double* dii = ddiav + (size_t)i * bii;
      for (unsigned q = fb; q < fe; q++) matpr(fval + (size_t)q * bs, dii, bi, bj);
    }
    __syncthreads();
  }
}

unsigned grid(long long n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

hipError_t launch_ms_blk_matvec(const MsBlkDev& a, double alpha, const double* val, const double* x, double beta,
                                double* y, hipStream_t st) {
  if (a.nrow == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ms_blk_matvec, dim3(grid(a.nrow)), dim3(kT), 0, st, a.nrow, a.bi, a.bj, a.rowptr, a.colind,
                     val, alpha, x, beta, y);
  return hipGetLastError();
}

hipError_t launch_ms_blk_bc(const MsBlkDev& a, double* val, const int* bc_i, const int* bc_j, hipStream_t st) {
  if (a.nrow == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ms_blk_bc, dim3(grid(a.nrow)), dim3(kT), 0, st, a.nrow, a.bi, a.bj, a.rowptr, a.colind, val,
                     bc_i, bc_j);
  return hipGetLastError();
}

hipError_t launch_ms_dia_lowup(const MsDiaDev& d, const MsBlkDev& low, const double* lval, const MsBlkDev& up,
                               const double* uval, hipStream_t st) {
  if (d.nrow == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ms_dia_lowup, dim3(grid(d.nrow)), dim3(kT), 0, st, d.nrow, d.b, low.bj, d.rowptr, d.colind,
                     d.crs, d.dia, low.rowptr, low.colind, lval, up.rowptr, up.colind, uval);
  return hipGetLastError();
}

hipError_t launch_ms_od_lowup(const MsBlkDev& f, double* fval, const MsBlkDev& low, const double* lval,
                              const MsBlkDev& up, const double* uval, hipStream_t st) {
  if (f.nrow == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ms_od_lowup, dim3(grid(f.nrow)), dim3(kT), 0, st, f.nrow, f.bi, low.bj, f.bj, f.rowptr,
                     f.colind, fval, low.rowptr, low.colind, lval, up.rowptr, up.colind, uval);
  return hipGetLastError();
}

hipError_t launch_ms_od_low(const MsBlkDev& f, double* fval, const MsDiaDev& d, hipStream_t st) {
  if (f.nrow == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ms_od_low, dim3(grid(f.nrow)), dim3(kT), 0, st, f.nrow, f.bi, f.bj, f.rowptr, f.colind, fval,
                     d.rowptr, d.colind, d.diaind, d.crs);
  return hipGetLastError();
}

hipError_t launch_ms_od_up(int nlev, const int* lptr, const int* lrows, const MsBlkDev& f, double* fval,
                           const MsDiaDev& d, hipStream_t st) {
  if (f.nrow == 0 || nlev == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ms_od_up, dim3(1), dim3(1024), 0, st, nlev, lptr, lrows, f.bi, f.bj, f.rowptr, f.colind, fval,
                     d.rowptr, d.colind, d.diaind, d.crs, d.dia);
  return hipGetLastError();
}

}  // namespace dfm
