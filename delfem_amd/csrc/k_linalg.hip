// Vector and matrix kernels behind the handle operations of the C-ABI (the reference's CLinearSystem
// MATVEC/DOT/AXPY/COPY/SCAL, src/ls/linearsystem.cpp:250-415, and CVector_Blk, src/matvec/vector_blk.cpp),
// the block-CRS SpMV (CMatDia_BlkCrs::MatVec, src/matvec/matdia_blkcrs.cpp:921-1094) and the ILU(0) numeric
// factorisation (CMatDiaFrac_BlkCrs::DoILUDecomp, src/matvec/matdiafrac_blkcrs.cpp:774-1107).
#include "dev_common.hpp"
#include "dfm_internal.hpp"
#include "kernels.hpp"

namespace dfm {

__global__ void k_fill(double* x, double v, long long m) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_fill_f(float* x, float v, long long m) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_copy(const double* x, double* y, long long m) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
    y[i] = x[i];
}
__global__ void k_scal(double* x, double a, long long m) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
    x[i] *= a;
}
__global__ void k_axpy(double a, const double* x, double* y, long long m) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}
__global__ void k_to_f32(const double* x, float* y, long long m) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

static dim3 stream_grid(long long m) {
  long long b = (m + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

hipError_t launch_fill(double* x, double v, long long m, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, stream_grid(m), dim3(256), 0, st, x, v, m);
  return hipGetLastError();
}
hipError_t launch_fill_f(float* x, float v, long long m, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f, stream_grid(m), dim3(256), 0, st, x, v, m);
  return hipGetLastError();
}
// dst block map[q] = src block q (bs doubles per block): matrix values into the ILU(k) fill pattern
__global__ void k_scatter_blocks(long long nblk, int bs, const unsigned* __restrict__ map, const double* __restrict__ src,
                                 double* __restrict__ dst) {
  const long long m = nblk * bs;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x) {
    const long long q = i / bs;
    dst[(long long)map[q] * bs + (i - q * bs)] = src[i];
  }
}
hipError_t launch_scatter_blocks(long long nblk, int bs, const unsigned* map, const double* src, double* dst,
                                 hipStream_t st) {
  if (nblk <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_blocks, stream_grid(nblk * bs), dim3(256), 0, st, nblk, bs, map, src, dst);
  return hipGetLastError();
}
hipError_t launch_copy(const double* x, double* y, long long m, hipStream_t st) {
  if (m <= 0 || x == y) return hipSuccess;
  hipLaunchKernelGGL(k_copy, stream_grid(m), dim3(256), 0, st, x, y, m);
  return hipGetLastError();
}
hipError_t launch_scal(double* x, double a, long long m, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scal, stream_grid(m), dim3(256), 0, st, x, a, m);
  return hipGetLastError();
}
hipError_t launch_axpy(double a, const double* x, double* y, long long m, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy, stream_grid(m), dim3(256), 0, st, a, x, y, m);
  return hipGetLastError();
}
hipError_t launch_to_f32(const double* x, float* y, long long m, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_f32, stream_grid(m), dim3(256), 0, st, x, y, m);
  return hipGetLastError();
}

template <int NT>
__global__ void __launch_bounds__(NT) k_dot_part(const double* x, const double* y, long long m, double* partial) {
  double v[1] = {0.0};
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT) v[0] += x[i] * y[i];
  produce<NT, 1>(partial, v);
}

template <int NT>
__global__ void __launch_bounds__(NT) k_dot_finish(const double* partial, int G, double* out) {
  double t[1];
  consume<NT, 1>(partial, G, t);
  if (threadIdx.x == 0) out[0] = t[0];
}

// deterministic dot product: fixed grid of partials, then one block sums them in a fixed order
hipError_t launch_dot(const double* x, const double* y, long long m, double* partial, unsigned* counter, double* out,
                      hipStream_t st) {
  (void)counter;
  const unsigned nb = vec_blocks(m, 256);
  hipLaunchKernelGGL(k_dot_part<256>, dim3(nb), dim3(256), 0, st, x, y, m, partial);
  hipLaunchKernelGGL(k_dot_finish<256>, dim3(1), dim3(256), 0, st, partial, (int)nb, out);
  return hipGetLastError();
}

// y = alpha*A*x + beta*y (beta == 0 overwrites y)
template <int B, int TPR>
__global__ void __launch_bounds__(256) k_spmv(int n, const unsigned* __restrict__ rp, const unsigned* __restrict__ ci,
                                              const double* __restrict__ crs, const double* __restrict__ dia,
                                              double alpha, const double* __restrict__ x, double beta, double* y) {
  const int row = (blockIdx.x * 256 + threadIdx.x) / TPR;
  const int lane = threadIdx.x % TPR;
  if (row >= n) return;
  double acc[B];
  spmv_row<B, TPR, double, double, double>(rp[row], rp[row + 1], ci, crs, x, lane, acc);
  if (lane == 0) {
    dia_row<B, double, double, double>(dia, x, row, acc);
#pragma unroll
    for (int r = 0; r < B; r++) {
      double* yr = &y[(size_t)row * B + r];
      *yr = (beta == 0.0) ? alpha * acc[r] : beta * (*yr) + alpha * acc[r];
    }
  }
}

int spmv_tpr(long long ncrs, int n) {
  const double avg = n > 0 ? (double)ncrs / n : 0.0;
  if (avg <= 4.0) return 4;
  if (avg <= 8.0) return 8;
  if (avg <= 16.0) return 16;
  return 32;
}

hipError_t launch_spmv(int n, int len, int tpr, const unsigned* rowptr, const unsigned* colind,
                           const double* crs, const double* dia, double alpha, const double* x, double beta,
                           double* y, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const long long threads = (long long)n * tpr;
  const dim3 grid((unsigned)((threads + 255) / 256));
#define DFM_SPMV(B, T) \
  if (len == B && tpr == T) { hipLaunchKernelGGL((k_spmv<B, T>), grid, dim3(256), 0, st, n, rowptr, colind, crs, dia, alpha, x, beta, y); return hipGetLastError(); }
  DFM_SPMV(1, 4) DFM_SPMV(1, 8) DFM_SPMV(1, 16) DFM_SPMV(1, 32)
  DFM_SPMV(2, 4) DFM_SPMV(2, 8) DFM_SPMV(2, 16) DFM_SPMV(2, 32)
  DFM_SPMV(3, 4) DFM_SPMV(3, 8) DFM_SPMV(3, 16) DFM_SPMV(3, 32)
#undef DFM_SPMV
  return hipErrorInvalidValue;
}

// ||b - A x||^2 and ||b||^2 in one pass over the matrix (no residual vector): a fixed grid of at most kMaxPartial
// blocks walks the block rows (TPR lanes per row as k_spmv), each block writes its two partial sums
// (slot[blockIdx.x], slot[kMaxPartial + blockIdx.x]); one block then sums them in block order into out[0..1].
template <int B, int TPR>
__global__ void __launch_bounds__(256) k_true_res_part(int n, const unsigned* __restrict__ rp,
                                                       const unsigned* __restrict__ ci, const double* __restrict__ crs,
                                                       const double* __restrict__ dia, const double* __restrict__ x,
                                                       const double* __restrict__ b, double* part) {
  constexpr int RPB = 256 / TPR;  // rows per block and sweep
  const int lane = threadIdx.x % TPR;
  double v[2] = {0.0, 0.0};
  for (int r0 = (int)blockIdx.x * RPB; r0 < n; r0 += (int)gridDim.x * RPB) {
    const int row = r0 + (int)threadIdx.x / TPR;
    if (row < n) {
      double acc[B];
      spmv_row<B, TPR, double, double, double>(rp[row], rp[row + 1], ci, crs, x, lane, acc);
      if (lane == 0) {
        dia_row<B, double, double, double>(dia, x, row, acc);
#pragma unroll
        for (int k = 0; k < B; k++) {
          const double bi = b[(size_t)row * B + k];
          const double t = bi - acc[k];
          v[0] += t * t;
          v[1] += bi * bi;
        }
      }
    }
  }
  produce<256, 2>(part, v);
}

__global__ void __launch_bounds__(256) k_true_res_finish(const double* part, int G, double* out) {
  double t[2];
  consume<256, 2>(part, G, t);
  if (threadIdx.x == 0) {
    out[0] = t[0];
    out[1] = t[1];
  }
}

hipError_t launch_true_residual(int n, int len, int tpr, const unsigned* rowptr, const unsigned* colind,
                                const double* crs, const double* dia, const double* x, const double* b, double* part,
                                double* out, hipStream_t st) {
  if (n <= 0) return hipErrorInvalidValue;
  const long long rows_per_block = 256 / tpr;
  const unsigned G = (unsigned)std::min<long long>(kMaxPartial, (n + rows_per_block - 1) / rows_per_block);
#define DFM_TRES(B, T)                                                                                          \
  if (len == B && tpr == T) {                                                                                   \
    hipLaunchKernelGGL((k_true_res_part<B, T>), dim3(G), dim3(256), 0, st, n, rowptr, colind, crs, dia, x, b, part); \
    hipLaunchKernelGGL(k_true_res_finish, dim3(1), dim3(256), 0, st, (const double*)part, (int)G, out);         \
    return hipGetLastError();                                                                                  \
  }
  DFM_TRES(1, 4) DFM_TRES(1, 8) DFM_TRES(1, 16) DFM_TRES(1, 32)
  DFM_TRES(2, 4) DFM_TRES(2, 8) DFM_TRES(2, 16) DFM_TRES(2, 32)
  DFM_TRES(3, 4) DFM_TRES(3, 8) DFM_TRES(3, 16) DFM_TRES(3, 32)
#undef DFM_TRES
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------------------------------
// ILU(0) numeric factorisation, row-wise IKJ exactly as DoILUDecomp: for row i and each k < i of its lower
// part in ascending order, A_ij -= A_ik * U_kj for every j > k of row k that is in row i's pattern (or is
// i itself: the diagonal block); then D_i <- D_i^-1 (closed forms for 1x1, 2x2, 3x3; pivots with
// |det| <= 1e-30 are left uninverted and counted); then U_ij <- D_i^-1 U_ij. Rows run in the forward
// schedule: a row starts when the rows it reads are final.

template <int B>
__device__ inline void sub_prod(double* pij, const double* pik, const double* pkj) {
  if (B == 1) {
    pij[0] -= pik[0] * pkj[0];
  } else if (B == 2) {
    pij[0] -= pik[0] * pkj[0] + pik[1] * pkj[2];
    pij[1] -= pik[0] * pkj[1] + pik[1] * pkj[3];
    pij[2] -= pik[2] * pkj[0] + pik[3] * pkj[2];
    pij[3] -= pik[2] * pkj[1] + pik[3] * pkj[3];
  } else {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      pij[a * 3 + 0] -= pik[a * 3 + 0] * pkj[0] + pik[a * 3 + 1] * pkj[3] + pik[a * 3 + 2] * pkj[6];
      pij[a * 3 + 1] -= pik[a * 3 + 0] * pkj[1] + pik[a * 3 + 1] * pkj[4] + pik[a * 3 + 2] * pkj[7];
      pij[a * 3 + 2] -= pik[a * 3 + 0] * pkj[2] + pik[a * 3 + 1] * pkj[5] + pik[a * 3 + 2] * pkj[8];
    }
  }
}

template <int B>
__device__ inline bool invert_dia(double* p) {
  if (B == 1) {
    const double v = p[0];
    if (fabs(v) > 1.0e-30) { p[0] = 1.0 / v; return true; }
    return false;
  } else if (B == 2) {
    const double det = p[0] * p[3] - p[1] * p[2];
    if (fabs(det) > 1.0e-30) {
      const double inv_det = 1.0 / det;
      const double t = p[0];
      p[0] = inv_det * p[3];
      p[1] = -inv_det * p[1];
      p[2] = -inv_det * p[2];
      p[3] = inv_det * t;
      return true;
    }
    return false;
  } else {
    const double* a = p;
    const double det = a[0] * a[4] * a[8] + a[3] * a[7] * a[2] + a[6] * a[1] * a[5] - a[0] * a[7] * a[5] -
                       a[6] * a[4] * a[2] - a[3] * a[1] * a[8];
    if (!(fabs(det) > 1.0e-30)) return false;
    const double inv_det = 1.0 / det;
    double t[9];
#pragma unroll
    for (int k = 0; k < 9; k++) t[k] = p[k];
    p[0] = inv_det * (t[4] * t[8] - t[5] * t[7]);
    p[1] = inv_det * (t[2] * t[7] - t[1] * t[8]);
    p[2] = inv_det * (t[1] * t[5] - t[2] * t[4]);
    p[3] = inv_det * (t[5] * t[6] - t[3] * t[8]);
    p[4] = inv_det * (t[0] * t[8] - t[2] * t[6]);
    p[5] = inv_det * (t[2] * t[3] - t[0] * t[5]);
    p[6] = inv_det * (t[3] * t[7] - t[4] * t[6]);
    p[7] = inv_det * (t[1] * t[6] - t[0] * t[7]);
    p[8] = inv_det * (t[0] * t[4] - t[1] * t[3]);
    return true;
  }
}

template <int B>
__device__ inline void scale_u(double* pij, const double* pii) {
  double t[B * B];
#pragma unroll
  for (int k = 0; k < B * B; k++) t[k] = pij[k];
  if (B == 1) {
    pij[0] = t[0] * pii[0];
  } else if (B == 2) {
    pij[0] = pii[0] * t[0] + pii[1] * t[2];
    pij[1] = pii[0] * t[1] + pii[1] * t[3];
    pij[2] = pii[2] * t[0] + pii[3] * t[2];
    pij[3] = pii[2] * t[1] + pii[3] * t[3];
  } else {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      pij[a * 3 + 0] = pii[a * 3 + 0] * t[0] + pii[a * 3 + 1] * t[3] + pii[a * 3 + 2] * t[6];
      pij[a * 3 + 1] = pii[a * 3 + 0] * t[1] + pii[a * 3 + 1] * t[4] + pii[a * 3 + 2] * t[7];
      pij[a * 3 + 2] = pii[a * 3 + 0] * t[2] + pii[a * 3 + 1] * t[5] + pii[a * 3 + 2] * t[8];
    }
  }
}

template <int B>
__global__ void __launch_bounds__(64) k_ilu_factor(int gbase, const int* __restrict__ grp, const int* __restrict__ rowlist,
                                                   const uint8_t* __restrict__ lev, int nlev,
                                                   const unsigned* __restrict__ rp, const unsigned* __restrict__ di,
                                                   const unsigned* __restrict__ ci, double* fcrs, double* fdia,
                                                   int* nsing) {
  constexpr int BS = B * B;
  const int g = gbase + blockIdx.x;
  const int start = grp[2 * g], end = grp[2 * g + 1];
  const int lane = threadIdx.x;
  const bool active = lane < end - start;
  const int i = active ? (rowlist ? rowlist[start + lane] : start + lane) : 0;
  const int my = active ? (int)lev[i] : -1;
  for (int l = 0; l < nlev; l++) {
    if (my == l) {
      const unsigned r0 = rp[i], r1 = rp[i + 1], d = di[i];
      double* pii = &fdia[(size_t)i * BS];
      for (unsigned ik = r0; ik < d; ik++) {
        const unsigned k = ci[ik];
        const double* pik = &fcrs[(size_t)ik * BS];
        unsigned qi = r0;
        for (unsigned kj = di[k]; kj < rp[k + 1]; kj++) {
          const unsigned j = ci[kj];
          const double* pkj = &fcrs[(size_t)kj * BS];
          if (j == (unsigned)i) {
            sub_prod<B>(pii, pik, pkj);
          } else {
            while (qi < r1 && ci[qi] < j) qi++;
            if (qi < r1 && ci[qi] == j) sub_prod<B>(&fcrs[(size_t)qi * BS], pik, pkj);
          }
        }
      }
      if (!invert_dia<B>(pii)) atomicAdd(nsing, 1);
      for (unsigned ij = d; ij < r1; ij++) scale_u<B>(&fcrs[(size_t)ij * BS], pii);
    }
    __syncthreads();
  }
}

// A b x b FP64 block (row-major, b*b doubles at blk*b*b) into registers with 16-byte loads. 3x3 blocks start at
// an 8-byte boundary only: the ten doubles from the 16-byte boundary at or below the block are read (one past its
// end at most: the arrays carry two doubles of padding) and the nine of the block picked out.
template <int B>
__device__ inline void load_block(const double* base, size_t blk, double* v) {
  if (B == 1) {
    v[0] = base[blk];
  } else if (B == 2) {
    const double2* p = (const double2*)(base + blk * 4);
    const double2 a = p[0], b = p[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
  } else {
    const size_t s0 = blk * 9;
    const double2* p = (const double2*)(base + (s0 & ~(size_t)1));
    const double2 t0 = p[0], t1 = p[1], t2 = p[2], t3 = p[3], t4 = p[4];
    if (s0 & 1) {
      v[0] = t0.y; v[1] = t1.x; v[2] = t1.y; v[3] = t2.x; v[4] = t2.y; v[5] = t3.x; v[6] = t3.y; v[7] = t4.x; v[8] = t4.y;
    } else {
      v[0] = t0.x; v[1] = t0.y; v[2] = t1.x; v[3] = t1.y; v[4] = t2.x; v[5] = t2.y; v[6] = t3.x; v[7] = t3.y; v[8] = t4.x;
    }
  }
}

// Numeric ILU factorisation, one wavefront per forward group, its rows spread over the lanes one block per lane
// (FactorPlan): a lane keeps its block in registers, applies its updates A_ij -= A_ik A_kj in ascending k with
// A_ik handed over at step q by the lane of the row's lower entry q once that entry is final (every update of A_ik
// comes from a k' < k, so from an earlier step), inverts the diagonal, scales its upper block by it and writes it
// once. The expressions are the per-row kernel's (the reference's) and each block sees the same updates in the same
// order, so the factor is bit-identical to it. Passes of a later intra-group level start after a barrier that makes
// the rows of earlier levels visible. a_crs/a_dia: the initial values (the matrix itself for ILU(0), else the
// scattered factor, in place). Measured on the 3D elasticity system of 96^3 hexes (2.74 M dofs, 4 phases): 1.63 ms
// against 13.6 ms for the per-row kernel; what bounds it now is the L1 tag traffic of the per-lane 72-byte block
// accesses (each wave instruction touches up to 64 cache lines), not HBM.
template <int B>
__global__ void __launch_bounds__(64) k_ilu_rows(int gbase, const int* __restrict__ gpass, const uint4* __restrict__ pmeta,
                                                 const unsigned* __restrict__ etab, const unsigned* __restrict__ utab,
                                                 const unsigned* __restrict__ ukj, const uint8_t* __restrict__ ustep,
                                                 const double* a_crs, const double* a_dia, double* fcrs, double* fdia,
                                                 int* nsing) {
  constexpr int BS = B * B;
  const int g = gbase + blockIdx.x;
  const int lane = threadIdx.x;
  const int p0 = gpass[g], p1 = gpass[g + 1];
  for (int p = p0; p < p1; p++) {
    const uint4 m0 = pmeta[2 * (size_t)p], m1 = pmeta[2 * (size_t)p + 1];
    if (m1.x >> 16) __syncthreads();  // a later level: the rows of the earlier ones were written
    const unsigned long long smask = (unsigned long long)m0.y << 32 | m0.x;
    const unsigned long long dmask = (unsigned long long)m0.w << 32 | m0.z;
    const int nlanes = (int)(m1.x & 255u), steps = (int)((m1.x >> 8) & 255u);
    const bool act = lane < nlanes;
    const unsigned long long upto = smask & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const int rstart = 63 - __clzll(upto | 1ull);                       // first lane of this lane's row
    const int dl = rstart + __ffsll((long long)(dmask >> rstart)) - 1;  // its diagonal lane
    const bool dia = act && lane == dl, up = act && lane > dl;
    const size_t li = (size_t)p * 64 + lane;
    const unsigned e = etab[li];  // off-diagonal entry, or the row (diagonal)
    unsigned u = utab[li];
    const unsigned ue = utab[li + 1];
    double acc[BS];
    if (act) load_block<B>(dia ? a_dia : a_crs, e, acc);
    else {
#pragma unroll
      for (int t = 0; t < BS; t++) acc[t] = 0.0;
    }
    // the lane's update (kj, step) and the one after it are held ahead, so that the block A_kj of the next update
    // is requested as soon as the current one is applied, without first waiting for its index
    unsigned nkj = u < ue ? ukj[u] : 0u;
    int nst = u < ue ? (int)ustep[u] : 255;
    unsigned nkj2 = u + 1 < ue ? ukj[u + 1] : 0u;
    int nst2 = u + 1 < ue ? (int)ustep[u + 1] : 255;
    double akj[BS];
    load_block<B>(fcrs, nkj, akj);
    for (int q = 0; q < steps; q++) {
      double aik[BS];
#pragma unroll
      for (int t = 0; t < BS; t++) aik[t] = __shfl(acc[t], rstart + q);
      if (nst == q) {
        sub_prod<B>(acc, aik, akj);
        u++;
        nkj = nkj2;
        nst = nst2;
        load_block<B>(fcrs, nkj, akj);
        nkj2 = u + 1 < ue ? ukj[u + 1] : 0u;
        nst2 = u + 1 < ue ? (int)ustep[u + 1] : 255;
      }
    }
    if (dia && !invert_dia<B>(acc)) atomicAdd(nsing, 1);
    double dinv[BS];
#pragma unroll
    for (int t = 0; t < BS; t++) dinv[t] = __shfl(acc[t], dl);
    if (up) scale_u<B>(acc, dinv);
    if (act) {
      double* out = dia ? fdia : fcrs;
#pragma unroll
      for (int t = 0; t < BS; t++) out[(size_t)e * BS + t] = acc[t];
    }
  }
}

template <int B, typename T>
__global__ void __launch_bounds__(256) k_slab_fill(long long nent, const unsigned* __restrict__ src,
                                                   const double* __restrict__ f, T* __restrict__ val) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= nent) return;
  const long long r = i / 64;
  const int lane = (int)(i % 64);
  const unsigned q = src[i];
#pragma unroll
  for (int e = 0; e < B * B; e++) val[(r * (B * B) + e) * 64 + lane] = q != ~0u ? (T)f[(size_t)q * (B * B) + e] : T(0);
}

template <typename T>
static hipError_t slab_fill_T(int len, long long nent, const unsigned* src, const double* f, T* val, hipStream_t st) {
  if (nent <= 0) return hipSuccess;
  const dim3 g((unsigned)((nent + 255) / 256));
  if (len == 1) hipLaunchKernelGGL((k_slab_fill<1, T>), g, dim3(256), 0, st, nent, src, f, val);
  else if (len == 2) hipLaunchKernelGGL((k_slab_fill<2, T>), g, dim3(256), 0, st, nent, src, f, val);
  else if (len == 3) hipLaunchKernelGGL((k_slab_fill<3, T>), g, dim3(256), 0, st, nent, src, f, val);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}
// Values of a sweep's packed entries from the FP64 factor: one wavefront per group, each lane placing its
// entries of both classes where the sweep kernels read them (jagged rows element-major, the rest per entry).
template <int B, typename T>
__global__ void __launch_bounds__(256) k_sweep_fill(int ngroup, const int* __restrict__ gmeta,
                                                    const unsigned* __restrict__ scnt, const unsigned* __restrict__ src,
                                                    const double* __restrict__ f, T* __restrict__ val) {
  constexpr int BS = B * B;
  const int g = (int)blockIdx.x * 4 + wave_in_block();
  if (g >= ngroup) return;
  const int* m = gmeta + 4 * (size_t)g;
  const int base = m[0], wo = m[1] & 255, wi = (m[1] >> 8) & 255, nout = m[3] >> 8;
  const unsigned c = scnt[(size_t)g * 64 + threadIdx.x % 64];
  const int cnt[2] = {(int)(c & 255u), (int)((c >> 8) & 255u)};
  const int wid[2] = {wo, wi};
  for (int cls = 0; cls < 2; cls++) {
    const int n = cnt[cls], w = wid[cls];
    int r = cls == 0 ? base : base + nout;
    for (int k = 0; k < kJag && k < w; k++) {
      const unsigned long long msk = __ballot(n > k);
      const int len = __popcll(msk), pos = lanes_below(msk);
      if (k < n) {
        const size_t q = src[(size_t)(r + pos)];
#pragma unroll
        for (int e = 0; e < BS; e++) val[(size_t)r * BS + (size_t)e * len + pos] = (T)f[q * BS + e];
      }
      r += len;
    }
    if (w > kJag) {
      const int s0 = r + wave_scan_u8(n > kJag ? (unsigned)(n - kJag) : 0u);
      for (int k = kJag; k < n; k++) {
        const size_t i = (size_t)s0 + (k - kJag), q = src[i];
#pragma unroll
        for (int e = 0; e < BS; e++) val[i * BS + e] = (T)f[q * BS + e];
      }
    }
  }
}

template <typename T>
static hipError_t sweep_fill_T(int len, int ngroup, const int* gmeta, const unsigned* scnt, const unsigned* src,
                               const double* f, T* val, hipStream_t st) {
  if (ngroup <= 0) return hipSuccess;
  const dim3 g((unsigned)((ngroup + 3) / 4));
  if (len == 1) hipLaunchKernelGGL((k_sweep_fill<1, T>), g, dim3(256), 0, st, ngroup, gmeta, scnt, src, f, val);
  else if (len == 2) hipLaunchKernelGGL((k_sweep_fill<2, T>), g, dim3(256), 0, st, ngroup, gmeta, scnt, src, f, val);
  else if (len == 3) hipLaunchKernelGGL((k_sweep_fill<3, T>), g, dim3(256), 0, st, ngroup, gmeta, scnt, src, f, val);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}
hipError_t launch_sweep_fill(int len, int ngroup, const int* gmeta, const unsigned* scnt, const unsigned* src,
                             const double* f, double* val, hipStream_t st) {
  return sweep_fill_T<double>(len, ngroup, gmeta, scnt, src, f, val, st);
}
hipError_t launch_sweep_fill(int len, int ngroup, const int* gmeta, const unsigned* scnt, const unsigned* src,
                             const double* f, float* val, hipStream_t st) {
  return sweep_fill_T<float>(len, ngroup, gmeta, scnt, src, f, val, st);
}

hipError_t launch_slab_fill(int len, long long nent, const unsigned* src, const double* f, double* val, hipStream_t st) {
  return slab_fill_T<double>(len, nent, src, f, val, st);
}
hipError_t launch_slab_fill(int len, long long nent, const unsigned* src, const double* f, float* val, hipStream_t st) {
  return slab_fill_T<float>(len, nent, src, f, val, st);
}

hipError_t launch_ilu_rows(int len, const FactorPlanDev& fp, const int* phase_gptr, int nphase, const double* a_crs,
                           const double* a_dia, double* fcrs, double* fdia, int* nsing, hipStream_t st) {
  for (int p = 0; p < nphase; p++) {
    const int g0 = phase_gptr[p], ng = phase_gptr[p + 1] - g0;
    if (ng == 0) continue;
    const uint4* pm = (const uint4*)fp.pmeta;
    if (len == 1)
      hipLaunchKernelGGL(k_ilu_rows<1>, dim3(ng), dim3(64), 0, st, g0, fp.gpass, pm, fp.etab, fp.utab, fp.ukj,
                         fp.ustep, a_crs, a_dia, fcrs, fdia, nsing);
    else if (len == 2)
      hipLaunchKernelGGL(k_ilu_rows<2>, dim3(ng), dim3(64), 0, st, g0, fp.gpass, pm, fp.etab, fp.utab, fp.ukj,
                         fp.ustep, a_crs, a_dia, fcrs, fdia, nsing);
    else if (len == 3)
      hipLaunchKernelGGL(k_ilu_rows<3>, dim3(ng), dim3(64), 0, st, g0, fp.gpass, pm, fp.etab, fp.utab, fp.ukj,
                         fp.ustep, a_crs, a_dia, fcrs, fdia, nsing);
    else
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_ilu_factor(int n, int len, const unsigned* rowptr, const unsigned* colind, const unsigned* diaind,
                             double* fcrs, double* fdia, const SweepDev& fw, int* nsing, hipStream_t st) {
  (void)n;
  for (int p = 0; p < fw.nphase; p++) {
    const int g0 = fw.phase_gptr[p], ng = fw.phase_gptr[p + 1] - g0;
    if (ng == 0) continue;
    const int nl = fw.phase_nlev[p];
    if (len == 1)
      hipLaunchKernelGGL(k_ilu_factor<1>, dim3(ng), dim3(64), 0, st, g0, fw.grp, fw.rowlist, fw.lev, nl, rowptr, diaind,
                         colind, fcrs, fdia, nsing);
    else if (len == 2)
      hipLaunchKernelGGL(k_ilu_factor<2>, dim3(ng), dim3(64), 0, st, g0, fw.grp, fw.rowlist, fw.lev, nl, rowptr, diaind,
                         colind, fcrs, fdia, nsing);
    else if (len == 3)
      hipLaunchKernelGGL(k_ilu_factor<3>, dim3(ng), dim3(64), 0, st, g0, fw.grp, fw.rowlist, fw.lev, nl, rowptr, diaind,
                         colind, fcrs, fdia, nsing);
    else
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace dfm
