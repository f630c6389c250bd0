// Device building blocks shared by k_linalg.hip and solver.hip:
//  - deterministic reductions (wavefront shuffles + LDS tree, per-block partials reduced by the consuming
//    kernel: the partition and the summation tree are fixed, so results do not depend on scheduling),
//  - the block-CRS SpMV row (TPR lanes per block row, diagonal block kept apart as in the reference),
//  - the ILU(0) phase kernels (one wavefront per group of <= 64 rows, intra-group levels through LDS).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dfm {

constexpr int kWave = 64;

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Deterministic reductions without fences. A producer kernel writes one partial per block (plain stores,
// fixed grid); the kernel that needs the total reduces all partials itself at its start, every block in the
// same fixed order, so every block obtains the identical value. Kernel boundaries make the partials
// visible; no block ever waits on another, and no device-scope fence (an L2 write-back on a multi-XCD part)
// is issued per block.
constexpr int kMaxPartial = 2048;  // producer grids are capped at this many blocks

// Sum of K values over the block; the totals are returned in every thread. NT = block size.
template <int NT, int K>
__device__ inline void block_allsum(double (&v)[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
  if (NT == kWave) return;
  __shared__ double sh[K][NT / kWave];
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0)
#pragma unroll
    for (int k = 0; k < K; k++) sh[k][w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    double t = sh[k][0];
#pragma unroll
    for (int j = 1; j < NT / kWave; j++) t += sh[k][j];
    v[k] = t;
  }
  __syncthreads();
}

// The block's K partial sums into slot[k*kMaxPartial + blockIdx.x].
template <int NT, int K>
__device__ inline void produce(double* slot, double (&v)[K]) {
  block_allsum<NT, K>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; k++) slot[k * kMaxPartial + blockIdx.x] = v[k];
}

// Totals of the G partials of a previous kernel, identical in every thread of every block.
template <int NT, int K>
__device__ inline void consume(const double* __restrict__ slot, int G, double (&tot)[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < G; b += NT) s += slot[k * kMaxPartial + b];
    tot[k] = s;
  }
  block_allsum<NT, K>(tot);
}

inline unsigned vec_blocks(long long m, int nt) {
  long long b = (m + nt - 1) / nt;
  if (b < 1) b = 1;
  if (b > kMaxPartial) b = kMaxPartial;
  return (unsigned)b;
}

// ---------------------------------------------------------------------------------------------------------
// SpMV row: acc[r] = sum over the row's off-diagonal blocks (lanes lane, lane+TPR, ...), reduced across the
// TPR lanes (result in every lane of the row group), then the caller adds the diagonal block.
template <int B, int TPR, typename T, typename A, typename VX>
__device__ inline void spmv_row(unsigned q0, unsigned q1, const unsigned* __restrict__ ci, const T* __restrict__ crs,
                                const VX* __restrict__ x, int lane, A (&acc)[B]) {
#pragma unroll
  for (int r = 0; r < B; r++) acc[r] = A(0);
  for (unsigned q = q0 + lane; q < q1; q += TPR) {
    const unsigned j = ci[q];
    const T* a = crs + (size_t)q * (B * B);
    A xj[B];
#pragma unroll
    for (int s = 0; s < B; s++) xj[s] = (A)x[(size_t)j * B + s];
#pragma unroll
    for (int r = 0; r < B; r++) {
      A t = (A)a[r * B] * xj[0];
#pragma unroll
      for (int s = 1; s < B; s++) t += (A)a[r * B + s] * xj[s];
      acc[r] += t;
    }
  }
#pragma unroll
  for (int o = TPR / 2; o > 0; o >>= 1)
#pragma unroll
    for (int r = 0; r < B; r++) acc[r] += __shfl_xor(acc[r], o, TPR);
}

template <int B, typename T, typename A, typename VX>
__device__ inline void dia_row(const T* __restrict__ dia, const VX* __restrict__ x, int i, A (&acc)[B]) {
  const T* d = dia + (size_t)i * (B * B);
  A xi[B];
#pragma unroll
  for (int s = 0; s < B; s++) xi[s] = (A)x[(size_t)i * B + s];
#pragma unroll
  for (int r = 0; r < B; r++) {
    A t = (A)d[r * B] * xi[0];
#pragma unroll
    for (int s = 1; s < B; s++) t += (A)d[r * B + s] * xi[s];
    acc[r] += t;
  }
}

// ---------------------------------------------------------------------------------------------------------
// ILU(0) sweeps. Forward (ForwardSubstitution, src/matvec/matdiafrac_blkcrs.cpp:332-507):
//   t = b_i - sum_{j<i} L_ij x_j (ascending j), x_i = D_i^-1 t.
// Backward (BackwardSubstitution, :509-669): x_i -= sum_{j>i} U_ij x_j (D^-1 already folded into U).
// One block of 64 threads per group; lane = row. Entries [rp, lsplit) / [usplit, rp+1) reference rows of
// earlier phases (final in global memory); [lsplit, di) / [di, usplit) reference rows of the same group,
// resolved level by level through LDS.

template <int B, typename T, typename A>
__device__ inline void blk_sub(A (&t)[B], const T* __restrict__ a, const A (&v)[B]) {
#pragma unroll
  for (int r = 0; r < B; r++) {
    A s = (A)a[r * B] * v[0];
#pragma unroll
    for (int c = 1; c < B; c++) s += (A)a[r * B + c] * v[c];
    t[r] -= s;
  }
}

// MAXIN in-group entries per row are loaded into registers before the level loop, so that a level costs
// LDS reads and arithmetic only; rows with more in-group entries read the rest from global memory.
template <int B>
struct InGroup {
  static constexpr int MAXIN = (B == 2) ? 4 : (B == 3 ? 7 : 8);
};

// GPB groups per workgroup, one wavefront per group: few, larger workgroups keep every group of a phase
// resident at once (the per-CU workgroup limit, not registers, bounds residency of one-wave workgroups).
constexpr int kGPB = 4;

template <int B, typename T, typename A, typename VI, typename VO>
__global__ void __launch_bounds__(64 * kGPB) k_ilu_fwd(int gbase, int gend, const int* __restrict__ grp, const int* __restrict__ rowlist,
                                                const uint8_t* __restrict__ lev, int nlev,
                                                const unsigned* __restrict__ rp, const unsigned* __restrict__ lsplit,
                                                const unsigned* __restrict__ di, const unsigned* __restrict__ ci,
                                                const T* __restrict__ fcrs, const T* __restrict__ fdia,
                                                const VI* b, VO* x, const int* __restrict__ done) {
  constexpr int BS = B * B;
  constexpr int MAXIN = InGroup<B>::MAXIN;
  if (done && *done) return;
  const int wv = threadIdx.x / 64;
  const int g = gbase + blockIdx.x * kGPB + wv;
  const bool gok = g < gend;
  const int start = gok ? grp[2 * g] : 0, end = gok ? grp[2 * g + 1] : 0;
  const int lane = threadIdx.x % 64;
  const bool active = lane < end - start;
  const int row = active ? (rowlist ? rowlist[start + lane] : start + lane) : 0;
  __shared__ A xsb[kGPB][64 * B];
  A* xs = xsb[wv];
  A t[B], D[BS];
  int jl[MAXIN];
  A lv[MAXIN][BS];
  unsigned ls = 0, d = 0;
  if (active) {
#pragma unroll
    for (int r = 0; r < B; r++) t[r] = (A)b[(size_t)row * B + r];
    ls = lsplit[row];
    d = di[row];
    // in-group entries and the diagonal block first (independent of the sweep), then rows of earlier phases
#pragma unroll
    for (int k = 0; k < MAXIN; k++) {
      if (ls + k < d) {
        jl[k] = (int)ci[ls + k] - start;
#pragma unroll
        for (int e = 0; e < BS; e++) lv[k][e] = (A)fcrs[(size_t)(ls + k) * BS + e];
      }
    }
#pragma unroll
    for (int e = 0; e < BS; e++) D[e] = (A)fdia[(size_t)row * BS + e];
    // rows of earlier phases: all column indices, then all gathers, then the products in ascending order
    const unsigned q0 = rp[row];
    unsigned jo[MAXIN];
#pragma unroll
    for (int k = 0; k < MAXIN; k++)
      if (q0 + k < ls) jo[k] = ci[q0 + k];
    A xo[MAXIN][B];
#pragma unroll
    for (int k = 0; k < MAXIN; k++)
      if (q0 + k < ls)
#pragma unroll
        for (int c = 0; c < B; c++) xo[k][c] = (A)x[(size_t)jo[k] * B + c];
#pragma unroll
    for (int k = 0; k < MAXIN; k++)
      if (q0 + k < ls) blk_sub<B>(t, fcrs + (size_t)(q0 + k) * BS, xo[k]);
    for (unsigned q = q0 + MAXIN; q < ls; q++) {
      const unsigned j = ci[q];
      A v[B];
#pragma unroll
      for (int c = 0; c < B; c++) v[c] = (A)x[(size_t)j * B + c];
      blk_sub<B>(t, fcrs + (size_t)q * BS, v);
    }
  }
  const int my = active ? (int)lev[row] : -1;
  for (int l = 0; l < nlev; l++) {
    if (my == l) {
#pragma unroll
      for (int k = 0; k < MAXIN; k++) {
        if (ls + k < d) {
          A v[B];
#pragma unroll
          for (int c = 0; c < B; c++) v[c] = xs[jl[k] * B + c];
#pragma unroll
          for (int r = 0; r < B; r++) {
            A s = lv[k][r * B] * v[0];
#pragma unroll
            for (int c = 1; c < B; c++) s += lv[k][r * B + c] * v[c];
            t[r] -= s;
          }
        }
      }
      for (unsigned q = ls + MAXIN; q < d; q++) {
        const int j = (int)ci[q] - start;
        A v[B];
#pragma unroll
        for (int c = 0; c < B; c++) v[c] = xs[j * B + c];
        blk_sub<B>(t, fcrs + (size_t)q * BS, v);
      }
      A y[B];
#pragma unroll
      for (int r = 0; r < B; r++) {
        A s = D[r * B] * t[0];
#pragma unroll
        for (int c = 1; c < B; c++) s += D[r * B + c] * t[c];
        y[r] = s;
      }
#pragma unroll
      for (int r = 0; r < B; r++) {
        xs[lane * B + r] = y[r];
        t[r] = y[r];
      }
    }
    // only LDS traffic inside the level loop: the barrier's fence then has no global store to wait for
    if (nlev > 1) __syncthreads();
  }
  if (active)
#pragma unroll
    for (int r = 0; r < B; r++) x[(size_t)row * B + r] = (VO)t[r];
}

template <int B, typename T, typename A, typename VX>
__global__ void __launch_bounds__(64 * kGPB) k_ilu_bwd(int gbase, int gend, const int* __restrict__ grp, const int* __restrict__ rowlist,
                                                const uint8_t* __restrict__ lev, int nlev,
                                                const unsigned* __restrict__ rp, const unsigned* __restrict__ usplit,
                                                const unsigned* __restrict__ di, const unsigned* __restrict__ ci,
                                                const T* __restrict__ fcrs, VX* x, const int* __restrict__ done) {
  constexpr int BS = B * B;
  constexpr int MAXIN = InGroup<B>::MAXIN;
  if (done && *done) return;
  const int wv = threadIdx.x / 64;
  const int g = gbase + blockIdx.x * kGPB + wv;
  const bool gok = g < gend;
  const int start = gok ? grp[2 * g] : 0, end = gok ? grp[2 * g + 1] : 0;
  const int lane = threadIdx.x % 64;
  const bool active = lane < end - start;
  const int row = active ? (rowlist ? rowlist[start + lane] : start + lane) : 0;
  __shared__ A xsb[kGPB][64 * B];
  A* xs = xsb[wv];
  A t[B];
  int jl[MAXIN];
  A uv[MAXIN][BS];
  unsigned d = 0, us = 0;
  if (active) {
#pragma unroll
    for (int r = 0; r < B; r++) t[r] = (A)x[(size_t)row * B + r];
    d = di[row];
    us = usplit[row];
    const unsigned qe = rp[row + 1];
#pragma unroll
    for (int k = 0; k < MAXIN; k++) {
      if (d + k < us) {
        jl[k] = (int)ci[d + k] - start;
#pragma unroll
        for (int e = 0; e < BS; e++) uv[k][e] = (A)fcrs[(size_t)(d + k) * BS + e];
      }
    }
    // rows of later phases first (with no intra-group entries this is the reference's ascending order):
    // all column indices, then all gathers, then the products
    unsigned jo[MAXIN];
#pragma unroll
    for (int k = 0; k < MAXIN; k++)
      if (us + k < qe) jo[k] = ci[us + k];
    A xo[MAXIN][B];
#pragma unroll
    for (int k = 0; k < MAXIN; k++)
      if (us + k < qe)
#pragma unroll
        for (int c = 0; c < B; c++) xo[k][c] = (A)x[(size_t)jo[k] * B + c];
#pragma unroll
    for (int k = 0; k < MAXIN; k++)
      if (us + k < qe) blk_sub<B>(t, fcrs + (size_t)(us + k) * BS, xo[k]);
    for (unsigned q = us + MAXIN; q < qe; q++) {
      const unsigned j = ci[q];
      A v[B];
#pragma unroll
      for (int c = 0; c < B; c++) v[c] = (A)x[(size_t)j * B + c];
      blk_sub<B>(t, fcrs + (size_t)q * BS, v);
    }
  }
  const int my = active ? (int)lev[row] : -1;
  for (int l = 0; l < nlev; l++) {
    if (my == l) {
#pragma unroll
      for (int k = 0; k < MAXIN; k++) {
        if (d + k < us) {
          A v[B];
#pragma unroll
          for (int c = 0; c < B; c++) v[c] = xs[jl[k] * B + c];
#pragma unroll
          for (int r = 0; r < B; r++) {
            A s = uv[k][r * B] * v[0];
#pragma unroll
            for (int c = 1; c < B; c++) s += uv[k][r * B + c] * v[c];
            t[r] -= s;
          }
        }
      }
      for (unsigned q = d + MAXIN; q < us; q++) {
        const int j = (int)ci[q] - start;
        A v[B];
#pragma unroll
        for (int c = 0; c < B; c++) v[c] = xs[j * B + c];
        blk_sub<B>(t, fcrs + (size_t)q * BS, v);
      }
#pragma unroll
      for (int r = 0; r < B; r++) xs[lane * B + r] = t[r];
    }
    if (nlev > 1) __syncthreads();
  }
  if (active)
#pragma unroll
    for (int r = 0; r < B; r++) x[(size_t)row * B + r] = (VX)t[r];
}

}  // namespace dfm
