// Device-resident Krylov loops: LsSol::Solve_PCG and Solve_PBiCGSTAB (src/ls/solver_ls_iter.cpp:100-172,
// :291-408) with every scalar kept in device memory, and the mixed-precision PCG with reliable updates.
//
// Every kernel first reads the done flag, so the host enqueues iterations in chunks and polls one flag per
// chunk. Dot products are produced as per-block partials and reduced by the kernel that consumes them
// (dev_common.hpp): deterministic, no fences, no atomics. A scalar written by block 0 of a kernel is never
// read by that kernel (values that change per iteration live in slots indexed by the iteration's parity).
// Vector updates keep the reference's operation order (SCAL then AXPY with 1.0, COPY then AXPY).
#include <cmath>
#include <cstdlib>

#include "dev_common.hpp"
#include "dfm_internal.hpp"
#include "kernels.hpp"
#include "solver.hpp"

namespace dfm {

constexpr int NT = 256;

__global__ void k_solver_init(double* sc, int* fl, double tol2, int maxit, double delta2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int k = 0; k < kNScal; k++) sc[k] = 0.0;
    sc[S_TOL2] = tol2;
    sc[S_DELTA2] = delta2;
    fl[F_DONE] = 0;
    fl[F_IT] = -1;
    fl[F_ITOUT] = maxit;
    fl[F_MAXIT] = maxit;
    fl[F_NREL] = 0;
    fl[F_PCUR] = 0;
    fl[F_PNEXT] = 0;
    fl[F_STALL] = 0;
  }
}

// The loop status into a pinned host slot: the flags (8 ints) and the scalars, then, after a system-scope fence,
// the chunk's sequence number at int 15, so that the host takes the slot as the chunk's only when it reads that
// number. A kernel on the loop's own queue, rather than a copy, which would wait for another engine.
__global__ void __launch_bounds__(64) k_status_out(const int* fl, const double* sc, int* hi, double* hd, int seq) {
  const int t = threadIdx.x;
  if (t < 8) hi[t] = fl[t];
  if (t < kNScal) hd[t] = sc[t];
  __threadfence_system();
  __syncthreads();
  if (t == 0) __hip_atomic_store(hi + 15, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T>
__global__ void __launch_bounds__(NT) k_vfill(T* x, T v, long long m, const int* done) {
  if (done && *done) return;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT) x[i] = v;
}

template <typename T>
__global__ void __launch_bounds__(NT) k_vcopy(const T* x, T* y, long long m, const int* done) {
  if (done && *done) return;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT) y[i] = x[i];
}

// partials of (a, b)
template <typename T>
__global__ void __launch_bounds__(NT) k_vdot(const T* a, const T* b, long long m, double* slot, const int* done) {
  if (done && *done) return;
  double v[1] = {0.0};
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT)
    v[0] += (double)a[i] * (double)b[i];
  produce<NT, 1>(slot, v);
}

// PCG start (solver_ls_iter.cpp:115-133): ||r0||^2 from partials; guard -> converged with ratio 0 at 0.
__global__ void __launch_bounds__(NT) k_pcg_start(const double* slot, int G, double guard, double* sc, int* fl) {
  double t[1];
  consume<NT, 1>(slot, G, t);
  if (threadIdx.x == 0) {
    sc[S_RR] = t[0];
    if (t[0] < guard) {
      sc[S_CONV] = 0.0;
      fl[F_ITOUT] = 0;
      fl[F_DONE] = 1;
    } else {
      sc[S_INV0] = 1.0 / t[0];
      sc[S_RMAX] = t[0];
      sc[S_TPREV] = t[0];
    }
  }
}

// rho_0 = (r, z) from partials
__global__ void __launch_bounds__(NT) k_pcg_rho0(const double* slot, int G, double* sc, const int* fl) {
  if (fl[F_DONE]) return;
  double t[1];
  consume<NT, 1>(slot, G, t);
  if (threadIdx.x == 0) sc[S_RZ] = t[0];
}

// q = A p, partials of (p, q). Block 0 advances the iteration counter (read only by later kernels).
// Mixed precision (REL): when a reliable update is pending, the same row sweep also forms the FP64 true
// residual b - A x with the FP64 matrix, stores it (rounded) as the new FP32 residual and produces the
// partials of its squared norm.
template <int B, int TPR, typename T, bool REL>
__global__ void __launch_bounds__(NT) k_pcg_spmv(int n, const unsigned* __restrict__ rp, const unsigned* __restrict__ ci,
                                                 const T* __restrict__ crs, const T* __restrict__ dia,
                                                 const T* __restrict__ p, T* __restrict__ q, double* slot,
                                                 const double* __restrict__ crs64, const double* __restrict__ dia64,
                                                 const double* __restrict__ x64, const double* __restrict__ b64,
                                                 T* __restrict__ r, double* slot_true, int* fl) {
  if (fl[F_DONE]) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) fl[F_IT] = fl[F_IT] + 1;
  const bool rel = REL && fl[F_PCUR];
  constexpr int RPB = NT / TPR;
  const int lane = threadIdx.x % TPR;
  double v[1] = {0.0}, vt[1] = {0.0};
  for (int row = blockIdx.x * RPB + threadIdx.x / TPR; row - (int)(threadIdx.x / TPR) < n; row += gridDim.x * RPB) {
    if (row < n) {
      T acc[B];
      spmv_row<B, TPR, T, T, T>(rp[row], rp[row + 1], ci, crs, p, lane, acc);
      if (lane == 0) {
        dia_row<B, T, T, T>(dia, p, row, acc);
#pragma unroll
        for (int k = 0; k < B; k++) {
          q[(size_t)row * B + k] = acc[k];
          v[0] += (double)p[(size_t)row * B + k] * (double)acc[k];
        }
      }
      if (REL && rel) {
        double a64[B];
        spmv_row<B, TPR, double, double, double>(rp[row], rp[row + 1], ci, crs64, x64, lane, a64);
        if (lane == 0) {
          dia_row<B, double, double, double>(dia64, x64, row, a64);
#pragma unroll
          for (int k = 0; k < B; k++) {
            const double t = b64[(size_t)row * B + k] - a64[k];
            r[(size_t)row * B + k] = (T)t;
            vt[0] += t * t;
          }
        }
      }
    }
  }
  produce<NT, 1>(slot, v);
  if (REL && rel) produce<NT, 1>(slot_true, vt);
}

// The same over the SELL-64 copy of the matrix: one slice (64 rows, one per lane) per wavefront, kSellWaves
// wavefronts per block; the row sums follow CRS order with the diagonal block last, as MatVec.
constexpr int kSellWaves = 16;
static_assert(kSellWaves == kSellSlicesPerWG && kWPB == kSweepWavesPerWG, "launch shapes of the host schedule");

template <int B, typename T, bool REL>
__global__ void __launch_bounds__(64 * kSellWaves)
    k_pcg_spmv_sell(int n, int npos, const int* __restrict__ visit, const int* __restrict__ sptr,
                    const unsigned* __restrict__ col,
                    const T* __restrict__ val, const T* __restrict__ dia, const T* __restrict__ p, T* __restrict__ q,
                    double* slot, const double* __restrict__ val64, const double* __restrict__ dia64,
                    const double* __restrict__ x64, const double* __restrict__ b64, T* __restrict__ r, double* slot_true,
                    int* fl) {
  constexpr int CH = SellChunk<B>::CH;
  if (fl[F_DONE]) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) fl[F_IT] = fl[F_IT] + 1;
  const bool rel = REL && fl[F_PCUR];
  const int wave = wave_in_block(), lane = threadIdx.x % 64;
  double v[1] = {0.0}, vt[1] = {0.0};
  for (int pos = (int)blockIdx.x * kSellWaves + wave; pos < npos; pos += (int)gridDim.x * kSellWaves) {
    const int sl = visit[pos];
    if (sl < 0) continue;
    const int r0 = sptr[sl], w = sptr[sl + 1] - r0;
    const int row = sl * 64 + lane;
    T acc[B];
    sell_row<B, CH, T, T, T>(col, val, r0, w, lane, p, acc);
    if (row < n) {
      dia_row<B, T, T, T>(dia, p, row, acc);
#pragma unroll
      for (int k = 0; k < B; k++) {
        q[(size_t)row * B + k] = acc[k];
        v[0] += (double)p[(size_t)row * B + k] * (double)acc[k];
      }
    }
    if (REL && rel) {
      double a64[B];
      sell_row<B, CH, double, double, double>(col, val64, r0, w, lane, x64, a64);
      if (row < n) {
        dia_row<B, double, double, double>(dia64, x64, row, a64);
#pragma unroll
        for (int k = 0; k < B; k++) {
          const double t = b64[(size_t)row * B + k] - a64[k];
          r[(size_t)row * B + k] = (T)t;
          vt[0] += t * t;
        }
      }
    }
  }
  produce<64 * kSellWaves, 1>(slot, v);
  if (REL && rel) produce<64 * kSellWaves, 1>(slot_true, vt);
}

// The mixed loop's SpMV with the direction update folded in: p = beta p + z (solver_ls_iter.cpp:160-166) and
// q = A p formed as beta q + A z, which is A (beta p + z) by linearity (the rounding differs from forming A p), so
// the SpMV gathers z alone and no separate pass over p is launched. Partials of (p, q). When a reliable update is
// due: r = b - A (x_hi + x_lo) in FP64 (the forward phases then fold x_lo into x_hi). q and p are read and written
// only for the lane's own rows; z, x_hi and x_lo are not written here.
template <int B, typename T, bool REL>
__global__ void __launch_bounds__(64 * kSellWaves)
    k_pcg_spmv_zq(int n, int npos, const int* __restrict__ visit, const int* __restrict__ sptr,
                  const unsigned* __restrict__ col, const T* __restrict__ val, const T* __restrict__ dia,
                  const T* __restrict__ z, T* __restrict__ p, T* __restrict__ q, const double* __restrict__ sc,
                  double* slot, const double* __restrict__ val64, const double* __restrict__ dia64,
                  const double* __restrict__ x64, const float* __restrict__ xlo, const double* __restrict__ b64,
                  T* __restrict__ r, double* slot_true, int* fl) {
  constexpr int CH = SellChunk<B>::CH;
  if (fl[F_DONE]) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) fl[F_IT] = fl[F_IT] + 1;
  const bool rel = REL && fl[F_PCUR];
  const T bt = (T)sc[S_BETA];
  const int wave = wave_in_block(), lane = threadIdx.x % 64;
  double v[1] = {0.0}, vt[1] = {0.0};
  for (int pos = (int)blockIdx.x * kSellWaves + wave; pos < npos; pos += (int)gridDim.x * kSellWaves) {
    const int sl = visit[pos];
    if (sl < 0) continue;
    const int r0 = sptr[sl], w = sptr[sl + 1] - r0;
    const int row = sl * 64 + lane;
    T acc[B];
    sell_row<B, CH, T, T, T>(col, val, r0, w, lane, z, acc);
    if (row < n) {
      dia_row<B, T, T, T>(dia, z, row, acc);
#pragma unroll
      for (int k = 0; k < B; k++) {
        const size_t i = (size_t)row * B + k;
        const T pn = p[i] * bt + z[i];
        const T qn = q[i] * bt + acc[k];
        p[i] = pn;
        q[i] = qn;
        v[0] += (double)pn * (double)qn;
      }
    }
    if (REL && rel) {
      const SumVec xs{x64, xlo};
      double a64[B];
      sell_row<B, CH, double, double, double>(col, val64, r0, w, lane, xs, a64);
      if (row < n) {
        dia_row<B, double, double, double>(dia64, xs, row, a64);
#pragma unroll
        for (int k = 0; k < B; k++) {
          const double t = b64[(size_t)row * B + k] - a64[k];
          r[(size_t)row * B + k] = (T)t;
          vt[0] += t * t;
        }
      }
    }
  }
  produce<64 * kSellWaves, 1>(slot, v);
  if (REL && rel) produce<64 * kSellWaves, 1>(slot_true, vt);
}

// ---- PCG iteration with the vector work folded into the sweeps ---------------------------------------------
// The update r -= alpha q, x += alpha p touches each row once, so every forward phase does it for its own rows
// (solver_ls_iter.cpp:139-144) and yields the partials of (r, r); every backward phase yields the partials
// of (r, z) for the rows it finalises. The first forward phase forms alpha from the SpMV's partials; a
// one-workgroup kernel forms the judgement with beta from the sweep partials. Sweep partials: one per workgroup,
// phase after phase.

// One-workgroup reductions of per-workgroup partials: every thread sums a fixed strided subset of each list with
// four independent accumulators (loads in flight together), then the workgroup sums; the order is fixed, so the
// totals are reproducible. Lists with G = 0 give 0.
constexpr int kRedNT = 1024;

template <int K>
__device__ inline void sum_partials(const double* const (&slot)[K], const int (&G)[K], double (&tot)[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int b = (int)threadIdx.x;
    for (; b + 3 * kRedNT < G[k]; b += 4 * kRedNT) {
      s0 += slot[k][b];
      s1 += slot[k][b + kRedNT];
      s2 += slot[k][b + 2 * kRedNT];
      s3 += slot[k][b + 3 * kRedNT];
    }
    for (; b < G[k]; b += kRedNT) s0 += slot[k][b];
    tot[k] = (s0 + s1) + (s2 + s3);
  }
  block_allsum<kRedNT, K>(tot);
}

// alpha = rho / (p, q) (solver_ls_iter.cpp:137-138) inside the first forward phase: every workgroup sums the
// SpMV's (p, q) partials in the same fixed order (so all hold the same alpha), workgroup 0 stores it for the later
// phases. Saves the launch of a separate reduction kernel between the SpMV and the sweeps.
__device__ inline double phase_alpha(const double* __restrict__ slot_pq, int G, double* sc, const int* fl) {
  __shared__ double red[kWPB];
  double s0 = 0.0, s1 = 0.0;
  int b = (int)threadIdx.x;
  for (; b + 64 * kWPB < G; b += 2 * 64 * kWPB) {
    s0 += slot_pq[b];
    s1 += slot_pq[b + 64 * kWPB];
  }
  if (b < G) s0 += slot_pq[b];
  const double v = wave_sum(s0 + s1);
  if (threadIdx.x % 64 == 0) red[wave_in_block()] = v;
  __syncthreads();
  double pq = red[0];
#pragma unroll
  for (int k = 1; k < kWPB; k++) pq += red[k];
  const double alpha = sc[S_RZ + (fl[F_IT] & 1)] / pq;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[S_ALPHA] = alpha;
  return alpha;
}

// x_lo of a row before alpha p is added; when a reliable update is due (the SpMV has just formed the true residual
// from x_hi + x_lo) x_lo is folded into the FP64 solution x_hi first and restarts from 0.
template <typename T>
__device__ inline T fold_xlo(double* xhi, bool fold, const T* x, size_t i) {
  if (!fold) return x[i];
  xhi[i] += (double)x[i];
  return T(0);
}

// forward phase with the update of its rows: r_i -= alpha q_i, x_i += alpha p_i, then z = L-sweep of r; the new
// r and x are kept in registers and stored at the end, so every load of the phase is issued up front.
// q and z may be one buffer: a lane reads q of its rows before the sweep writes z there, and the sweep reads z
// only of rows finished in earlier phases or in LDS.
template <int B, int MO, int MI, typename TV, typename T>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(SweepOccupancy<B, MO, MI, T>::kWaves)))
    k_fwd_update(int gbase, int gend, SlabArgs sa, const TV* __restrict__ val, const TV* __restrict__ dval, T* r,
                 const T* q, const T* p, T* x, T* z, double* sc, const int* fl, double* slot_rr,
                 const double* __restrict__ slot_pq, int Gpq, double* xhi) {
  __shared__ T xsb[kWPB][64 * B];
  if (fl[F_DONE]) return;
  const int w = wave_in_block();
  const int g = gbase + (int)blockIdx.x * kWPB + w;
  double rr = 0.0;
  const double alpha = slot_pq ? phase_alpha(slot_pq, Gpq, sc, fl) : sc[S_ALPHA];
  const bool fold = xhi && fl[F_PCUR];
  if (g < gend) {
    const T ma = (T)(-alpha), pa = (T)alpha;
    T rn[B], xn[B];
    SlabSlot ss;
    T t[B];
    sweep_fwd_core<B, MO, MI, TV, T, T>(
        sa, g, val, dval,
        [&](const SlabSlot& s, T(&u)[B]) {
#pragma unroll
          for (int k = 0; k < B; k++) {
            const size_t i = (size_t)s.row * B + k;
            rn[k] = r[i] + ma * q[i];
            xn[k] = fold_xlo(xhi, fold, x, i) + pa * p[i];
            u[k] = rn[k];
          }
        },
        z, nullptr, xsb[w], ss, t);
    if (ss.active)
#pragma unroll
      for (int k = 0; k < B; k++) {
        const size_t i = (size_t)ss.row * B + k;
        r[i] = rn[k];
        x[i] = xn[k];
        rr += (double)rn[k] * (double)rn[k];
      }
  }
  double v[1] = {rr};
  block_allsum<64 * kWPB, 1>(v);
  if (threadIdx.x == 0) slot_rr[blockIdx.x] = v[0];
}

// backward phase, then the partials of (r, z) of its rows
template <int B, int MO, int MI, typename TV, typename T>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(SweepOccupancy<B, MO, MI, T>::kWaves)))
    k_bwd_rz(int gbase, int gend, SlabArgs sa, const TV* __restrict__ val, const T* __restrict__ r, T* z,
             const int* fl, double* slot_rz) {
  __shared__ T xsb[kWPB][64 * B];
  if (fl[F_DONE]) return;
  const int w = wave_in_block();
  const int g = gbase + (int)blockIdx.x * kWPB + w;
  double rz = 0.0;
  if (g < gend) {
    SlabSlot ss;
    T t[B];
    sweep_bwd_core<B, MO, MI, TV, T, T>(sa, g, val, z, nullptr, xsb[w], ss, t, [&](const SlabSlot& s, T(&u)[B]) {
#pragma unroll
      for (int k = 0; k < B; k++) u[k] = z[(size_t)s.row * B + k];
    });
    if (ss.active)
#pragma unroll
      for (int k = 0; k < B; k++) rz += (double)r[(size_t)ss.row * B + k] * (double)t[k];
  }
  double v[1] = {rz};
  block_allsum<64 * kWPB, 1>(v);
  if (threadIdx.x == 0) slot_rz[blockIdx.x] = v[0];
}

// The last forward phase fused with the first backward phase when they hold the same tiles: per tile, the
// forward sweep with the r/x update, then the backward sweep starting from the forward values the lanes hold (the
// backward groups of this phase have no entries outside their tile); (r, r) and (r, z) partials of the rows.
template <int B, int MO, int MI, typename TV, typename T>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(SweepOccupancy<B, MO, MI, T>::kWaves)))
    k_fwd_bwd_tile(int gbase, int gend, SlabArgs fa, const TV* __restrict__ fval, const TV* __restrict__ fdval,
                   SlabArgs ba, const TV* __restrict__ bval, const int* __restrict__ bgrp, T* r, const T* q,
                   const T* p, T* x, T* z, double* sc, const int* fl, double* slot_rr, double* slot_rz,
                   const double* __restrict__ slot_pq, int Gpq, double* xhi) {
  __shared__ T xsb[kWPB][64 * B];
  if (fl[F_DONE]) return;
  const int w = wave_in_block();
  const int k = (int)blockIdx.x * kWPB + w;
  const int g = gbase + k;
  double v[2] = {0.0, 0.0};
  const double alpha = slot_pq ? phase_alpha(slot_pq, Gpq, sc, fl) : sc[S_ALPHA];
  const bool fold = xhi && fl[F_PCUR];
  if (g < gend) {
    const T ma = (T)(-alpha), pa = (T)alpha;
    T rn[B], xn[B];
    SlabSlot ss;
    T t[B];
    sweep_fwd_core<B, MO, MI, TV, T, T>(
        fa, g, fval, fdval,
        [&](const SlabSlot& s, T(&u)[B]) {
#pragma unroll
          for (int c = 0; c < B; c++) {
            const size_t i = (size_t)s.row * B + c;
            rn[c] = r[i] + ma * q[i];
            xn[c] = fold_xlo(xhi, fold, x, i) + pa * p[i];
            u[c] = rn[c];
          }
        },
        z, nullptr, xsb[w], ss, t);
    if (ss.active)
#pragma unroll
      for (int c = 0; c < B; c++) {
        const size_t i = (size_t)ss.row * B + c;
        r[i] = rn[c];
        x[i] = xn[c];
        v[0] += (double)rn[c] * (double)rn[c];
      }
    wave_lds_sync();
    SlabSlot sb;
    T u[B];
    sweep_bwd_core<B, MO, MI, TV, T, T>(ba, bgrp[k], bval, z, nullptr, xsb[w], sb, u, [&](const SlabSlot&, T(&a)[B]) {
#pragma unroll
      for (int c = 0; c < B; c++) a[c] = t[c];
    });
    if (sb.active)
#pragma unroll
      for (int c = 0; c < B; c++) v[1] += (double)rn[c] * (double)u[c];
  }
  block_allsum<64 * kWPB, 2>(v);
  if (threadIdx.x == 0) {
    slot_rr[blockIdx.x] = v[0];
    slot_rz[blockIdx.x] = v[1];
  }
}

// Judgement and beta in one workgroup. FP64 (solver_ls_iter.cpp:146-166): on the recursive residual. Mixed: on
// the true residual when one was formed this iteration, else the recursive residual only schedules the next
// reliable update; the scheduled update becomes current for the next SpMV.
template <bool MX>
__global__ void __launch_bounds__(kRedNT) k_pcg_judge_beta(const double* slot_rr, int Grr, const double* slot_true,
                                                           int Gt, const double* slot_rz, int Grz, double* sc, int* fl) {
  if (fl[F_DONE]) return;
  const double* const sl[3] = {slot_rr, slot_rz, slot_true};
  const int g[3] = {Grr, Grz, MX && fl[F_PCUR] ? Gt : 0};
  double tot[3];
  sum_partials<3>(sl, g, tot);
  const double rr[1] = {tot[0]}, rz[1] = {tot[1]}, rt[1] = {tot[2]};
  if (threadIdx.x) return;
  const int it = fl[F_IT];
  sc[S_RR] = rr[0];
  if (!MX) {
    const double ratio2 = rr[0] * sc[S_INV0];
    if (ratio2 < sc[S_TOL2] || it + 1 >= fl[F_MAXIT]) {
      sc[S_CONV] = sqrt(ratio2);
      if (ratio2 < sc[S_TOL2]) {
        fl[F_ITOUT] = it;
        fl[F_DONE] = 1;
      } else {
        fl[F_DONE] = 2;
      }
      return;
    }
  } else {
    double rmax = sc[S_RMAX + (it & 1)];
    if (fl[F_PCUR]) {
      const double ratio2 = rt[0] * sc[S_INV0];
      if (ratio2 < sc[S_TOL2]) {
        sc[S_CONV] = sqrt(ratio2);
        sc[S_RR] = rt[0];
        fl[F_ITOUT] = it - 1;
        fl[F_DONE] = 1;
        return;
      }
      if (mx_slow(rt[0], sc[S_RRTRIG], sc[S_TPREV], sc[S_SLOW], sc[S_TBEST], fl[F_STALL])) {
        sc[S_CONV] = sqrt(ratio2);
        sc[S_RR] = rt[0];
        fl[F_DONE] = 3;
        return;
      }
      rmax = rt[0];
    }
    // the last iteration max_iter allows: stop, unless the recursive residual says converged, which schedules
    // the true-residual check of this iteration's solution in one more iteration (the solution reported is
    // still this one, index it < max_iter, as in the reference's counting)
    if (it + 1 >= fl[F_MAXIT] && !(it + 1 == fl[F_MAXIT] && rr[0] * sc[S_INV0] < sc[S_TOL2])) {
      sc[S_CONV] = sqrt(rr[0] * sc[S_INV0]);
      fl[F_DONE] = 2;
      return;
    }
    const bool due = rr[0] * sc[S_INV0] < sc[S_TOL2] || rr[0] < sc[S_DELTA2] * rmax ||
                     it - (int)sc[S_LASTUPD] >= kForceUpdate;
    sc[S_RMAX + ((it + 1) & 1)] = due ? rr[0] : fmax(rmax, rr[0]);
    fl[F_PNEXT] = due ? 1 : 0;
    if (due) {
      fl[F_NREL] = fl[F_NREL] + 1;
      sc[S_LASTUPD] = (double)it;
      sc[S_RRTRIG] = rr[0];
    }
  }
  sc[S_BETA] = rz[0] / sc[S_RZ + (it & 1)];
  sc[S_RZ + ((it + 1) & 1)] = rz[0];
  if (MX) fl[F_PCUR] = fl[F_PNEXT];
}

// p = beta p; p += z; mixed: the FP64 solution takes the partial solution when the next SpMV forms the true
// residual
template <typename T, bool MX>
__global__ void __launch_bounds__(NT) k_pcg_pupdate(long long m, T* p, const T* z, const double* sc, const int* fl,
                                                     float* xlo, double* x64) {
  if (fl[F_DONE]) return;
  const T bt = (T)sc[S_BETA];
  const bool fold = MX && fl[F_PCUR];
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT) {
    const T t = p[i] * bt;
    p[i] = t + z[i];
    if (MX && fold) {
      x64[i] += (double)xlo[i];
      xlo[i] = 0.0f;
    }
  }
}

// ---- mixed precision: FP32 PCG with reliable updates ------------------------------------------------------
__global__ void __launch_bounds__(NT) k_mx_to32(long long m, const double* a, float* b) {
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT) b[i] = (float)a[i];
}
__global__ void __launch_bounds__(NT) k_mx_final_fold(long long m, float* xlo, double* x) {
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT) {
    x[i] += (double)xlo[i];
    xlo[i] = 0.0f;
  }
}

// ---- PBiCGSTAB ------------------------------------------------------------------------------------------
// y = A x with partials of (y, w) and (u, w) [mode 0: w = r2, u = r] or (y, y) and (w, y) [mode 1: w = s].
template <int B, int TPR>
__global__ void __launch_bounds__(NT) k_bi_spmv(int n, const unsigned* __restrict__ rp, const unsigned* __restrict__ ci,
                                                const double* __restrict__ crs, const double* __restrict__ dia,
                                                const double* __restrict__ x, double* __restrict__ y,
                                                const double* __restrict__ w, const double* __restrict__ u, int mode,
                                                double* slot, int* fl) {
  if (fl[F_DONE]) return;
  if (mode == 0 && blockIdx.x == 0 && threadIdx.x == 0) fl[F_IT] = fl[F_IT] + 1;
  constexpr int RPB = NT / TPR;
  const int lane = threadIdx.x % TPR;
  double v[2] = {0.0, 0.0};
  for (int row = blockIdx.x * RPB + threadIdx.x / TPR; row - (int)(threadIdx.x / TPR) < n; row += gridDim.x * RPB) {
    if (row < n) {
      double acc[B];
      spmv_row<B, TPR, double, double, double>(rp[row], rp[row + 1], ci, crs, x, lane, acc);
      if (lane == 0) {
        dia_row<B, double, double, double>(dia, x, row, acc);
#pragma unroll
        for (int r = 0; r < B; r++) {
          const size_t k = (size_t)row * B + r;
          y[k] = acc[r];
          if (mode == 0) {
            v[0] += acc[r] * w[k];  // (AMp, r2)
            v[1] += u[k] * w[k];    // (r, r2)
          } else {
            v[0] += acc[r] * acc[r];  // (AMs, AMs)
            v[1] += w[k] * acc[r];    // (s, AMs)
          }
        }
      }
    }
  }
  produce<NT, 2>(slot, v);
}

// alpha = (r, r2) / (AMp, r2); s = r; s += (-alpha) AMp
__global__ void __launch_bounds__(NT) k_bi_s(long long m, const double* r, const double* amp, double* s,
                                             const double* slot2, int G2, const int* fl) {
  if (fl[F_DONE]) return;
  double t[2];
  consume<NT, 2>(slot2, G2, t);
  const double alpha = t[1] / t[0];
  const double ma = -alpha;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT) s[i] = r[i] + ma * amp[i];
}

// omega = (s, AMs) / (AMs, AMs); x += alpha Mp; x += omega Ms; r = s; r += (-omega) AMs; partials (r,r), (r,r2)
__global__ void __launch_bounds__(NT) k_bi_update(long long m, double* x, double* r, const double* mp, const double* ms,
                                                  const double* s, const double* ams, const double* r2,
                                                  const double* slot2, int G2, const double* slot5, int G5,
                                                  double* slot6, const int* fl) {
  if (fl[F_DONE]) return;
  double t2[2], t5[2];
  consume<NT, 2>(slot2, G2, t2);
  consume<NT, 2>(slot5, G5, t5);
  const double alpha = t2[1] / t2[0];
  const double omega = t5[1] / t5[0];
  double v[2] = {0.0, 0.0};
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT) {
    double xi = x[i] + alpha * mp[i];
    xi = xi + omega * ms[i];
    x[i] = xi;
    const double ri = s[i] + (-omega) * ams[i];
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += ri * r2[i];
  }
  produce<NT, 2>(slot6, v);
}

// judgement on (r, r) (:360-368); beta = (r,r2)_new * alpha / ((r,r2) * omega); p = beta p; p += r;
// p += (-beta*omega) AMp
__global__ void __launch_bounds__(NT) k_bi_pupdate(long long m, double* p, const double* r, const double* amp,
                                                   const double* slot2, int G2, const double* slot5, int G5,
                                                   const double* slot6, int G6, double* sc, int* fl) {
  if (fl[F_DONE]) return;
  double t2[2], t5[2], t6[2];
  consume<NT, 2>(slot6, G6, t6);
  const int it = fl[F_IT];
  const double sqc = t6[0] * sc[S_INV0];
  if (sqc < sc[S_TOL2] || it + 2 >= fl[F_MAXIT]) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      if (sqc < sc[S_TOL2]) {
        sc[S_CONV] = sqrt(t6[0] * sc[S_INV0]);
        fl[F_ITOUT] = it;
        fl[F_DONE] = 1;
      } else {
        fl[F_DONE] = 2;  // iitr = it + 1 was the last one below max_iter
      }
    }
    return;
  }
  consume<NT, 2>(slot2, G2, t2);
  consume<NT, 2>(slot5, G5, t5);
  const double alpha = t2[1] / t2[0];
  const double omega = t5[1] / t5[0];
  const double beta = t6[1] * alpha / (t2[1] * omega);
  const double bo = -beta * omega;
  for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < m; i += (long long)gridDim.x * NT) {
    double t = p[i] * beta;
    t = t + r[i];
    p[i] = t + bo * amp[i];
  }
}

// =========================================================================================================
// host side

static dim3 vgrid(long long m) { return dim3(vec_blocks(m, NT)); }

static unsigned spmv_grid(const Sys& s, int tpr) {
  const long long rows_per_block = NT / tpr;
  return vec_blocks((s.n + rows_per_block - 1) / rows_per_block, 1);
}

// z = M^-1 src over the sweep slabs: TV the factor values (FP64 or the FP32 copy), V the vectors and the
// arithmetic.
template <int B, typename TV, typename V>
static hipError_t precond_T(const Sys& s, const V* src, V* z, const int* done, hipStream_t st, int dirs = 3) {
  constexpr bool f32 = sizeof(TV) == 4;
  const Sweep& f = s.fwd;
  const SlabArgs fa{f.d_gmeta.p, f.d_srow.p, f.d_scnt.p, f.d_ecol.p};
  const TV* fval = f32 ? (const TV*)f.d_eval32.p : (const TV*)f.d_eval.p;
  const TV* fdval = f32 ? (const TV*)f.d_dval32.p : (const TV*)f.d_dval.p;
  for (int p = 0; p < f.nphase() && (dirs & 1); p++) {
    const int g0 = f.phase_gptr[p], ng = f.phase_gptr[p + 1] - g0;
    if (!ng) continue;
    with_slab_caps(f.phase_wo[p], f.phase_wi[p], B == 3 && sizeof(V) == 8 && sizeof(TV) == 8, [&]<int MO, int MI>() {
      hipLaunchKernelGGL((k_sweep_fwd<B, MO, MI, TV, V, V, V>), dim3((ng + kWPB - 1) / kWPB), dim3(64 * kWPB), 0, st,
                         g0, g0 + ng, fa, fval, fdval, src, z, done);
    });
  }
  const Sweep& b = s.bwd;
  const SlabArgs ba{b.d_gmeta.p, b.d_srow.p, b.d_scnt.p, b.d_ecol.p};
  const TV* bval = f32 ? (const TV*)b.d_eval32.p : (const TV*)b.d_eval.p;
  for (int p = 0; p < b.nphase() && (dirs & 2); p++) {
    const int g0 = b.phase_gptr[p], ng = b.phase_gptr[p + 1] - g0;
    if (!ng) continue;
    with_slab_caps(b.phase_wo[p], b.phase_wi[p], B == 3 && sizeof(V) == 8 && sizeof(TV) == 8, [&]<int MO, int MI>() {
      hipLaunchKernelGGL((k_sweep_bwd<B, MO, MI, TV, V, V>), dim3((ng + kWPB - 1) / kWPB), dim3(64 * kWPB), 0, st, g0,
                         g0 + ng, ba, bval, z, done);
    });
  }
  return hipGetLastError();
}

hipError_t precond_apply64_dirs(const Sys& s, double* z, int dirs, hipStream_t st) {
  if (s.len == 1) return precond_T<1, double, double>(s, z, z, nullptr, st, dirs);
  if (s.len == 2) return precond_T<2, double, double>(s, z, z, nullptr, st, dirs);
  if (s.len == 3) return precond_T<3, double, double>(s, z, z, nullptr, st, dirs);
  return hipErrorInvalidValue;
}

hipError_t precond_apply64(const Sys& s, const double* src, double* z, const int* done, hipStream_t st) {
  if (s.len == 1) return precond_T<1, double, double>(s, src, z, done, st);
  if (s.len == 2) return precond_T<2, double, double>(s, src, z, done, st);
  if (s.len == 3) return precond_T<3, double, double>(s, src, z, done, st);
  return hipErrorInvalidValue;
}

// FP64 vectors and arithmetic, FP32 factor values
static hipError_t precond_apply_f32ilu(const Sys& s, const double* src, double* z, const int* done, hipStream_t st) {
  if (s.len == 1) return precond_T<1, float, double>(s, src, z, done, st);
  if (s.len == 2) return precond_T<2, float, double>(s, src, z, done, st);
  if (s.len == 3) return precond_T<3, float, double>(s, src, z, done, st);
  return hipErrorInvalidValue;
}

static hipError_t precond_apply32(const Sys& s, const float* src, float* z, const int* done, hipStream_t st) {
  if (s.len == 1) return precond_T<1, float, float>(s, src, z, done, st);
  if (s.len == 2) return precond_T<2, float, float>(s, src, z, done, st);
  if (s.len == 3) return precond_T<3, float, float>(s, src, z, done, st);
  return hipErrorInvalidValue;
}

static hipError_t precond_apply_d(const Sys& s, const PcgBufs& v, const double* src, double* z, const int* done,
                                  hipStream_t st) {
  return v.f32ilu ? precond_apply_f32ilu(s, src, z, done, st) : precond_apply64(s, src, z, done, st);
}

// Workgroups of the SELL SpMV over npos positions (kSellWaves per workgroup, a multiple of 8 so that a strided
// grid keeps positions on their XCD): one position per wavefront for smaller systems; from 512 workgroups on, two
// per wavefront (measured on a 1M-row Poisson system: 14.4 -> 12.8 us FP32, the chip still full and fewer
// workgroups draining at the end).
static unsigned sell_grid(int npos) {
  unsigned G = (unsigned)std::min<long long>(kMaxPartial, npos / kSellWaves);
  if (G >= 512) G = G / 2 / 8 * 8;
  return G;
}

template <typename T, bool REL>
static hipError_t pcg_spmv(const Sys& s, const T* crs, const T* dia, const T* p, T* q, double* slot, unsigned& G,
                           const double* x64, const double* b64, T* r, double* slot_true, hipStream_t st) {
  if (s.sell.nslice() > 0) {
    // kMaxPartial is a multiple of 8, so a strided grid keeps every position on the XCD it was placed for
    const int npos = (int)s.sell.visit.size();
    G = sell_grid(npos);
    const T* sval = sizeof(T) == 4 ? (const T*)s.sell.d_val32.p : (const T*)s.sell.d_val.p;
#define DFM_SS(B)                                                                                              \
  if (s.len == B) {                                                                                            \
    hipLaunchKernelGGL((k_pcg_spmv_sell<B, T, REL>), dim3(G), dim3(64 * kSellWaves), 0, st, s.n, npos,         \
                       s.sell.d_visit.p, s.sell.d_sptr.p, s.sell.d_col.p, sval, dia, p, q, slot, s.sell.d_val.p,  \
                       s.d_dia.p, x64, b64, r, slot_true, s.d_iflag.p);                                         \
    return hipGetLastError();                                                                                  \
  }
    DFM_SS(1) DFM_SS(2) DFM_SS(3)
#undef DFM_SS
    return hipErrorInvalidValue;
  }
  const int tpr = spmv_tpr(s.ncrs, s.n);
  G = spmv_grid(s, tpr);
#define DFM_SD(B, TP)                                                                                          \
  if (s.len == B && tpr == TP) {                                                                               \
    hipLaunchKernelGGL((k_pcg_spmv<B, TP, T, REL>), dim3(G), dim3(NT), 0, st, s.n, s.d_rowptr.p, s.d_colind.p, \
                       crs, dia, p, q, slot, s.d_crs.p, s.d_dia.p, x64, b64, r, slot_true, s.d_iflag.p);       \
    return hipGetLastError();                                                                                  \
  }
  DFM_SD(1, 4) DFM_SD(1, 8) DFM_SD(1, 16) DFM_SD(1, 32)
  DFM_SD(2, 4) DFM_SD(2, 8) DFM_SD(2, 16) DFM_SD(2, 32)
  DFM_SD(3, 4) DFM_SD(3, 8) DFM_SD(3, 16) DFM_SD(3, 32)
#undef DFM_SD
  return hipErrorInvalidValue;
}

static hipError_t bi_spmv(const Sys& s, const double* x, double* y, const double* w, const double* u, int mode,
                          double* slot, unsigned& G, hipStream_t st) {
  const int tpr = spmv_tpr(s.ncrs, s.n);
  G = spmv_grid(s, tpr);
#define DFM_BS(B, TP)                                                                                         \
  if (s.len == B && tpr == TP) {                                                                              \
    hipLaunchKernelGGL((k_bi_spmv<B, TP>), dim3(G), dim3(NT), 0, st, s.n, s.d_rowptr.p, s.d_colind.p,        \
                       s.d_crs.p, s.d_dia.p, x, y, w, u, mode, slot, s.d_iflag.p);                            \
    return hipGetLastError();                                                                                 \
  }
  DFM_BS(1, 4) DFM_BS(1, 8) DFM_BS(1, 16) DFM_BS(1, 32)
  DFM_BS(2, 4) DFM_BS(2, 8) DFM_BS(2, 16) DFM_BS(2, 32)
  DFM_BS(3, 4) DFM_BS(3, 8) DFM_BS(3, 16) DFM_BS(3, 32)
#undef DFM_BS
  return hipErrorInvalidValue;
}

static double* slot(Sys& s, int k) { return s.d_partial.p + (size_t)k * kMaxPartial; }

static int poll(Sys& s, hipStream_t st) {
  DFM_HIP(hipMemcpyAsync(s.h_pinned, s.d_iflag.p, sizeof(int) * 8, hipMemcpyDeviceToHost, st));
  DFM_HIP(hipMemcpyAsync(s.h_pinned_d, s.d_scal.p, sizeof(double) * kNScal, hipMemcpyDeviceToHost, st));
  DFM_HIP(hipStreamSynchronize(st));
  return DFM_OK;
}

// sweep partial slots: one per workgroup of each phase
static int sweep_blocks(const Sweep& w) {
  int nb = 0;
  for (int p = 0; p < w.nphase(); p++) nb += (w.phase_gptr[p + 1] - w.phase_gptr[p] + kWPB - 1) / kWPB;
  return nb;
}

template <typename T, typename TV>
static hipError_t fused_sweeps(Sys& s, const PcgBufs& v, const TV* fval, const TV* fdval, const TV* bval,
                               double* part_rr, double* part_rz, const double* slot_pq, int Gpq, double* xhi,
                               hipStream_t st) {
  int* fl = s.d_iflag.p;
  double* sc = s.d_scal.p;
  T* r = (T*)v.r;
  T* z = (T*)v.z;
  const Sweep& f = s.fwd;
  const SlabArgs fa{f.d_gmeta.p, f.d_srow.p, f.d_scnt.p, f.d_ecol.p};
  const Sweep& b = s.bwd;
  const SlabArgs ba{b.d_gmeta.p, b.d_srow.p, b.d_scnt.p, b.d_ecol.p};
  const bool fuse = !f.fuse_bgrp.empty();
  int off = 0;
  for (int p = 0; p < f.nphase(); p++) {
    const int g0 = f.phase_gptr[p], ng = f.phase_gptr[p + 1] - g0, nb = (ng + kWPB - 1) / kWPB;
    if (!ng) continue;
    const bool narrow = sizeof(T) == 8 && sizeof(TV) == 8;
    // the first launched forward phase forms alpha from the SpMV partials
    const double* apq = slot_pq;
    slot_pq = nullptr;
    if (fuse && p == f.nphase() - 1) {
      const int wo = std::max(f.phase_wo[p], b.phase_wo[0]), wi = std::max(f.phase_wi[p], b.phase_wi[0]);
#define DFM_FB(B)                                                                                                  \
  if (s.len == B)                                                                                                  \
    with_slab_caps(wo, wi, B == 3 && narrow, [&]<int MO, int MI>() {                                               \
      hipLaunchKernelGGL((k_fwd_bwd_tile<B, MO, MI, TV, T>), dim3(nb), dim3(64 * kWPB), 0, st, g0, g0 + ng, fa, fval, \
                         fdval, ba, bval, (const int*)f.d_fuse_bgrp.p, r, (const T*)v.q, (const T*)v.p, (T*)v.x, z,  \
                         sc, fl, part_rr + off, part_rz, apq, Gpq, xhi);                                           \
    });
      DFM_FB(1) DFM_FB(2) DFM_FB(3)
#undef DFM_FB
      off += nb;
      continue;
    }
#define DFM_FU(B)                                                                                                    \
  if (s.len == B)                                                                                                    \
    with_slab_caps(f.phase_wo[p], f.phase_wi[p], B == 3 && narrow, [&]<int MO, int MI>() {                           \
      hipLaunchKernelGGL((k_fwd_update<B, MO, MI, TV, T>), dim3(nb), dim3(64 * kWPB), 0, st, g0, g0 + ng, fa, fval,  \
                         fdval, r, (const T*)v.q, (const T*)v.p, (T*)v.x, z, sc, fl, part_rr + off, apq, Gpq, xhi);  \
    });
    DFM_FU(1) DFM_FU(2) DFM_FU(3)
#undef DFM_FU
    off += nb;
  }
  off = 0;
  for (int p = 0; p < b.nphase(); p++) {
    const int g0 = b.phase_gptr[p], ng = b.phase_gptr[p + 1] - g0, nb = (ng + kWPB - 1) / kWPB;
    if (!ng) continue;
    if (fuse && p == 0) {  // done with the last forward phase
      off += nb;
      continue;
    }
    const bool narrow = sizeof(T) == 8 && sizeof(TV) == 8;
#define DFM_BZ(B)                                                                                                  \
  if (s.len == B)                                                                                                  \
    with_slab_caps(b.phase_wo[p], b.phase_wi[p], B == 3 && narrow, [&]<int MO, int MI>() {                         \
      hipLaunchKernelGGL((k_bwd_rz<B, MO, MI, TV, T>), dim3(nb), dim3(64 * kWPB), 0, st, g0, g0 + ng, ba, bval,     \
                         (const T*)r, z, fl, part_rz + off);                                                       \
    });
    DFM_BZ(1) DFM_BZ(2) DFM_BZ(3)
#undef DFM_BZ
    off += nb;
  }
  return hipGetLastError();
}

static hipError_t pcg_spmv_zq(Sys& s, const PcgBufs& v, unsigned& G, hipStream_t st) {
  const int npos = (int)s.sell.visit.size();
  G = sell_grid(npos);
#define DFM_ZQ(B)                                                                                                 \
  if (s.len == B) {                                                                                               \
    hipLaunchKernelGGL((k_pcg_spmv_zq<B, float, true>), dim3(G), dim3(64 * kSellWaves), 0, st, s.n, npos,         \
                       s.sell.d_visit.p, s.sell.d_sptr.p, s.sell.d_col.p, s.sell.d_val32.p, (const float*)v.dia,   \
                       (const float*)v.z, (float*)v.p, (float*)v.q, (const double*)s.d_scal.p, slot(s, P_A),       \
                       s.sell.d_val.p, s.d_dia.p, v.x64, (const float*)v.x, v.b64, (float*)v.r, slot(s, P_C),      \
                       s.d_iflag.p);                                                                              \
    return hipGetLastError();                                                                                     \
  }
  DFM_ZQ(1) DFM_ZQ(2) DFM_ZQ(3)
#undef DFM_ZQ
  return hipErrorInvalidValue;
}

template <typename T>
static int pcg_iteration(Sys& s, const PcgBufs& v, bool mixed, hipStream_t st) {
  const long long m = (long long)s.n * s.len;
  const dim3 vg = vgrid(m);
  unsigned Gs = 0;
  int* fl = s.d_iflag.p;
  double* sc = s.d_scal.p;
  T* r = (T*)v.r;
  T* z = (T*)v.z;
  T* p = (T*)v.p;
  T* q = (T*)v.q;
  const int nbf = sweep_blocks(s.fwd), nbb = sweep_blocks(s.bwd);
  double* part_rr = s.d_part_sw.p;
  double* part_rz = s.d_part_sw.p + nbf;
  hipError_t e;
  // mixed, scalar systems over the SELL copy: the direction update rides in the SpMV and x_lo is folded in the
  // forward phases (for 2x2 and 3x3 blocks the strided own-row traffic costs the SpMV more than the separate pass)
  const bool zq = mixed && s.sell.nslice() > 0 && s.len == 1;
  if (zq)
    e = pcg_spmv_zq(s, v, Gs, st);
  else if (mixed)
    e = pcg_spmv<T, true>(s, (const T*)v.crs, (const T*)v.dia, p, q, slot(s, P_A), Gs, v.x64, v.b64, r, slot(s, P_C), st);
  else
    e = pcg_spmv<T, false>(s, (const T*)v.crs, (const T*)v.dia, p, q, slot(s, P_A), Gs, nullptr, nullptr, nullptr,
                           nullptr, st);
  if (e != hipSuccess) return DFM_ERR_HIP;
  const double* pq = slot(s, P_A);
  if (mixed)
    e = fused_sweeps<float, float>(s, v, s.fwd.d_eval32.p, s.fwd.d_dval32.p, s.bwd.d_eval32.p, part_rr, part_rz, pq,
                                   (int)Gs, zq ? v.x64 : nullptr, st);
  else if (v.f32ilu)
    e = fused_sweeps<double, float>(s, v, s.fwd.d_eval32.p, s.fwd.d_dval32.p, s.bwd.d_eval32.p, part_rr, part_rz, pq,
                                    (int)Gs, nullptr, st);
  else
    e = fused_sweeps<double, double>(s, v, s.fwd.d_eval.p, s.fwd.d_dval.p, s.bwd.d_eval.p, part_rr, part_rz, pq,
                                     (int)Gs, nullptr, st);
  if (e != hipSuccess) return DFM_ERR_HIP;
  if (mixed) {
    hipLaunchKernelGGL(k_pcg_judge_beta<true>, dim3(1), dim3(kRedNT), 0, st, (const double*)part_rr, nbf,
                       (const double*)slot(s, P_C), (int)Gs, (const double*)part_rz, nbb, sc, fl);
    if (!zq)
      hipLaunchKernelGGL((k_pcg_pupdate<T, true>), vg, dim3(NT), 0, st, m, p, (const T*)z, (const double*)sc,
                         (const int*)fl, (float*)v.x, v.x64);
  } else {
    hipLaunchKernelGGL(k_pcg_judge_beta<false>, dim3(1), dim3(kRedNT), 0, st, (const double*)part_rr, nbf,
                       (const double*)nullptr, 0, (const double*)part_rz, nbb, sc, fl);
    hipLaunchKernelGGL((k_pcg_pupdate<T, false>), vg, dim3(NT), 0, st, m, p, (const T*)z, (const double*)sc,
                       (const int*)fl, (float*)nullptr, (double*)nullptr);
  }
  return hipGetLastError() == hipSuccess ? DFM_OK : DFM_ERR_HIP;
}

// Every device pointer an iteration's launches take, and the generation of the schedule whose launch shapes
// the capture baked in: a captured graph is valid while these are unchanged (re-ordering, a new fill level,
// re-factoring or growing a work buffer may move them).
static std::vector<const void*> graph_key(const Sys& s, const PcgBufs& v) {
  std::vector<const void*> k{v.crs, v.dia, v.r, v.z, v.p, v.q, v.x, v.x64, v.b64, s.d_crs.p, s.d_dia.p,
                             s.d_rowptr.p, s.d_colind.p, s.d_partial.p, s.d_part_sw.p, s.d_scal.p, s.d_iflag.p,
                             s.sell.d_sptr.p, s.sell.d_visit.p, s.sell.d_col.p, s.sell.d_val.p, s.sell.d_val32.p};
  for (const Sweep* w : {&s.fwd, &s.bwd})
    for (const void* p : {(const void*)w->d_gmeta.p, (const void*)w->d_srow.p, (const void*)w->d_scnt.p,
                          (const void*)w->d_ecol.p, (const void*)w->d_eval.p, (const void*)w->d_dval.p,
                          (const void*)w->d_eval32.p, (const void*)w->d_dval32.p})
      k.push_back(p);
  k.push_back((const void*)(size_t)v.f32ilu);
  k.push_back((const void*)s.fwd.d_fuse_bgrp.p);
  k.push_back((const void*)(size_t)s.struct_gen);
  return k;
}

// A chunk of `k` iterations captured once as a hipGraph and replayed (the kernels take every scalar from
// device memory, so a replay is exactly the same launches). Captured again when a pointer it holds changes.
template <typename F>
static int graph_chunk(Sys& s, int kind, std::vector<const void*> key, F enqueue_iteration, hipStream_t st) {
  if (s.gexec[kind] && s.gkey[kind] == key) return DFM_OK;
  if (s.gexec[kind]) {
    (void)hipGraphExecDestroy(s.gexec[kind]);
    s.gexec[kind] = nullptr;
  }
  hipGraph_t g = nullptr;
  DFM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = DFM_OK;
  for (int t = 0; t < s.gchunk && rc == DFM_OK; t++) rc = enqueue_iteration();
  hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != DFM_OK) return rc;
  DFM_HIP(e);
  e = hipGraphInstantiate(&s.gexec[kind], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  DFM_HIP(e);
  s.gkey[kind] = std::move(key);
  return DFM_OK;
}

// Iterations to enqueue before the next poll: from the decay rate of the residual between the last two
// polls, the estimated number left to reach the tolerance (plus a margin), within [lo, hi].
static unsigned next_chunk(double ratio2, double prev_ratio2, unsigned since, double tol2, unsigned lo, unsigned hi) {
  if (!(ratio2 > 0.0) || !(prev_ratio2 > ratio2) || since == 0) return lo;
  const double per = std::log(ratio2 / prev_ratio2) / since;  // < 0
  const double need = std::log(tol2 / ratio2) / per;
  double k = need * 1.05 + 4.0;
  if (!(k > lo)) k = lo;
  if (k > hi) k = hi;
  return (unsigned)k;
}

// x = 0; ||r0||^2 (guard); z = M^-1 r; p = z; rho = (r, z)   (solver_ls_iter.cpp:107-133)
template <typename T>
static int pcg_start_T(Sys& s, const PcgBufs& v, bool mixed, double tol, unsigned maxit, double guard, double* sc,
                       int* fl, hipStream_t st) {
  const long long m = (long long)s.n * s.len;
  const dim3 vg = vgrid(m);
  const int G = (int)vg.x;
  T* r = (T*)v.r;
  T* z = (T*)v.z;
  T* p = (T*)v.p;
  // reliable updates when the recursive residual has dropped by delta since the last one. Measured on 3D
  // elasticity systems (48^3 to 96^3 hexes, four material seeds, mixed loop to 1e-8): delta 0.05 gave the FP64
  // loop's iteration count within 1 % on four of six and 9-10 % fewer iterations than 0.1 on the other two (0.1
  // let an update land where it cost ~70 extra iterations); 0.03-0.04 were worse on some. Scalar systems were
  // insensitive. DFM_RELIABLE_DELTA overrides it.
  const char* dl = getenv("DFM_RELIABLE_DELTA");
  const double delta = dl ? atof(dl) : 0.05;
  hipLaunchKernelGGL(k_solver_init, dim3(1), dim3(64), 0, st, sc, fl, tol * tol, (int)maxit, delta * delta);
  hipLaunchKernelGGL(k_vfill<T>, vg, dim3(NT), 0, st, (T*)v.x, (T)0, m, (const int*)nullptr);
  if (mixed) hipLaunchKernelGGL(k_vfill<double>, vg, dim3(NT), 0, st, v.x64, 0.0, m, (const int*)nullptr);
  hipLaunchKernelGGL(k_vdot<T>, vg, dim3(NT), 0, st, (const T*)r, (const T*)r, m, slot(s, P_A), (const int*)nullptr);
  hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(NT), 0, st, slot(s, P_A), G, guard, sc, fl);
  hipError_t e = mixed ? precond_apply32(s, (const float*)r, (float*)z, fl, st)
                       : precond_apply_d(s, v, (const double*)r, (double*)z, fl, st);
  if (e != hipSuccess) return DFM_ERR_HIP;
  hipLaunchKernelGGL(k_vcopy<T>, vg, dim3(NT), 0, st, (const T*)z, p, m, fl);
  // the mixed loop forms q as beta q + A z (beta = 0 in the first iteration): q starts finite
  if (mixed) hipLaunchKernelGGL(k_vfill<T>, vg, dim3(NT), 0, st, (T*)v.q, (T)0, m, (const int*)nullptr);
  hipLaunchKernelGGL(k_vdot<T>, vg, dim3(NT), 0, st, (const T*)r, (const T*)z, m, slot(s, P_A), fl);
  hipLaunchKernelGGL(k_pcg_rho0, dim3(1), dim3(NT), 0, st, slot(s, P_A), G, sc, fl);
  DFM_HIP(hipGetLastError());
  return DFM_OK;
}

// warm start: inv0 from ||b||^2 (slot_b), ||r||^2 (slot_r) of r = b - A x; converged at once when below tol
__global__ void __launch_bounds__(NT) k_pcg_start_warm(const double* slot_b, const double* slot_r, int G, double* sc,
                                                       int* fl) {
  double bb[1], rr[1];
  consume<NT, 1>(slot_b, G, bb);
  consume<NT, 1>(slot_r, G, rr);
  if (threadIdx.x == 0) {
    sc[S_INV0] = 1.0 / bb[0];
    sc[S_RR] = rr[0];
    sc[S_RMAX] = rr[0];
    if (rr[0] * sc[S_INV0] < sc[S_TOL2]) {
      sc[S_CONV] = sqrt(rr[0] * sc[S_INV0]);
      fl[F_ITOUT] = 0;
      fl[F_DONE] = 1;
    }
  }
}

int pcg_start_warm(Sys& s, const PcgBufs& v, double tol, unsigned maxit, double* sc, int* fl, hipStream_t st) {
  const long long m = (long long)s.n * s.len;
  const dim3 vg = vgrid(m);
  const int G = (int)vg.x;
  double* r = (double*)v.r;
  double* z = (double*)v.z;
  double* p = (double*)v.p;
  hipLaunchKernelGGL(k_solver_init, dim3(1), dim3(64), 0, st, sc, fl, tol * tol, (int)maxit, 0.01);
  hipLaunchKernelGGL(k_vdot<double>, vg, dim3(NT), 0, st, (const double*)r, (const double*)r, m, slot(s, P_B),
                     (const int*)nullptr);
  DFM_HIP(launch_spmv(s.n, s.len, spmv_tpr(s.ncrs, s.n), s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p, -1.0,
                      (const double*)v.x, 1.0, r, st));
  hipLaunchKernelGGL(k_vdot<double>, vg, dim3(NT), 0, st, (const double*)r, (const double*)r, m, slot(s, P_A),
                     (const int*)nullptr);
  hipLaunchKernelGGL(k_pcg_start_warm, dim3(1), dim3(NT), 0, st, (const double*)slot(s, P_B),
                     (const double*)slot(s, P_A), G, sc, fl);
  if (precond_apply_d(s, v, (const double*)r, z, fl, st) != hipSuccess) return DFM_ERR_HIP;
  hipLaunchKernelGGL(k_vcopy<double>, vg, dim3(NT), 0, st, (const double*)z, p, m, (const int*)fl);
  hipLaunchKernelGGL(k_vdot<double>, vg, dim3(NT), 0, st, (const double*)r, (const double*)z, m, slot(s, P_A),
                     (const int*)fl);
  hipLaunchKernelGGL(k_pcg_rho0, dim3(1), dim3(NT), 0, st, slot(s, P_A), G, sc, fl);
  DFM_HIP(hipGetLastError());
  return DFM_OK;
}

int pcg_start(Sys& s, const PcgBufs& v, bool mixed, double tol, unsigned maxit, double* sc, int* fl, hipStream_t st) {
  return mixed ? pcg_start_T<float>(s, v, true, tol, maxit, 1.0e-30, sc, fl, st)
               : pcg_start_T<double>(s, v, false, tol, maxit, 1.0e-30, sc, fl, st);
}

PcgBufs pcg_bufs(Sys& s, bool mixed) {
  const long long m = (long long)s.n * s.len;
  if (!mixed) return PcgBufs{s.d_crs.p, s.d_dia.p, s.d_res.p, s.d_work.p + m, s.d_work.p, s.d_work.p + m, s.d_upd.p,
                             nullptr, nullptr};
  // FP32: r, z, p, q, xlo; FP64: the right-hand side (d_rhs) and the solution (the update handle)
  float* w = s.d_work32.p;
  return PcgBufs{s.d_crs32.p, s.d_dia32.p, w, w + m, w + 2 * m, w + 3 * m, w + 4 * m, s.d_upd.p, s.d_rhs.p};
}

int pcg_prepare(Sys& s, const PcgBufs& v, bool mixed, hipStream_t st) {
  const long long m = (long long)s.n * s.len;
  if (mixed) hipLaunchKernelGGL(k_mx_to32, vgrid(m), dim3(NT), 0, st, m, s.d_rhs.p, (float*)v.r);
  DFM_HIP(hipGetLastError());
  return DFM_OK;
}

int pcg_finish(Sys& s, const PcgBufs& v, bool mixed, int done, hipStream_t st) {
  const long long m = (long long)s.n * s.len;
  if (!mixed || done == 3) return DFM_OK;  // stalled: the residual handle keeps the right-hand side for FP64
  // stopped by max_iter: the FP64 solution takes the partial solution not yet folded (when converged, the
  // FP64 solution is the one whose true residual was confirmed)
  if (done != 1) hipLaunchKernelGGL(k_mx_final_fold, vgrid(m), dim3(NT), 0, st, m, (float*)v.x, v.x64);
  // the residual handle receives the FP64 true residual b - A x (the reference leaves its final residual)
  DFM_HIP(launch_copy(s.d_rhs.p, s.d_res.p, m, st));
  DFM_HIP(launch_spmv(s.n, s.len, spmv_tpr(s.ncrs, s.n), s.d_rowptr.p, s.d_colind.p, s.d_crs.p, s.d_dia.p, -1.0,
                      s.d_upd.p, 1.0, s.d_res.p, st));
  DFM_HIP(hipGetLastError());
  return DFM_OK;
}

template <typename T>
static int pcg_run(Sys& s, const PcgBufs& v, bool mixed, double tol, unsigned maxit, double& conv, unsigned& iter,
                   bool& converged, bool warm = false) {
  hipStream_t st = s.ctx->stream;
  int* fl = s.d_iflag.p;
  double* sc = s.d_scal.p;
  {
    const size_t need = (size_t)(sweep_blocks(s.fwd) + sweep_blocks(s.bwd));
    if (s.d_part_sw.n < need) DFM_HIP(s.d_part_sw.alloc(need));
  }
  {
    const int rc = warm ? pcg_start_warm(s, v, tol, maxit, sc, fl, st) : pcg_start(s, v, mixed, tol, maxit, sc, fl, st);
    if (rc != DFM_OK) return rc;
  }
  const int gk = mixed ? 1 : (v.f32ilu ? 3 : 0);
  const bool use_graph = getenv("DFM_NO_GRAPH") == nullptr;
  if (use_graph) {
    const int rc = graph_chunk(s, gk, graph_key(s, v), [&]() { return pcg_iteration<T>(s, v, mixed, st); }, st);
    if (rc != DFM_OK) return rc;
  }
  // Chunks of iterations, each followed by a status write into one of two pinned slots (k_status_out) and an
  // event. The next chunk is enqueued before the status of the current one is waited for, so the device never
  // idles while the host decides, and a chunk boundary costs one small kernel: the chunks can stay short (one or
  // two graphs: half the iterations estimated to remain, at most two graphs), which bounds the no-op iterations
  // (kernels that see the done flag and return, about 11 us each on the 2.7 M-dof elasticity system) queued past
  // convergence to the two chunks in flight. Longer chunks sized from the early decay rate overshot by hundreds
  // of iterations: preconditioned CG converges faster later than its first iterations predict.
  for (hipEvent_t& ev : s.poll_ev)
    if (!ev) DFM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  const unsigned gc = use_graph ? (unsigned)s.gchunk : 1u;
  unsigned ksz[2] = {0, 0};
  int seq[2] = {0, 0};
  auto enqueue_chunk = [&](unsigned k, int slot) -> int {
    unsigned t = 0;
    if (use_graph)
      for (; t + gc <= k; t += gc) DFM_HIP(hipGraphLaunch(s.gexec[gk], st));
    for (; t < k; t++) {
      const int rc = pcg_iteration<T>(s, v, mixed, st);
      if (rc != DFM_OK) return rc;
    }
    seq[slot] = ++s.poll_seq;
    hipLaunchKernelGGL(k_status_out, dim3(1), dim3(64), 0, st, (const int*)fl, (const double*)sc, s.h_pinned + 16 * slot,
                       s.h_pinned_d + kNScal * slot, seq[slot]);
    DFM_HIP(hipGetLastError());
    DFM_HIP(hipEventRecord(s.poll_ev[slot], st));
    ksz[slot] = k;
    return DFM_OK;
  };
  auto whole = [&](double k) {
    const unsigned lo = std::max(gc, 16u), hi = 2 * lo;
    const unsigned u = (unsigned)std::min((double)hi, std::max((double)lo, k));
    return std::max(lo, u / gc * gc);
  };
  // cur: the slot of the oldest chunk whose status is unread; inflight: such chunks (at most two); next: the size
  // of the chunk to queue ahead, 0 when the chunk in flight is expected to end the solve (then it is waited for
  // before anything more is queued)
  unsigned launched = 0, since = 0, next = whole(0.0);
  double prev_ratio2 = 1.0;
  int cur = 0, inflight = 0;
  // the mixed loop may run one iteration past max_iter: the true-residual check of a solution that converged in
  // the last allowed iteration (k_pcg_judge_beta)
  const unsigned cap = mixed && maxit > 0 ? maxit + 1 : maxit;
  for (;;) {
    if (inflight < 2 && launched < cap && (inflight == 0 || next > 0)) {
      const unsigned k = std::min(whole((double)next), cap - launched);
      const int rc = enqueue_chunk(k, (cur + inflight) & 1);
      if (rc != DFM_OK) return rc;
      launched += k;
      inflight++;
    }
    if (inflight == 0) break;  // max_iter reached, every status read
    DFM_HIP(hipEventSynchronize(s.poll_ev[cur]));
    const int* f = s.h_pinned + 16 * cur;
    const double* d = s.h_pinned_d + kNScal * cur;
    if (__atomic_load_n(f + 15, __ATOMIC_ACQUIRE) != seq[cur]) {
      // the slot does not show this chunk's status yet: read it with a copy (everything queued runs first)
      const int rc = poll(s, st);
      if (rc != DFM_OK) return rc;
      f = s.h_pinned;
      d = s.h_pinned_d;
    }
    inflight--;
    if (f[F_DONE]) break;
    since += ksz[cur];
    const double ratio2 = d[S_RR] * d[S_INV0];
    const double left = (double)next_chunk(ratio2, prev_ratio2, since, tol * tol, 0, 1u << 30) -
                        (inflight ? (double)ksz[cur ^ 1] : 0.0);
    next = left > 0.0 ? whole(0.5 * left) : 0u;
    if (ratio2 < prev_ratio2) {
      prev_ratio2 = ratio2;
      since = 0;
    }
    cur ^= 1;
  }
  {
    // everything queued has run: the final status into slot 0, where the callers read it
    const int rc = poll(s, st);
    if (rc != DFM_OK) return rc;
  }
  const int done = s.h_pinned[F_DONE];
  {
    const int rc = pcg_finish(s, v, mixed, done, st);
    if (rc != DFM_OK) return rc;
  }
  converged = done == 1;
  // not converged within max_iter: the ratio of the last residual seen (solver_ls_iter.cpp:168-171)
  conv = done ? s.h_pinned_d[S_CONV] : std::sqrt(s.h_pinned_d[S_RR] * s.h_pinned_d[S_INV0]);
  iter = converged ? (unsigned)s.h_pinned[F_ITOUT] : maxit;
  return DFM_OK;
}

int solve_pcg(Sys& s, int precision, double* conv_ratio, unsigned* iteration) {
  const double tol = *conv_ratio;
  const unsigned maxit = *iteration;
  hipStream_t st = s.ctx->stream;
  hipEvent_t e0, e1;
  DFM_HIP(hipEventCreate(&e0));
  DFM_HIP(hipEventCreate(&e1));
  DFM_HIP(hipEventRecord(e0, st));
  double conv = 0;
  unsigned it = 0;
  bool ok = false;
  int rc;
  const bool mixed = precision == DFM_MIXED;
  PcgBufs v = pcg_bufs(s, mixed);
  v.f32ilu = precision == DFM_FP32_ILU;
  s.last_fallback = 0;
  rc = pcg_prepare(s, v, mixed, st);
  if (rc == DFM_OK)
    rc = mixed ? pcg_run<float>(s, v, true, tol, maxit, conv, it, ok) : pcg_run<double>(s, v, false, tol, maxit, conv, it, ok);
  if (rc != DFM_OK) return rc;
  int nrel = s.h_pinned[F_NREL];
  unsigned before = 0;  // iterations of a mixed-precision loop that was stopped
  // iterations executed by all loops of this solve (a loop that did not converge ran its whole budget)
  unsigned total = (unsigned)(s.h_pinned[F_IT] + 1);  // the device's count (a mixed check may run one past it)
  if (mixed && s.h_pinned[F_DONE] == 3) {
    // the FP32 matrix limits progress (cond(A) eps32 not small): continue from the FP64 solution with FP64
    // vectors and matrix, keeping the FP32 factor (the preconditioner's precision does not bound the attainable
    // accuracy); max_iter bounds both loops together
    s.last_fallback = 1;
    before = (unsigned)s.h_pinned[F_IT] + 1;
    total = before;
    PcgBufs v64 = pcg_bufs(s, false);
    v64.f32ilu = true;
    if (before < maxit) {
      rc = pcg_run<double>(s, v64, false, tol, maxit - before, conv, it, ok, true);
      if (rc != DFM_OK) return rc;
      total += ok ? it + 1 : maxit - before;
    } else {
      ok = false;
    }
  }
  // The FP64 loop (DFM_FP64_TRUE; with the FP32 factor DFM_FP32_ILU, or a mixed solve that fell back to it) stops on
  // its recursive residual: its true residual is checked and the loop restarted from its solution (r = b - A x, the
  // recursive target halved, still relative to ||b||) until the true residual meets the tolerance, so that
  // conv_ratio is the true relative residual in every mode but DFM_FP64 (the reference's own test). A restart
  // that does not lower the true residual by a fifth ends the solve unconverged: the residual of b - A x in
  // FP64 has reached what that arithmetic resolves for this system.
  if (ok && (precision == DFM_FP32_ILU || precision == DFM_FP64_TRUE || s.last_fallback)) {
    PcgBufs v64 = pcg_bufs(s, false);
    v64.f32ilu = precision != DFM_FP64_TRUE;
    const long long m = (long long)s.n * s.len;
    double prev = 0.0;
    for (int restart = 0;; restart++) {
      double rel = 0.0;
      rc = sys_true_residual(s, st, &rel);
      if (rc != DFM_OK) return rc;
      conv = rel;
      if (rel <= tol) break;
      before += it + 1;
      if (before >= maxit || restart == kMaxRestarts || (restart > 0 && rel > kRestartGain * prev)) {
        ok = false;
        break;
      }
      prev = rel;
      DFM_HIP(launch_copy(s.d_rhs.p, s.d_res.p, m, st));  // the warm start forms b - A x from the right-hand side
      double c2 = 0.0;
      rc = pcg_run<double>(s, v64, false, 0.5 * tol, maxit - before, c2, it, ok, true);
      if (rc != DFM_OK) return rc;
      total += ok ? it + 1 : maxit - before;
      if (!ok) {
        rc = sys_true_residual(s, st, &conv);
        if (rc != DFM_OK) return rc;
        break;
      }
    }
  }
  DFM_HIP(hipEventRecord(e1, st));
  DFM_HIP(hipEventSynchronize(e1));
  float ms = 0;
  DFM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *conv_ratio = conv;
  if (ok) *iteration = it + before;
  const unsigned its = total;
  s.last_ms = ms;
  s.last_inner = its;
  s.last_ms_iter = its ? ms / its : 0.0;
  s.last_nrel = nrel;
  return ok ? 1 : 0;
}

int solve_pbicgstab(Sys& s, bool f32ilu, double* conv_ratio, unsigned* iteration) {
  const long long m = (long long)s.n * s.len;
  const double tol = *conv_ratio;
  const unsigned maxit = *iteration;
  hipStream_t st = s.ctx->stream;
  const dim3 vg = vgrid(m);
  const int G = (int)vg.x;
  int* fl = s.d_iflag.p;
  double* sc = s.d_scal.p;
  double* r = s.d_res.p;
  double* x = s.d_upd.p;
  double* vs = s.d_work.p;
  double* vMs = vs + m;
  double* vAMs = vs + 2 * m;
  double* vp = vs + 3 * m;
  double* vMp = vs + 4 * m;
  double* vAMp = vs + 5 * m;
  double* vr2 = vs + 6 * m;
  hipEvent_t e0, e1;
  DFM_HIP(hipEventCreate(&e0));
  DFM_HIP(hipEventCreate(&e1));
  DFM_HIP(hipEventRecord(e0, st));
  hipLaunchKernelGGL(k_solver_init, dim3(1), dim3(64), 0, st, sc, fl, tol * tol, (int)maxit, 0.0);
  hipLaunchKernelGGL(k_vdot<double>, vg, dim3(NT), 0, st, r, r, m, slot(s, P_A), (const int*)nullptr);
  hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(NT), 0, st, slot(s, P_A), G, 1.0e-60, sc, fl);
  int rc = poll(s, st);
  if (rc != DFM_OK) return rc;
  if (s.h_pinned[F_DONE]) {
    // guard: conv = sqrt(sq0), iteration 0, the update untouched (solver_ls_iter.cpp:305-313)
    *conv_ratio = std::sqrt(s.h_pinned_d[S_RR]);
    *iteration = 0;
    return 1;
  }
  hipLaunchKernelGGL(k_vfill<double>, vg, dim3(NT), 0, st, x, 0.0, m, fl);
  hipLaunchKernelGGL(k_vcopy<double>, vg, dim3(NT), 0, st, r, vr2, m, fl);
  hipLaunchKernelGGL(k_vcopy<double>, vg, dim3(NT), 0, st, r, vp, m, fl);
  // the reference loop counts iitr from 1 (:328); F_IT counts from 0, iitr = F_IT + 1
  unsigned launched = 1, chunk = 8;
  while (launched < maxit) {
    const unsigned k = std::min(chunk, maxit - launched);
    for (unsigned t = 0; t < k; t++) {
      unsigned G2 = 0, G5 = 0;
      if ((f32ilu ? precond_apply_f32ilu(s, vp, vMp, fl, st) : precond_apply64(s, vp, vMp, fl, st)) != hipSuccess)
        return DFM_ERR_HIP;
      if (bi_spmv(s, vMp, vAMp, vr2, r, 0, slot(s, P_A), G2, st) != hipSuccess) return DFM_ERR_HIP;
      hipLaunchKernelGGL(k_bi_s, vg, dim3(NT), 0, st, m, r, vAMp, vs, slot(s, P_A), (int)G2, fl);
      if ((f32ilu ? precond_apply_f32ilu(s, vs, vMs, fl, st) : precond_apply64(s, vs, vMs, fl, st)) != hipSuccess)
        return DFM_ERR_HIP;
      if (bi_spmv(s, vMs, vAMs, vs, nullptr, 1, slot(s, P_B), G5, st) != hipSuccess) return DFM_ERR_HIP;
      hipLaunchKernelGGL(k_bi_update, vg, dim3(NT), 0, st, m, x, r, vMp, vMs, vs, vAMs, vr2, slot(s, P_A), (int)G2,
                         slot(s, P_B), (int)G5, slot(s, P_C), fl);
      hipLaunchKernelGGL(k_bi_pupdate, vg, dim3(NT), 0, st, m, vp, r, vAMp, slot(s, P_A), (int)G2, slot(s, P_B),
                         (int)G5, slot(s, P_C), G, sc, fl);
    }
    launched += k;
    rc = poll(s, st);
    if (rc != DFM_OK) return rc;
    if (s.h_pinned[F_DONE]) break;
    chunk = std::min(chunk * 2, 128u);
  }
  DFM_HIP(hipGetLastError());
  const bool converged = s.h_pinned[F_DONE] == 1;
  if (converged) {
    *conv_ratio = s.h_pinned_d[S_CONV];
    *iteration = (unsigned)s.h_pinned[F_ITOUT] + 1;
  } else {
    *iteration = maxit;  // the reference returns true with num_iter = max_iter and the ratio untouched
  }
  DFM_HIP(hipEventRecord(e1, st));
  DFM_HIP(hipEventSynchronize(e1));
  float ms = 0;
  DFM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const unsigned its = converged ? (unsigned)s.h_pinned[F_ITOUT] + 1 : maxit;
  s.last_ms = ms;
  s.last_inner = its;
  s.last_ms_iter = its ? ms / its : 0.0;
  return 1;
}

// Average time of one launch of a hot kernel, timed with HIP events on the system's stream (the stream the
// kernels run on). Algorithmic bytes: SpMV (ncrs + n) b^2 s_v + 4 ncrs + 4 (n+1) + 2 n b s_x; ILU apply
// (forward + backward) (ncrs + n) b^2 s_v + 4 ncrs + 4 (n+1) + 4 n b s_x.
int bench_kernel(Sys& s, int which, int reps, double* avg_ms, double* bytes) {
  hipStream_t st = s.ctx->stream;
  const long long m = (long long)s.n * s.len;
  int* fl = s.d_iflag.p;
  double* sc = s.d_scal.p;
  hipLaunchKernelGGL(k_solver_init, dim3(1), dim3(64), 0, st, sc, fl, 0.0, 1 << 30, 0.0);
  float* x32 = s.d_work32.p;
  float* y32 = s.d_work32.p + m;
  hipLaunchKernelGGL(k_mx_to32, vgrid(m), dim3(NT), 0, st, m, s.d_rhs.p, x32);
  unsigned G = 0;
  // 4, 5: the sweep phases of the PCG iteration (forward phases with alpha and the r/x update and (r, r), backward
  // phases with (r, z)) in the mixed (FP32) and FP64 loops, on the solve's own buffers (alpha = 0 after the reset
  // above: r and x keep their values)
  const PcgBufs v32 = pcg_bufs(s, true), v64 = pcg_bufs(s, false);
  const size_t need = (size_t)(sweep_blocks(s.fwd) + sweep_blocks(s.bwd));
  if ((which == 4 || which == 5) && s.d_part_sw.n < need) DFM_HIP(s.d_part_sw.alloc(need));
  double* prr = s.d_part_sw.p;
  double* prz = s.d_part_sw.p + sweep_blocks(s.fwd);
  // the first forward phase forms alpha from SpMV partials: one SpMV provides them ((x, Ax) > 0, so alpha = 0)
  unsigned Gpq = 0;
  if (which == 4 && pcg_spmv<float, false>(s, s.d_crs32.p, s.d_dia32.p, x32, y32, slot(s, P_A), Gpq, nullptr, nullptr,
                                           nullptr, nullptr, st) != hipSuccess)
    return DFM_ERR_HIP;
  if (which == 5 && pcg_spmv<double, false>(s, s.d_crs.p, s.d_dia.p, s.d_rhs.p, s.d_work.p, slot(s, P_A), Gpq, nullptr,
                                            nullptr, nullptr, nullptr, st) != hipSuccess)
    return DFM_ERR_HIP;
  auto launch = [&]() -> hipError_t {
    switch (which) {
      case 0: return pcg_spmv<double, false>(s, s.d_crs.p, s.d_dia.p, s.d_rhs.p, s.d_work.p, slot(s, P_A), G, nullptr, nullptr, nullptr, nullptr, st);
      case 1: return pcg_spmv<float, false>(s, s.d_crs32.p, s.d_dia32.p, x32, y32, slot(s, P_A), G, nullptr, nullptr, nullptr, nullptr, st);
      case 2: return precond_apply64(s, s.d_rhs.p, s.d_work.p, fl, st);
      case 3: return precond_apply32(s, x32, y32, fl, st);
      case 4:
        return fused_sweeps<float, float>(s, v32, s.fwd.d_eval32.p, s.fwd.d_dval32.p, s.bwd.d_eval32.p, prr, prz,
                                          slot(s, P_A), (int)Gpq, nullptr, st);
      default:
        return fused_sweeps<double, double>(s, v64, s.fwd.d_eval.p, s.fwd.d_dval.p, s.bwd.d_eval.p, prr, prz,
                                            slot(s, P_A), (int)Gpq, nullptr, st);
    }
  };
  for (int w = 0; w < 3; w++)
    if (launch() != hipSuccess) return DFM_ERR_HIP;
  hipEvent_t e0, e1;
  DFM_HIP(hipEventCreate(&e0));
  DFM_HIP(hipEventCreate(&e1));
  DFM_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; r++)
    if (launch() != hipSuccess) return DFM_ERR_HIP;
  DFM_HIP(hipEventRecord(e1, st));
  DFM_HIP(hipEventSynchronize(e1));
  float ms = 0;
  DFM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (avg_ms) *avg_ms = ms / reps;
  // algorithmic bytes: matrix or factor values and column indices once, plus the vectors each pass must move
  // (SpMV: x read, y written; apply: b read, z written forward, z read and written backward; fused phases: r, q,
  // p, x read, r, x, z written forward, z read and written and r read backward)
  const double sv = (which == 0 || which == 2 || which == 5) ? 8.0 : 4.0;
  const double sx = sv;
  const double b2 = (double)s.bs;
  const double fn = (double)s.fcolind.size();
  const double mat = which < 2 ? ((double)s.ncrs + s.n) * b2 * sv + 4.0 * s.ncrs + 4.0 * (s.n + 1)
                               : (fn + s.n) * b2 * sv + 4.0 * fn + 4.0 * (s.n + 1);
  double nvec = which < 2 ? 2.0 : which < 4 ? 4.0 : 10.0;
  if (which >= 4 && !s.fwd.fuse_bgrp.empty()) {
    // rows of the fused last/first phase: z is written once and not read back
    const Sweep& f = s.fwd;
    long long rows = 0;
    for (int g = f.phase_gptr[f.nphase() - 1]; g < f.phase_gptr[f.nphase()]; g++) rows += f.grp[2 * g + 1] - f.grp[2 * g];
    nvec -= 3.0 * (double)rows / (double)s.n;
  }
  if (bytes) *bytes = mat + nvec * (double)m * sx;
  return DFM_OK;
}

}  // namespace dfm
