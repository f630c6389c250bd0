// Segmented PCG: many independent linear systems ("segments") advanced together, one launch per kernel for
// all running segments of one block size. Each segment keeps its own Krylov scalars, done flag and iteration
// count (LsSol::Solve_PCG semantics per segment, src/ls/solver_ls_iter.cpp:100-172; the mixed-precision form
// with reliable updates as in solver.hip).
//
// Launch geometry. Row-parallel kernels walk a block table: SpMV blocks cover RS = 256/TPR rows of one
// segment, vector blocks cover DV dofs of one segment. The ILU sweeps walk, per phase, a table of (segment,
// group) pairs, one wavefront each. Reductions are per segment: every block writes one partial into its own
// slot; a small kernel with one workgroup per segment then sums exactly that segment's partials in block order
// and does the segment's scalar work (alpha; the convergence judgement; beta), so the row-parallel kernels read
// one scalar instead of re-reducing. The partition of a segment does not depend on the other segments, so its
// results are bit-identical alone or in any mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <tuple>
#include <type_traits>
#include <vector>

#include "dev_common.hpp"
#include "dfm_internal.hpp"
#include "kernels.hpp"
#include "solver.hpp"

namespace dfm {

namespace {

constexpr int NT = 256;
constexpr int VPT = 4;            // dofs per thread of the vector kernels
constexpr int DV = NT * VPT;      // dofs per vector block

struct SegDev {
  int n, m;
  const void* dia;      // diagonal blocks of the Krylov SpMV (FP32 in mixed mode)
  const double* dia64;  // FP64 diagonal blocks (true residual, mixed mode)
  SlabArgs fa, ba;
  const void* fval;
  const void* fdval;
  const void* bval;
  void *r, *z, *p, *q, *x;  // x: partial solution (mixed) or the solution (FP64)
  double* x64;
  const double* b64;
  double* sc;
  int* fl;
  int ps0, ps1;  // SpMV blocks of the segment
  int pv0, pv1;  // vector blocks of the segment
  int pr0, pr1;  // sweep partials of the segment: forward (r, r) [pr0, pr1), backward (r, z) [pz0, pz1)
  int pz0, pz1;
  int nslice;           // SELL-64 copy of the matrix: slices, slab rows per slice, columns, values
  const int* sptr;
  const unsigned* scol;
  const void* sval;     // values of the Krylov SpMV (FP32 in mixed mode)
  const double* sval64; // FP64 values (true residual, mixed mode)
};

constexpr int kBSellWaves = 16;  // SELL slices (one per wavefront) per SpMV workgroup

// scalar slots of the batch loop beyond those of pcg_start (rho stays in S_RZ, the reliable-update maximum in
// S_RMAX: only the per-segment reduction kernels write them, one workgroup per segment)

// block-wide sum of the partials slot[b0..b1), identical in every thread
__device__ inline double consume_range(const double* __restrict__ slot, int b0, int b1) {
  double t[1] = {0.0};
  for (int b = b0 + (int)threadIdx.x; b < b1; b += NT) t[0] += slot[b];
  block_allsum<NT, 1>(t);
  return t[0];
}

// q = A p, partials of (p, q); mixed, when a reliable update is due: r = b - A64 x64 with its partials. Over the
// SELL-64 copy of each system: one slice (64 rows, one per lane) per wavefront, row sums in CRS order with the
// diagonal block last (MatVec, src/matvec/matdia_blkcrs.cpp:921-1094). TM: matrix values, T: vectors. The mixed
// loop of scalar systems (ZQ) folds the direction update in as the single-system SpMV does: p = beta p + z,
// q = beta q + A z, and its true residual at a reliable update is b - A (x_hi + x_lo) (the forward phases fold x_lo
// afterwards).
template <int B, typename TM, typename T, bool REL>
__global__ void __launch_bounds__(64 * kBSellWaves) k_b_spmv(const SegDev* __restrict__ segs,
                                                             const int* __restrict__ bseg, double* slot_pq,
                                                             double* slot_true) {
  constexpr int CH = SellChunk<B>::CH;
  constexpr bool ZQ = REL && B == 1;
  const int b = blockIdx.x;
  const SegDev& d = segs[bseg[b]];
  int* fl = d.fl;
  if (fl[F_DONE]) return;
  if (b == d.ps0 && threadIdx.x == 0) fl[F_IT] = fl[F_IT] + 1;  // the segment's iteration counter
  const bool rel = REL && fl[F_PCUR];
  const int wave = wave_in_block(), lane = (int)(threadIdx.x % 64);
  const int sl = (b - d.ps0) * kBSellWaves + wave;
  T* p = (T*)d.p;
  T* q = (T*)d.q;
  const T* z = (const T*)d.z;
  const T bt = ZQ ? (T)d.sc[S_BETA] : T(0);
  double v[2] = {0.0, 0.0};
  if (sl < d.nslice) {
    const int r0 = d.sptr[sl], w = d.sptr[sl + 1] - r0;
    const int row = sl * 64 + lane;
    T acc[B];
    sell_row<B, CH, TM, T, T>(d.scol, (const TM*)d.sval, r0, w, lane, ZQ ? z : (const T*)p, acc);
    if (row < d.n) {
      dia_row<B, TM, T, T>((const TM*)d.dia, ZQ ? z : (const T*)p, row, acc);
#pragma unroll
      for (int k = 0; k < B; k++) {
        const size_t i = (size_t)row * B + k;
        if (ZQ) {
          const T pn = p[i] * bt + z[i];
          const T qn = q[i] * bt + acc[k];
          p[i] = pn;
          q[i] = qn;
          v[0] += (double)pn * (double)qn;
        } else {
          q[i] = acc[k];
          v[0] += (double)p[i] * (double)acc[k];
        }
      }
    }
    if (REL && rel) {
      // ZQ: x_lo not folded yet; otherwise folded by the p update (x_lo = 0)
      const SumVec xs{d.x64, (const float*)d.x};
      double a64[B];
      sell_row<B, CH, double, double, double>(d.scol, d.sval64, r0, w, lane, xs, a64);
      if (row < d.n) {
        dia_row<B, double, double, double>(d.dia64, xs, row, a64);
#pragma unroll
        for (int k = 0; k < B; k++) {
          const double t = d.b64[(size_t)row * B + k] - a64[k];
          ((T*)d.r)[(size_t)row * B + k] = (T)t;
          v[1] += t * t;
        }
      }
    }
  }
  block_allsum<64 * kBSellWaves, 2>(v);
  if (threadIdx.x == 0) {
    slot_pq[b] = v[0];
    if (REL && rel) slot_true[b] = v[1];
  }
}

// per segment: convergence judgement, then beta. FP64: on the recursive residual (:146-155). Mixed: on the true
// residual when one was formed this iteration (the solution is then the previous iteration's), else the
// recursive residual only schedules the next reliable update (it dropped by delta, or looks converged). Then
// beta = rho_new / rho (:158-159), and the scheduled reliable update becomes current.
template <bool MX>
__global__ void __launch_bounds__(NT) k_b_judge_beta(const SegDev* __restrict__ segs, const double* slot_rr,
                                                     const double* slot_rz, const double* slot_true) {
  const SegDev& d = segs[blockIdx.x];
  int* fl = d.fl;
  double* sc = d.sc;
  if (fl[F_DONE]) return;
  const double rr = consume_range(slot_rr, d.pr0, d.pr1);
  const double rz = consume_range(slot_rz, d.pz0, d.pz1);
  const double rt = (MX && fl[F_PCUR]) ? consume_range(slot_true, d.ps0, d.ps1) : 0.0;
  if (threadIdx.x) return;
  const int it = fl[F_IT];
  sc[S_RR] = rr;
  if (!MX) {
    const double ratio2 = rr * sc[S_INV0];
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
    sc[S_BETA] = rz / sc[S_RZ];
    sc[S_RZ] = rz;
    return;
  }
  double rmax = sc[S_RMAX];
  if (fl[F_PCUR]) {
    const double ratio2 = rt * sc[S_INV0];
    if (ratio2 < sc[S_TOL2]) {
      sc[S_CONV] = sqrt(ratio2);
      sc[S_RR] = rt;
      fl[F_ITOUT] = it - 1;
      fl[F_DONE] = 1;
      return;
    }
    if (mx_slow(rt, sc[S_RRTRIG], sc[S_TPREV], sc[S_SLOW], sc[S_TBEST], fl[F_STALL])) {
      sc[S_CONV] = sqrt(ratio2);
      sc[S_RR] = rt;
      fl[F_DONE] = 3;
      return;
    }
    rmax = rt;
  }
  // the last iteration max_iter allows: stop, unless the recursive residual says converged (its true-residual
  // check then runs in one more iteration, the solution reported being this one)
  if (it + 1 >= fl[F_MAXIT] && !(it + 1 == fl[F_MAXIT] && rr * sc[S_INV0] < sc[S_TOL2])) {
    sc[S_CONV] = sqrt(rr * sc[S_INV0]);
    fl[F_DONE] = 2;
    return;
  }
  const bool due = rr * sc[S_INV0] < sc[S_TOL2] || rr < sc[S_DELTA2] * rmax ||
                   it - (int)sc[S_LASTUPD] >= kForceUpdate;
  sc[S_RMAX] = due ? rr : fmax(rmax, rr);
  fl[F_PNEXT] = due ? 1 : 0;
  if (due) {
    fl[F_NREL] = fl[F_NREL] + 1;
    sc[S_LASTUPD] = (double)it;
    sc[S_RRTRIG] = rr;
  }
  sc[S_BETA] = rz / sc[S_RZ];
  sc[S_RZ] = rz;
  fl[F_PCUR] = fl[F_PNEXT];
}

// TF: factor values (FP64 or the FP32 copy); T: vectors and arithmetic; MO/MI: register-held entries per class.
// One wavefront per (segment, tile) entry. Forward: the tile rows' r -= alpha q, x += alpha p (registers, stored
// at the end), the forward sweep from the new r, and the wavefront's (r, r) partial at the entry's slot (each
// segment's slots are contiguous, so its reduction does not depend on the other segments).
// In the first forward phase (slot_pq given) every wavefront forms its segment's alpha = rho / (p, q)
// (solver_ls_iter.cpp:137-138) from the segment's SpMV partials, in a fixed order, and the wavefront of the
// segment's tile 0 stores it for the later phases.
template <int B, int MO, int MI, typename TF, typename T>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(SweepOccupancy<B, MO, MI, T>::kWaves)))
    k_b_fwd(const SegDev* __restrict__ segs, const int2* __restrict__ ent, const int* __restrict__ eslot, int count,
        double* slot_rr, const double* __restrict__ slot_pq) {
  __shared__ T xsb[kWPB][64 * B];
  const int w = wave_in_block();
  const int e = (int)blockIdx.x * kWPB + w;
  if (e >= count) return;
  const int2 en = ent[e];
  const SegDev& d = segs[en.x];
  if (d.fl[F_DONE]) return;
  double alpha;
  if (slot_pq) {
    double pq = 0.0;
    for (int b = d.ps0 + (int)(threadIdx.x % 64); b < d.ps1; b += 64) pq += slot_pq[b];
    alpha = d.sc[S_RZ] / wave_sum(pq);
    if (en.y == 0 && threadIdx.x % 64 == 0) d.sc[S_ALPHA] = alpha;
  } else {
    alpha = d.sc[S_ALPHA];
  }
  const T ma = (T)(-alpha), pa = (T)alpha;
  // scalar mixed segments: x is x_lo; at a reliable update it is folded into x_hi before alpha p is added
  const bool fold = B == 1 && sizeof(T) == 4 && d.fl[F_PCUR];
  T* r = (T*)d.r;
  T* x = (T*)d.x;
  const T* q = (const T*)d.q;
  const T* p = (const T*)d.p;
  T rn[B], xn[B];
  SlabSlot ss;
  T t[B];
  sweep_fwd_core<B, MO, MI, TF, T, T>(
      d.fa, en.y, (const TF*)d.fval, (const TF*)d.fdval,
      [&](const SlabSlot& sl, T(&u)[B]) {
#pragma unroll
        for (int c = 0; c < B; c++) {
          const size_t i = (size_t)sl.row * B + c;
          rn[c] = r[i] + ma * q[i];
          T xo = x[i];
          if (fold) {
            d.x64[i] += (double)xo;
            xo = T(0);
          }
          xn[c] = xo + pa * p[i];
          u[c] = rn[c];
        }
      },
      (T*)d.z, nullptr, xsb[w], ss, t);
  double v = 0.0;
  if (ss.active)
#pragma unroll
    for (int c = 0; c < B; c++) {
      const size_t i = (size_t)ss.row * B + c;
      r[i] = rn[c];
      x[i] = xn[c];
      v += (double)rn[c] * (double)rn[c];
    }
  v = wave_sum(v);
  if (threadIdx.x % 64 == 0) slot_rr[eslot[e]] = v;
}

// Backward sweep of one (segment, tile) entry and the wavefront's (r, z) partial at the entry's slot.
template <int B, int MO, int MI, typename TF, typename T>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(SweepOccupancy<B, MO, MI, T>::kWaves)))
    k_b_bwd(const SegDev* __restrict__ segs, const int2* __restrict__ ent, const int* __restrict__ eslot, int count,
        double* slot_rz) {
  __shared__ T xsb[kWPB][64 * B];
  const int w = wave_in_block();
  const int e = (int)blockIdx.x * kWPB + w;
  if (e >= count) return;
  const int2 en = ent[e];
  const SegDev& d = segs[en.x];
  if (d.fl[F_DONE]) return;
  T* z = (T*)d.z;
  const T* r = (const T*)d.r;
  SlabSlot ss;
  T t[B];
  sweep_bwd_core<B, MO, MI, TF, T, T>(d.ba, en.y, (const TF*)d.bval, z, nullptr, xsb[w], ss, t,
                                      [&](const SlabSlot& sl, T(&u)[B]) {
#pragma unroll
                                        for (int c = 0; c < B; c++) u[c] = z[(size_t)sl.row * B + c];
                                      });
  double v = 0.0;
  if (ss.active)
#pragma unroll
    for (int c = 0; c < B; c++) v += (double)r[(size_t)ss.row * B + c] * (double)t[c];
  v = wave_sum(v);
  if (threadIdx.x % 64 == 0) slot_rz[eslot[e]] = v;
}

// p = beta p; p += z (:160-166); mixed: the FP64 solution takes the partial solution when an update is due
// (scalar mixed systems form p in the SpMV and fold in the forward phases instead)
template <typename T, bool MX>
__global__ void __launch_bounds__(NT) k_b_pupdate(const SegDev* __restrict__ segs, const int* __restrict__ bseg) {
  const int b = blockIdx.x;
  const SegDev& d = segs[bseg[b]];
  if (d.fl[F_DONE]) return;
  const T bt = (T)d.sc[S_BETA];
  const bool fold = MX && d.fl[F_PCUR];
  T* p = (T*)d.p;
  const T* z = (const T*)d.z;
  const long long i0 = (long long)(b - d.pv0) * DV + threadIdx.x;
#pragma unroll
  for (int j = 0; j < VPT; j++) {
    const long long i = i0 + (long long)j * NT;
    if (i < d.m) {
      const T t = p[i] * bt;
      p[i] = t + z[i];
      if (MX && fold) {
        float* xlo = (float*)d.x;
        d.x64[i] += (double)xlo[i];
        xlo[i] = 0.0f;
      }
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------
// host side

struct BatchSeg {
  Sys* s = nullptr;
  int slot = -1;
  int mode = DFM_MIXED;  // DFM_FP64, DFM_MIXED (a stalled one restarts as DFM_FP32_ILU), DFM_FP32_ILU
  bool verify = false;   // the true residual decides at exit (DFM_FP32_ILU, DFM_FP64_TRUE, a stalled mixed one)
  bool fell_back = false;
  PcgBufs v{};
  double tol = 0;
  unsigned maxit = 0;
  int state = 0;  // 0 running, 1 converged, 2 stopped at max_iter
  double conv = 0;
  unsigned iters = 0;
  unsigned before = 0;  // iterations before the running loop (a mixed loop before its switch, earlier restarts)
  int restarts = 0;     // restarts of the FP64 loop after its true residual was checked
  double prev_rel = 0;  // the true residual that led to the last restart
  int nrel = 0;
  double t_add = 0, t_done = 0;  // host clock (s)
};

struct Batch {
  double t_rebuild = 0, t_poll = 0;  // host seconds spent in rebuild() and poll()
  long long n_rebuild = 0, n_restart = 0;
  Ctx* ctx = nullptr;
  int prec = DFM_MIXED;  // mode of added systems (DFM_MIXED_AUTO: by block size)
  int cap = 0;
  DBuf<double> d_sc;  // kNScal per slot
  DBuf<int> d_fl;     // 8 per slot
  int* h_fl = nullptr;
  double* h_sc = nullptr;
  std::vector<int> free_slots;
  std::vector<BatchSeg> segs;
  bool dirty = true;
  // one iteration of every group as a hipGraph, valid for the tables of one rebuild (gen), and the argument
  // values its kernel nodes were made from
  hipGraphExec_t gexec = nullptr;
  std::vector<std::shared_ptr<void>> gargs;
  long long gen = 0, gexec_gen = -1;
  int gen_steps = 0;  // steps run with the tables of this generation
  struct Group {
    int B = 0;
    std::vector<int> members;  // indices into segs
    DBuf<SegDev> d_segs;
    DBuf<int> d_sblk, d_vblk;
    DBuf<int2> d_fent, d_bent;
    DBuf<int> d_fslot, d_bslot;  // partial slot of each entry (segment-contiguous)
    std::vector<int> fptr, bptr;  // per phase offsets into the entry tables
    std::vector<int> fwo, fwi, bwo, bwi;  // per phase widest classes over the segments
    int nsb = 0, nvb = 0;
    int mode = 0;
    DBuf<double> d_pq, d_true, d_rr, d_rz;
  } grp[9];  // (block size, mode): index (B - 1) + 3 * mode
  ~Batch() {
    for (auto& g : grp) {
      g.d_segs.release(); g.d_sblk.release(); g.d_vblk.release(); g.d_fent.release(); g.d_bent.release();
      g.d_fslot.release(); g.d_bslot.release();
      g.d_pq.release(); g.d_true.release(); g.d_rr.release(); g.d_rz.release();
    }
    if (gexec) (void)hipGraphExecDestroy(gexec);
    d_sc.release();
    d_fl.release();
    if (h_fl) (void)hipHostFree(h_fl);
    if (h_sc) (void)hipHostFree(h_sc);
  }
};

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
static void grow(DBuf<T>& b, size_t n, hipError_t& e) {
  if (e != hipSuccess) return;
  if (b.n < n) e = b.alloc(std::max<size_t>(n, 1));
}

static int rebuild_tables(Batch& bt);
static int rebuild(Batch& bt) {
  const double t0 = now_s();
  bt.gen++;  // the tables (and the pointers a captured iteration holds) change
  bt.gen_steps = 0;
  const int rc = rebuild_tables(bt);
  bt.t_rebuild += now_s() - t0;
  bt.n_rebuild++;
  return rc;
}

static int rebuild_tables(Batch& bt) {
  hipStream_t st = bt.ctx->stream;
  for (auto& g : bt.grp) g.members.clear();
  for (int k = 0; k < (int)bt.segs.size(); k++)
    if (bt.segs[k].state == 0) bt.grp[bt.segs[k].s->len - 1 + 3 * bt.segs[k].mode].members.push_back(k);
  for (int gi = 0; gi < 9; gi++) {
    Batch::Group& g = bt.grp[gi];
    g.B = gi % 3 + 1;
    g.mode = gi / 3;
    std::vector<SegDev> h;
    std::vector<int> sblk, vblk;
    std::vector<std::vector<int2>> fph, bph;
    std::vector<std::vector<int>> fsl, bsl;  // per phase: partial slot of each entry
    int nfslot = 0, nbslot = 0;
    std::vector<int> fwo, fwi, bwo, bwi;
    for (int k : g.members) {
      BatchSeg& sg = bt.segs[k];
      Sys& s = *sg.s;
      SegDev d{};
      const int si = (int)h.size();
      d.n = s.n;
      d.m = s.n * s.len;
      d.dia = sg.v.dia;
      d.dia64 = s.d_dia.p;
      d.fa = SlabArgs{s.fwd.d_gmeta.p, s.fwd.d_srow.p, s.fwd.d_scnt.p, s.fwd.d_ecol.p};
      d.ba = SlabArgs{s.bwd.d_gmeta.p, s.bwd.d_srow.p, s.bwd.d_scnt.p, s.bwd.d_ecol.p};
      const bool f32 = sg.mode != DFM_FP64;
      d.fval = f32 ? (const void*)s.fwd.d_eval32.p : (const void*)s.fwd.d_eval.p;
      d.fdval = f32 ? (const void*)s.fwd.d_dval32.p : (const void*)s.fwd.d_dval.p;
      d.bval = f32 ? (const void*)s.bwd.d_eval32.p : (const void*)s.bwd.d_eval.p;
      d.r = sg.v.r;
      d.z = sg.v.z;
      d.p = sg.v.p;
      d.q = sg.v.q;
      d.x = sg.v.x;
      d.x64 = sg.v.x64;
      d.b64 = sg.v.b64;
      d.sc = bt.d_sc.p + kNScal * (size_t)sg.slot;
      d.fl = bt.d_fl.p + 8 * (size_t)sg.slot;
      d.nslice = s.sell.nslice();
      d.sptr = s.sell.d_sptr.p;
      d.scol = s.sell.d_col.p;
      d.sval = sg.mode == DFM_MIXED ? (const void*)s.sell.d_val32.p : (const void*)s.sell.d_val.p;
      d.sval64 = s.sell.d_val.p;
      d.ps0 = (int)sblk.size();
      for (int sl = 0; sl < d.nslice; sl += kBSellWaves) sblk.push_back(si);
      d.ps1 = (int)sblk.size();
      d.pv0 = (int)vblk.size();
      for (long long i = 0; i < (long long)d.m; i += DV) vblk.push_back(si);
      d.pv1 = (int)vblk.size();
      d.pr0 = nfslot;
      d.pz0 = nbslot;
      for (int dir = 0; dir < 2; dir++) {
        const Sweep& w = dir == 0 ? s.fwd : s.bwd;
        auto& ph = dir == 0 ? fph : bph;
        auto& psl = dir == 0 ? fsl : bsl;
        int& nslot = dir == 0 ? nfslot : nbslot;
        auto& po = dir == 0 ? fwo : bwo;
        auto& pi = dir == 0 ? fwi : bwi;
        if ((int)ph.size() < w.nphase()) ph.resize(w.nphase());
        if ((int)psl.size() < w.nphase()) psl.resize(w.nphase());
        if ((int)po.size() < w.nphase()) {
          po.resize(w.nphase(), 0);
          pi.resize(w.nphase(), 0);
        }
        for (int p = 0; p < w.nphase(); p++) {
          po[p] = std::max(po[p], w.phase_wo[p]);
          pi[p] = std::max(pi[p], w.phase_wi[p]);
        }
        for (int p = 0; p < w.nphase(); p++)
          for (int q = w.phase_gptr[p]; q < w.phase_gptr[p + 1]; q++) {
            ph[p].push_back(make_int2(si, q));
            psl[p].push_back(nslot++);
          }
      }
      d.pr1 = nfslot;
      d.pz1 = nbslot;
      h.push_back(d);
    }
    g.nsb = (int)sblk.size();
    g.nvb = (int)vblk.size();
    g.fwo = fwo;
    g.fwi = fwi;
    g.bwo = bwo;
    g.bwi = bwi;
    if (h.empty()) continue;
    std::vector<int2> fe, be;
    std::vector<int> fs, bs;
    g.fptr.assign(1, 0);
    g.bptr.assign(1, 0);
    for (size_t p = 0; p < fph.size(); p++) {
      fe.insert(fe.end(), fph[p].begin(), fph[p].end());
      fs.insert(fs.end(), fsl[p].begin(), fsl[p].end());
      g.fptr.push_back((int)fe.size());
    }
    for (size_t p = 0; p < bph.size(); p++) {
      be.insert(be.end(), bph[p].begin(), bph[p].end());
      bs.insert(bs.end(), bsl[p].begin(), bsl[p].end());
      g.bptr.push_back((int)be.size());
    }
    hipError_t e = hipSuccess;
    grow(g.d_segs, h.size(), e);
    grow(g.d_sblk, sblk.size(), e);
    grow(g.d_vblk, vblk.size(), e);
    grow(g.d_fent, fe.size(), e);
    grow(g.d_bent, be.size(), e);
    grow(g.d_fslot, fs.size(), e);
    grow(g.d_bslot, bs.size(), e);
    grow(g.d_pq, sblk.size(), e);
    grow(g.d_true, sblk.size(), e);
    grow(g.d_rr, (size_t)nfslot, e);
    grow(g.d_rz, (size_t)nbslot, e);
    DFM_HIP(e);
    DFM_HIP(hipMemcpyAsync(g.d_segs.p, h.data(), h.size() * sizeof(SegDev), hipMemcpyHostToDevice, st));
    DFM_HIP(hipMemcpyAsync(g.d_sblk.p, sblk.data(), sblk.size() * sizeof(int), hipMemcpyHostToDevice, st));
    DFM_HIP(hipMemcpyAsync(g.d_vblk.p, vblk.data(), vblk.size() * sizeof(int), hipMemcpyHostToDevice, st));
    if (!fe.empty()) DFM_HIP(hipMemcpyAsync(g.d_fent.p, fe.data(), fe.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    if (!be.empty()) DFM_HIP(hipMemcpyAsync(g.d_bent.p, be.data(), be.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    if (!fs.empty()) DFM_HIP(hipMemcpyAsync(g.d_fslot.p, fs.data(), fs.size() * sizeof(int), hipMemcpyHostToDevice, st));
    if (!bs.empty()) DFM_HIP(hipMemcpyAsync(g.d_bslot.p, bs.data(), bs.size() * sizeof(int), hipMemcpyHostToDevice, st));
    // the tables are host temporaries: the copies must finish before they go out of scope
    DFM_HIP(hipStreamSynchronize(st));
  }
  bt.dirty = false;
  return DFM_OK;
}

// Where an iteration's kernels go: launched on a stream, or appended as kernel nodes to a graph (each depending
// on the one before), built explicitly rather than by stream capture, so that other threads' device-wide calls (a
// server's setup workers allocate and free) cannot invalidate it.
struct Launcher {
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;
  hipError_t err = hipSuccess;
  // the argument values and their pointer arrays, kept for the life of the graph made from them
  std::vector<std::shared_ptr<void>> keep;
  template <typename... KA>
  struct Args {
    std::tuple<KA...> vals;
    void* ptrs[sizeof...(KA) > 0 ? sizeof...(KA) : 1];
  };
  template <typename... KA, typename... A>
  void operator()(void (*kern)(KA...), dim3 grid, dim3 block, A... args) {
    if (err != hipSuccess) return;
    if (!graph) {
      hipLaunchKernelGGL(kern, grid, block, 0, st, args...);
      err = hipGetLastError();
      return;
    }
    auto a = std::make_shared<Args<KA...>>(Args<KA...>{std::tuple<KA...>(static_cast<KA>(args)...), {}});
    std::apply([&](auto&... v) {
      int k = 0;
      ((a->ptrs[k++] = (void*)&v), ...);
    }, a->vals);
    keep.push_back(a);
    hipKernelNodeParams p{};
    p.func = (void*)kern;
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = 0;
    p.kernelParams = a->ptrs;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    err = hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &p);
    if (err == hipSuccess) last = node;
  }
};

// T: vectors; MX: reliable-update loop; TF: factor values
template <int B, typename T, bool MX, typename TF>
static hipError_t iterate_group(Batch::Group& g, Launcher& L) {
  using TM = std::conditional_t<MX, float, double>;
  const SegDev* segs = g.d_segs.p;
  const dim3 nseg((unsigned)g.members.size());
  if (g.nsb)
    L(k_b_spmv<B, TM, T, MX>, dim3(g.nsb), dim3(64 * kBSellWaves), segs, (const int*)g.d_sblk.p, g.d_pq.p,
      g.d_true.p);
  for (size_t p = 0; p + 1 < g.fptr.size(); p++) {
    const int c = g.fptr[p + 1] - g.fptr[p];
    if (c)
      with_slab_caps(g.fwo[p], g.fwi[p], B == 3 && sizeof(T) == 8 && sizeof(TF) == 8, [&]<int MO, int MI>() {
        L(k_b_fwd<B, MO, MI, TF, T>, dim3((c + kWPB - 1) / kWPB), dim3(64 * kWPB), segs,
          (const int2*)g.d_fent.p + g.fptr[p], (const int*)g.d_fslot.p + g.fptr[p], c, g.d_rr.p,
          p == 0 ? (const double*)g.d_pq.p : (const double*)nullptr);
      });
  }
  for (size_t p = 0; p + 1 < g.bptr.size(); p++) {
    const int c = g.bptr[p + 1] - g.bptr[p];
    if (c)
      with_slab_caps(g.bwo[p], g.bwi[p], B == 3 && sizeof(T) == 8 && sizeof(TF) == 8, [&]<int MO, int MI>() {
        L(k_b_bwd<B, MO, MI, TF, T>, dim3((c + kWPB - 1) / kWPB), dim3(64 * kWPB), segs,
          (const int2*)g.d_bent.p + g.bptr[p], (const int*)g.d_bslot.p + g.bptr[p], c, g.d_rz.p);
      });
  }
  L(k_b_judge_beta<MX>, nseg, dim3(NT), segs, (const double*)g.d_rr.p, (const double*)g.d_rz.p,
    (const double*)g.d_true.p);
  // the scalar mixed SpMV forms p itself
  if (!(MX && B == 1)) L(k_b_pupdate<T, MX>, dim3(g.nvb), dim3(NT), segs, (const int*)g.d_vblk.p);
  return L.err;
}

// A mixed-precision segment whose true residual stalled restarts as an FP64 solve of the same system with the
// FP32 factor (its residual handle still holds the right-hand side).
// The FP32 matrix limits progress: the system continues from its FP64 solution with FP64 vectors and matrix and
// the FP32 factor (as solve_pcg does); max_iter bounds both loops together.
static int restart_fp64(Batch& bt, BatchSeg& sg) {
  Sys& s = *sg.s;
  sg.mode = DFM_FP32_ILU;
  sg.verify = true;
  sg.fell_back = true;
  sg.nrel = bt.h_fl[8 * sg.slot + F_NREL];
  sg.before = (unsigned)bt.h_fl[8 * sg.slot + F_IT] + 1;
  int rc = sys_solve_setup(s, DFM_FP32_ILU);
  if (rc != DFM_OK) return rc;
  if (s.ctx != bt.ctx) DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  s.prepared = -1;
  sg.v = pcg_bufs(s, false);
  sg.v.f32ilu = true;
  const unsigned left = sg.maxit > sg.before ? sg.maxit - sg.before : 0;
  rc = pcg_start_warm(s, sg.v, sg.tol, left, bt.d_sc.p + kNScal * (size_t)sg.slot, bt.d_fl.p + 8 * (size_t)sg.slot,
                      bt.ctx->stream);
  bt.dirty = true;
  return rc;
}

static int finish_seg(Batch& bt, BatchSeg& sg) {
  const int* fl = bt.h_fl + 8 * sg.slot;
  const double* sc = bt.h_sc + kNScal * sg.slot;
  const int done = fl[F_DONE];
  if (done == 3) return restart_fp64(bt, sg);
  sg.state = done == 1 ? 1 : 2;
  sg.conv = done ? sc[S_CONV] : std::sqrt(sc[S_RR] * sc[S_INV0]);
  sg.iters = done == 1 ? (unsigned)fl[F_ITOUT] + sg.before : sg.maxit;
  if (done != 1 && sg.restarts > 0) {
    // a restarted loop ran out of iterations: report the true residual of its solution
    const int rc = sys_true_residual(*sg.s, bt.ctx->stream, &sg.conv);
    if (rc != DFM_OK) return rc;
  }
  if (done == 1 && sg.verify) {
    // the FP64 loop with the FP32 factor stopped on its recursive residual: the true residual decides, and the
    // loop restarts from its solution (r = b - A x) while it is above the tolerance, as solve_pcg does
    Sys& s = *sg.s;
    double rel = 0.0;
    int rc = sys_true_residual(s, bt.ctx->stream, &rel);
    if (rc != DFM_OK) return rc;
    sg.conv = rel;
    const unsigned used = sg.iters + 1;
    if (rel > sg.tol) {
      if (used >= sg.maxit || sg.restarts == kMaxRestarts || (sg.restarts > 0 && rel > kRestartGain * sg.prev_rel)) {
        sg.state = 2;
        sg.iters = sg.maxit;
      } else {
        sg.restarts++;
        bt.n_restart++;
        sg.prev_rel = rel;
        sg.before = used;
        sg.state = 0;
        DFM_HIP(launch_copy(s.d_rhs.p, s.d_res.p, (long long)s.n * s.len, bt.ctx->stream));
        rc = pcg_start_warm(s, sg.v, 0.5 * sg.tol, sg.maxit - used, bt.d_sc.p + kNScal * (size_t)sg.slot,
                            bt.d_fl.p + 8 * (size_t)sg.slot, bt.ctx->stream);
        bt.dirty = true;
        return rc;
      }
    }
  }
  if (!sg.fell_back) sg.nrel = fl[F_NREL];
  sg.t_done = now_s();
  const int rc = pcg_finish(*sg.s, sg.v, sg.mode == DFM_MIXED, done, bt.ctx->stream);
  bt.dirty = true;
  return rc;
}

static int poll_flags(Batch& bt);
static int poll(Batch& bt) {
  const double t0 = now_s();
  const int rc = poll_flags(bt);
  bt.t_poll += now_s() - t0;
  return rc;
}

static int poll_flags(Batch& bt) {
  hipStream_t st = bt.ctx->stream;
  DFM_HIP(hipMemcpyAsync(bt.h_fl, bt.d_fl.p, sizeof(int) * 8 * bt.cap, hipMemcpyDeviceToHost, st));
  DFM_HIP(hipMemcpyAsync(bt.h_sc, bt.d_sc.p, sizeof(double) * kNScal * bt.cap, hipMemcpyDeviceToHost, st));
  DFM_HIP(hipStreamSynchronize(st));
  bool any = false;
  for (auto& sg : bt.segs)
    if (sg.state == 0 && bt.h_fl[8 * sg.slot + F_DONE]) {
      const int rc = finish_seg(bt, sg);
      if (rc != DFM_OK) return rc;
      any = sg.state != 0 || any;
    }
  // finished systems are read through their own context's stream: complete the batch stream's work on them
  if (any) DFM_HIP(hipStreamSynchronize(st));
  return DFM_OK;
}

static int find_seg(Batch& bt, Sys* s) {
  for (int k = 0; k < (int)bt.segs.size(); k++)
    if (bt.segs[k].s == s) return k;
  return -1;
}

void batch_stats(const Batch* bt, double* out) {
  out[0] = bt->t_rebuild;
  out[1] = (double)bt->n_rebuild;
  out[2] = bt->t_poll;
  out[3] = (double)bt->n_restart;
}

void batch_detach(Batch* bt, Sys* s) {
  if (!bt) return;
  const int k = find_seg(*bt, s);
  if (k < 0) return;
  (void)hipStreamSynchronize(bt->ctx->stream);
  bt->free_slots.push_back(bt->segs[k].slot);
  bt->segs.erase(bt->segs.begin() + k);
  s->batch = nullptr;
  bt->dirty = true;
}

}  // namespace dfm

using namespace dfm;

struct dfm_batch_s {
  Batch b;
  dfm_ctx_s* owner = nullptr;
};

Batch* dfm::batch_of(dfm_batch b) { return &b->b; }

extern "C" {

int dfm_batch_create(dfm_ctx ctx, int precision, int max_systems, dfm_batch* out) {
  DFM_CHECK(ctx && out, DFM_ERR_ARG, "dfm_batch_create: null argument");
  DFM_CHECK(precision == DFM_FP64 || precision == DFM_MIXED || precision == DFM_FP32_ILU ||
                precision == DFM_MIXED_AUTO || precision == DFM_FP64_TRUE, DFM_ERR_ARG,
            "dfm_batch_create: bad precision");
  DFM_CHECK(max_systems > 0 && max_systems <= (1 << 20), DFM_ERR_ARG, "dfm_batch_create: bad max_systems");
  DFM_HIP(hipSetDevice(ctx->c.device));
  auto* h = new dfm_batch_s;
  Batch& bt = h->b;
  bt.ctx = &ctx->c;
  bt.prec = precision;
  bt.cap = max_systems;
  if (bt.d_sc.alloc((size_t)kNScal * bt.cap) != hipSuccess || bt.d_fl.alloc((size_t)8 * bt.cap) != hipSuccess ||
      hipHostMalloc((void**)&bt.h_fl, sizeof(int) * 8 * bt.cap) != hipSuccess ||
      hipHostMalloc((void**)&bt.h_sc, sizeof(double) * kNScal * bt.cap) != hipSuccess) {
    delete h;
    set_error("dfm_batch_create: out of memory");
    return DFM_ERR_NOMEM;
  }
  // flags of free slots read as done, so no stale slot is ever taken for a running one
  std::vector<int> f((size_t)8 * bt.cap, 0);
  for (int k = 0; k < bt.cap; k++) f[8 * (size_t)k + F_DONE] = 3;
  if (hipMemcpy(bt.d_fl.p, f.data(), f.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    delete h;
    set_error("dfm_batch_create: copy failed");
    return DFM_ERR_HIP;
  }
  for (int k = bt.cap - 1; k >= 0; k--) bt.free_slots.push_back(k);
  ctx->c.refs++;
  h->owner = ctx;
  *out = h;
  return DFM_OK;
}

int dfm_batch_destroy(dfm_batch b) {
  if (!b) return DFM_OK;
  Batch& bt = b->b;
  (void)hipStreamSynchronize(bt.ctx->stream);
  for (auto& sg : bt.segs) sg.s->batch = nullptr;
  dfm_ctx owner = b->owner;
  delete b;
  if (owner) ctx_release(owner);
  return DFM_OK;
}

int dfm_batch_add(dfm_batch b, dfm_sys sys, double tol, unsigned max_iter) {
  DFM_CHECK(b && sys, DFM_ERR_ARG, "dfm_batch_add: null argument");
  Batch& bt = b->b;
  Sys& s = sys->s;
  DFM_CHECK(s.ctx->device == bt.ctx->device, DFM_ERR_ARG, "dfm_batch_add: the system is on another device");
  DFM_CHECK(!s.batch, DFM_ERR_STATE, "dfm_batch_add: the system is already in a batch");
  DFM_CHECK(!bt.free_slots.empty(), DFM_ERR_STATE, "dfm_batch_add: batch full");
  DFM_CHECK(max_iter > 0, DFM_ERR_ARG, "dfm_batch_add: max_iter must be positive");
  const int prec = resolve_precision(s, bt.prec);
  int rc = DFM_OK;
  if (s.prepared != prec) {
    rc = sys_solve_setup(s, prec);
    if (rc != DFM_OK) return rc;
    // prepared on another context's stream: the data must be complete before the batch stream reads it
    if (s.ctx != bt.ctx) DFM_HIP(hipStreamSynchronize(s.ctx->stream));
  }
  s.prepared = -1;  // the right-hand side copy is consumed by this solve
  BatchSeg sg;
  sg.s = &s;
  sg.slot = bt.free_slots.back();
  sg.mode = prec == DFM_FP64_TRUE ? DFM_FP64 : prec;  // the FP64 kernels; the check at exit is the difference
  sg.verify = prec == DFM_FP32_ILU || prec == DFM_FP64_TRUE;
  sg.v = pcg_bufs(s, prec == DFM_MIXED);
  sg.v.f32ilu = prec == DFM_FP32_ILU;
  sg.tol = tol;
  sg.maxit = max_iter;
  hipStream_t st = bt.ctx->stream;
  rc = pcg_prepare(s, sg.v, prec == DFM_MIXED, st);
  if (rc == DFM_OK)
    rc = pcg_start(s, sg.v, prec == DFM_MIXED, tol, max_iter, bt.d_sc.p + kNScal * (size_t)sg.slot,
                   bt.d_fl.p + 8 * (size_t)sg.slot, st);
  if (rc != DFM_OK) return rc;
  bt.free_slots.pop_back();
  sg.t_add = now_s();
  bt.segs.push_back(sg);
  s.batch = &bt;
  bt.dirty = true;
  return DFM_OK;
}

int dfm_batch_step(dfm_batch b, int iterations, int* n_running) {
  DFM_CHECK(b && iterations >= 0, DFM_ERR_ARG, "dfm_batch_step: bad argument");
  Batch& bt = b->b;
  hipStream_t st = bt.ctx->stream;
  int rc = DFM_OK;
  if (bt.dirty) {
    // segments whose start already decided them (zero right-hand side) leave before the first iteration
    rc = poll(bt);
    if (rc != DFM_OK) return rc;
    rc = rebuild(bt);
    if (rc != DFM_OK) return rc;
  }
  // one iteration of every group (the kernels read every scalar and flag from device memory, so repeating the
  // same launches is the next iteration)
  auto iteration = [&](Launcher& L) -> hipError_t {
    for (auto& g : bt.grp) {
      if (g.members.empty()) continue;
      hipError_t e;
      if (g.mode == DFM_MIXED)
        e = g.B == 1 ? iterate_group<1, float, true, float>(g, L)
                     : g.B == 2 ? iterate_group<2, float, true, float>(g, L) : iterate_group<3, float, true, float>(g, L);
      else if (g.mode == DFM_FP32_ILU)
        e = g.B == 1 ? iterate_group<1, double, false, float>(g, L)
                     : g.B == 2 ? iterate_group<2, double, false, float>(g, L)
                                : iterate_group<3, double, false, float>(g, L);
      else
        e = g.B == 1 ? iterate_group<1, double, false, double>(g, L)
                     : g.B == 2 ? iterate_group<2, double, false, double>(g, L)
                                : iterate_group<3, double, false, double>(g, L);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  // The iteration as a graph, built once per membership (rebuild) from the second step of that membership on (a
  // serving stream whose members change nearly every step would pay the build for one step's launches); the
  // previous step ended with a poll, so the stream is idle when an old graph is destroyed.
  const bool use_graph = iterations > 1 && bt.gen_steps++ > 0 && getenv("DFM_NO_GRAPH") == nullptr;
  if (use_graph && bt.gexec_gen != bt.gen) {
    if (bt.gexec) {
      (void)hipGraphExecDestroy(bt.gexec);
      bt.gexec = nullptr;
    }
    bt.gargs.clear();
    Launcher L;
    DFM_HIP(hipGraphCreate(&L.graph, 0));
    const hipError_t eb = iteration(L);
    hipError_t ei = eb;
    if (eb == hipSuccess) ei = hipGraphInstantiate(&bt.gexec, L.graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(L.graph);
    DFM_HIP(ei);
    bt.gargs = std::move(L.keep);
    bt.gexec_gen = bt.gen;
  }
  Launcher plain;
  plain.st = st;
  for (int t = 0; t < iterations; t++) {
    if (use_graph) DFM_HIP(hipGraphLaunch(bt.gexec, st));
    else DFM_HIP(iteration(plain));
  }
  rc = poll(bt);
  if (rc != DFM_OK) return rc;
  int nr = 0;
  for (auto& sg : bt.segs) nr += sg.state == 0;
  if (n_running) *n_running = nr;
  return DFM_OK;
}

int dfm_batch_status(dfm_batch b, dfm_sys sys, int* state, double* conv_ratio, unsigned* iteration,
                     double* seconds) {
  DFM_CHECK(b && sys, DFM_ERR_ARG, "dfm_batch_status: null argument");
  Batch& bt = b->b;
  const int k = find_seg(bt, &sys->s);
  DFM_CHECK(k >= 0, DFM_ERR_ARG, "dfm_batch_status: the system is not in this batch");
  const BatchSeg& sg = bt.segs[k];
  if (state) *state = sg.state;
  if (conv_ratio) *conv_ratio = sg.conv;
  if (iteration) *iteration = sg.iters;
  if (seconds) *seconds = sg.state ? sg.t_done - sg.t_add : now_s() - sg.t_add;
  return DFM_OK;
}

int dfm_batch_info(dfm_batch b, dfm_sys sys, int* reliable_updates, int* fell_back) {
  DFM_CHECK(b && sys, DFM_ERR_ARG, "dfm_batch_info: null argument");
  Batch& bt = b->b;
  const int k = find_seg(bt, &sys->s);
  DFM_CHECK(k >= 0, DFM_ERR_ARG, "dfm_batch_info: the system is not in this batch");
  if (reliable_updates) *reliable_updates = bt.segs[k].nrel;
  if (fell_back) *fell_back = bt.segs[k].fell_back ? 1 : 0;
  return DFM_OK;
}

int dfm_batch_remove(dfm_batch b, dfm_sys sys) {
  DFM_CHECK(b && sys, DFM_ERR_ARG, "dfm_batch_remove: null argument");
  Batch& bt = b->b;
  DFM_CHECK(find_seg(bt, &sys->s) >= 0, DFM_ERR_ARG, "dfm_batch_remove: the system is not in this batch");
  batch_detach(&bt, &sys->s);
  return DFM_OK;
}

}  // extern "C"
