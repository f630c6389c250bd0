// Device building blocks shared by k_linalg.hip and solver.hip:
//  - deterministic reductions (wavefront shuffles + LDS tree, per-block partials reduced by the consuming
//    kernel: the partition and the summation tree are fixed, so results do not depend on scheduling),
//  - the block-CRS SpMV row (TPR lanes per block row, diagonal block kept apart as in the reference),
//  - the ILU(0) phase kernels (one wavefront per group of <= 64 rows, intra-group levels through LDS).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

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
static_assert(kMaxPartial % 8 == 0, "strided grids keep workgroup positions on their XCD");

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

// Mixed-precision stagnation test, called with each FP64 true residual rt (squared norm) that did not meet the
// tolerance: progress means at least halving the norm against the best one so far. With an FP32 matrix the
// recurrence solves A32 x = b; when cond(A) * eps32 is not small the true residual of A64 stops falling while
// reliable updates keep coming. Returns true after kStallUpdates updates in a row without progress.
constexpr int kStallUpdates = 6;
// A reliable update is forced after this many iterations without one, so that a recursive residual that has
// stopped falling (an FP32 plateau) still reaches the stagnation test.
constexpr int kForceUpdate = 200;
__device__ inline bool stall_update(double rt, double& tbest, int& stall) {
  if (tbest == 0.0 || rt < 0.25 * tbest) {
    tbest = rt;
    stall = 0;
    return false;
  }
  stall = stall + 1;
  return stall >= kStallUpdates;
}

// FP32-matrix progress test at a reliable update (squared norms): rt, the FP64 true residual; rrec, the recursive
// residual that scheduled the update (same iterate); tprev, the true residual at the previous update (||r0||^2
// before the first; updated here). A cycle is slow when the FP32 matrix held the true residual well above the
// recursive one (rt > kGap2 rrec), or when the true residual gained less than 4x over the cycle (the FP32
// recurrence crawls: the cycles of a healthy run gain about 10x, delta = 0.1). kSlowCycles slow cycles in a row,
// or kStallUpdates updates without halving the best true residual, end the FP32-matrix loop: a cycle must gain
// about 5x for the FP32 matrix and vectors to pay off against FP64 vectors with the FP32 factor.
constexpr double kGap2 = 4.0;
constexpr double kProg2 = 1.0 / 16.0;
constexpr int kSlowCycles = 2;
__device__ inline bool mx_slow(double rt, double rrec, double& tprev, double& slow, double& tbest, int& stall) {
  const bool st = stall_update(rt, tbest, stall);
  const bool cycle_slow = rt > kGap2 * rrec || rt > kProg2 * tprev;
  tprev = rt;
  slow = cycle_slow ? slow + 1.0 : 0.0;
  return st || slow >= (double)kSlowCycles;
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

// x: a vector pointer, or any value indexable like one (SumVec)
template <int B, typename T, typename A, typename VX, typename XP = const VX*>
__device__ inline void dia_row(const T* __restrict__ dia, XP x, int i, A (&acc)[B]) {
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

// SELL-64 SpMV row (one row per lane): acc = sum over the slice's w entries of the lane, in CRS order; CH
// entries' columns and values are loaded together, then their x gathered, then accumulated. Padding entries
// (value 0, column 0) add exact zeros. The caller adds the diagonal block last, as MatVec does
// (src/matvec/matdia_blkcrs.cpp:1008-1025).
template <int B, int CH, typename T, typename A, typename VX, typename XP = const VX*>
__device__ inline void sell_row(const unsigned* __restrict__ col, const T* __restrict__ val, int r0, int w, int lane,
                                XP x, A (&acc)[B]) {
#pragma unroll
  for (int r = 0; r < B; r++) acc[r] = A(0);
  for (int k0 = 0; k0 < w; k0 += CH) {
    unsigned j[CH];
    A a[CH][B * B], xv[CH][B];
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (k0 + k < w) {
        const size_t sr = (size_t)(r0 + k0 + k);
        j[k] = col[sr * 64 + lane];
#pragma unroll
        for (int e = 0; e < B * B; e++) a[k][e] = (A)val[(sr * (B * B) + e) * 64 + lane];
      }
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (k0 + k < w)
#pragma unroll
        for (int c = 0; c < B; c++) xv[k][c] = (A)x[(size_t)j[k] * B + c];
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (k0 + k < w)
#pragma unroll
        for (int r = 0; r < B; r++) {
          A t = a[k][r * B] * xv[k][0];
#pragma unroll
          for (int c = 1; c < B; c++) t += a[k][r * B + c] * xv[k][c];
          acc[r] += t;
        }
  }
}

// x_hi + x_lo (FP64 solution plus FP32 partial solution) read like one FP64 vector by the SpMV helpers
struct SumVec {
  const double* __restrict__ hi;
  const float* __restrict__ lo;
  __device__ double operator[](size_t i) const { return hi[i] + (double)lo[i]; }
};

template <int B>
struct SellChunk {
  static constexpr int CH = B == 1 ? 8 : 4;
};

// ---------------------------------------------------------------------------------------------------------
// ILU(0) sweeps over the slab layout. Forward (ForwardSubstitution, src/matvec/matdiafrac_blkcrs.cpp:332-507):
//   t = b_i - sum_{j<i} L_ij x_j (ascending j), x_i = D_i^-1 t.
// Backward (BackwardSubstitution, :509-669): x_i -= sum_{j>i} U_ij x_j (ascending j; D^-1 folded into U).
// One wavefront per group, lane = row, WPB independent wavefronts per workgroup. Entries referencing rows of
// earlier phases are final in global memory: their columns, values and gathered x are loaded before the level
// loop. In-group entries are resolved level by level through the wavefront's own LDS row, with only a
// wavefront-scope fence between levels. Colour-major orderings put the outside entries of a forward row
// below its in-group ones and those of a backward row above them, so each class is applied in the position
// the ascending order gives it: the sweeps reproduce the reference's floating-point order exactly.

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

// t -= a v for a block held in registers (row-major b x b)
template <int B, typename A>
__device__ inline void blk_sub_reg(A (&t)[B], const A (&a)[B * B], const A (&v)[B]) {
#pragma unroll
  for (int r = 0; r < B; r++) {
    A s = a[r * B] * v[0];
#pragma unroll
    for (int c = 1; c < B; c++) s += a[r * B + c] * v[c];
    t[r] -= s;
  }
}

// MO / MI: entries per row of the two classes held in registers from the start of the kernel (template
// parameters chosen per phase from its widest group: the level loop costs in proportion to them); further
// entries are read from the slab when they are applied.
inline int slab_cap(int w) { return w <= 4 ? 4 : 8; }

constexpr int kWPB = 4;  // wavefronts (groups) per workgroup of the sweep kernels

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct SlabArgs {
  const int* gmeta;        // 4 ints per group
  const int* srow;         // row per slot (indirect groups)
  const unsigned* scnt;    // nout | nin << 8 | level << 16 per slot
  const unsigned* ecol;    // per entry
};

// Per-slot state common to both sweeps. The group descriptor is wave-uniform (scalar loads); every vector load
// below depends only on the slot, so all of them are in flight together. bo / bi: first entry of the group's
// outside and in-group classes. Empty lanes map to row 0 and have no entries.
struct SlabSlot {
  int g, lane, bo, bi, wo, wi, nlev, row, no, ni, lv;
  bool active;
};

// Exclusive prefix sum over the wavefront of 8-bit lane values, from one ballot per bit (all lanes of the
// wavefront call it).
__device__ inline int wave_scan_u8(unsigned v) {
  int s = 0;
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const unsigned long long m = __ballot((v >> b) & 1u);
    s += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u)) << b;
  }
  return s;
}

// Lanes below the calling one in mask m.
__device__ inline int lanes_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// g: the group (wave-uniform); returns false when the system's done flag is set.
__device__ inline bool slab_slot(const SlabArgs& sa, int g, const int* done, SlabSlot& ss) {
  ss.g = g;
  const int* m = sa.gmeta + 4 * (size_t)g;
  const int m0 = m[0], m1 = m[1], row0 = m[2], m3 = m[3];
  if (done && *done) return false;
  const int nrows = m3 & 255;
  ss.bo = m0;
  ss.bi = m0 + (m3 >> 8);
  ss.wo = m1 & 255;
  ss.wi = (m1 >> 8) & 255;
  ss.nlev = m1 >> 16;
  ss.lane = (int)(threadIdx.x % 64);
  ss.active = ss.lane < nrows;
  const size_t slot = (size_t)g * 64 + ss.lane;
  ss.row = row0 >= 0 ? (ss.active ? row0 + ss.lane : row0) : (ss.active ? sa.srow[slot] : 0);
  const unsigned c = sa.scnt[slot];
  ss.no = (int)(c & 255u);
  ss.ni = (int)((c >> 8) & 255u);
  ss.lv = ss.active ? (int)(c >> 16) : -1;
  return true;
}

__device__ inline int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64)); }

// Entries per lane and class stored in jagged rows. A class of a group's entries (see Sweep) starts with its
// jagged rows: row k < kJag holds the k-th entry of every lane that has one, in lane order, with the row's b*b
// values element-major (value e of the row's p-th entry at (row start)*bs + e*(row length) + p), so a wavefront
// reads a row as one contiguous segment per value; a lane finds its place in the row from a ballot of the lanes'
// counts. Entries k >= kJag follow, lane after lane, each entry's values together: a lane finds them from its
// first one alone, also inside branches that only part of the wavefront takes.
constexpr int kJag = 4;

// Where the lane's class entries k >= kJag lie.
struct ClassTail {
  int n;    // the lane's count
  int csr;  // the lane's first entry with k >= kJag
};

// How a lane treats register-held entries it does not have (measured choice): with 1x1 blocks it loads the first
// entry of the row (or of the class) anyway, so that every load is unconditional and in flight at once; with
// larger blocks, where such loads would fetch b*b values each, it loads nothing and holds column 0 and zeros.
// Either way what it holds is valid to use, and the arithmetic, predicated by its count, discards it.
template <int B>
struct SlabLoads {
  static constexpr bool kAll = B == 1;
  using Idx = std::conditional_t<B == 1, unsigned, size_t>;  // entry index arithmetic (32-bit keeps 1x1 lean)
};

// Whole wavefront: loads the register-held entries k < MX of one class (base: its first entry; w: the
// wave-uniform width; n: the lane's count).
template <int B, int MX, typename T>
__device__ inline void slab_load_class(const unsigned* __restrict__ ecol, const T* __restrict__ val, int base, int w,
                                       int n, unsigned (&col)[MX], T (&v)[MX][B * B], ClassTail& ct) {
  static_assert(MX >= kJag, "the jagged rows are register-held");
  constexpr int BS = B * B;
  constexpr bool all = SlabLoads<B>::kAll;
  if (!all) {
#pragma unroll
    for (int k = 0; k < MX; k++)
      if (k < w) {
        col[k] = 0u;
#pragma unroll
        for (int q = 0; q < BS; q++) v[k][q] = T(0);
      }
  }
  int r = base;
#pragma unroll
  for (int k = 0; k < kJag; k++)
    if (k < w) {
      const unsigned long long m = __ballot(n > k);
      const unsigned len = (unsigned)__popcll(m), pos = k < n ? (unsigned)lanes_below(m) : 0u;
      if (all) {
        // a wave-uniform row base plus a 32-bit lane offset
        const unsigned* crow = ecol + r;
        const T* vrow = val + (size_t)r * BS;
        col[k] = crow[pos];
#pragma unroll
        for (int q = 0; q < BS; q++) v[k][q] = vrow[(unsigned)q * len + pos];
      } else if (k < n) {
        col[k] = ecol[(size_t)r + pos];
#pragma unroll
        for (int q = 0; q < BS; q++) v[k][q] = val[(size_t)r * BS + (size_t)q * len + pos];
      }
      r += (int)len;
    }
  ct.n = n;
  ct.csr = r + (w > kJag ? wave_scan_u8(n > kJag ? (unsigned)(n - kJag) : 0u) : 0);
  using Idx = typename SlabLoads<B>::Idx;
#pragma unroll
  for (int k = kJag; k < MX; k++)
    if (k < w && (all || k < n)) {
      const Idx e = k < n ? (Idx)ct.csr + (Idx)(k - kJag) : (Idx)base;
      col[k] = ecol[e];
#pragma unroll
      for (int q = 0; q < BS; q++) v[k][q] = val[e * (Idx)BS + (Idx)q];
    }
}

// Calls f(column, values) for the lane's class entries k = MX .. n-1 in ascending k (safe in divergent code).
template <int B, int MX, typename T, typename A, typename F>
__device__ inline void class_tail(const unsigned* __restrict__ ecol, const T* __restrict__ val, const ClassTail& ct,
                                  F&& f) {
  constexpr int BS = B * B;
  using Idx = typename SlabLoads<B>::Idx;
  for (int k = MX; k < ct.n; k++) {
    const Idx e = (Idx)ct.csr + (Idx)(k - kJag);
    A a[BS];
#pragma unroll
    for (int q = 0; q < BS; q++) a[q] = (A)val[e * (Idx)BS + (Idx)q];
    f(ecol[e], a);
  }
}

// x of the register-held outside entries (as slab_load_class: all of them with 1x1 blocks, else those the lane
// has, zeros for the others)
template <int B, int MX, typename A, typename VX>
__device__ inline void slab_gather(const VX* x, int w, int n, const unsigned (&col)[MX], A (&xo)[MX][B]) {
  constexpr bool all = SlabLoads<B>::kAll;
#pragma unroll
  for (int k = 0; k < MX; k++)
    if (k < w) {
      if (!all)
#pragma unroll
        for (int c = 0; c < B; c++) xo[k][c] = A(0);
      if (all || k < n)
#pragma unroll
        for (int c = 0; c < B; c++) xo[k][c] = (A)x[(size_t)col[k] * B + c];
    }
}

// t -= sum over the lane's outside entries beyond the register-held ones, x from global memory (final)
template <int B, int MX, typename T, typename A, typename VX>
__device__ inline void slab_tail_global(const unsigned* __restrict__ ecol, const T* __restrict__ val,
                                        const ClassTail& ct, const VX* x, A (&t)[B]) {
  class_tail<B, MX, T, A>(ecol, val, ct, [&](unsigned j, const A (&a)[B * B]) {
    A v[B];
#pragma unroll
    for (int c = 0; c < B; c++) v[c] = (A)x[(size_t)j * B + c];
    blk_sub_reg<B, A>(t, a, v);
  });
}

// t -= v y where the lane's own count says the entry exists (k < n); a select, not a branch, so the
// wavefront never splits around the loads that feed it.
template <int B, typename A, typename TV>
__device__ inline void blk_sub_if(bool on, A (&t)[B], const TV (&a)[B * B], const A (&v)[B]) {
#pragma unroll
  for (int r = 0; r < B; r++) {
    A s = (A)a[r * B] * v[0];
#pragma unroll
    for (int c = 1; c < B; c++) s += (A)a[r * B + c] * v[c];
    const A u = t[r] - s;
    t[r] = on ? u : t[r];
  }
}

// t -= sum over the in-group entries of the lane (values of this wavefront's rows from LDS), ascending. All
// reads of the register-held entries are issued together (every column held is a valid lane).
template <int B, int MX, typename T, typename A>
__device__ inline void slab_in_group(const unsigned* __restrict__ ecol, const T* __restrict__ val, int w, int n,
                                     const unsigned (&col)[MX], const T (&v)[MX][B * B], const ClassTail& ct,
                                     const A* xs, A (&t)[B]) {
  A y[MX][B];
#pragma unroll
  for (int k = 0; k < MX; k++)
    if (k < w)
#pragma unroll
      for (int c = 0; c < B; c++) y[k][c] = xs[col[k] * B + c];
#pragma unroll
  for (int k = 0; k < MX; k++)
    if (k < w) blk_sub_if<B, A, T>(k < n, t, v[k], y[k]);
  class_tail<B, MX, T, A>(ecol, val, ct, [&](unsigned j, const A (&a)[B * B]) {
    A z[B];
#pragma unroll
    for (int c = 0; c < B; c++) z[c] = xs[j * B + c];
    blk_sub_reg<B, A>(t, a, z);
  });
}

// Forward sweep of group g by the calling wavefront: x = D^-1 (b - L x) for the group's rows. `rhs(ss, t)`
// fills t with the lane's right-hand side (ss.row; also called for lanes without a row, whose row index is then
// a valid one); on return t holds the lane's solution, which x has received. xs: the wavefront's LDS row
// (64*B values). Returns false when the done flag is set.
template <int B, int MO, int MI, typename T, typename A, typename VO, typename Rhs>
__device__ inline bool sweep_fwd_core(const SlabArgs& sa, int g, const T* __restrict__ val, const T* __restrict__ dval,
                                      Rhs&& rhs, VO* x, const int* __restrict__ done, A* xs, SlabSlot& ss,
                                      A (&t)[B]) {
  constexpr int BS = B * B;
  if (!slab_slot(sa, g, done, ss)) return false;
  A D[BS];
  T vo[MO][BS], vi[MI][BS];
  A xo[MO][B];
  unsigned jo[MO], li[MI];
  rhs(ss, t);
#pragma unroll
  for (int e = 0; e < BS; e++) D[e] = (A)dval[((size_t)ss.g * BS + e) * 64 + ss.lane];
  ClassTail to, ti;
  slab_load_class<B, MO, T>(sa.ecol, val, ss.bo, ss.wo, ss.no, jo, vo, to);
  slab_load_class<B, MI, T>(sa.ecol, val, ss.bi, ss.wi, ss.ni, li, vi, ti);
  slab_gather<B, MO, A, VO>(x, ss.wo, ss.no, jo, xo);
#pragma unroll
  for (int k = 0; k < MO; k++)
    if (k < ss.wo) blk_sub_if<B, A, T>(k < ss.no, t, vo[k], xo[k]);
  slab_tail_global<B, MO, T, A, VO>(sa.ecol, val, to, x, t);
  for (int l = 0; l < ss.nlev; l++) {
    if (ss.lv == l) {
      slab_in_group<B, MI, T, A>(sa.ecol, val, ss.wi, ss.ni, li, vi, ti, xs, t);
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
        t[r] = y[r];
        xs[ss.lane * B + r] = y[r];
      }
    }
    if (ss.nlev > 1) wave_lds_sync();
  }
  if (ss.active)
#pragma unroll
    for (int r = 0; r < B; r++) x[(size_t)ss.row * B + r] = (VO)t[r];
  return true;
}

// The forward sweep with the right-hand side b (b and x may alias).
template <int B, int MO, int MI, typename T, typename A, typename VI, typename VO>
__device__ inline void sweep_fwd_wave(const SlabArgs& sa, int g, const T* __restrict__ val, const T* __restrict__ dval,
                                      const VI* b, VO* x, const int* __restrict__ done, A* xs) {
  SlabSlot ss;
  A t[B];
  sweep_fwd_core<B, MO, MI, T, A, VO>(
      sa, g, val, dval,
      [&](const SlabSlot& s, A(&u)[B]) {
#pragma unroll
        for (int r = 0; r < B; r++) u[r] = (A)b[(size_t)s.row * B + r];
      },
      x, done, xs, ss, t);
}

// Backward sweep of group g by the calling wavefront, in place on x: x_i -= U_in x_in (in-group, ascending) then
// U_out x_out. `init(ss, t)` fills t with the lane's starting value (x of its row, or values the caller holds).
// On return t holds the lane's result (for lanes with a row). Returns false when the done flag is set.
template <int B, int MO, int MI, typename T, typename A, typename VX, typename Init>
__device__ inline bool sweep_bwd_core(const SlabArgs& sa, int g, const T* __restrict__ val, VX* x,
                                      const int* __restrict__ done, A* xs, SlabSlot& ss, A (&t)[B], Init&& init) {
  constexpr int BS = B * B;
  if (!slab_slot(sa, g, done, ss)) return false;
  T vo[MO][BS], vi[MI][BS];
  A xo[MO][B];
  unsigned jo[MO], li[MI];
  init(ss, t);
  ClassTail to, ti;
  slab_load_class<B, MO, T>(sa.ecol, val, ss.bo, ss.wo, ss.no, jo, vo, to);
  slab_load_class<B, MI, T>(sa.ecol, val, ss.bi, ss.wi, ss.ni, li, vi, ti);
  slab_gather<B, MO, A, VX>(x, ss.wo, ss.no, jo, xo);
  for (int l = 0; l < ss.nlev; l++) {
    if (ss.lv == l) {
      slab_in_group<B, MI, T, A>(sa.ecol, val, ss.wi, ss.ni, li, vi, ti, xs, t);
#pragma unroll
      for (int k = 0; k < MO; k++)
        if (k < ss.wo) blk_sub_if<B, A, T>(k < ss.no, t, vo[k], xo[k]);
      slab_tail_global<B, MO, T, A, VX>(sa.ecol, val, to, x, t);
#pragma unroll
      for (int r = 0; r < B; r++) xs[ss.lane * B + r] = t[r];
    }
    if (ss.nlev > 1) wave_lds_sync();
  }
  if (ss.active)
#pragma unroll
    for (int r = 0; r < B; r++) x[(size_t)ss.row * B + r] = (VX)t[r];
  return true;
}

template <int B, int MO, int MI, typename T, typename A, typename VX>
__device__ inline void sweep_bwd_wave(const SlabArgs& sa, int g, const T* __restrict__ val, VX* x,
                                      const int* __restrict__ done, A* xs) {
  SlabSlot ss;
  A t[B];
  sweep_bwd_core<B, MO, MI, T, A, VX>(sa, g, val, x, done, xs, ss, t, [&](const SlabSlot& s, A(&u)[B]) {
#pragma unroll
    for (int r = 0; r < B; r++) u[r] = (A)x[(size_t)s.row * B + r];
  });
}

// Wavefronts per SIMD the sweep kernels are compiled for (the register budget follows; T: the arithmetic type):
// scalar systems are latency-bound and gain from every wavefront in flight (measured: the FP64 PCG phases of a
// 2D Poisson system at 5 instead of 6 waves took 10 % longer); block systems keep the compiler's choice.
template <int B, int MO, int MI, typename T>
struct SweepOccupancy {
  static constexpr int kWaves = B != 1 ? 1 : sizeof(T) == 4 ? 8 : MO + MI <= 8 ? 6 : 5;
};

// One phase of one system: groups [gbase, gend), kWPB groups per workgroup.
template <int B, int MO, int MI, typename T, typename A, typename VI, typename VO>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(SweepOccupancy<B, MO, MI, A>::kWaves)))
    k_sweep_fwd(int gbase, int gend, SlabArgs sa, const T* __restrict__ val, const T* __restrict__ dval, const VI* b,
        VO* x, const int* __restrict__ done) {
  __shared__ A xsb[kWPB][64 * B];
  const int w = wave_in_block();
  const int g = gbase + (int)blockIdx.x * kWPB + w;
  if (g >= gend) return;
  sweep_fwd_wave<B, MO, MI, T, A, VI, VO>(sa, g, val, dval, b, x, done, xsb[w]);
}

template <int B, int MO, int MI, typename T, typename A, typename VX>
__global__ void __launch_bounds__(64 * kWPB) __attribute__((amdgpu_waves_per_eu(SweepOccupancy<B, MO, MI, A>::kWaves)))
    k_sweep_bwd(int gbase, int gend, SlabArgs sa, const T* __restrict__ val, VX* x, const int* __restrict__ done) {
  __shared__ A xsb[kWPB][64 * B];
  const int w = wave_in_block();
  const int g = gbase + (int)blockIdx.x * kWPB + w;
  if (g >= gend) return;
  sweep_bwd_wave<B, MO, MI, T, A, VX>(sa, g, val, x, done, xsb[w]);
}

// Launch helper: calls f.template operator()<MO, MI>() for the caps of a phase's widths (wo, wi). FP64
// arithmetic on 3x3 blocks keeps caps of 4 (registers: wider caps drop occupancy to one wave per SIMD).
template <typename F>
inline void with_slab_caps(int wo, int wi, bool narrow, F&& f) {
  const int co = narrow ? 4 : slab_cap(wo), ci = narrow ? 4 : slab_cap(wi);
  if (co == 4 && ci == 4) f.template operator()<4, 4>();
  else if (co == 4) f.template operator()<4, 8>();
  else if (ci == 4) f.template operator()<8, 4>();
  else f.template operator()<8, 8>();
}

}  // namespace dfm
