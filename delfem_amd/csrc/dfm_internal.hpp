// Internal declarations of libdfm_gpu.so: device buffers, the system object and its ILU schedule.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "dfm_gpu.h"

namespace dfm {

void set_error(const std::string& msg);

#define DFM_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) {                                                                        \
      ::dfm::set_error(std::string(#call) + ": " + hipGetErrorString(e_) + " at " + __FILE__ + ":" + \
                       std::to_string(__LINE__));                                                  \
      return DFM_ERR_HIP;                                                                          \
    }                                                                                              \
  } while (0)

#define DFM_CHECK(cond, code, msg) \
  do {                             \
    if (!(cond)) {                 \
      ::dfm::set_error(msg);       \
      return code;                 \
    }                              \
  } while (0)

// Device allocation with byte accounting (the serving admission budget reads it).
// Device allocations go through a per-device cache: blocks are sized in classes (<= 12.5 % above the request)
// and a block released while a CacheRelease guard is alive (a system being destroyed after its stream was
// synchronised, so no queued kernel can still touch it) is kept for reuse instead of returned with hipFree, which
// stalls the whole device. Releases outside a guard (a buffer regrown while work may be queued) free at once.
// The cache holds at most a fixed number of bytes per device; beyond that the largest blocks are freed.
struct DevAlloc {
  static hipError_t alloc(void** p, size_t bytes);
  static void release(void* p);
  static hipError_t alloc_host(void** p, size_t bytes);  // pinned host memory, same caching rules
  static void release_host(void* p);
};
struct CacheRelease {
  CacheRelease();
  ~CacheRelease();
};

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return DevAlloc::alloc((void**)&p, count * sizeof(T));
  }
  void release() {
    if (p) DevAlloc::release(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // the handle itself + one per live system; systems are created and destroyed from several threads (the
  // server's workers and its client), so the count is atomic
  std::atomic<int> refs{1};
  bool destroyed = false;  // dfm_ctx_destroy called; freed when the last system goes
  // element-store buffer of the assemblies run on this context's stream (they run one after another there)
  DBuf<double> ebuf;
};

// Host-side schedule of one triangular sweep: phases run one after another (kernel boundaries); a phase is
// a set of independent groups; a group (<= 64 rows, one wavefront) is either a contiguous row range of the
// device numbering with intra-group levels, or a list of independent rows (indirect mode).
struct Sweep {
  std::vector<int> phase_gptr;  // [nphase+1] group index ranges, in execution order
  std::vector<int> phase_nlev;  // [nphase] number of intra-group levels (1 = no intra-group dependency)
  std::vector<int> grp;         // [2*ngroup] (start, end) per group in execution order: row ids
                                // (contiguous mode) or offsets into rowlist (indirect mode)
  std::vector<int> rowlist;     // indirect mode: rows of the groups, grp indexes into it; else empty
  std::vector<uint8_t> lev;     // [n] intra-group level of each row for this sweep's direction
  DBuf<int> d_grp, d_rowlist;
  DBuf<uint8_t> d_lev;
  int nphase() const { return (int)phase_nlev.size(); }

  // Packed layout of the factor for the sweep kernels (one 64-lane slot per row, lane = position in its group).
  // Group k (execution order) owns entries [off, off + its count): first the class of entries referencing rows of
  // earlier phases ("outside"), then the in-group class (the column stored as the source lane); each lane's
  // entries of a class in ascending column order. A class is stored as jagged rows k < 4 (row k: the k-th entry of
  // every lane that has one, in lane order; the row's b*b values element-major: value e of its p-th entry at
  // (row start)*bs + e*(row length) + p) followed by the entries k >= 4 lane after lane (values of an entry
  // together, row-major). Nothing is padded.
  std::vector<int> gmeta;        // [4*ngroup] off, wo | wi << 8 | nlev << 16, first row (contiguous) or -1,
                                 // nrows | (outside entries of the group) << 8
  std::vector<int> srow;         // [64*ngroup] device row of each slot, -1 for an empty lane
  std::vector<unsigned> scnt;    // [64*ngroup] nout | nin << 8 | level << 16
  std::vector<unsigned> ecol;    // [nent] column (outside) or source lane (in-group)
  std::vector<unsigned> esrc;    // [nent] factor block index of the entry
  long long nent = 0;
  std::vector<int> phase_wo, phase_wi;  // [nphase] widest classes of the phase's groups
  DBuf<int> d_gmeta, d_srow;
  DBuf<unsigned> d_scnt, d_ecol, d_esrc;
  DBuf<double> d_eval, d_dval;   // FP64 values (d_dval: forward sweep only, D^-1 per slot, value e of slot
                                 // (g, l) at (g*bs + e)*64 + l)
  DBuf<float> d_eval32, d_dval32;
  // forward sweep only: when the last forward phase and the first backward phase hold the same tiles, the
  // backward group of each group of the last forward phase (in its order); empty otherwise
  std::vector<int> fuse_bgrp;
  DBuf<int> d_fuse_bgrp;
  void release_dev() {
    d_grp.release(); d_rowlist.release(); d_lev.release();
    d_gmeta.release(); d_srow.release(); d_scnt.release(); d_ecol.release(); d_esrc.release();
    d_eval.release(); d_dval.release(); d_eval32.release(); d_dval32.release(); d_fuse_bgrp.release();
  }
};

// Launch shapes the host schedule is laid out for: sweep kernels run kSweepWavesPerWG groups per workgroup, the
// SELL SpMV kSellSlicesPerWG slices per workgroup.
constexpr int kSweepWavesPerWG = 4;
constexpr int kSellSlicesPerWG = 16;

// SELL-64 copy of the off-diagonal blocks for the SpMV of the Krylov loops: slice s = device rows
// [64 s, 64 s + 64), one row per lane; the slice's entries are slab rows [sptr[s], sptr[s+1]) in CRS order per
// lane (padding: column 0, no source), value e of entry (slab row r, lane l) at (r*bs + e)*64 + l.
// visit[i]: the slice the SpMV handles at position i (workgroup i / kSellSlicesPerWG), -1 for none; the slices
// are placed so that each XCD works on one spatial strip of the mesh.
struct Sell {
  std::vector<int> sptr;          // [nslice+1]
  std::vector<unsigned> col, src; // [64*nslab]; src = crs block index or ~0u
  std::vector<int> visit;         // [npos], npos a multiple of 8 * kSellSlicesPerWG
  long long nslab = 0;
  DBuf<int> d_sptr, d_visit;
  DBuf<unsigned> d_col, d_src;
  DBuf<double> d_val;
  DBuf<float> d_val32;
  int nslice() const { return sptr.empty() ? 0 : (int)sptr.size() - 1; }
  void release_dev() {
    d_sptr.release(); d_visit.release(); d_col.release(); d_src.release(); d_val.release(); d_val32.release();
  }
};

// Lane plan of the numeric ILU factorisation (k_ilu_rows). Every factor row is held by consecutive lanes of one
// wavefront, one lane per block in the row's order: its lower entries, the diagonal, its upper entries. The
// forward schedule's groups keep their phases and their intra-group levels; the rows of one level of a group are
// packed into passes of at most 64 lanes. Each lane has the list of its updates A_ij -= A_ik A_kj as (factor
// entry kj, step = position of k among the row's lower entries), in ascending k, the order the reference's IKJ
// loop applies them (src/matvec/matdiafrac_blkcrs.cpp:774-1107); at step q the lane of the row's lower entry q
// hands its final A_ik to the row's lanes. Lane data and lists are stored in pass order.
struct FactorPlan {
  bool ok = false;                // every row fits one wavefront and the lists fit 32-bit offsets
  std::vector<int> gpass;         // [ngroup+1] pass range of each forward group (execution order)
  std::vector<unsigned> pmeta;    // [8*npass] row-start lane mask (lo, hi), diagonal lane mask (lo, hi),
                                  // lanes | steps << 8 | (first pass of a later level) << 16, 3 unused
  std::vector<unsigned> etab;     // [64*npass] per lane: its off-diagonal factor entry, or its row (diagonal)
  std::vector<unsigned> utab;     // [64*npass + 1] per lane: start of its updates (the next lane's start ends them)
  std::vector<unsigned> ukj;      // the factor entries kj of the updates
  std::vector<uint8_t> ustep;     // the step of each update
  DBuf<int> d_gpass;
  DBuf<unsigned> d_pmeta, d_etab, d_utab, d_ukj;
  DBuf<uint8_t> d_ustep;
  void release_dev() {
    d_gpass.release(); d_pmeta.release(); d_etab.release(); d_utab.release(); d_ukj.release(); d_ustep.release();
  }
};

struct Sys {
  Ctx* ctx = nullptr;
  int ndim = 0, n = 0, len = 0, bs = 0, npoel = 0, nelem = 0;
  long long ncrs = 0;

  // mesh (original numbering)
  std::vector<double> coords;
  std::vector<int> lnods;
  std::vector<unsigned> rowptr0, colind0;  // pattern, original numbering
  std::vector<int> elsup_ptr, elsup;       // node -> elements, ascending element index

  // ordering: order[new] = old, inv[old] = new
  std::vector<int> order, inv;
  bool natural = true;
  int ncolors = 0;
  std::vector<int> color_of;  // per device row (MC/BMC), empty for natural / external orderings
  std::vector<int> tile_of;   // per device row (BMC)

  // device-numbering pattern of the permuted matrix
  std::vector<unsigned> rowptr, colind, diaind;
  unsigned max_rowlen = 0;  // the most off-diagonal blocks of a row
  std::vector<long long> perm_crs;  // original crs index -> device crs index
  // element-store lists of the assembly (ElemStore, kernels.hpp): per device entry the (e*NNO + a)*NNO + c of the
  // elements holding both of its nodes in ascending e; per node-element incidence (elsup order) e*NNO + a; empty
  // where the indices do not fit 32 bits (the per-row gather runs instead)
  std::vector<unsigned> eptr, eidx, eloc;
  // ILU factor pattern (device numbering): the matrix pattern for ILU(0), the level-of-fill pattern for ILU(k);
  // lsplit/usplit split each factor row into entries outside / inside its sweep group
  int fill_level = 0;  // 0: ILU(0); k > 0: ILU(k); -1: no limit
  std::vector<unsigned> frowptr, fcolind, fdiaind, lsplit, usplit;
  std::vector<unsigned> a2f;  // matrix entry -> factor entry (empty when the patterns coincide)
  Sweep fwd, bwd;
  Sell sell;
  FactorPlan fplan;

  // device arrays
  DBuf<unsigned> d_rowptr, d_colind, d_diaind, d_lsplit, d_usplit;
  DBuf<unsigned> d_frowptr, d_fcolind, d_fdiaind, d_a2f;
  DBuf<int> d_order, d_inv, d_elsup_ptr, d_elsup, d_lnods, d_bc;
  DBuf<unsigned> d_eptr, d_eidx, d_eloc;
  // the original pattern and, per original entry, its device entry (linear elasticity's row-group assembly)
  DBuf<unsigned> d_rowptr0, d_colind0, d_permq;
  DBuf<int> d_elnod;  // per node-element incidence the element's nodes (4 ints), built at the first such assembly
  DBuf<double> d_coords, d_u, d_v;
  DBuf<double> d_a;   // ACCELERATION field, allocated on first use (Newmark)
  DBuf<double> d_v0;  // previous velocity of the backward-Euler loop, allocated on first use
  DBuf<double> d_crs, d_dia;         // matrix, device numbering, FP64
  DBuf<float> d_crs32, d_dia32;      // FP32 copy (mixed precision)
  DBuf<double> d_fcrs, d_fdia;       // ILU factor FP64 (factor pattern)
  DBuf<double> d_res, d_upd, d_rhs;  // handles -1, -2 and a copy of the assembled right-hand side
  std::vector<DBuf<double>*> d_tmp;  // handles 0..
  DBuf<double> d_work;               // solver work vectors
  DBuf<float> d_work32;
  DBuf<double> d_partial;            // per-block partial sums
  DBuf<double> d_part_sw;            // per-workgroup partial sums of the sweep phases (forward, then backward)
  DBuf<unsigned> d_counter;          // last-block counters
  DBuf<double> d_scal;               // device scalars of the Krylov loops
  DBuf<int> d_iflag;                 // device int scalars (done flag, iteration)
  int* h_pinned = nullptr;           // host-pinned mirror of the done flag + iteration: two slots of 16 ints
  double* h_pinned_d = nullptr;      // and of the device scalars: two slots of kNScal doubles
  hipEvent_t poll_ev[2] = {nullptr, nullptr};  // the PCG loop's status writes, one per slot (made on first use)
  int poll_seq = 0;                            // sequence number of the last status write

  // Dirichlet flags in the original numbering (d_bc holds them in the device numbering)
  std::vector<int> bc0;

  // saved stiffness (CLinearSystem_Save): the external force and the boundary coupling K_b as scalar terms per
  // device dof, in the order the reference's boundary matrix visits them
  DBuf<double> d_force;     // f, device numbering
  DBuf<unsigned> d_bptr;    // [n*len + 1] term ranges per device dof
  DBuf<unsigned> d_bcol;    // u index (original numbering) of each term
  DBuf<long long> d_bsrc;   // matrix value of each term: diagonal block entry (< n*bs) or n*bs + crs entry
  DBuf<double> d_bval;      // K_b values, gathered before the Dirichlet zeroing
  long long nbterm = 0;
  int saved = 0;  // SAVE_NONE, SAVE_PLAIN (CLinearSystem_Save), SAVE_DIA_NEWMARK / SAVE_DIA_NEWMARKBETA
  double save_dt = 0.0, save_gamma = 0.0, save_beta = 0.0;  // the lumped-mass saved systems' Newmark parameters

  // resident assembly inputs
  DBuf<double> d_par;
  int npar = 0, pstride = 0;
  DBuf<int> d_lptr;
  DBuf<double> d_loads;
  bool has_loads = false;
  double load_scale = 1.0;

  // captured chunks of Krylov iterations (hipGraph), per loop kind: 0 PCG FP64, 1 PCG mixed, 2 PBiCGSTAB,
  // 3 PCG FP64 with the FP32 factor
  hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<const void*> gkey[4];  // every device pointer the graph was captured with
  int gchunk = 16;
  unsigned long long struct_gen = 0;  // bumped whenever the pattern, ordering, fill level or schedule is rebuilt

  // the kernel arguments of the last assembly (timed again by dfm_sys_bench_kernel)
  struct AsmArgs* last_asm = nullptr;
  bool assembled = false, factored = false, factored32 = false;
  double last_ms = 0.0, last_ms_iter = 0.0;
  unsigned last_inner = 0;
  int last_nrel = 0;  // reliable updates in the last mixed solve
  int last_fallback = 0;  // the last mixed solve stalled and was repeated in FP64
  struct Batch* batch = nullptr;  // the batch the system is being solved in, if any
  int prepared = -1;              // precision sys_solve_setup last prepared a solve for (-1: none pending)

  ~Sys();
};

// Number of kind-specific scalars (glob) an element kind reads: diffusion (dt, gamma), the Newmark kinds
// (dt, gamma, beta, is_initial; the 3D linear one has no is_initial), backward Euler (dt, is_initial), none for
// the static kinds. Assembly and the server's request copy both go by it.
inline int kind_glob_count(int kind) {
  switch (kind) {
    case DFM_DIFFUSION2D: return 2;
    case DFM_SOLID2D_NEWMARK: case DFM_STVK3D_NEWMARK: case DFM_STVK2D_NEWMARK: return 4;
    case DFM_SOLID3D_NEWMARK: return 3;
    case DFM_SOLID2D_BEULER: case DFM_STVK2D_BEULER: return 2;
    default: return 0;
  }
}

// The precision a solve of this system runs in: DFM_MIXED_AUTO resolved by block size (capi.cpp).
int resolve_precision(const Sys& s, int precision);
// Buffers, FP32 copies and the right-hand side copy for a solve of the given precision (capi.cpp).
int sys_solve_setup(Sys& s, int precision);
// ||b - A x|| / ||b|| in FP64 (b the assembled right-hand side, x the update handle), synchronous on st (capi.cpp).
int sys_true_residual(Sys& s, hipStream_t st, double* rel);
// A vector handle (caller's numbering) read on st, the stream whose work last wrote it (capi.cpp).
int sys_download(Sys& s, int iv, double* out, hipStream_t st);
// The device vector of a handle (-1 residual, -2 update, >= 0 temporaries), nullptr for a bad handle (capi.cpp).
double* sys_vec_ptr(Sys& s, int iv);
// FinalizeMarge of host-installed values: Dirichlet treatment by the flags of dfm_sys_set_bc (capi.cpp).
int sys_apply_bc(Sys& s);
// The ILU factor in two halves (capi.cpp): the matrix values copied (ILU(0)) or scattered (ILU(k)) into the factor
// arrays, then the numeric factorisation, in place when prefilled, with its sweep slabs; DFM_ERR_SINGULAR as
// dfm_prec_factor. Between the two a caller may update the factor arrays (multi-segment systems).
int sys_factor_prefill(Sys& s);
int sys_factor_run(Sys& s, bool prefilled);
// z := M^-1 z restricted to the forward (dirs bit 0) and/or backward (bit 1) substitution, FP64 (solver.hip).
hipError_t precond_apply64_dirs(const Sys& s, double* z, int dirs, hipStream_t st);
// Take a system out of the batch it is solved in (batch.hip).
void batch_detach(struct Batch* b, Sys* s);
// Host time of a batch: seconds in table rebuilds, rebuilds, seconds in polls, true-residual restarts (batch.hip).
void batch_stats(const struct Batch* b, double* out);
struct Batch* batch_of(dfm_batch b);

// host symbolic routines (host_symbolic.cpp)
void build_pattern(int nnode, int nelem, int npoel, const int* lnods, std::vector<unsigned>& rowptr,
                   std::vector<unsigned>& colind, std::vector<int>& elsup_ptr, std::vector<int>& elsup);
void build_device_pattern(Sys& s);  // from order/inv: rowptr, colind, diaind, perm_crs, element-store lists
void build_elem_lists(Sys& s);      // the element-store lists for the current device numbering
void build_schedule(Sys& s);        // SELL copy of the pattern, then build_sweeps
void build_sweeps(Sys& s);          // factor pattern, fwd/bwd sweeps, lsplit/usplit, slabs, factor plan
void make_multicolor(const Sys& s, std::vector<int>& order, std::vector<int>& color_of_new, int& ncolors);
// COrdering_Blk's orderings of a block pattern (DFM_ORDER_RCM / _RCM2 / _AMD), new -> old as the reference stores
// them; false where the reference stops on an assertion (host_symbolic.cpp)
bool make_reference_ordering(int kind, int n, const unsigned* rowptr, const unsigned* colind, std::vector<int>& order);
void make_blockcolor(const Sys& s, int tile, std::vector<int>& order, std::vector<int>& color_of_new,
                     std::vector<int>& tile_of_new, int& ncolors);

// mesh generation (host_mesh.cpp)
uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t idx);
double rng_u01(uint64_t seed, uint64_t stream, uint64_t idx);
double rng_normal(uint64_t seed, uint64_t stream, uint64_t idx);

}  // namespace dfm

struct dfm_ctx_s {
  dfm::Ctx c;
};
struct dfm_sys_s {
  dfm::Sys s;
  dfm_ctx_s* owner = nullptr;
};

// Drop one reference of a context; the last one frees it (capi.cpp).
void ctx_release(dfm_ctx ctx);
