// Launchers of the device kernels (k_assemble.hip, k_linalg.hip). All run on the given stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dfm {

// ---- assembly (k_assemble.hip) ----
struct AsmArgs {
  int kind, ndim, npoel, len, n;
  const double* coords;  // original node numbering
  const int* lnods;
  const int* elsup_ptr;  // node -> elements (ascending element index), original numbering
  const int* elsup;
  const int* order;      // device row -> original node
  const int* inv;        // original node -> device row
  const double* params;
  int npar, pstride;
  double glob0, glob1, glob2, glob3;  // kind-specific scalars (diffusion: dt, gamma; Newmark: dt, gamma, beta, initial)
  const double* u;  // field values (original numbering), may be null
  const double* v;
  const double* acc;  // ACCELERATION field (Newmark) or the previous velocity (backward Euler), original numbering
  const int* load_ptr;   // node -> surface-load entries (original numbering), may be null
  const double* loads;   // [nentries*len]
  double load_scale;     // multiplies the surface loads (load stepping); 1 = as given
  const unsigned* rowptr;  // device numbering pattern
  const unsigned* rowptr0;  // original numbering pattern, and per original entry its device entry (may be null)
  const unsigned* colind0;
  const unsigned* permq;
  const int* elnod;  // per node-element incidence (elsup order) the element's nodes, 4 ints (unused -1); may be null
  const unsigned* colind;
  double* crs;
  double* dia;
  double* res;
  bool save;  // the *_SaveStiffMat element loops: external force into res, no internal force
};
// The element store of an assembly: ebuf holds nelem * NNO rows of NNO*B*B + B doubles; per off-diagonal entry q the
// list eidx[eptr[q] .. eptr[q+1]) of (e*NNO + a)*NNO + c in ascending element order; per node-element incidence
// (elsup order) eloc = e*NNO + a. Null: the per-row gather, into a matrix and residual zeroed beforehand.
// lpr > 0 (the vector kinds): the row-group kernel, lpr lanes per node, runs instead and needs neither lists nor
// buffer.
struct ElemStore {
  int nelem;
  int lpr;
  long long ncrs;
  const unsigned* eptr;
  const unsigned* eidx;
  const unsigned* eloc;
  double* ebuf;
};
hipError_t launch_assemble(const AsmArgs& a, const ElemStore* es, hipStream_t st);
hipError_t launch_set_bc(int n, int len, const unsigned* rowptr, const unsigned* colind, double* crs, double* dia,
                         double* res, const int* bc, hipStream_t st);
hipError_t launch_update_field(int n, int len, const int* order, const double* upd, double* u, hipStream_t st);
hipError_t launch_update_field_newmark1(int n, int len, const int* order, const double* upd, double gdt, double dt,
                                        int initial, double* u, double* v, hipStream_t st);
hipError_t launch_update_field_beuler(int n, int len, const int* order, const double* upd, double dt, int initial,
                                      double* u, double* v, hipStream_t st);
hipError_t launch_update_field_newmark(int n, int len, const int* order, const double* upd, double gamma, double beta,
                                       double dt, int initial, double* u, double* v, double* a, hipStream_t st);
// saved stiffness: K_b values from the matrix before the Dirichlet zeroing (src < ndia: diagonal blocks, else
// off-diagonal blocks at src - ndia); MakeResidual; UpdateValueOfField of CLinearSystem_Save
hipError_t launch_gather_terms(long long nterm, const long long* src, long long ndia, const double* dia,
                               const double* crs, double* val, hipStream_t st);
// saved-system kinds of MakeResidual / UpdateValueOfField (Sys::saved)
enum { SAVE_NONE = 0, SAVE_PLAIN = 1, SAVE_DIA_NEWMARK = 2, SAVE_DIA_NEWMARKBETA = 3 };
// internal element kinds of the lumped-mass saved systems' matrix (dfm_sys_assemble_save_dia; not public kinds)
constexpr int KIND_DIFFUSION2D_SAVEDIA = 100;
constexpr int KIND_SOLID2D_SAVEDIA = 101;
hipError_t launch_make_residual(int n, int len, int mode, double c0, double c1, const int* order,
                                const unsigned* bptr, const unsigned* bcol, const double* bval, const double* force,
                                const double* u, const double* v, const double* a, const int* bc, double* res,
                                double* upd, hipStream_t st);
hipError_t launch_update_field_savedia(int n, int len, int mode, double dt, double gamma, double beta,
                                       const int* order, const int* bc, double* upd, double* u, double* v, double* a,
                                       hipStream_t st);
hipError_t launch_update_field_save(int n, int len, const int* order, const int* bc, double* upd, double* u,
                                    hipStream_t st);

// ---- linear algebra (k_linalg.hip) ----
hipError_t launch_fill(double* x, double v, long long m, hipStream_t st);
hipError_t launch_fill_f(float* x, float v, long long m, hipStream_t st);
hipError_t launch_copy(const double* x, double* y, long long m, hipStream_t st);
hipError_t launch_scatter_blocks(long long nblk, int bs, const unsigned* map, const double* src, double* dst,
                                 hipStream_t st);
hipError_t launch_scal(double* x, double a, long long m, hipStream_t st);
hipError_t launch_axpy(double a, const double* x, double* y, long long m, hipStream_t st);
hipError_t launch_to_f32(const double* x, float* y, long long m, hipStream_t st);

// deterministic reductions: result written to out[0] (device)
hipError_t launch_dot(const double* x, const double* y, long long m, double* partial, unsigned* counter,
                      double* out, hipStream_t st);

// y = alpha*A*x + beta*y, device numbering, FP64 values
int spmv_tpr(long long ncrs, int n);
hipError_t launch_spmv(int n, int len, int tpr, const unsigned* rowptr, const unsigned* colind, const double* crs,
                       const double* dia, double alpha, const double* x, double beta, double* y, hipStream_t st);

// ||b - A x||^2 -> out[0], ||b||^2 -> out[1] (device numbering, FP64) without a residual vector: partial sums of a
// fixed grid of at most kMaxPartial blocks in part[0 .. 2*kMaxPartial), then one block sums them in block order.
hipError_t launch_true_residual(int n, int len, int tpr, const unsigned* rowptr, const unsigned* colind,
                                const double* crs, const double* dia, const double* x, const double* b, double* part,
                                double* out, hipStream_t st);

// ILU(0)
struct SweepDev {
  int nphase;
  const int* phase_gptr;  // host array
  const int* phase_nlev;  // host array
  const int* grp;         // device (start,end) pairs
  const int* rowlist;     // device or null
  const uint8_t* lev;     // device
};
hipError_t launch_ilu_factor(int n, int len, const unsigned* rowptr, const unsigned* colind,
                             const unsigned* diaind, double* fcrs, double* fdia, const SweepDev& fwd,
                             int* nsing, hipStream_t st);
// The lane-parallel factorisation over the forward phases (FactorPlan; device arrays).
struct FactorPlanDev {
  const int* gpass;
  const unsigned* pmeta;
  const unsigned* etab;
  const unsigned* utab;
  const unsigned* ukj;
  const uint8_t* ustep;
};
// a_crs/a_dia and fcrs/fdia carry two doubles of padding (3x3 blocks are read with aligned 16-byte loads).
hipError_t launch_ilu_rows(int len, const FactorPlanDev& fp, const int* phase_gptr, int nphase, const double* a_crs,
                           const double* a_dia, double* fcrs, double* fdia, int* nsing, hipStream_t st);
// Values of a sweep's packed entries (jagged rows, then per-entry) from the FP64 factor through src (the factor
// block of each entry). T = double or float.
hipError_t launch_sweep_fill(int len, int ngroup, const int* gmeta, const unsigned* scnt, const unsigned* src,
                             const double* f, double* val, hipStream_t st);
hipError_t launch_sweep_fill(int len, int ngroup, const int* gmeta, const unsigned* scnt, const unsigned* src,
                             const double* f, float* val, hipStream_t st);
// 64-lane slab values (SELL copy, per-slot diagonal): val[(r*bs + e)*64 + l] = f[src[r*64 + l]*bs + e], 0 where
// src = ~0u. nent = 64 * slab rows. T = double or float.
hipError_t launch_slab_fill(int len, long long nent, const unsigned* src, const double* f, double* val, hipStream_t st);
hipError_t launch_slab_fill(int len, long long nent, const unsigned* src, const double* f, float* val, hipStream_t st);
// Krylov device loops (solver.hip) use the kernels below through solver-specific launchers.
}  // namespace dfm
