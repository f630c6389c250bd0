// Host entry points of the device Krylov loops (solver.hip).
#pragma once

#include "dfm_internal.hpp"

namespace dfm {

// scalar slots (doubles)
enum {
  S_RZ = 0,     // rho = (r, z), two parity slots 0/1
  S_INV0 = 2,   // 1 / ||r0||^2
  S_TOL2 = 3,   // tolerance^2
  S_CONV = 4,   // achieved ratio
  S_RMAX = 5,   // reliable updates: max ||r||^2 since the last update, two parity slots 5/6
  S_DELTA2 = 7, // reliable updates: trigger ratio^2
  S_RR = 8,     // last ||r||^2 seen by the convergence test
  S_ALPHA = 9,  // alpha of the current iteration
  S_BETA = 10,  // beta of the current iteration
  S_TBEST = 11, // mixed precision: smallest true ||r||^2 that counted as progress
  S_LASTUPD = 12, // mixed precision: iteration of the last reliable update
  S_RRTRIG = 13,  // mixed precision: recursive ||r||^2 that scheduled the pending reliable update
  S_SLOW = 14,    // mixed precision: slow reliable-update cycles in a row
  S_TPREV = 16,   // mixed precision: true ||r||^2 at the previous reliable update (||r0||^2 before the first)
  S_TMP = 22,     // two slots of scratch for host-side dot products outside solves
  kNScal = 24     // doubles per system
};
// int flags
// F_DONE: 0 running, 1 converged, 2 max_iter, 3 mixed precision stalled (the FP32 matrix cannot reach the
// tolerance: the solve is repeated in FP64); F_STALL: reliable updates in a row without progress
enum { F_DONE = 0, F_IT = 1, F_ITOUT = 2, F_MAXIT = 3, F_STALL = 4, F_NREL = 5, F_PCUR = 6, F_PNEXT = 7 };
// partial slots, kMaxPartial doubles per value
enum { P_A = 0, P_B = 2, P_C = 4, P_D = 6, P_NSLOT = 8 };

// Buffers of one PCG run: FP64 loop (x = the FP64 solution), or the FP32 loop with reliable updates
// (x = the FP32 partial solution, x64/b64 the FP64 solution and right-hand side).
struct PcgBufs {
  const void* crs;
  const void* dia;
  void* r;
  void* z;
  void* p;
  void* q;
  void* x;
  double* x64;
  const double* b64;
  bool f32ilu = false;  // FP64 loop with the FP32 factor copy as preconditioner
};

// Start of a PCG solve on the system's own buffers with scalars/flags at sc/fl (kNScal doubles, 8 ints): x = 0,
// guard on ||r0||, z = M^-1 r, p = z, rho. pcg_prepare first (mixed: r = (float) rhs). pcg_finish after the
// loop stopped with flag `done` (1 converged, 2 max_iter): final fold (mixed) and the residual handle.
// pcg_start_warm: the FP64 loop continued from the FP64 solution a stopped mixed-precision loop left (done 3):
// the residual handle still holds b; r = b - A x, and the tolerance stays relative to ||b||.
PcgBufs pcg_bufs(Sys& s, bool mixed);
int pcg_prepare(Sys& s, const PcgBufs& v, bool mixed, hipStream_t st);
int pcg_start(Sys& s, const PcgBufs& v, bool mixed, double tol, unsigned maxit, double* sc, int* fl, hipStream_t st);
int pcg_start_warm(Sys& s, const PcgBufs& v, double tol, unsigned maxit, double* sc, int* fl, hipStream_t st);
int pcg_finish(Sys& s, const PcgBufs& v, bool mixed, int done, hipStream_t st);

// Restarts of an FP64 loop with the FP32 factor after its true residual was checked: at most kMaxRestarts, each
// with half the tolerance as recursive target, and only while every restart lowers the true residual below
// kRestartGain times the previous one.
constexpr int kMaxRestarts = 8;
constexpr double kRestartGain = 0.8;
int solve_pcg(Sys& s, int precision, double* conv_ratio, unsigned* iteration);
// f32ilu: the preconditioner applied with the FP32 copy of the factor (FP64 vectors and matrix)
int solve_pbicgstab(Sys& s, bool f32ilu, double* conv_ratio, unsigned* iteration);
// z = M^-1 src with the FP64 factor (src may equal z); kernels skip when *done != 0 (done may be null)
int bench_kernel(Sys& s, int which, int reps, double* avg_ms, double* bytes);
hipError_t precond_apply64(const Sys& s, const double* src, double* z, const int* done, hipStream_t st);

}  // namespace dfm
