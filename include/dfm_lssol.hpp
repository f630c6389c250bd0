// dfm_lssol.hpp — C++ host layer over the C-ABI mirroring the reference's solver seam and its Krylov loops.
//
// The interfaces have the names and meaning of LsSol::ILinearSystem_Sol / ILinearSystemPreconditioner_Sol
// (include/delfem/ls/linearsystem_interface_solver.h:39-82 in the reference): vector handles -1 residual,
// -2 update, >= 0 temporaries. The loops take the steps of LsSol::Solve_CG, Solve_PCG, Solve_BiCGSTAB and
// Solve_PBiCGSTAB (src/ls/solver_ls_iter.cpp:31-97, :100-172, :177-286, :291-408), including their return
// values and in/out arguments, so host code written against the reference runs unchanged. Each vector
// operation is one call into the device library (host-driven loop); dfm_solve_pcg / dfm_solve_pbicgstab are
// the device-resident forms of the preconditioned loops.
//
// dfm::matsol holds the matrix-level twins (MatVec::Sol::Solve_*, src/matvec/solver_mat_iter.cpp) over the same seam.
// DeviceLinearSystem adapts a dfm_sys (and its ILU(0) preconditioner) to the seam.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>

#include "dfm_gpu.h"

namespace dfm {
namespace lssol {

class ILinearSystem_Sol {
 public:
  virtual ~ILinearSystem_Sol() = default;
  virtual unsigned int GetTmpVectorArySize() const = 0;
  virtual bool ReSizeTmpVecSolver(unsigned int size_new) = 0;
  virtual double DOT(int iv1, int iv2) = 0;                         // {v1} . {v2}
  virtual bool COPY(int iv1, int iv2) = 0;                          // {v2} := {v1}
  virtual bool SCAL(double alpha, int iv1) = 0;                     // {v1} := alpha {v1}
  virtual bool AXPY(double alpha, int iv1, int iv2) = 0;            // {v2} := alpha {v1} + {v2}
  virtual bool MATVEC(double alpha, int iv1, double beta, int iv2) = 0;  // {v2} := alpha A {v1} + beta {v2}
};

class ILinearSystemPreconditioner_Sol : public ILinearSystem_Sol {
 public:
  virtual bool SolvePrecond(int iv) = 0;  // {v} := M^-1 {v}
};

// Unpreconditioned CG: returns true also when not converged (conv_ratio then keeps its input value).
inline bool Solve_CG(double& conv_ratio, unsigned int& num_iter, ILinearSystem_Sol& ls) {
  const unsigned int max_iter = num_iter;
  const double tol = conv_ratio;
  if (ls.GetTmpVectorArySize() < 2) ls.ReSizeTmpVecSolver(2);
  const int ix = -2, ir = -1, ip = 0, iap = 1;
  ls.SCAL(0.0, ix);
  double sq = ls.DOT(ir, ir);
  if (sq < 1.0e-60) {
    num_iter = 0;
    conv_ratio = std::sqrt(sq);
    return true;
  }
  const double sq_inv0 = 1.0 / sq;
  ls.COPY(ir, ip);
  for (num_iter = 0; num_iter < max_iter; num_iter++) {
    ls.MATVEC(1.0, ip, 0.0, iap);
    const double alpha = sq / ls.DOT(ip, iap);
    ls.AXPY(alpha, ip, ix);
    ls.AXPY(-alpha, iap, ir);
    const double sq_new = ls.DOT(ir, ir);
    if (sq_new * sq_inv0 < tol * tol) {
      conv_ratio = std::sqrt(sq_new * sq_inv0);
      break;
    }
    const double beta = sq_new / sq;
    sq = sq_new;
    ls.SCAL(beta, ip);
    ls.AXPY(1.0, ir, ip);
  }
  return true;
}

// Preconditioned CG: iteration = 0-based index at convergence; false (and the last ratio) when not converged.
inline bool Solve_PCG(double& conv_ratio, unsigned int& iteration, ILinearSystemPreconditioner_Sol& ls) {
  const double tol = conv_ratio;
  const unsigned int mx_iter = iteration;
  if (ls.GetTmpVectorArySize() < 2) ls.ReSizeTmpVecSolver(2);
  const int ix = -2, ir = -1, ip = 0, iz = 1;
  ls.SCAL(0.0, ix);
  double sq_inv0;
  {
    const double sq0 = ls.DOT(ir, ir);
    if (sq0 < 1.0e-30) {
      conv_ratio = 0.0;
      iteration = 0;
      return true;
    }
    sq_inv0 = 1.0 / sq0;
  }
  ls.COPY(ir, iz);
  ls.SolvePrecond(iz);
  ls.COPY(iz, ip);
  double rz = ls.DOT(ir, iz);
  for (unsigned int it = 0; it < mx_iter; it++) {
    ls.MATVEC(1.0, ip, 0.0, iz);
    const double alpha = rz / ls.DOT(ip, iz);
    ls.AXPY(-alpha, iz, ir);
    ls.AXPY(alpha, ip, ix);
    const double sq = ls.DOT(ir, ir);
    if (sq * sq_inv0 < tol * tol) {
      conv_ratio = std::sqrt(sq * sq_inv0);
      iteration = it;
      return true;
    }
    ls.COPY(ir, iz);
    ls.SolvePrecond(iz);
    const double rz_new = ls.DOT(ir, iz);
    const double beta = rz_new / rz;
    rz = rz_new;
    ls.SCAL(beta, ip);
    ls.AXPY(1.0, iz, ip);
  }
  conv_ratio = std::sqrt(ls.DOT(ir, ir) * sq_inv0);
  return false;
}

// Unpreconditioned BiCGSTAB: loop from 1; iteration = max_iter unless converged; always returns true.
inline bool Solve_BiCGSTAB(double& conv_ratio, unsigned int& num_iter, ILinearSystem_Sol& ls) {
  const unsigned int max_iter = num_iter;
  const double tol = conv_ratio;
  if (ls.GetTmpVectorArySize() < 5) ls.ReSizeTmpVecSolver(5);
  const int ix = -2, ir = -1, is = 0, ias = 1, ip = 2, iap = 3, ir2 = 4;
  ls.SCAL(0.0, ix);
  double sq_inv0;
  {
    const double sq0 = ls.DOT(ir, ir);
    if (sq0 < 1.0e-60) {
      conv_ratio = std::sqrt(sq0);
      num_iter = 0;
      return true;
    }
    sq_inv0 = 1.0 / sq0;
  }
  ls.COPY(ir, ir2);
  ls.COPY(ir, ip);
  double r_r2 = ls.DOT(ir, ir2);
  num_iter = max_iter;
  for (unsigned int it = 1; it < max_iter; it++) {
    ls.MATVEC(1.0, ip, 0.0, iap);
    const double alpha = r_r2 / ls.DOT(iap, ir2);
    ls.COPY(ir, is);
    ls.AXPY(-alpha, iap, is);
    ls.MATVEC(1.0, is, 0.0, ias);
    const double den = ls.DOT(ias, ias);
    const double num = ls.DOT(ias, is);
    const double omega = num / den;
    ls.AXPY(alpha, ip, ix);
    ls.AXPY(omega, is, ix);
    ls.COPY(is, ir);
    ls.AXPY(-omega, ias, ir);
    const double sqc = ls.DOT(ir, ir) * sq_inv0;
    if (sqc < tol * tol) {
      conv_ratio = std::sqrt(sqc);
      num_iter = it;
      break;
    }
    const double tmp1 = ls.DOT(ir, ir2);
    const double beta = (tmp1 * alpha) / (r_r2 * omega);
    r_r2 = tmp1;
    ls.SCAL(beta, ip);
    ls.AXPY(1.0, ir, ip);
    ls.AXPY(-beta * omega, iap, ip);
  }
  return true;
}

// Preconditioned BiCGSTAB: loop from 1; iteration = max_iter unless converged; always returns true.
inline bool Solve_PBiCGSTAB(double& conv_ratio, unsigned int& num_iter, ILinearSystemPreconditioner_Sol& ls) {
  const double tol = conv_ratio;
  const unsigned int max_iter = num_iter;
  if (ls.GetTmpVectorArySize() < 7) ls.ReSizeTmpVecSolver(7);
  const int ix = -2, ir = -1, is = 0, ims = 1, iams = 2, ip = 3, imp = 4, iamp = 5, ir2 = 6;
  double sq_inv0;
  {
    const double sq0 = ls.DOT(ir, ir);
    if (sq0 < 1.0e-60) {
      conv_ratio = std::sqrt(sq0);
      num_iter = 0;
      return true;
    }
    sq_inv0 = 1.0 / sq0;
  }
  ls.SCAL(0.0, ix);
  ls.COPY(ir, ir2);
  ls.COPY(ir, ip);
  num_iter = max_iter;
  for (unsigned int it = 1; it < max_iter; it++) {
    ls.COPY(ip, imp);
    ls.SolvePrecond(imp);
    const double r_r2 = ls.DOT(ir, ir2);
    ls.MATVEC(1.0, imp, 0.0, iamp);
    const double alpha = r_r2 / ls.DOT(iamp, ir2);
    ls.COPY(ir, is);
    ls.AXPY(-alpha, iamp, is);
    ls.COPY(is, ims);
    ls.SolvePrecond(ims);
    ls.MATVEC(1.0, ims, 0.0, iams);
    const double den = ls.DOT(iams, iams);
    const double num = ls.DOT(is, iams);
    const double omega = num / den;
    ls.AXPY(alpha, imp, ix);
    ls.AXPY(omega, ims, ix);
    ls.COPY(is, ir);
    ls.AXPY(-omega, iams, ir);
    const double sq = ls.DOT(ir, ir);
    if (sq * sq_inv0 < tol * tol) {
      conv_ratio = std::sqrt(sq * sq_inv0);
      num_iter = it;
      return true;
    }
    const double tmp1 = ls.DOT(ir, ir2);
    const double beta = tmp1 * alpha / (r_r2 * omega);
    ls.SCAL(beta, ip);
    ls.AXPY(1.0, ir, ip);
    ls.AXPY(-beta * omega, iamp, ip);
  }
  return true;
}

}  // namespace lssol

// The matrix-level twins MatVec::Sol::Solve_CG / Solve_PCG / Solve_BiCGSTAB / Solve_PBiCGSTAB
// (src/matvec/solver_mat_iter.cpp:37-423), which take the matrix, the residual vector (in: right-hand side, out:
// final residual) and the solution vector directly; here the matrix is the seam's, the residual is handle -1 and the
// solution handle -2. Their console output is not reproduced.
namespace matsol {

// MatVec::Sol::Solve_CG (:37-113): loop from 1; the convergence test reads the residual norm of the iteration
// before the update just made (so it stops one step late), and the solution takes alpha p only after the test.
// Not converged: iteration = max, conv_ratio keeps its input; returns true either way.
inline bool Solve_CG(double& conv_ratio, unsigned int& iteration, lssol::ILinearSystem_Sol& ls) {
  const double tol = conv_ratio;
  const unsigned int mx_iter = iteration;
  if (ls.GetTmpVectorArySize() < 2) ls.ReSizeTmpVecSolver(2);
  const int iu = -2, ir = -1, ip = 0, iap = 1;
  ls.SCAL(0.0, iu);
  double sq = ls.DOT(ir, ir);
  if (sq < 1.0e-30) {
    conv_ratio = 0.0;
    iteration = 0;
    return true;
  }
  const double sq_inv = 1.0 / sq;
  ls.COPY(ir, ip);
  iteration = mx_iter;
  for (unsigned int it = 1; it < mx_iter; it++) {
    ls.MATVEC(1.0, ip, 0.0, iap);
    const double alpha = sq / ls.DOT(ip, iap);
    ls.AXPY(-alpha, iap, ir);
    const double sq1 = ls.DOT(ir, ir);
    if (sq * sq_inv < tol * tol) {
      conv_ratio = std::sqrt(sq * sq_inv);
      iteration = it;
      return true;
    }
    ls.AXPY(alpha, ip, iu);
    const double beta = sq1 / sq;
    sq = sq1;
    ls.SCAL(beta, ip);
    ls.AXPY(1.0, ir, ip);
  }
  return true;
}

// MatVec::Sol::Solve_PCG (:115-184): the steps of LsSol::Solve_PCG (its (r, z), recomputed at the top of each
// iteration, is bit for bit the one LsSol carries); not converged it returns false with conv_ratio and iteration
// as they came in.
inline bool Solve_PCG(double& conv_ratio, unsigned int& iteration, lssol::ILinearSystemPreconditioner_Sol& ls) {
  const double tol = conv_ratio;
  const unsigned int mx = iteration;
  const bool ok = lssol::Solve_PCG(conv_ratio, iteration, ls);
  if (!ok) {
    conv_ratio = tol;
    iteration = mx;
  }
  return ok;
}

// MatVec::Sol::Solve_BiCGSTAB / Solve_PBiCGSTAB (:191-299, :305-423): the steps of the LsSol loops but with the
// guard ||r0||^2 < 1e-30, which returns conv_ratio 0 (the solution zeroed first).
inline bool zero_rhs_guard(double& conv_ratio, unsigned int& iteration, lssol::ILinearSystem_Sol& ls) {
  if (ls.DOT(-1, -1) >= 1.0e-30) return false;
  ls.SCAL(0.0, -2);
  conv_ratio = 0.0;
  iteration = 0;
  return true;
}
inline bool Solve_BiCGSTAB(double& conv_ratio, unsigned int& iteration, lssol::ILinearSystem_Sol& ls) {
  if (zero_rhs_guard(conv_ratio, iteration, ls)) return true;
  return lssol::Solve_BiCGSTAB(conv_ratio, iteration, ls);
}
inline bool Solve_PBiCGSTAB(double& conv_ratio, unsigned int& iteration, lssol::ILinearSystemPreconditioner_Sol& ls) {
  if (zero_rhs_guard(conv_ratio, iteration, ls)) return true;
  lssol::Solve_PBiCGSTAB(conv_ratio, iteration, ls);
  return true;
}

}  // namespace matsol

namespace lssol {

// The seam over one device system: every operation is one C-ABI call (errors become exceptions, as the
// reference's contract violations assert). The system is not owned.
class DeviceLinearSystem : public ILinearSystemPreconditioner_Sol {
 public:
  explicit DeviceLinearSystem(dfm_sys sys) : sys_(sys) {}
  unsigned int GetTmpVectorArySize() const override {
    int n = 0;
    check(dfm_sys_ntmp(sys_, &n));
    return (unsigned int)n;
  }
  bool ReSizeTmpVecSolver(unsigned int n) override { return dfm_sys_resize_tmp(sys_, (int)n) == DFM_OK; }
  double DOT(int a, int b) override {
    double d = 0.0;
    check(dfm_sys_dot(sys_, a, b, &d));
    return d;
  }
  bool COPY(int a, int b) override { return dfm_sys_copy(sys_, a, b) == DFM_OK; }
  bool SCAL(double s, int a) override { return dfm_sys_scal(sys_, s, a) == DFM_OK; }
  bool AXPY(double s, int a, int b) override { return dfm_sys_axpy(sys_, s, a, b) == DFM_OK; }
  bool MATVEC(double a, int x, double b, int y) override { return dfm_sys_matvec(sys_, a, x, b, y) == DFM_OK; }
  bool SolvePrecond(int iv) override { return dfm_prec_apply(sys_, iv) == DFM_OK; }
  dfm_sys handle() const { return sys_; }

 private:
  static void check(int rc) {
    if (rc < 0) throw std::runtime_error(dfm_last_error());
  }
  dfm_sys sys_;
};

}  // namespace lssol
}  // namespace dfm
