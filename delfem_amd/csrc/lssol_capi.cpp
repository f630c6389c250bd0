// C-ABI entry points of the host-driven Krylov loops (include/dfm_lssol.hpp): the reference's unpreconditioned
// Solve_CG / Solve_BiCGSTAB, and Solve_PCG / Solve_PBiCGSTAB driven from the host over the seam (one device
// call per vector operation) beside their device-resident forms, and the matrix-level MatVec::Sol twins.
#include <exception>
#include <string>

#include "dfm_gpu.h"
#include "dfm_internal.hpp"
#include "dfm_lssol.hpp"

namespace {

template <typename F>
int run(dfm_sys sys, double* conv_ratio, unsigned* iteration, bool need_factor, const char* who, F&& f) {
  if (!sys || !conv_ratio || !iteration) {
    dfm::set_error(std::string(who) + ": null argument");
    return DFM_ERR_ARG;
  }
  if (!sys->s.assembled || (need_factor && !sys->s.factored)) {
    dfm::set_error(std::string(who) + (need_factor ? ": assemble and factor first" : ": assemble first"));
    return DFM_ERR_STATE;
  }
  if (sys->s.batch) {
    dfm::set_error(std::string(who) + ": the system is being solved in a batch");
    return DFM_ERR_STATE;
  }
  try {
    dfm::lssol::DeviceLinearSystem ls(sys);
    return f(ls) ? 1 : 0;
  } catch (const std::exception& e) {
    dfm::set_error(std::string(who) + ": " + e.what());
    return DFM_ERR_HIP;
  }
}

}  // namespace

extern "C" {

int dfm_solve_cg(dfm_sys sys, double* conv_ratio, unsigned* iteration) {
  return run(sys, conv_ratio, iteration, false, "dfm_solve_cg",
             [&](dfm::lssol::DeviceLinearSystem& ls) { return dfm::lssol::Solve_CG(*conv_ratio, *iteration, ls); });
}

int dfm_solve_bicgstab(dfm_sys sys, double* conv_ratio, unsigned* iteration) {
  return run(sys, conv_ratio, iteration, false, "dfm_solve_bicgstab", [&](dfm::lssol::DeviceLinearSystem& ls) {
    return dfm::lssol::Solve_BiCGSTAB(*conv_ratio, *iteration, ls);
  });
}

int dfm_solve_pcg_host(dfm_sys sys, double* conv_ratio, unsigned* iteration) {
  return run(sys, conv_ratio, iteration, true, "dfm_solve_pcg_host",
             [&](dfm::lssol::DeviceLinearSystem& ls) { return dfm::lssol::Solve_PCG(*conv_ratio, *iteration, ls); });
}

int dfm_solve_pbicgstab_host(dfm_sys sys, double* conv_ratio, unsigned* iteration) {
  return run(sys, conv_ratio, iteration, true, "dfm_solve_pbicgstab_host", [&](dfm::lssol::DeviceLinearSystem& ls) {
    return dfm::lssol::Solve_PBiCGSTAB(*conv_ratio, *iteration, ls);
  });
}

int dfm_solve_mat_cg(dfm_sys sys, double* conv_ratio, unsigned* iteration) {
  return run(sys, conv_ratio, iteration, false, "dfm_solve_mat_cg",
             [&](dfm::lssol::DeviceLinearSystem& ls) { return dfm::matsol::Solve_CG(*conv_ratio, *iteration, ls); });
}

int dfm_solve_mat_pcg(dfm_sys sys, double* conv_ratio, unsigned* iteration) {
  return run(sys, conv_ratio, iteration, true, "dfm_solve_mat_pcg",
             [&](dfm::lssol::DeviceLinearSystem& ls) { return dfm::matsol::Solve_PCG(*conv_ratio, *iteration, ls); });
}

int dfm_solve_mat_bicgstab(dfm_sys sys, double* conv_ratio, unsigned* iteration) {
  return run(sys, conv_ratio, iteration, false, "dfm_solve_mat_bicgstab", [&](dfm::lssol::DeviceLinearSystem& ls) {
    return dfm::matsol::Solve_BiCGSTAB(*conv_ratio, *iteration, ls);
  });
}

int dfm_solve_mat_pbicgstab(dfm_sys sys, double* conv_ratio, unsigned* iteration) {
  return run(sys, conv_ratio, iteration, true, "dfm_solve_mat_pbicgstab", [&](dfm::lssol::DeviceLinearSystem& ls) {
    return dfm::matsol::Solve_PBiCGSTAB(*conv_ratio, *iteration, ls);
  });
}

}  // extern "C"
