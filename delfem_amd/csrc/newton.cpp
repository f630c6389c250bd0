// Load-stepped Newton for geometrically nonlinear solids, many instances advanced together. Host C++ over the
// library's own C-ABI, following CEqn_Solid3D_Linear::Solve, geometric-nonlinear branch
// (src/femeqn/eqnsys_solid.cpp:174-209):
//   per Newton iteration: MakeLinearSystem (tangent + residual, Dirichlet, ILU factor) -> ||r||;
//   iteration 0: stop if ||r|| < 1e-20, else ||r_0|| = ||r||; stop when ||r|| < tol * ||r_0||;
//   else Solve_PCG and u += du (UpdateValueOfField), at most max_newton iterations.
// Load stepping (the traction ramped over nsteps) wraps the loop. Every instance runs its own Newton iteration
// count; the linear solves of one Newton iteration of all instances still running form one batch.
#include <cmath>
#include <string>
#include <vector>

#include "dfm_gpu.h"

namespace dfm {
void set_error(const std::string& msg);
}

extern "C" int dfm_newton_batch(dfm_ctx ctx, int nsys, const dfm_sys* sys, int kind, const double* glob, int nsteps,
                                const double* load_scale, int max_newton, double newton_tol, int precision,
                                double lin_tol, unsigned lin_maxit, int* newton_iters, unsigned* lin_iters,
                                int* lin_unconverged) {
  if (!ctx || nsys <= 0 || !sys || !load_scale || nsteps <= 0 || max_newton <= 0 || !(newton_tol > 0)) {
    dfm::set_error("dfm_newton_batch: bad argument");
    return DFM_ERR_ARG;
  }
  dfm_batch b = nullptr;
  int rc = dfm_batch_create(ctx, precision, nsys, &b);
  if (rc != DFM_OK) return rc;
  std::vector<double> ini(nsys, 0.0);
  std::vector<char> active(nsys, 0);
  std::vector<int> in_batch;
  for (int i = 0; i < nsys; i++) {
    if (lin_iters) lin_iters[i] = 0;
    if (lin_unconverged) lin_unconverged[i] = 0;
  }
  auto fail = [&](int code) {
    dfm_batch_destroy(b);
    return code;
  };
  for (int step = 1; step <= nsteps; step++) {
    for (int i = 0; i < nsys; i++) {
      rc = dfm_sys_set_load_scale(sys[i], load_scale[i] * (double)step / (double)nsteps);
      if (rc != DFM_OK) return fail(rc);
      active[i] = 1;
      if (newton_iters) newton_iters[(size_t)i * nsteps + step - 1] = max_newton;
    }
    for (int it = 0; it < max_newton; it++) {
      in_batch.clear();
      for (int i = 0; i < nsys; i++) {
        if (!active[i]) continue;
        double norm = 0.0;
        rc = dfm_sys_assemble_resident(sys[i], kind, glob, &norm);
        if (rc != DFM_OK) return fail(rc);
        bool stop = false;
        if (it == 0) {
          if (norm < 1.0e-20) stop = true;
          ini[i] = norm;
        }
        if (!stop && norm < ini[i] * newton_tol) stop = true;
        if (stop) {
          active[i] = 0;
          if (newton_iters) newton_iters[(size_t)i * nsteps + step - 1] = it;
          continue;
        }
        rc = dfm_prec_factor(sys[i]);
        if (rc != DFM_OK) return fail(rc);
        rc = dfm_batch_add(b, sys[i], lin_tol, lin_maxit);
        if (rc != DFM_OK) return fail(rc);
        in_batch.push_back(i);
      }
      if (in_batch.empty()) break;
      int running = (int)in_batch.size();
      while (running > 0) {
        rc = dfm_batch_step(b, 16, &running);
        if (rc != DFM_OK) return fail(rc);
      }
      for (int i : in_batch) {
        int state = 0;
        unsigned iters = 0;
        rc = dfm_batch_status(b, sys[i], &state, nullptr, &iters, nullptr);
        if (rc != DFM_OK) return fail(rc);
        // Solve_PCG's result is used whether or not it converged, as the reference does
        if (lin_iters) lin_iters[i] += state == 1 ? iters + 1 : iters;
        if (lin_unconverged && state != 1) lin_unconverged[i]++;
        rc = dfm_batch_remove(b, sys[i]);
        if (rc != DFM_OK) return fail(rc);
        rc = dfm_sys_update_field(sys[i]);
        if (rc != DFM_OK) return fail(rc);
      }
    }
  }
  return fail(DFM_OK);
}
