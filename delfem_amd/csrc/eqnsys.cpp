// Equation systems: CEqnSystem_Solid2D, CEqn_Solid3D_Linear, CEqnSystem_Scalar2D and CEqn_Scalar3D
// (include/delfem/eqnsys_solid.h, eqnsys_scalar.h, eqnsys.h) over the library's device linear systems. Each keeps
// the reference's state machine: a linear system and a preconditioner that exist or not (pLS, pPrec), the
// "values cleared" flags its setters raise, and the Solve flows of src/femeqn/eqnsys_solid.cpp:175-247 (3D solid),
// :671-761 (2D solid), src/femeqn/eqnsys_scalar.cpp:294-359 (2D scalar) and :554-584 (3D scalar): element loop,
// ILU(1) for solids and ILU(0) for scalars, PCG to 1e-6 in at most 1000 iterations, the field update of the mode,
// and a Newton loop (10 steps in 2D, 40 in 3D, relative residual 1e-6) under geometrical nonlinearity.
//
// The reference's field world is replaced by the mesh: element groups play its element arrays (one CEqn_* each,
// elements taken group by group as the reference's arrays hold them), boundary element lists its fixed element
// arrays. The field values live on the host between linear systems and on the device inside one.
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "dfm_gpu.h"
#include "dfm_internal.hpp"

namespace {

constexpr double kConvRatio = 1.0e-6;  // every Solve of the reference's equation systems
constexpr unsigned kMaxIter = 1000;

struct Eqn {
  int type = 0, ndim = 2, len = 1, nnode = 0, ngroup = 1;
  dfm_ctx ctx = nullptr;
  std::vector<double> coords;
  std::vector<int> lnods, group;  // grouped element order
  std::vector<int> flags;         // fixed dofs
  std::vector<double> u, v, a;    // field values (host copies while no linear system holds them)
  // per group: solids (lambda, myu, rho, gx, gy, gz), scalars (alpha, capacity, source)
  std::vector<std::vector<double>> gp;
  bool stationary = true, save = false, nonlin = false;
  double dt = 0.1, gamma = 0.6, beta = 0.3025;  // CEqnSystem's constructor
  dfm_sys sys = nullptr;                        // pLS
  int sys_reps = 1;                             // element passes the system was made with
  bool has_prec = false;                        // pPrec
  bool cleared_ls = true, cleared_prec = true;
  std::vector<std::pair<unsigned, double>> itr;  // GetAry_ItrNormRes

  int lev() const { return (type == DFM_EQN_SOLID2D || type == DFM_EQN_SOLID3D) ? 1 : 0; }
  bool solid() const { return type == DFM_EQN_SOLID2D || type == DFM_EQN_SOLID3D; }
  long long mdof() const { return (long long)nnode * len; }
};

int fail(int code, const std::string& msg) {
  dfm::set_error(msg);
  return code;
}

#define EQN_TRY(call)              \
  do {                             \
    const int rc_ = (call);        \
    if (rc_ != DFM_OK) return rc_; \
  } while (0)

// the field values back to the host and the system gone (ClearLinearSystem; the reference's fields outlive it)
int drop_ls(Eqn& e) {
  if (!e.sys) return DFM_OK;
  EQN_TRY(dfm_sys_get_fields(e.sys, e.u.data(), e.v.data(), e.a.data()));
  dfm_sys_destroy(e.sys);
  e.sys = nullptr;
  e.has_prec = false;
  return DFM_OK;
}

// CEqn_Solid2D::AddLinSys_Save (src/femeqn/eqnsys_solid.cpp:425-432) names no element array, so in the saved mode
// each group's equation assembles every element of the field with its own material: one pass per group
int save_passes(const Eqn& e) { return (e.type == DFM_EQN_SOLID2D && e.save) ? e.ngroup : 1; }

// InitializeLinearSystem: the pattern of the field's elements, the fixed dofs, the fill level of the preconditioner
int init_ls(Eqn& e) {
  EQN_TRY(drop_ls(e));
  const int reps = save_passes(e), npoel = e.ndim + 1;
  const int nelem = (int)(e.lnods.size() / npoel);
  std::vector<int> ln;
  ln.reserve(e.lnods.size() * reps);
  for (int r = 0; r < reps; r++) ln.insert(ln.end(), e.lnods.begin(), e.lnods.end());
  EQN_TRY(dfm_sys_create_mesh(e.ctx, e.ndim, e.nnode, e.coords.data(), nelem * reps, ln.data(), e.len, &e.sys));
  e.sys_reps = reps;
  EQN_TRY(dfm_sys_set_bc(e.sys, e.flags.data()));
  EQN_TRY(dfm_sys_set_field(e.sys, e.u.data(), e.v.data()));
  if (e.solid() && !e.stationary) EQN_TRY(dfm_sys_set_acceleration(e.sys, e.a.data()));
  EQN_TRY(dfm_prec_set_fill_level(e.sys, e.lev()));
  e.has_prec = false;
  return DFM_OK;
}

// per-element parameters of the element loop in use (grouped order; repeated per pass in the saved 2D mode)
std::vector<double> elem_params(const Eqn& e, int* npar) {
  const size_t nelem = e.group.size();
  std::vector<double> p;
  if (e.solid()) {
    *npar = e.ndim == 2 ? 5 : 6;
    const int reps = e.sys_reps;
    p.reserve(nelem * reps * *npar);
    for (int r = 0; r < reps; r++)
      for (size_t k = 0; k < nelem; k++) {
        const std::vector<double>& g = e.gp[reps > 1 ? r : (e.type == DFM_EQN_SOLID2D ? e.group[k] : 0)];
        p.insert(p.end(), g.begin(), g.begin() + *npar);
      }
    return p;
  }
  const bool diffusion = e.type == DFM_EQN_SCALAR2D && !e.stationary;
  *npar = diffusion ? 3 : 2;
  p.reserve(nelem * *npar);
  for (size_t k = 0; k < nelem; k++) {
    const std::vector<double>& g = e.gp[e.type == DFM_EQN_SCALAR2D ? e.group[k] : 0];
    if (diffusion) p.insert(p.end(), {g[1], g[0], g[2]});  // capacity, alpha, source
    else p.insert(p.end(), {g[0], g[2]});                   // alpha, source
  }
  return p;
}

// MakeLinearSystem: InitializeMarge, the element loop of the mode, FinalizeMarge; the residual norm in *norm
int make_ls(Eqn& e, bool initial, double* norm) {
  if (!e.sys) EQN_TRY(init_ls(e));
  int npar = 0;
  const std::vector<double> p = elem_params(e, &npar);
  e.cleared_ls = false;
  *norm = 0.0;
  if (e.save) {
    if (!e.stationary)
      return fail(DFM_ERR_STATE, "equation system: the saved stiffness of a dynamic system "
                                 "(CLinearSystem_SaveDiaM_Newmark*) is not supported");
    const int kind = e.type == DFM_EQN_SOLID2D ? DFM_SOLID2D : e.type == DFM_EQN_SOLID3D ? DFM_SOLID3D : DFM_POISSON;
    return dfm_sys_assemble_save(e.sys, kind, p.data(), npar, npar);
  }
  int kind;
  double glob[4] = {0.0, 0.0, 0.0, 0.0};
  if (e.solid() && !e.stationary) {
    glob[0] = e.dt; glob[1] = e.gamma; glob[2] = e.beta; glob[3] = initial ? 1.0 : 0.0;
    if (e.type == DFM_EQN_SOLID2D) kind = e.nonlin ? DFM_STVK2D_NEWMARK : DFM_SOLID2D_NEWMARK;
    else kind = e.nonlin ? DFM_STVK3D_NEWMARK : DFM_SOLID3D_NEWMARK;
  } else if (e.solid()) {
    if (e.type == DFM_EQN_SOLID2D) kind = e.nonlin ? DFM_STVK2D : DFM_SOLID2D;
    else kind = e.nonlin ? DFM_STVK3D : DFM_SOLID3D;
  } else if (e.type == DFM_EQN_SCALAR2D && !e.stationary) {
    kind = DFM_DIFFUSION2D;
    glob[0] = e.dt; glob[1] = e.gamma;
  } else {
    kind = DFM_POISSON;
  }
  return dfm_sys_assemble(e.sys, kind, p.data(), npar, npar, glob, 0, nullptr, nullptr, norm);
}

int make_prec(Eqn& e) {
  EQN_TRY(dfm_prec_factor(e.sys));
  e.has_prec = true;
  e.cleared_prec = false;
  return DFM_OK;
}

int pcg(Eqn& e) {
  double conv = kConvRatio;
  unsigned it = kMaxIter;
  const int rc = dfm_solve_pcg(e.sys, DFM_FP64, &conv, &it);
  if (rc < 0) return rc;
  e.itr.emplace_back(it, conv);
  return DFM_OK;
}

int residual_save(Eqn& e) {
  double norm = 0.0;
  return dfm_sys_make_residual(e.sys, &norm);
}

// the field update of the mode (UpdateValueOfField / _NewmarkBeta / _Newmark, src/femls/linearsystem_field.cpp)
int update(Eqn& e, bool initial) {
  if (e.save) return dfm_sys_update_field_save(e.sys);
  if (e.stationary) return dfm_sys_update_field(e.sys);
  if (e.solid()) return dfm_sys_update_field_newmark(e.sys, e.gamma, e.beta, e.dt, initial ? 1 : 0);
  return dfm_sys_update_field_newmark1(e.sys, e.gamma, e.dt, initial ? 1 : 0);
}

int solve_solid(Eqn& e) {
  e.itr.clear();
  const bool d3 = e.type == DFM_EQN_SOLID3D;
  const int max_newton = d3 ? 40 : 10;
  double norm = 0.0;
  if (e.nonlin) {
    if (!e.sys || (d3 && !e.has_prec)) EQN_TRY(init_ls(e));
    double ini = 0.0;
    for (int it = 0; it < max_newton; it++) {
      EQN_TRY(make_ls(e, it == 0, &norm));
      EQN_TRY(make_prec(e));
      if (it == 0) {
        if (norm < 1.0e-20) break;
        ini = norm;
      }
      if (norm < ini * 1.0e-6) break;
      EQN_TRY(pcg(e));
      EQN_TRY(update(e, it == 0));
    }
    return DFM_OK;
  }
  if (d3) {
    // CEqn_Solid3D_Linear::MakeLinearSystem sets the preconditioner's values every time
    if (!e.sys || !e.has_prec) {
      EQN_TRY(init_ls(e));
      EQN_TRY(make_ls(e, true, &norm));
      EQN_TRY(make_prec(e));
    } else if (!e.save) {
      EQN_TRY(make_ls(e, true, &norm));
      EQN_TRY(make_prec(e));
    }
    if (e.save) EQN_TRY(residual_save(e));
  } else {
    if (!e.sys) {
      EQN_TRY(init_ls(e));
      EQN_TRY(make_ls(e, true, &norm));
    } else if (!e.save) {
      EQN_TRY(make_ls(e, true, &norm));
    }
    if (e.save) {
      if (e.cleared_ls) EQN_TRY(make_ls(e, true, &norm));
      EQN_TRY(residual_save(e));
    }
    if (!e.has_prec || !e.save || e.cleared_prec) EQN_TRY(make_prec(e));
  }
  EQN_TRY(pcg(e));
  return update(e, true);
}

int solve_scalar(Eqn& e) {
  double norm = 0.0;
  if (e.type == DFM_EQN_SCALAR3D) {
    if (!e.sys) EQN_TRY(init_ls(e));
    EQN_TRY(make_ls(e, true, &norm));
    EQN_TRY(make_prec(e));
    e.itr.clear();
    EQN_TRY(pcg(e));
    return update(e, true);
  }
  // CEqnSystem_Scalar2D: its MakeLinearSystem also sets the preconditioner's values
  if (!e.sys || !e.has_prec) {
    EQN_TRY(init_ls(e));
    EQN_TRY(make_ls(e, true, &norm));
    EQN_TRY(make_prec(e));
  } else if (!e.save) {
    EQN_TRY(make_ls(e, true, &norm));
    EQN_TRY(make_prec(e));
  }
  if (e.save) {
    if (e.cleared_ls || e.cleared_prec) {
      EQN_TRY(make_ls(e, true, &norm));
      EQN_TRY(make_prec(e));
    }
    EQN_TRY(residual_save(e));
  }
  e.itr.clear();
  EQN_TRY(pcg(e));
  return update(e, true);
}

}  // namespace

struct dfm_eqn_s {
  Eqn e;
};

extern "C" {

int dfm_eqn_create(dfm_ctx ctx, int type, int nnode, const double* coords, int nelem, const int* lnods,
                   const int* elem_group, int ngroup, dfm_eqn* out) {
  if (!ctx || !coords || !lnods || !out) return fail(DFM_ERR_ARG, "dfm_eqn_create: null argument");
  if (type < DFM_EQN_SOLID2D || type > DFM_EQN_SCALAR3D) return fail(DFM_ERR_ARG, "dfm_eqn_create: unknown type");
  if (nnode <= 0 || nelem <= 0 || ngroup < 1) return fail(DFM_ERR_ARG, "dfm_eqn_create: empty mesh");
  if (type == DFM_EQN_SOLID3D || type == DFM_EQN_SCALAR3D) {
    if (ngroup != 1) return fail(DFM_ERR_ARG, "dfm_eqn_create: the 3D equations hold one material (ngroup 1)");
  }
  dfm_eqn h = new dfm_eqn_s;
  Eqn& e = h->e;
  e.type = type;
  e.ctx = ctx;
  e.ndim = (type == DFM_EQN_SOLID3D || type == DFM_EQN_SCALAR3D) ? 3 : 2;
  e.len = type == DFM_EQN_SOLID2D ? 2 : type == DFM_EQN_SOLID3D ? 3 : 1;
  e.nnode = nnode;
  e.ngroup = ngroup;
  const int npoel = e.ndim + 1;
  e.coords.assign(coords, coords + (size_t)nnode * e.ndim);
  // the element arrays: group by group, element order kept inside each group
  std::vector<int> g(nelem, 0);
  if (elem_group)
    for (int k = 0; k < nelem; k++) {
      if (elem_group[k] < 0 || elem_group[k] >= ngroup) {
        delete h;
        return fail(DFM_ERR_ARG, "dfm_eqn_create: element group out of range");
      }
      g[k] = elem_group[k];
    }
  std::vector<int> perm(nelem);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return g[x] < g[y]; });
  e.lnods.resize((size_t)nelem * npoel);
  e.group.resize(nelem);
  for (int k = 0; k < nelem; k++) {
    e.group[k] = g[perm[k]];
    for (int c = 0; c < npoel; c++) {
      const int nd = lnods[(size_t)perm[k] * npoel + c];
      if (nd < 0 || nd >= nnode) {
        delete h;
        return fail(DFM_ERR_ARG, "dfm_eqn_create: node index out of range");
      }
      e.lnods[(size_t)k * npoel + c] = nd;
    }
  }
  e.flags.assign(e.mdof(), 0);
  e.u.assign(e.mdof(), 0.0);
  e.v.assign(e.mdof(), 0.0);
  e.a.assign(e.mdof(), 0.0);
  if (e.type == DFM_EQN_SOLID2D) {
    // CEqnSystem_Solid2D's defaults: Young 1, Poisson 0, the plane-stress flag set, rho 1, no gravity
    const double lambda = 1.0 * 0.0 / ((1.0 + 0.0) * (1 - 2.0 * 0.0)), myu = 1.0 / (2.0 * (1.0 + 0.0));
    e.gp.assign(ngroup, {lambda, myu, 1.0, 0.0, 0.0, 0.0});
  } else if (e.type == DFM_EQN_SOLID3D) {
    e.gp.assign(1, {0.0, 0.0, 1.0, 0.0, 0.0, 0.0});
    e.stationary = false;  // CEqn_Solid3D_Linear's constructor
  } else {
    e.gp.assign(ngroup, {1.0, 1.0, 0.0});  // alpha, capacity, source
  }
  *out = h;
  return DFM_OK;
}

int dfm_eqn_destroy(dfm_eqn eq) {
  if (!eq) return DFM_OK;
  if (eq->e.sys) dfm_sys_destroy(eq->e.sys);
  delete eq;
  return DFM_OK;
}

int dfm_eqn_add_fix(dfm_eqn eq, int nbnd, int nbno, const int* bnd_lnods, int idof) {
  if (!eq || (nbnd > 0 && !bnd_lnods)) return fail(DFM_ERR_ARG, "dfm_eqn_add_fix: null argument");
  Eqn& e = eq->e;
  if (idof >= e.len) return fail(DFM_ERR_ARG, "dfm_eqn_add_fix: component out of range");
  EQN_TRY(drop_ls(e));
  for (long long k = 0; k < (long long)nbnd * nbno; k++) {
    const int nd = bnd_lnods[k];
    if (nd < 0 || nd >= e.nnode) return fail(DFM_ERR_ARG, "dfm_eqn_add_fix: node index out of range");
    // AddFixElemAry: every component of the fixed nodes set to zero (value, velocity, acceleration)
    for (int d = 0; d < e.len; d++) {
      const size_t i = (size_t)nd * e.len + d;
      e.u[i] = e.v[i] = e.a[i] = 0.0;
      if (idof < 0 || idof == d) e.flags[i] = 1;
    }
  }
  return DFM_OK;
}

int dfm_eqn_clear_fix(dfm_eqn eq) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_clear_fix: null argument");
  EQN_TRY(drop_ls(eq->e));
  std::fill(eq->e.flags.begin(), eq->e.flags.end(), 0);
  return DFM_OK;
}

static int check_group(const Eqn& e, int group, const char* name) {
  if (group >= (int)e.gp.size()) return fail(DFM_ERR_ARG, std::string(name) + ": group out of range");
  return DFM_OK;
}

// a group's setter goes through SetEquation (src/femeqn/eqnsys_solid.cpp, eqnsys_scalar.cpp), which drops the
// linear system and the preconditioner; the system's setters flag their values
static int after_set(Eqn& e, int group) {
  if (group >= 0 && (e.type == DFM_EQN_SOLID2D || e.type == DFM_EQN_SCALAR2D)) EQN_TRY(drop_ls(e));
  e.cleared_ls = e.cleared_prec = true;
  return DFM_OK;
}

int dfm_eqn_set_young_poisson(dfm_eqn eq, int group, double young, double poisson, int is_plane_stress) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_set_young_poisson: null argument");
  Eqn& e = eq->e;
  if (!e.solid()) return fail(DFM_ERR_STATE, "dfm_eqn_set_young_poisson: not a solid");
  EQN_TRY(check_group(e, group, "dfm_eqn_set_young_poisson"));
  double lambda = young * poisson / ((1.0 + poisson) * (1 - 2.0 * poisson));
  const double myu = young / (2.0 * (1.0 + poisson));
  // CEqn_Solid2D::SetYoungPoisson (include/delfem/eqnsys_solid.h:190-196): the plane-stress reduction of lambda
  // applies when the flag is false
  if (e.type == DFM_EQN_SOLID2D && !is_plane_stress) lambda = 2 * lambda * myu / (lambda + 2 * myu);
  for (int g = 0; g < (int)e.gp.size(); g++)
    if (group < 0 || g == group) {
      e.gp[g][0] = lambda;
      e.gp[g][1] = myu;
    }
  return after_set(e, group);
}

int dfm_eqn_set_rho(dfm_eqn eq, int group, double rho) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_set_rho: null argument");
  Eqn& e = eq->e;
  if (e.type != DFM_EQN_SOLID2D) return fail(DFM_ERR_STATE, "dfm_eqn_set_rho: a 2D solid only (CEqn_Solid2D)");
  EQN_TRY(check_group(e, group, "dfm_eqn_set_rho"));
  for (int g = 0; g < (int)e.gp.size(); g++)
    if (group < 0 || g == group) e.gp[g][2] = rho;
  return after_set(e, group);
}

int dfm_eqn_set_gravitation(dfm_eqn eq, int group, double gx, double gy, double gz) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_set_gravitation: null argument");
  Eqn& e = eq->e;
  if (!e.solid()) return fail(DFM_ERR_STATE, "dfm_eqn_set_gravitation: not a solid");
  EQN_TRY(check_group(e, group, "dfm_eqn_set_gravitation"));
  for (int g = 0; g < (int)e.gp.size(); g++)
    if (group < 0 || g == group) {
      e.gp[g][3] = gx;
      e.gp[g][4] = gy;
      if (e.type == DFM_EQN_SOLID3D) e.gp[g][5] = gz;
    }
  return after_set(e, group);
}

int dfm_eqn_set_scalar(dfm_eqn eq, int which, int group, double value) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_set_scalar: null argument");
  Eqn& e = eq->e;
  if (e.solid()) return fail(DFM_ERR_STATE, "dfm_eqn_set_scalar: not a scalar equation");
  if (which < 0 || which > 2 || (e.type == DFM_EQN_SCALAR3D && which == 1))
    return fail(DFM_ERR_ARG, "dfm_eqn_set_scalar: which is 0 alpha, 1 capacity (2D), 2 source");
  EQN_TRY(check_group(e, group, "dfm_eqn_set_scalar"));
  for (int g = 0; g < (int)e.gp.size(); g++)
    if (group < 0 || g == group) e.gp[g][which] = value;
  return after_set(e, group);
}

int dfm_eqn_set_flag(dfm_eqn eq, int which, int on) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_set_flag: null argument");
  Eqn& e = eq->e;
  const bool b = on != 0;
  if (which == DFM_EQN_NONLINEAR) {
    if (!e.solid()) return fail(DFM_ERR_STATE, "dfm_eqn_set_flag: geometrical nonlinearity of a solid only");
    if (e.type == DFM_EQN_SOLID2D) {
      e.nonlin = b;
      if (b) {
        e.save = false;
        EQN_TRY(drop_ls(e));
      }
    } else if (b && !e.nonlin) {
      e.nonlin = true;
      e.save = false;
    } else if (!b) {
      e.nonlin = false;
    }
    e.cleared_ls = e.cleared_prec = true;
  } else if (which == DFM_EQN_STATIONARY) {
    if (e.type == DFM_EQN_SCALAR3D) return fail(DFM_ERR_STATE, "dfm_eqn_set_flag: CEqn_Scalar3D is stationary");
    if (e.type == DFM_EQN_SOLID3D) {
      if (b != e.stationary) {
        e.stationary = b;
        if (!b) e.save = false;
        e.cleared_ls = e.cleared_prec = true;
      }
    } else {
      e.stationary = b;
      EQN_TRY(drop_ls(e));
    }
  } else if (which == DFM_EQN_SAVE_STIFF_MAT) {
    if (e.type == DFM_EQN_SCALAR3D) return fail(DFM_ERR_STATE, "dfm_eqn_set_flag: CEqn_Scalar3D has no saved mode");
    if (e.type == DFM_EQN_SOLID3D) {
      // SetSaveStiffMat / UnSetSaveStiffMat (include/delfem/eqnsys_solid.h:101-112); the linear system is made
      // again in the new mode (the reference keeps its old one, of the other class)
      if (b && !(e.save || e.nonlin || !e.stationary)) {
        e.save = true;
        EQN_TRY(drop_ls(e));
        e.cleared_ls = e.cleared_prec = true;
      } else if (!b && e.save) {
        e.save = false;
        EQN_TRY(drop_ls(e));
        e.cleared_ls = e.cleared_prec = true;
      }
    } else {
      e.save = b && !(e.type == DFM_EQN_SOLID2D && e.nonlin);
      EQN_TRY(drop_ls(e));
    }
  } else {
    return fail(DFM_ERR_ARG, "dfm_eqn_set_flag: unknown flag");
  }
  return DFM_OK;
}

int dfm_eqn_set_time(dfm_eqn eq, double dt, double gamma, double beta) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_set_time: null argument");
  Eqn& e = eq->e;
  // CEqnSystem::SetTimeIntegrationParameter (include/delfem/eqnsys.h:88-93)
  e.dt = dt;
  e.gamma = gamma;
  e.beta = beta < 0.0 ? 0.25 * (e.gamma + 0.5) * (e.gamma + 0.5) : beta;
  return DFM_OK;
}

int dfm_eqn_solve(dfm_eqn eq) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_solve: null argument");
  Eqn& e = eq->e;
  return e.solid() ? solve_solid(e) : solve_scalar(e);
}

int dfm_eqn_itr_norm_res(dfm_eqn eq, int* count, unsigned* iters, double* norms, int cap) {
  if (!eq || !count) return fail(DFM_ERR_ARG, "dfm_eqn_itr_norm_res: null argument");
  const Eqn& e = eq->e;
  *count = (int)e.itr.size();
  for (int k = 0; k < *count && k < cap; k++) {
    if (iters) iters[k] = e.itr[k].first;
    if (norms) norms[k] = e.itr[k].second;
  }
  return DFM_OK;
}

int dfm_eqn_get_field(dfm_eqn eq, double* value, double* velocity, double* acceleration) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_get_field: null argument");
  Eqn& e = eq->e;
  if (e.sys) EQN_TRY(dfm_sys_get_fields(e.sys, e.u.data(), e.v.data(), e.a.data()));
  const size_t bytes = sizeof(double) * e.mdof();
  if (value) std::memcpy(value, e.u.data(), bytes);
  if (velocity) std::memcpy(velocity, e.v.data(), bytes);
  if (acceleration) std::memcpy(acceleration, e.a.data(), bytes);
  return DFM_OK;
}

int dfm_eqn_set_field(dfm_eqn eq, const double* value, const double* velocity, const double* acceleration) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_set_field: null argument");
  Eqn& e = eq->e;
  if (e.sys) EQN_TRY(dfm_sys_get_fields(e.sys, e.u.data(), e.v.data(), e.a.data()));
  const size_t bytes = sizeof(double) * e.mdof();
  if (value) std::memcpy(e.u.data(), value, bytes);
  if (velocity) std::memcpy(e.v.data(), velocity, bytes);
  if (acceleration) std::memcpy(e.a.data(), acceleration, bytes);
  if (e.sys) {
    EQN_TRY(dfm_sys_set_field(e.sys, e.u.data(), e.v.data()));
    if (e.solid() && !e.stationary) EQN_TRY(dfm_sys_set_acceleration(e.sys, e.a.data()));
  }
  return DFM_OK;
}

int dfm_eqn_clear(dfm_eqn eq, int what) {
  if (!eq) return fail(DFM_ERR_ARG, "dfm_eqn_clear: null argument");
  Eqn& e = eq->e;
  // ClearValueLinearSystem (1), ClearValuePreconditioner (2), ClearLinearSystemPreconditioner (4)
  if (what & 1) e.cleared_ls = true;
  if (what & 2) e.cleared_prec = true;
  if (what & 4) EQN_TRY(drop_ls(e));
  return DFM_OK;
}

int dfm_eqn_system(dfm_eqn eq, dfm_sys* sys) {
  if (!eq || !sys) return fail(DFM_ERR_ARG, "dfm_eqn_system: null argument");
  *sys = eq->e.sys;
  return DFM_OK;
}

}  // extern "C"
