// dfm_eqnsys.hpp — the reference's equation-system classes over libdfm_gpu (header only, C++11).
//
// dfm::Eqn::CEqnSystem_Solid2D, CEqn_Solid3D_Linear, CEqnSystem_Scalar2D and CEqn_Scalar3D keep the names, setters
// and Solve semantics of include/delfem/eqnsys_solid.h, eqnsys_scalar.h and eqnsys.h in the reference: ILU(1) for
// solids and ILU(0) for scalars, PCG to 1e-6 in at most 1000 iterations, GetAry_ItrNormRes, SetYoungPoisson with
// the reference's plane-stress flag (true keeps the 3D lambda), SetSaveStiffMat, Newton under geometrical
// nonlinearity, Newmark time stepping. The field world is replaced by a mesh: element groups stand for the element
// arrays (GetEquation / SetEquation take the group index where the reference takes an element-array id), boundary
// element lists for the fixed element arrays. The work runs on the GPU through the dfm_eqn_* entry points.
#ifndef DFM_EQNSYS_HPP
#define DFM_EQNSYS_HPP

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dfm_gpu.h"

namespace dfm {
namespace Eqn {

// A P1 mesh: nnode nodes of ndim coordinates, nelem triangles (2D) or tetrahedra (3D); elem_group[k] the element
// array of element k (empty: one array).
struct Mesh {
  int ndim = 2;
  std::vector<double> coords;
  std::vector<int> lnods;
  std::vector<int> elem_group;
  int ngroup = 1;
  int nnode() const { return (int)(coords.size() / ndim); }
  int nelem() const { return (int)(lnods.size() / (ndim + 1)); }
};

// Boundary elements (lines in 2D, triangles in 3D) whose nodes AddFixElemAry fixes.
struct ElemAry {
  int nnoel = 2;
  std::vector<int> lnods;
};

inline void check(int rc) {
  if (rc < 0) throw std::runtime_error(dfm_last_error());
}

class CEqnSystem {
 public:
  CEqnSystem(const CEqnSystem&) = delete;
  CEqnSystem& operator=(const CEqnSystem&) = delete;
  virtual ~CEqnSystem() {
    if (h_) dfm_eqn_destroy(h_);
  }
  virtual bool Solve() {
    check(dfm_eqn_solve(h_));
    return true;
  }
  // the (iterations, conv_ratio) of each PCG solve of the last Solve (include/delfem/eqnsys.h:69)
  std::vector<std::pair<unsigned int, double> > GetAry_ItrNormRes() const {
    int n = 0;
    check(dfm_eqn_itr_norm_res(h_, &n, nullptr, nullptr, 0));
    std::vector<unsigned> it(n);
    std::vector<double> nr(n);
    check(dfm_eqn_itr_norm_res(h_, &n, it.data(), nr.data(), n));
    std::vector<std::pair<unsigned int, double> > out;
    for (int k = 0; k < n; k++) out.push_back(std::make_pair(it[k], nr[k]));
    return out;
  }
  void SetTimeIntegrationParameter(double dt, double gamma = 0.6, double beta = -1.0) {
    check(dfm_eqn_set_time(h_, dt, gamma, beta));
  }
  bool AddFixElemAry(const ElemAry& ea, int idof = -1) {
    check(dfm_eqn_add_fix(h_, (int)(ea.lnods.size() / ea.nnoel), ea.nnoel, ea.lnods.data(), idof));
    return true;
  }
  void ClearFixElemAry() { check(dfm_eqn_clear_fix(h_)); }
  void ClearValueLinearSystem() { check(dfm_eqn_clear(h_, 1)); }
  void ClearValuePreconditioner() { check(dfm_eqn_clear(h_, 2)); }
  void ClearValueLinearSystemPreconditioner() { check(dfm_eqn_clear(h_, 3)); }
  void ClearLinearSystemPreconditioner() { check(dfm_eqn_clear(h_, 4)); }
  // field values (nnode * dofs per node each; VALUE, VELOCITY, ACCELERATION)
  void GetField(std::vector<double>* value, std::vector<double>* velocity = nullptr,
                std::vector<double>* acceleration = nullptr) const {
    for (std::vector<double>* v : {value, velocity, acceleration})
      if (v) v->assign((size_t)nnode_ * len_, 0.0);
    check(dfm_eqn_get_field(h_, value ? value->data() : nullptr, velocity ? velocity->data() : nullptr,
                            acceleration ? acceleration->data() : nullptr));
  }
  void SetField(const double* value, const double* velocity = nullptr, const double* acceleration = nullptr) {
    check(dfm_eqn_set_field(h_, value, velocity, acceleration));
  }
  dfm_eqn handle() const { return h_; }

 protected:
  CEqnSystem(dfm_ctx ctx, int type, const Mesh& m) : nnode_(m.nnode()) {
    len_ = type == DFM_EQN_SOLID2D ? 2 : type == DFM_EQN_SOLID3D ? 3 : 1;
    check(dfm_eqn_create(ctx, type, m.nnode(), m.coords.data(), m.nelem(), m.lnods.data(),
                         m.elem_group.empty() ? nullptr : m.elem_group.data(), m.ngroup, &h_));
  }
  dfm_eqn h_ = nullptr;
  int nnode_ = 0, len_ = 1;
};

// The material of one element array of a 2D solid (include/delfem/eqnsys_solid.h:155-245)
class CEqn_Solid2D {
 public:
  explicit CEqn_Solid2D(unsigned int id_ea = 0) : id_ea_(id_ea) {}
  void SetYoungPoisson(double young, double poisson, bool is_plane_stress) {
    young_ = young; poisson_ = poisson; plane_ = is_plane_stress; has_yp_ = true;
  }
  void SetRho(double rho) { rho_ = rho; has_rho_ = true; }
  void SetGravitation(double g_x, double g_y) { gx_ = g_x; gy_ = g_y; has_g_ = true; }
  unsigned int GetIdEA() const { return id_ea_; }

 private:
  friend class CEqnSystem_Solid2D;
  unsigned int id_ea_;
  double young_ = 1.0, poisson_ = 0.0, rho_ = 1.0, gx_ = 0.0, gy_ = 0.0;
  bool plane_ = true, has_yp_ = false, has_rho_ = false, has_g_ = false;
};

class CEqnSystem_Solid2D : public CEqnSystem {
 public:
  CEqnSystem_Solid2D(dfm_ctx ctx, const Mesh& m) : CEqnSystem(ctx, DFM_EQN_SOLID2D, m) {}
  void SetYoungPoisson(double young, double poisson, bool is_plane_stress) {
    check(dfm_eqn_set_young_poisson(h_, -1, young, poisson, is_plane_stress ? 1 : 0));
  }
  void SetRho(double rho) { check(dfm_eqn_set_rho(h_, -1, rho)); }
  void SetGravitation(double g_x, double g_y) { check(dfm_eqn_set_gravitation(h_, -1, g_x, g_y, 0.0)); }
  void SetGeometricalNonlinear(bool is_nonlin) { check(dfm_eqn_set_flag(h_, DFM_EQN_NONLINEAR, is_nonlin)); }
  void SetStationary(bool is_stationary) { check(dfm_eqn_set_flag(h_, DFM_EQN_STATIONARY, is_stationary)); }
  void SetSaveStiffMat(bool is_save) { check(dfm_eqn_set_flag(h_, DFM_EQN_SAVE_STIFF_MAT, is_save)); }
  CEqn_Solid2D GetEquation(unsigned int id_ea) const { return CEqn_Solid2D(id_ea); }
  // the values set on eqn go to its element array (through SetEquation, which drops the linear system)
  bool SetEquation(const CEqn_Solid2D& eqn) {
    const int g = (int)eqn.id_ea_;
    if (eqn.has_yp_) check(dfm_eqn_set_young_poisson(h_, g, eqn.young_, eqn.poisson_, eqn.plane_ ? 1 : 0));
    if (eqn.has_rho_) check(dfm_eqn_set_rho(h_, g, eqn.rho_));
    if (eqn.has_g_) check(dfm_eqn_set_gravitation(h_, g, eqn.gx_, eqn.gy_, 0.0));
    return true;
  }
};

class CEqn_Solid3D_Linear : public CEqnSystem {
 public:
  CEqn_Solid3D_Linear(dfm_ctx ctx, const Mesh& m) : CEqnSystem(ctx, DFM_EQN_SOLID3D, m) {}
  void SetYoungPoisson(double young, double poisson) {
    check(dfm_eqn_set_young_poisson(h_, -1, young, poisson, 1));
  }
  void SetGravitation(double g_x, double g_y, double g_z) { check(dfm_eqn_set_gravitation(h_, -1, g_x, g_y, g_z)); }
  void SetGeometricalNonLinear() { check(dfm_eqn_set_flag(h_, DFM_EQN_NONLINEAR, 1)); }
  void UnSetGeometricalNonLinear() { check(dfm_eqn_set_flag(h_, DFM_EQN_NONLINEAR, 0)); }
  void SetSaveStiffMat() { check(dfm_eqn_set_flag(h_, DFM_EQN_SAVE_STIFF_MAT, 1)); }
  void UnSetSaveStiffMat() { check(dfm_eqn_set_flag(h_, DFM_EQN_SAVE_STIFF_MAT, 0)); }
  void SetStationary() { check(dfm_eqn_set_flag(h_, DFM_EQN_STATIONARY, 1)); }
  void UnSetStationary() { check(dfm_eqn_set_flag(h_, DFM_EQN_STATIONARY, 0)); }
};

// The coefficients of one element array of a 2D scalar equation (include/delfem/eqnsys_scalar.h:60-100)
class CEqn_Scalar2D {
 public:
  explicit CEqn_Scalar2D(unsigned int id_ea = 0) : id_ea_(id_ea) {}
  void SetAlpha(double alpha) { v_[0] = alpha; has_[0] = true; }
  void SetCapacity(double capa) { v_[1] = capa; has_[1] = true; }
  void SetSource(double source) { v_[2] = source; has_[2] = true; }
  unsigned int GetIdEA() const { return id_ea_; }

 private:
  friend class CEqnSystem_Scalar2D;
  unsigned int id_ea_;
  double v_[3] = {1.0, 1.0, 0.0};
  bool has_[3] = {false, false, false};
};

class CEqnSystem_Scalar2D : public CEqnSystem {
 public:
  CEqnSystem_Scalar2D(dfm_ctx ctx, const Mesh& m) : CEqnSystem(ctx, DFM_EQN_SCALAR2D, m) {}
  void SetAlpha(double alpha) { check(dfm_eqn_set_scalar(h_, 0, -1, alpha)); }
  void SetCapacity(double capa) { check(dfm_eqn_set_scalar(h_, 1, -1, capa)); }
  void SetSource(double source) { check(dfm_eqn_set_scalar(h_, 2, -1, source)); }
  void SetStationary(bool is_stat) { check(dfm_eqn_set_flag(h_, DFM_EQN_STATIONARY, is_stat)); }
  void SetSaveStiffMat(bool is_save) { check(dfm_eqn_set_flag(h_, DFM_EQN_SAVE_STIFF_MAT, is_save)); }
  CEqn_Scalar2D GetEquation(unsigned int id_ea) const { return CEqn_Scalar2D(id_ea); }
  bool SetEquation(const CEqn_Scalar2D& eqn) {
    for (int k = 0; k < 3; k++)
      if (eqn.has_[k]) check(dfm_eqn_set_scalar(h_, k, (int)eqn.id_ea_, eqn.v_[k]));
    return true;
  }
};

class CEqn_Scalar3D : public CEqnSystem {
 public:
  CEqn_Scalar3D(dfm_ctx ctx, const Mesh& m) : CEqnSystem(ctx, DFM_EQN_SCALAR3D, m) {}
  void SetAlpha(double alpha) { check(dfm_eqn_set_scalar(h_, 0, -1, alpha)); }
  void SetSource(double source) { check(dfm_eqn_set_scalar(h_, 2, -1, source)); }
};

}  // namespace Eqn
}  // namespace dfm

#endif  // DFM_EQNSYS_HPP
