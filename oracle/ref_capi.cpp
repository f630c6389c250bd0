// Test-infrastructure driver: a C interface over the *reference* DelFEM (compiled from its own sources by
// oracle/Makefile into oracle/_ref/libdfm_ref.so). It exists only so that tests can pin the plain-C
// restatement (oracle/cpu_ref.c) against the real reference. Nothing in the product links it.
//
// It drives the reference exactly through its public API, in the order of the reference's own low-level
// usage (test_glut_lowlev/solid2d/main.cpp:183-298 there):
//   CFieldWorld::AddMesh(IMesh) -> MakeField_FieldElemDim -> CLinearSystem_Field::AddPattern_Field
//   -> CBCFlag::SetBC -> InitializeMarge -> Fem::Eqn::AddLinSys_* -> FinalizeMarge
//   -> CPreconditioner_ILU::SetLinearSystem/SetValue -> LsSol::Solve_PCG / Solve_PBiCGSTAB.
// Per-element materials are expressed with one element array per material group, which the reference
// supports natively (one (lambda,mu) or alpha per element array, include/delfem/eqnsys_solid.h:155-245).

#include <cassert>
#include <cstring>
#include <vector>

#include "delfem/field_world.h"
#include "delfem/field.h"
#include "delfem/mesh_interface.h"
#include "delfem/femls/linearsystem_field.h"
#include "delfem/femls/linearsystem_fieldsave.h"
#include "delfem/ls/linearsystem.h"
#include "delfem/ls/preconditioner.h"
#include "delfem/ls/solver_ls_iter.h"
#include "delfem/matvec/matdia_blkcrs.h"
#include "delfem/matvec/matdiafrac_blkcrs.h"
#include "delfem/matvec/ordering_blk.h"
#include "delfem/matvec/solver_mat_iter.h"
#include "delfem/matvec/vector_blk.h"
#include "delfem/matvec/bcflag_blk.h"
#include "delfem/femeqn/eqn_poisson.h"
#include "delfem/femeqn/eqn_diffusion.h"
#include "delfem/femeqn/eqn_linear_solid2d.h"
#include "delfem/femeqn/eqn_linear_solid3d.h"
#include "delfem/femeqn/eqn_st_venant.h"
#include "delfem/eqnsys_solid.h"
#include "delfem/eqnsys_scalar.h"

namespace {

// IMesh adapter: one element array per material group, element order preserved inside each group.
class CGroupMesh : public Msh::IMesh {
 public:
  unsigned int ndim;
  std::vector<double> coord;
  std::vector<std::vector<int> > groups;  // lnods per group
  Msh::MSH_TYPE type;
  std::vector<Msh::MSH_TYPE> types;  // per group where set (boundary groups of other element types)
  virtual unsigned int GetDimention() const { return ndim; }
  virtual void GetInfo(unsigned int id_msh, unsigned int& id_cad, unsigned int& id_before_ext,
                       unsigned int& inum_ext, int& ilayer) const {
    id_cad = id_msh; id_before_ext = 0; inum_ext = 0; ilayer = 0;
  }
  virtual void GetCoord(std::vector<double>& c) const { c = coord; }
  virtual Msh::MSH_TYPE GetConnectivity(unsigned int id_msh, std::vector<int>& lnods) const {
    lnods = groups[id_msh - 1];
    return types.empty() ? type : types[id_msh - 1];
  }
  virtual std::vector<unsigned int> GetAry_ID() const {
    std::vector<unsigned int> ids;
    for (unsigned int i = 0; i < groups.size(); i++) ids.push_back(i + 1);
    return ids;
  }
  virtual std::vector<unsigned int> GetIncludeElemIDAry(unsigned int) const {
    return std::vector<unsigned int>();
  }
};

struct RefSys {
  Fem::Field::CFieldWorld world;
  Fem::Ls::CLinearSystem_Field* ls = 0;  // a CLinearSystem_Save for the saved-stiffness path
  bool save = false;
  bool savedia = false;  // CLinearSystem_SaveDiaM_Newmark (scalar field) or _NewmarkBeta (vector field)
  MatVec::CVector_Blk* velo_pre = 0;  // the backward-Euler loop's previous velocity
  ~RefSys() { delete ls; delete velo_pre; }
  CGroupMesh mesh;
  unsigned int id_base, id_field;
  std::vector<unsigned int> id_ea;       // per group
  std::vector<unsigned int> id_partial;  // per group partial field (3D linear solid takes a field)
  int ndim, nnode, len;
};

MatVec::CMatDia_BlkCrs& Mat(RefSys* s) { return s->ls->m_ls.GetMatrix(0); }

}  // namespace

extern "C" {

// kind of field: len = 1 (scalar) or ndim (vector). Element groups are 0..ngroup-1 per element.
// mode bit 0: the linear system is a CLinearSystem_Save (src/femls/linearsystem_fieldsave.cpp); bit 1: the field
// also has ACCELERATION values (Newmark); bit 2: the linear system is a CLinearSystem_SaveDiaM_Newmark (len 1) or
// CLinearSystem_SaveDiaM_NewmarkBeta (len 2; set bit 1 too). Without bit 1 the field is the one the other drivers use (VALUE |
// VELOCITY): the reference's matrix-level loops read uninitialised heap memory, whose contents depend on it.
void* ref_sys_create2(int ndim, int nnode, const double* coords, int nelem, const int* lnods,
                      int ngroup, const int* elem_group, int len, int mode) {
  RefSys* s = new RefSys;
  s->save = (mode & 5) != 0;
  s->savedia = (mode & 4) != 0;
  if (s->savedia && len == 1) s->ls = new Fem::Ls::CLinearSystem_SaveDiaM_Newmark;
  else if (s->savedia) s->ls = new Fem::Ls::CLinearSystem_SaveDiaM_NewmarkBeta;
  else if (s->save) s->ls = new Fem::Ls::CLinearSystem_Save;
  else s->ls = new Fem::Ls::CLinearSystem_Field;
  s->ndim = ndim; s->nnode = nnode; s->len = len;
  const int npoel = (ndim == 2) ? 3 : 4;
  s->mesh.ndim = ndim;
  s->mesh.type = (ndim == 2) ? Msh::TRI : Msh::TET;
  s->mesh.coord.assign(coords, coords + (size_t)nnode * ndim);
  s->mesh.groups.resize(ngroup);
  for (int ie = 0; ie < nelem; ie++) {
    const int g = elem_group ? elem_group[ie] : 0;
    for (int k = 0; k < npoel; k++) s->mesh.groups[g].push_back(lnods[(size_t)ie * npoel + k]);
  }
  s->id_base = s->world.AddMesh(s->mesh);
  Fem::Field::FIELD_TYPE ft = Fem::Field::SCALAR;
  if (len == 2) ft = Fem::Field::VECTOR2;
  if (len == 3) ft = Fem::Field::VECTOR3;
  s->id_field = s->world.MakeField_FieldElemDim(s->id_base, ndim, ft,
                                                Fem::Field::VALUE | Fem::Field::VELOCITY |
                                                    ((mode & 2) ? Fem::Field::ACCELERATION : 0),
                                                Fem::Field::CORNER);
  const Fem::Field::CField& field = s->world.GetField(s->id_field);
  s->id_ea = field.GetAryIdEA();
  assert((int)s->id_ea.size() == ngroup);
  s->ls->AddPattern_Field(s->id_field, s->world);
  if (ndim == 3 && len == 3 && ngroup > 1) {
    for (int g = 0; g < ngroup; g++) {
      s->id_partial.push_back(s->world.GetPartialField(s->id_field, s->id_ea[g]));
    }
  }
  return s;
}

void* ref_sys_create(int ndim, int nnode, const double* coords, int nelem, const int* lnods,
                     int ngroup, const int* elem_group, int len) {
  return ref_sys_create2(ndim, nnode, coords, nelem, lnods, ngroup, elem_group, len, 0);
}

void ref_sys_destroy(void* h) { delete (RefSys*)h; }

int ref_sys_ncrs(void* h) { return (int)Mat((RefSys*)h).NCrs(); }

void ref_sys_pattern(void* h, unsigned int* rowptr, unsigned int* colind) {
  RefSys* s = (RefSys*)h;
  const MatVec::CMatDia_BlkCrs& m = Mat(s);
  unsigned int pos = 0;
  rowptr[0] = 0;
  for (int i = 0; i < s->nnode; i++) {
    unsigned int np = 0;
    const unsigned int* cols = m.GetPtrIndPSuP(i, np);
    for (unsigned int k = 0; k < np; k++) colind[pos + k] = cols[k];
    pos += np;
    rowptr[i + 1] = pos;
  }
}

// nodal field values: u = VALUE (n*len), v = VELOCITY (n*len) or NULL
void ref_sys_set_field(void* h, const double* u, const double* v) {
  RefSys* s = (RefSys*)h;
  Fem::Field::CField& field = s->world.GetField(s->id_field);
  Fem::Field::CNodeAry::CNodeSeg& ns_u =
      field.GetNodeSeg(Fem::Field::CORNER, true, s->world, Fem::Field::VALUE);
  Fem::Field::CNodeAry::CNodeSeg& ns_v =
      field.GetNodeSeg(Fem::Field::CORNER, true, s->world, Fem::Field::VELOCITY);
  for (int i = 0; i < s->nnode; i++) {
    for (int d = 0; d < s->len; d++) {
      ns_u.SetValue(i, d, u ? u[(size_t)i * s->len + d] : 0.0);
      ns_v.SetValue(i, d, v ? v[(size_t)i * s->len + d] : 0.0);
    }
  }
}

// ACCELERATION field values (n*len)
void ref_sys_set_acc(void* h, const double* a) {
  RefSys* s = (RefSys*)h;
  Fem::Field::CField& field = s->world.GetField(s->id_field);
  Fem::Field::CNodeAry::CNodeSeg& ns_a =
      field.GetNodeSeg(Fem::Field::CORNER, true, s->world, Fem::Field::ACCELERATION);
  for (int i = 0; i < s->nnode; i++)
    for (int d = 0; d < s->len; d++) ns_a.SetValue(i, d, a ? a[(size_t)i * s->len + d] : 0.0);
}

void ref_sys_get_fields(void* h, double* u, double* v, double* a) {
  RefSys* s = (RefSys*)h;
  Fem::Field::CField& field = s->world.GetField(s->id_field);
  const Fem::Field::FIELD_DERIVATION_TYPE fdt[3] = {Fem::Field::VALUE, Fem::Field::VELOCITY,
                                                     Fem::Field::ACCELERATION};
  double* out[3] = {u, v, a};
  for (int k = 0; k < 3; k++) {
    if (!out[k]) continue;
    const Fem::Field::CNodeAry::CNodeSeg& ns = field.GetNodeSeg(Fem::Field::CORNER, true, s->world, fdt[k]);
    for (int i = 0; i < s->nnode; i++) ns.GetValue(i, &out[k][(size_t)i * s->len]);
  }
}

void ref_sys_set_velo_pre(void* h, const double* v0) {
  RefSys* s = (RefSys*)h;
  if (!s->velo_pre) s->velo_pre = new MatVec::CVector_Blk(s->nnode, s->len);
  for (int i = 0; i < s->nnode; i++)
    for (int d = 0; d < s->len; d++) s->velo_pre->SetValue(i, d, v0 ? v0[(size_t)i * s->len + d] : 0.0);
}

// CLinearSystem_Field::UpdateValueOfField_BackwardEular
void ref_sys_update_field_beuler(void* h, double dt, int is_initial) {
  RefSys* s = (RefSys*)h;
  s->ls->UpdateValueOfField_BackwardEular(dt, s->id_field, s->world, is_initial != 0);
}

// CLinearSystem_Field::UpdateValueOfField_NewmarkBeta
void ref_sys_update_field_newmark(void* h, double gamma, double beta, double dt, int is_initial) {
  RefSys* s = (RefSys*)h;
  s->ls->UpdateValueOfField_NewmarkBeta(gamma, beta, dt, s->id_field, s->world, is_initial != 0);
}

void ref_sys_get_field(void* h, double* u) {
  RefSys* s = (RefSys*)h;
  Fem::Field::CField& field = s->world.GetField(s->id_field);
  const Fem::Field::CNodeAry::CNodeSeg& ns_u =
      field.GetNodeSeg(Fem::Field::CORNER, true, s->world, Fem::Field::VALUE);
  for (int i = 0; i < s->nnode; i++) ns_u.GetValue(i, &u[(size_t)i * s->len]);
}

void ref_sys_set_bc(void* h, const int* flags) {
  RefSys* s = (RefSys*)h;
  s->ls->m_ls.ClearFixedBoundaryCondition();
  MatVec::CBCFlag& bc = s->ls->m_ls.GetBCFlag(0);
  for (int i = 0; i < s->nnode; i++)
    for (int d = 0; d < s->len; d++)
      if (flags[(size_t)i * s->len + d]) bc.SetBC(i, d);
}

// kind: 0 Poisson (alpha, source), 1 LinearSolid2D (lambda, myu, rho, gx, gy),
//       2 LinearSolid3D (lambda, myu, rho, gx, gy, gz), 3 StVenant2D (lambda, myu, rho, gx, gy),
//       4 StVenant3D (lambda, myu, rho, gx, gy, gz), 5 Diffusion (rho, alpha, source) with glob = (dt, gamma)
// params: per group, npar doubles each. Returns FinalizeMarge's residual norm in *res_norm.
int ref_sys_assemble(void* h, int kind, const double* params, int npar, const double* glob,
                     double* res_norm) {
  RefSys* s = (RefSys*)h;
  s->ls->InitializeMarge();
  const int ngroup = (int)s->id_ea.size();
  for (int g = 0; g < ngroup; g++) {
    const double* p = &params[(size_t)g * npar];
    const unsigned int id_ea = s->id_ea[g];
    bool ok = true;
    switch (kind) {
      case 0: ok = Fem::Eqn::AddLinSys_Poisson(*s->ls, p[0], p[1], s->world, s->id_field, id_ea); break;
      case 1:
        ok = Fem::Eqn::AddLinSys_LinearSolid2D_Static(*s->ls, p[0], p[1], p[2], p[3], p[4], s->world,
                                                      s->id_field, id_ea);
        break;
      case 2: {
        const unsigned int idf = s->id_partial.empty() ? s->id_field : s->id_partial[g];
        ok = Fem::Eqn::AddLinSys_LinearSolid3D_Static(*s->ls, p[0], p[1], p[2], p[3], p[4], p[5], s->world, idf);
        if (s->id_partial.empty()) g = ngroup;  // one call covers every element array
        break;
      }
      case 3:
        ok = Fem::Eqn::AddLinSys_StVenant2D_Static(*s->ls, p[0], p[1], p[2], p[3], p[4], s->world,
                                                   s->id_field, id_ea);
        break;
      case 4:
        ok = Fem::Eqn::AddLinSys_StVenant3D_Static(*s->ls, p[0], p[1], p[2], p[3], p[4], p[5], s->world,
                                                   s->id_field, id_ea);
        break;
      case 5:
        ok = Fem::Eqn::AddLinSys_Diffusion(glob[0], glob[1], *s->ls, p[0], p[1], p[2], s->world,
                                           s->id_field, id_ea);
        break;
      case 6:  // glob = (dt, gamma, beta, is_initial)
        ok = Fem::Eqn::AddLinSys_LinearSolid2D_NonStatic_NewmarkBeta(glob[0], glob[1], glob[2], *s->ls, p[0], p[1],
                                                                     p[2], p[3], p[4], s->world, s->id_field,
                                                                     glob[3] != 0.0, id_ea);
        break;
      case 9:  // glob = (dt, gamma, beta, is_initial)
        ok = Fem::Eqn::AddLinSys_StVenant3D_NonStatic_NewmarkBeta(glob[0], glob[1], glob[2], *s->ls, p[0], p[1], p[2],
                                                                  p[3], p[4], p[5], s->world, s->id_field,
                                                                  glob[3] != 0.0, id_ea);
        break;
      case 10:  // glob = (dt, gamma, beta, is_initial)
        ok = Fem::Eqn::AddLinSys_StVenant2D_NonStatic_NewmarkBeta(glob[0], glob[1], glob[2], *s->ls, p[0], p[1], p[2],
                                                                  p[3], p[4], s->world, s->id_field, glob[3] != 0.0,
                                                                  id_ea);
        break;
      case 11:  // glob = (dt, is_initial); the previous velocity from ref_sys_set_velo_pre
        ok = Fem::Eqn::AddLinSys_StVenant2D_NonStatic_BackwardEular(glob[0], *s->ls, p[0], p[1], p[2], p[3], p[4],
                                                                    s->world, s->id_field, *s->velo_pre,
                                                                    glob[1] != 0.0, id_ea);
        break;
      case 8:  // glob = (dt, is_initial); the previous velocity from ref_sys_set_velo_pre
        ok = Fem::Eqn::AddLinSys_LinearSolid2D_NonStatic_BackwardEular(glob[0], *s->ls, p[0], p[1], p[2], p[3], p[4],
                                                                       s->world, s->id_field, *s->velo_pre,
                                                                       glob[1] != 0.0, id_ea);
        break;
      case 7: {
        const unsigned int idf = s->id_partial.empty() ? s->id_field : s->id_partial[g];
        ok = Fem::Eqn::AddLinSys_LinearSolid3D_NonStatic_NewmarkBeta(glob[0], glob[1], glob[2], *s->ls, p[0], p[1],
                                                                     p[2], p[3], p[4], p[5], s->world, idf);
        if (s->id_partial.empty()) g = ngroup;
        break;
      }
      default: return -1;
    }
    if (!ok) return -2;
  }
  *res_norm = s->ls->FinalizeMarge();
  return 0;
}

// Matrix values: crs[ncrs*len*len] row-major blocks in CRS order, dia[n*len*len].
void ref_sys_get_matrix(void* h, double* crs, double* dia) {
  RefSys* s = (RefSys*)h;
  const MatVec::CMatDia_BlkCrs& m = Mat(s);
  const int bs = s->len * s->len;
  size_t pos = 0;
  for (int i = 0; i < s->nnode; i++) {
    unsigned int np = 0;
    const double* v = m.GetPtrValPSuP(i, np);
    if (np) std::memcpy(&crs[pos * bs], v, sizeof(double) * np * bs);
    pos += np;
    std::memcpy(&dia[(size_t)i * bs], m.GetPtrValDia(i), sizeof(double) * bs);
  }
}

// iv: -1 residual, -2 update
void ref_sys_get_vector(void* h, int iv, double* out) {
  RefSys* s = (RefSys*)h;
  const MatVec::CVector_Blk& v = s->ls->m_ls.GetVector(iv, 0);
  for (int i = 0; i < s->nnode; i++)
    for (int d = 0; d < s->len; d++) out[(size_t)i * s->len + d] = v.GetValue(i, d);
}

void ref_sys_set_vector(void* h, int iv, const double* in) {
  RefSys* s = (RefSys*)h;
  MatVec::CVector_Blk& v = s->ls->m_ls.GetVector(iv, 0);
  for (int i = 0; i < s->nnode; i++)
    for (int d = 0; d < s->len; d++) v.SetValue(i, d, in[(size_t)i * s->len + d]);
}

// y = alpha*A*x + beta*y through CLinearSystem::MATVEC on handles -1 (x) and -2 (y).
void ref_sys_matvec(void* h, double alpha, const double* x, double beta, double* y) {
  RefSys* s = (RefSys*)h;
  ref_sys_set_vector(h, -1, x);
  ref_sys_set_vector(h, -2, y);
  s->ls->m_ls.MATVEC(alpha, -1, beta, -2);
  ref_sys_get_vector(h, -2, y);
}

// ILU(0) numeric factor of the current matrix, optionally under an ordering (new->old, as in
// CPreconditioner_ILU::SetOrdering). Writes the factor pattern (rowptr/colind, in the new numbering) and
// values (L unscaled, D^-1, D^-1 U). Returns DoILUDecomp's result (1 ok, 0 failed).
int ref_ilu0_factor(void* h, const int* order, unsigned int* rowptr, unsigned int* colind, double* crs,
                    double* dia) {
  RefSys* s = (RefSys*)h;
  const MatVec::CMatDia_BlkCrs& m = Mat(s);
  MatVec::CMatDiaFrac_BlkCrs* f = 0;
  MatVec::COrdering_Blk ord;
  if (order) {
    std::vector<int> o(order, order + s->nnode);
    ord.SetOrdering(o);
    f = new MatVec::CMatDiaFrac_BlkCrs(0, m, ord);
    f->SetValue_Initialize(m, ord);
  } else {
    f = new MatVec::CMatDiaFrac_BlkCrs(0, m);
    f->SetValue_Initialize(m);
  }
  const bool ok = f->DoILUDecomp();
  const int bs = s->len * s->len;
  size_t pos = 0;
  rowptr[0] = 0;
  for (int i = 0; i < s->nnode; i++) {
    unsigned int np = 0;
    const unsigned int* cols = f->GetPtrIndPSuP(i, np);
    const double* v = f->GetPtrValPSuP(i, np);
    for (unsigned int k = 0; k < np; k++) colind[pos + k] = cols[k];
    if (np) std::memcpy(&crs[pos * bs], v, sizeof(double) * np * bs);
    pos += np;
    rowptr[i + 1] = (unsigned int)pos;
    std::memcpy(&dia[(size_t)i * bs], f->GetPtrValDia(i), sizeof(double) * bs);
  }
  delete f;
  return ok ? 1 : 0;
}

// ILU(lev) numeric factor of the current matrix in the natural order: CMatDiaFrac_BlkCrs(lev, A)
// (src/matvec/matdiafrac_blkcrs.cpp:227-287: MakePattern_Initialize, AddFracPtn, MakePatternFinalize), then
// SetValue_Initialize and DoILUDecomp. Writes the factor pattern and values when they fit in cap entries and
// returns the entry count, or -(entry count) when cap is too small; *ok = DoILUDecomp's result.
long ref_ilu_factor(void* h, int lev, long cap, unsigned int* rowptr, unsigned int* colind, double* crs, double* dia,
                    int* ok) {
  RefSys* s = (RefSys*)h;
  const MatVec::CMatDia_BlkCrs& m = Mat(s);
  MatVec::CMatDiaFrac_BlkCrs f(lev, m);
  f.SetValue_Initialize(m);
  *ok = f.DoILUDecomp() ? 1 : 0;
  const long ncrs = (long)f.NCrs();
  if (ncrs > cap) return -ncrs;
  const int bs = s->len * s->len;
  size_t pos = 0;
  rowptr[0] = 0;
  for (int i = 0; i < s->nnode; i++) {
    unsigned int np = 0;
    const unsigned int* cols = f.GetPtrIndPSuP(i, np);
    const double* v = f.GetPtrValPSuP(i, np);
    for (unsigned int k = 0; k < np; k++) colind[pos + k] = cols[k];
    if (np) std::memcpy(&crs[pos * bs], v, sizeof(double) * np * bs);
    pos += np;
    rowptr[i + 1] = (unsigned int)pos;
    std::memcpy(&dia[(size_t)i * bs], f.GetPtrValDia(i), sizeof(double) * bs);
  }
  return ncrs;
}

// z = M^-1 r with the reference's CPreconditioner_ILU (ILU(lev) or ordered ILU(0)).
int ref_ilu_apply(void* h, int lev, const int* order, const double* r, double* z) {
  RefSys* s = (RefSys*)h;
  LsSol::CPreconditioner_ILU prec;
  prec.SetFillInLevel(lev);
  if (order) prec.SetOrdering(std::vector<int>(order, order + s->nnode));
  prec.SetLinearSystem(s->ls->m_ls);
  if (!prec.SetValue(s->ls->m_ls)) return 0;
  if (s->ls->m_ls.GetTmpVectorArySize() < 1) s->ls->m_ls.ReSizeTmpVecSolver(1);
  MatVec::CVector_Blk& v = s->ls->m_ls.GetVector(0, 0);
  for (int i = 0; i < s->nnode; i++)
    for (int d = 0; d < s->len; d++) v.SetValue(i, d, r[(size_t)i * s->len + d]);
  prec.SolvePrecond(s->ls->m_ls, 0);
  for (int i = 0; i < s->nnode; i++)
    for (int d = 0; d < s->len; d++) z[(size_t)i * s->len + d] = v.GetValue(i, d);
  return 1;
}

// Solve with the reference's Krylov loops on the assembled system: method 0 Solve_PCG, 1 Solve_PBiCGSTAB.
// lev = ILU fill level (ignored with an ordering). Returns the solver's bool; conv/iter are in/out as
// in the reference. The update (handle -2) is left in the system (read it with ref_sys_get_vector).
// method: 0 Solve_PCG, 1 Solve_PBiCGSTAB (ILU(lev) under `order`), 2 Solve_CG, 3 Solve_BiCGSTAB (no preconditioner)
int ref_sys_solve(void* h, int method, int lev, const int* order, double* conv_ratio, unsigned int* iter) {
  RefSys* s = (RefSys*)h;
  if (method == 2) return LsSol::Solve_CG(*conv_ratio, *iter, s->ls->m_ls) ? 1 : 0;
  if (method == 3) return LsSol::Solve_BiCGSTAB(*conv_ratio, *iter, s->ls->m_ls) ? 1 : 0;
  LsSol::CPreconditioner_ILU prec;
  prec.SetFillInLevel(lev);
  if (order) prec.SetOrdering(std::vector<int>(order, order + s->nnode));
  prec.SetLinearSystem(s->ls->m_ls);
  if (!prec.SetValue(s->ls->m_ls)) return -1;
  LsSol::CLinearSystemPreconditioner lsp(s->ls->m_ls, prec);
  bool ok = false;
  if (method == 0) ok = LsSol::Solve_PCG(*conv_ratio, *iter, lsp);
  else ok = LsSol::Solve_PBiCGSTAB(*conv_ratio, *iter, lsp);
  return ok ? 1 : 0;
}

// COrdering_Blk's orderings of the system's matrix (kind 3 MakeOrdering_RCM, 4 MakeOrdering_RCM2, 5
// MakeOrdering_AMD), written as NewToOld(i) for every i.
void ref_ordering(void* h, int kind, int* order) {
  RefSys* s = (RefSys*)h;
  const MatVec::CMatDia_BlkCrs& m = Mat(s);
  MatVec::COrdering_Blk o;
  if (kind == 3) o.MakeOrdering_RCM(m);
  else if (kind == 4) o.MakeOrdering_RCM2(m);
  else o.MakeOrdering_AMD(m);
  for (unsigned int i = 0; i < m.NBlkMatCol(); i++) order[i] = o.NewToOld(i);
}

// The matrix-level twins MatVec::Sol::Solve_CG / _PCG / _BiCGSTAB / _PBiCGSTAB (method 0..3; src/matvec/
// solver_mat_iter.cpp) on the system's matrix with its residual (in: right-hand side, out: final residual) and
// update vectors; the preconditioned ones with CMatDiaFrac_BlkCrs ILU(0) in the natural order as the CPrecond_Blk.
int ref_mat_solve(void* h, int method, double* conv_ratio, unsigned int* iter) {
  RefSys* s = (RefSys*)h;
  const MatVec::CMatDia_BlkCrs& a = Mat(s);
  MatVec::CVector_Blk& r = s->ls->m_ls.GetVector(-1, 0);
  MatVec::CVector_Blk& u = s->ls->m_ls.GetVector(-2, 0);
  if (method == 0) return MatVec::Sol::Solve_CG(*conv_ratio, *iter, a, r, u) ? 1 : 0;
  if (method == 2) return MatVec::Sol::Solve_BiCGSTAB(*conv_ratio, *iter, a, r, u) ? 1 : 0;
  MatVec::CMatDiaFrac_BlkCrs frac(0, a);
  frac.SetValue_Initialize(a);
  if (!frac.DoILUDecomp()) return -1;
  if (method == 1) return MatVec::Sol::Solve_PCG(*conv_ratio, *iter, a, r, u, frac, frac) ? 1 : 0;
  return MatVec::Sol::Solve_PBiCGSTAB(*conv_ratio, *iter, a, r, u, frac, frac) ? 1 : 0;
}

// u_field += update (CLinearSystem_Field::UpdateValueOfField, VALUE)
void ref_sys_update_field(void* h) {
  RefSys* s = (RefSys*)h;
  s->ls->UpdateValueOfField(s->id_field, s->world, Fem::Field::VALUE);
}

// Saved stiffness on a system made with save = 1: InitializeMarge + the *_SaveStiffMat element loop +
// FinalizeMarge of CLinearSystem_Save. kind 0 Poisson (triangles), 1 LinearSolid2D, 2 LinearSolid3D.
int ref_sys_assemble_save(void* h, int kind, const double* params, int npar) {
  RefSys* s = (RefSys*)h;
  if (!s->save) return -3;
  Fem::Ls::CLinearSystem_Save& ls = *(Fem::Ls::CLinearSystem_Save*)s->ls;
  ls.InitializeMarge();
  const int ngroup = (int)s->id_ea.size();
  for (int g = 0; g < ngroup; g++) {
    const double* p = &params[(size_t)g * npar];
    const unsigned int id_ea = s->id_ea[g];
    bool ok = true;
    switch (kind) {
      case 0: ok = Fem::Eqn::AddLinSys_Poisson(ls, p[0], p[1], s->world, s->id_field, id_ea); break;
      case 1:
        ok = Fem::Eqn::AddLinSys_LinearSolid2D_Static_SaveStiffMat(ls, p[0], p[1], p[2], p[3], p[4], s->world,
                                                                   s->id_field, id_ea);
        break;
      case 2: {
        const unsigned int idf = s->id_partial.empty() ? s->id_field : s->id_partial[g];
        ok = Fem::Eqn::AddLinSys_LinearSolid3D_Static_SaveStiffMat(ls, p[0], p[1], p[2], p[3], p[4], p[5], s->world,
                                                                   idf);
        if (s->id_partial.empty()) g = ngroup;
        break;
      }
      default: return -1;
    }
    if (!ok) return -2;
  }
  ls.FinalizeMarge();
  return 0;
}

// CLinearSystem_Save::MakeResidual; returns its norm
double ref_sys_make_residual(void* h) {
  RefSys* s = (RefSys*)h;
  return ((Fem::Ls::CLinearSystem_Save*)s->ls)->MakeResidual(s->world);
}

// CLinearSystem_Save::GetForce of the field's corner segment (n*len)
void ref_sys_get_force(void* h, double* f) {
  RefSys* s = (RefSys*)h;
  const MatVec::CVector_Blk& v =
      ((Fem::Ls::CLinearSystem_Save*)s->ls)->GetForce(s->id_field, Fem::Field::CORNER, s->world);
  for (int i = 0; i < s->nnode; i++)
    for (int d = 0; d < s->len; d++) f[(size_t)i * s->len + d] = v.GetValue(i, d);
}

// CLinearSystem_Save::UpdateValueOfField (VALUE)
void ref_sys_update_field_save(void* h) {
  RefSys* s = (RefSys*)h;
  ((Fem::Ls::CLinearSystem_Save*)s->ls)->UpdateValueOfField(s->id_field, s->world, Fem::Field::VALUE);
}

// The lumped-mass saved systems on a system made with mode bit 2: InitializeMarge, SetNewmarkParameter, the element
// loop of each group, FinalizeMarge. kind 5: AddLinSys_Diffusion on CLinearSystem_SaveDiaM_Newmark
// (src/femeqn/eqn_diffusion.cpp:769-805), p = (rho, alpha, source); kind 1:
// AddLinSys_LinearSolid2D_NonStatic_Save_NewmarkBeta (src/femeqn/eqn_linear_solid2d.cpp:313-353),
// p = (lambda, myu, rho, gx, gy).
int ref_sys_assemble_savedia(void* h, int kind, const double* params, int npar, double dt, double gamma,
                             double beta) {
  RefSys* s = (RefSys*)h;
  if (!s->savedia) return -3;
  const int ngroup = (int)s->id_ea.size();
  if (kind == 5 && s->len == 1) {
    Fem::Ls::CLinearSystem_SaveDiaM_Newmark& ls = *(Fem::Ls::CLinearSystem_SaveDiaM_Newmark*)s->ls;
    ls.InitializeMarge();
    ls.SetNewmarkParameter(gamma, dt);
    for (int g = 0; g < ngroup; g++) {
      const double* p = &params[(size_t)g * npar];
      if (!Fem::Eqn::AddLinSys_Diffusion(ls, p[0], p[1], p[2], s->world, s->id_field, s->id_ea[g])) return -2;
    }
    ls.FinalizeMarge();
    return 0;
  }
  if (kind == 1 && s->len == 2) {
    Fem::Ls::CLinearSystem_SaveDiaM_NewmarkBeta& ls = *(Fem::Ls::CLinearSystem_SaveDiaM_NewmarkBeta*)s->ls;
    ls.InitializeMarge();
    ls.SetNewmarkParameter(beta, gamma, dt);
    for (int g = 0; g < ngroup; g++) {
      const double* p = &params[(size_t)g * npar];
      if (!Fem::Eqn::AddLinSys_LinearSolid2D_NonStatic_Save_NewmarkBeta(ls, p[0], p[1], p[2], p[3], p[4], s->world,
                                                                       s->id_field, s->id_ea[g]))
        return -2;
    }
    ls.FinalizeMarge();
    return 0;
  }
  return -1;
}

// UpdateValueOfField of the lumped-mass saved systems: VELOCITY (SaveDiaM_Newmark) or ACCELERATION (_NewmarkBeta)
void ref_sys_update_field_savedia(void* h) {
  RefSys* s = (RefSys*)h;
  s->ls->UpdateValueOfField(s->id_field, s->world,
                            s->len == 1 ? Fem::Field::VELOCITY : Fem::Field::ACCELERATION);
}

// The saved system's boundary matrix (GetMatrix_Boundary of the field's corner segment): its pattern (browptr
// [n+1], bcolind) and values; returns its block count (call with NULL arrays for the count alone).
long ref_sys_get_boundary(void* h, unsigned* browptr, unsigned* bcolind, double* bval) {
  RefSys* s = (RefSys*)h;
  MatVec::CMat_BlkCrs& b = ((Fem::Ls::CLinearSystem_Save*)s->ls)
                               ->GetMatrix_Boundary(s->id_field, Fem::Field::CORNER, s->id_field, Fem::Field::CORNER,
                                                    s->world);
  const unsigned bs = (unsigned)(s->len * s->len);
  long nb = 0;
  if (browptr) browptr[0] = 0;
  for (int i = 0; i < s->nnode; i++) {
    unsigned np = 0;
    const unsigned* ind = b.GetPtrIndPSuP((unsigned)i, np);
    const double* val = b.GetPtrValPSuP((unsigned)i, np);
    for (unsigned k = 0; k < np; k++) {
      if (bcolind) bcolind[nb + k] = ind[k];
      if (bval)
        for (unsigned t = 0; t < bs; t++) bval[(size_t)(nb + k) * bs + t] = val[(size_t)k * bs + t];
    }
    nb += np;
    if (browptr) browptr[i + 1] = (unsigned)nb;
  }
  return nb;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------------------
// The reference's own equation-system classes (CEqnSystem_Solid2D, CEqn_Solid3D_Linear, CEqnSystem_Scalar2D,
// CEqn_Scalar3D, src/femeqn/eqnsys_solid.cpp, eqnsys_scalar.cpp) on a mesh of element groups (one element array
// and so one CEqn_* per group) plus one boundary group (lines in 2D, triangles in 3D) that AddFixElemAry fixes,
// as the reference's examples fix CAD edges (test_glut/solid2d/main.cpp there).

namespace {
struct RefEqn {
  Fem::Field::CFieldWorld world;
  CGroupMesh mesh;
  unsigned int id_base = 0, id_ea_bnd = 0;
  std::vector<unsigned int> id_ea;  // per domain group
  int type = 0, ndim = 0, nnode = 0, len = 0;
  Fem::Eqn::CEqnSystem* eq = 0;
  ~RefEqn() { delete eq; }
  unsigned int id_field() const {
    switch (type) {
      case 0: return ((Fem::Eqn::CEqnSystem_Solid2D*)eq)->GetIdField_Disp();
      case 1: return ((Fem::Eqn::CEqn_Solid3D_Linear*)eq)->GetIdField_Disp();
      case 2: return ((Fem::Eqn::CEqnSystem_Scalar2D*)eq)->GetIdField_Value();
      default: return ((Fem::Eqn::CEqn_Scalar3D*)eq)->GetIdField_Value();
    }
  }
};
}  // namespace

extern "C" {

// type 0 CEqnSystem_Solid2D, 1 CEqn_Solid3D_Linear, 2 CEqnSystem_Scalar2D, 3 CEqn_Scalar3D. Domain elements in
// groups 0..ngroup-1 (elem_group null: one group); nbnd boundary elements of nbnd_no nodes each are fixed with
// AddFixElemAry(id_ea, world, idof) (idof -1: every component).
void* ref_eqn_create(int type, int nnode, const double* coords, int nelem, const int* lnods, int ngroup,
                     const int* elem_group, int nbnd, const int* bnd_lnods, int idof) {
  RefEqn* r = new RefEqn;
  r->type = type;
  r->ndim = (type == 1 || type == 3) ? 3 : 2;
  r->len = type == 0 ? 2 : type == 1 ? 3 : 1;
  r->nnode = nnode;
  const int npoel = r->ndim == 2 ? 3 : 4, nbno = r->ndim == 2 ? 2 : 3;
  r->mesh.ndim = r->ndim;
  r->mesh.type = r->ndim == 2 ? Msh::TRI : Msh::TET;
  r->mesh.coord.assign(coords, coords + (size_t)nnode * r->ndim);
  r->mesh.groups.resize(ngroup + 1);
  r->mesh.types.assign(ngroup, r->mesh.type);
  r->mesh.types.push_back(r->ndim == 2 ? Msh::BAR : Msh::TRI);
  for (int ie = 0; ie < nelem; ie++) {
    const int g = elem_group ? elem_group[ie] : 0;
    for (int k = 0; k < npoel; k++) r->mesh.groups[g].push_back(lnods[(size_t)ie * npoel + k]);
  }
  for (int ib = 0; ib < nbnd * nbno; ib++) r->mesh.groups[ngroup].push_back(bnd_lnods[ib]);
  r->id_base = r->world.AddMesh(r->mesh);
  const Fem::Field::CIDConvEAMshCad conv = r->world.GetIDConverter(r->id_base);
  for (int g = 0; g < ngroup; g++) r->id_ea.push_back(conv.GetIdEA_fromMsh(g + 1));
  r->id_ea_bnd = conv.GetIdEA_fromMsh(ngroup + 1);
  switch (type) {
    case 0: r->eq = new Fem::Eqn::CEqnSystem_Solid2D(r->id_base, r->world); break;
    case 1: r->eq = new Fem::Eqn::CEqn_Solid3D_Linear(r->id_base, r->world); break;
    case 2: {  // (the constructor taking a field is declared but not defined in the reference)
      Fem::Eqn::CEqnSystem_Scalar2D* e = new Fem::Eqn::CEqnSystem_Scalar2D;
      e->SetDomain_Field(r->id_base, r->world);
      r->eq = e;
      break;
    }
    default: {
      Fem::Eqn::CEqn_Scalar3D* e = new Fem::Eqn::CEqn_Scalar3D;
      e->SetDomain(r->id_base, r->world);
      r->eq = e;
    }
  }
  if (nbnd > 0) r->eq->AddFixElemAry(r->id_ea_bnd, r->world, idof);
  return r;
}

void ref_eqn_destroy(void* h) { delete (RefEqn*)h; }

// SetYoungPoisson: Solid2D with is_plane_stress, per group (group >= 0, through GetEquation / SetEquation) or for
// the system; Solid3D for the system.
void ref_eqn_set_young_poisson(void* h, int group, double young, double poisson, int is_plane_stress) {
  RefEqn* r = (RefEqn*)h;
  if (r->type == 1) {
    ((Fem::Eqn::CEqn_Solid3D_Linear*)r->eq)->SetYoungPoisson(young, poisson);
    return;
  }
  Fem::Eqn::CEqnSystem_Solid2D* e = (Fem::Eqn::CEqnSystem_Solid2D*)r->eq;
  if (group < 0) {
    e->SetYoungPoisson(young, poisson, is_plane_stress != 0);
    return;
  }
  Fem::Eqn::CEqn_Solid2D q = e->GetEquation(r->id_ea[group]);
  q.SetYoungPoisson(young, poisson, is_plane_stress != 0);
  e->SetEquation(q);
}

// Solid2D: SetRho / SetGravitation (system or group); Solid3D: SetGravitation
void ref_eqn_set_rho(void* h, int group, double rho) {
  RefEqn* r = (RefEqn*)h;
  Fem::Eqn::CEqnSystem_Solid2D* e = (Fem::Eqn::CEqnSystem_Solid2D*)r->eq;
  if (group < 0) {
    e->SetRho(rho);
    return;
  }
  Fem::Eqn::CEqn_Solid2D q = e->GetEquation(r->id_ea[group]);
  q.SetRho(rho);
  e->SetEquation(q);
}

void ref_eqn_set_gravitation(void* h, int group, double gx, double gy, double gz) {
  RefEqn* r = (RefEqn*)h;
  if (r->type == 1) {
    ((Fem::Eqn::CEqn_Solid3D_Linear*)r->eq)->SetGravitation(gx, gy, gz);
    return;
  }
  Fem::Eqn::CEqnSystem_Solid2D* e = (Fem::Eqn::CEqnSystem_Solid2D*)r->eq;
  if (group < 0) {
    e->SetGravitation(gx, gy);
    return;
  }
  Fem::Eqn::CEqn_Solid2D q = e->GetEquation(r->id_ea[group]);
  q.SetGravitation(gx, gy);
  e->SetEquation(q);
}

// which: 0 geometrical nonlinearity, 1 stationary, 2 saved stiffness matrix
void ref_eqn_set_flag(void* h, int which, int on) {
  RefEqn* r = (RefEqn*)h;
  if (r->type == 0) {
    Fem::Eqn::CEqnSystem_Solid2D* e = (Fem::Eqn::CEqnSystem_Solid2D*)r->eq;
    if (which == 0) e->SetGeometricalNonlinear(on != 0);
    if (which == 1) e->SetStationary(on != 0);
    if (which == 2) e->SetSaveStiffMat(on != 0);
  } else if (r->type == 1) {
    Fem::Eqn::CEqn_Solid3D_Linear* e = (Fem::Eqn::CEqn_Solid3D_Linear*)r->eq;
    if (which == 0) { if (on) e->SetGeometricalNonLinear(); else e->UnSetGeometricalNonLinear(); }
    if (which == 1) { if (on) e->SetStationary(); else e->UnSetStationary(); }
    if (which == 2) { if (on) e->SetSaveStiffMat(); else e->UnSetSaveStiffMat(); }
  } else if (r->type == 2) {
    Fem::Eqn::CEqnSystem_Scalar2D* e = (Fem::Eqn::CEqnSystem_Scalar2D*)r->eq;
    if (which == 1) e->SetStationary(on != 0);
    if (which == 2) e->SetSaveStiffMat(on != 0);
  }
}

// scalars: which 0 alpha, 1 capacity, 2 source (Scalar2D: system or group; Scalar3D: alpha and source)
void ref_eqn_set_scalar(void* h, int which, int group, double val) {
  RefEqn* r = (RefEqn*)h;
  if (r->type == 3) {
    Fem::Eqn::CEqn_Scalar3D* e = (Fem::Eqn::CEqn_Scalar3D*)r->eq;
    if (which == 0) e->SetAlpha(val);
    if (which == 2) e->SetSource(val);
    return;
  }
  Fem::Eqn::CEqnSystem_Scalar2D* e = (Fem::Eqn::CEqnSystem_Scalar2D*)r->eq;
  if (group < 0) {
    if (which == 0) e->SetAlpha(val);
    if (which == 1) e->SetCapacity(val);
    if (which == 2) e->SetSource(val);
    return;
  }
  Fem::Eqn::CEqn_Scalar2D q = e->GetEquation(r->id_ea[group]);
  if (which == 0) q.SetAlpha(val);
  if (which == 1) q.SetCapacity(val);
  if (which == 2) q.SetSource(val);
  e->SetEquation(q);
}

void ref_eqn_set_time(void* h, double dt, double gamma, double beta) {
  ((RefEqn*)h)->eq->SetTimeIntegrationParameter(dt, gamma, beta);
}

int ref_eqn_solve(void* h) {
  RefEqn* r = (RefEqn*)h;
  return r->eq->Solve(r->world) ? 1 : 0;
}

// GetAry_ItrNormRes into iters / norms (at most cap); returns the count
int ref_eqn_itr(void* h, unsigned int* iters, double* norms, int cap) {
  const std::vector<std::pair<unsigned int, double> >& a = ((RefEqn*)h)->eq->GetAry_ItrNormRes();
  for (int i = 0; i < (int)a.size() && i < cap; i++) {
    iters[i] = a[i].first;
    norms[i] = a[i].second;
  }
  return (int)a.size();
}

// VALUE / VELOCITY / ACCELERATION of the equation's field (n*len each; null to skip; a field without the
// derivative leaves its array untouched)
void ref_eqn_get_field(void* h, double* u, double* v, double* a) {
  RefEqn* r = (RefEqn*)h;
  Fem::Field::CField& field = r->world.GetField(r->id_field());
  const Fem::Field::FIELD_DERIVATION_TYPE fdt[3] = {Fem::Field::VALUE, Fem::Field::VELOCITY,
                                                     Fem::Field::ACCELERATION};
  double* out[3] = {u, v, a};
  for (int k = 0; k < 3; k++) {
    if (!out[k] || !(field.GetFieldDerivativeType() & fdt[k])) continue;
    const Fem::Field::CNodeAry::CNodeSeg& ns = field.GetNodeSeg(Fem::Field::CORNER, true, r->world, fdt[k]);
    for (int i = 0; i < r->nnode; i++) ns.GetValue(i, &out[k][(size_t)i * r->len]);
  }
}

}  // extern "C"
