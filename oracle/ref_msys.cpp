// Test-infrastructure driver of the *reference*'s multi-segment linear systems (compiled with the reference's own
// sources into oracle/_ref/libdfm_ref.so by oracle/Makefile). Nothing in the product links it.
//
// Two ways to make a system, both ending in a LsSol::CLinearSystem with several segments:
//  - from host arrays: CLinearSystem::AddLinSysSeg / AddMat_Dia / AddMat_NonDia (src/ls/linearsystem.cpp:138-201,
//    419-500) with the values written into the reference's own storage through GetPtrValPSuP / GetPtrValDia;
//  - the 2D Mooney-Rivlin P1b/P1 system the reference assembles itself: a displacement field with CORNER and BUBBLE
//    nodes and a CORNER lambda field (CLinearSystem_Field::AddPattern_Field(disp), AddPattern_Field(lambda, disp),
//    as test_glut/hyper3d/main.cpp:261-270 does in 3D) filled by Fem::Eqn::AddLinSys_Hyper2D_Static
//    (src/femeqn/eqn_hyper.cpp:82-112, 342-727).
// Then the reference's CPreconditioner_ILU (ILU(0), src/ls/preconditioner.cpp:57-335), CLinearSystem::MATVEC and
// LsSol::Solve_PCG / Solve_PBiCGSTAB run on it. The factor values are read by repeating CPreconditioner_ILU's
// SetLinearSystem / SetValue sequence with the same public classes (CMatDiaFrac_BlkCrs, CMatFrac_BlkCrs), whose
// factors CPreconditioner_ILU keeps private.
//
// Vectors and Dirichlet flags are concatenated over the segments (segment 0 first).

#include <cassert>
#include <cstring>
#include <vector>

#include "delfem/field_world.h"
#include "delfem/field.h"
#include "delfem/mesh_interface.h"
#include "delfem/femls/linearsystem_field.h"
#include "delfem/ls/linearsystem.h"
#include "delfem/ls/preconditioner.h"
#include "delfem/ls/solver_ls_iter.h"
#include "delfem/matvec/matdia_blkcrs.h"
#include "delfem/matvec/matdiafrac_blkcrs.h"
#include "delfem/matvec/matfrac_blkcrs.h"
#include "delfem/matvec/vector_blk.h"
#include "delfem/matvec/bcflag_blk.h"
#include "delfem/femeqn/eqn_hyper.h"
#include "delfem/indexed_array.h"

namespace {

class CTriMesh : public Msh::IMesh {
 public:
  std::vector<double> coord;
  std::vector<int> lnods;
  virtual unsigned int GetDimention() const { return 2; }
  virtual void GetInfo(unsigned int id_msh, unsigned int& id_cad, unsigned int& id_before_ext, unsigned int& inum_ext,
                       int& ilayer) const {
    id_cad = id_msh; id_before_ext = 0; inum_ext = 0; ilayer = 0;
  }
  virtual void GetCoord(std::vector<double>& c) const { c = coord; }
  virtual Msh::MSH_TYPE GetConnectivity(unsigned int, std::vector<int>& l) const {
    l = lnods;
    return Msh::TRI;
  }
  virtual std::vector<unsigned int> GetAry_ID() const { return std::vector<unsigned int>(1, 1); }
  virtual std::vector<unsigned int> GetIncludeElemIDAry(unsigned int) const { return std::vector<unsigned int>(); }
};

struct RefMs {
  LsSol::CLinearSystem own;
  Fem::Field::CFieldWorld world;
  CTriMesh mesh;
  Fem::Ls::CLinearSystem_Field* lsf = 0;
  LsSol::CLinearSystem* ls = 0;
  ~RefMs() { delete lsf; }
};

LsSol::CLinearSystem& L(void* h) { return *((RefMs*)h)->ls; }

MatVec::CVector_Blk& V(LsSol::CLinearSystem& ls, int iv, unsigned int i) { return ls.GetVector(iv, i); }

// The coupling matrices' block shape (uniform block lengths only)
int rows_of(LsSol::CLinearSystem& ls, unsigned int i) { return (int)ls.GetVector(-1, i).NBlk(); }
int len_of(LsSol::CLinearSystem& ls, unsigned int i) { return ls.GetVector(-1, i).Len(); }

void copy_factor(const MatVec::CMat_BlkCrs& f, int nrow, int bs, double* val) {
  size_t pos = 0;
  for (int r = 0; r < nrow; r++) {
    unsigned int np = 0;
    const double* v = f.GetPtrValPSuP(r, np);
    if (np) std::memcpy(&val[pos * bs], v, sizeof(double) * np * bs);
    pos += np;
  }
}

}  // namespace

extern "C" {

// An empty system of nseg segments, segment i: nblk[i] blocks of len[i] dofs.
void* ref_ms_create(int nseg, const int* nblk, const int* len) {
  RefMs* s = new RefMs;
  s->ls = &s->own;
  for (int i = 0; i < nseg; i++) s->own.AddLinSysSeg(nblk[i], len[i]);
  return s;
}

// AddMat_Dia (i == j: the off-diagonal pattern of a diagonal segment) or AddMat_NonDia (i != j) from a CRS pattern.
int ref_ms_add_pattern(void* h, int i, int j, const unsigned int* rowptr, const unsigned int* colind) {
  LsSol::CLinearSystem& ls = L(h);
  const int nrow = rows_of(ls, i);
  Com::CIndexedArray crs;
  crs.index.assign(rowptr, rowptr + nrow + 1);
  crs.array.assign(colind, colind + rowptr[nrow]);
  const bool ok = (i == j) ? ls.AddMat_Dia(i, crs) : ls.AddMat_NonDia(i, j, crs);
  return ok ? 1 : 0;
}

// InitializeMarge (src/ls/linearsystem.cpp:74-88): allocates and zeroes the matrices and vectors once the
// patterns are in.
void ref_ms_initialize(void* h) { L(h).InitializeMarge(); }

// The 2D Mooney-Rivlin P1b/P1 system at zero displacement and lambda, assembled by the reference; fixed: per mesh
// node, 1 where both displacement components of its CORNER node are fixed. Segments as AddPattern_Field makes them
// (displacement BUBBLE, displacement CORNER, lambda CORNER). Returns FinalizeMarge's residual norm in *res_norm.
void* ref_ms_hyper2d(int nnode, const double* coords, int nelem, const int* lnods, const int* fixed, double c1,
                     double c2, double rho, double gx, double gy, double* res_norm) {
  RefMs* s = new RefMs;
  s->mesh.coord.assign(coords, coords + (size_t)nnode * 2);
  s->mesh.lnods.assign(lnods, lnods + (size_t)nelem * 3);
  const unsigned int id_base = s->world.AddMesh(s->mesh);
  const unsigned int id_disp = s->world.MakeField_FieldElemDim(
      id_base, 2, Fem::Field::VECTOR2, Fem::Field::VALUE | Fem::Field::VELOCITY | Fem::Field::ACCELERATION,
      Fem::Field::CORNER | Fem::Field::BUBBLE);
  const unsigned int id_lambda = s->world.MakeField_FieldElemDim(
      id_base, 2, Fem::Field::SCALAR, Fem::Field::VALUE | Fem::Field::VELOCITY | Fem::Field::ACCELERATION,
      Fem::Field::CORNER);
  s->lsf = new Fem::Ls::CLinearSystem_Field;
  s->lsf->AddPattern_Field(id_disp, s->world);
  s->lsf->AddPattern_Field(id_lambda, id_disp, s->world);
  s->ls = &s->lsf->m_ls;
  // AddPattern_Field adds the BUBBLE segment before the CORNER one (src/femls/linearsystem_field.cpp:196-202)
  int icorner = -1;
  for (unsigned int i = 0; i < s->ls->GetNLinSysSeg(); i++)
    if (icorner < 0 && rows_of(*s->ls, i) == nnode && len_of(*s->ls, i) == 2) icorner = (int)i;
  assert(icorner >= 0);
  MatVec::CBCFlag& bc = s->ls->GetBCFlag(icorner);
  for (int n = 0; n < nnode; n++)
    if (fixed[n]) { bc.SetBC(n, 0); bc.SetBC(n, 1); }
  s->lsf->InitializeMarge();
  Fem::Eqn::AddLinSys_Hyper2D_Static(*s->lsf, c1, c2, rho, gx, gy, id_disp, id_lambda, s->world);
  *res_norm = s->lsf->FinalizeMarge();
  return s;
}

void ref_ms_destroy(void* h) { delete (RefMs*)h; }

// Number of segments; nblk / len per segment when non-NULL.
int ref_ms_info(void* h, int* nblk, int* len) {
  LsSol::CLinearSystem& ls = L(h);
  const int ns = (int)ls.GetNLinSysSeg();
  for (int i = 0; i < ns; i++) {
    if (nblk) nblk[i] = rows_of(ls, i);
    if (len) len[i] = len_of(ls, i);
  }
  return ns;
}

// Block count of matrix (i, j) (off-diagonal blocks for i == j), -1 when absent.
long ref_ms_ncrs(void* h, int i, int j) {
  LsSol::CLinearSystem& ls = L(h);
  if (!ls.IsMatrix(i, j)) return -1;
  return (i == j) ? (long)ls.GetMatrix(i).NCrs() : (long)ls.GetMatrix(i, j).NCrs();
}

void ref_ms_get_pattern(void* h, int i, int j, unsigned int* rowptr, unsigned int* colind) {
  LsSol::CLinearSystem& ls = L(h);
  const MatVec::CMat_BlkCrs& m =
      (i == j) ? (const MatVec::CMat_BlkCrs&)ls.GetMatrix(i) : (const MatVec::CMat_BlkCrs&)ls.GetMatrix(i, j);
  const int nrow = rows_of(ls, i);
  unsigned int pos = 0;
  rowptr[0] = 0;
  for (int r = 0; r < nrow; r++) {
    unsigned int np = 0;
    const unsigned int* c = m.GetPtrIndPSuP(r, np);
    for (unsigned int k = 0; k < np; k++) colind[pos + k] = c[k];
    pos += np;
    rowptr[r + 1] = pos;
  }
}

// Values of matrix (i, j): blocks row-major in CRS order; dia (i == j) the diagonal blocks.
void ref_ms_get_values(void* h, int i, int j, double* val, double* dia) {
  LsSol::CLinearSystem& ls = L(h);
  const int nrow = rows_of(ls, i);
  const int bs = len_of(ls, i) * len_of(ls, j);
  if (i == j) {
    MatVec::CMatDia_BlkCrs& m = ls.GetMatrix(i);
    copy_factor(m, nrow, bs, val);
    for (int r = 0; r < nrow; r++) std::memcpy(&dia[(size_t)r * bs], m.GetPtrValDia(r), sizeof(double) * bs);
  } else {
    copy_factor(ls.GetMatrix(i, j), nrow, bs, val);
  }
}

// Writes the values of matrix (i, j) into the reference's storage (layout of ref_ms_get_values).
void ref_ms_set_values(void* h, int i, int j, const double* val, const double* dia) {
  LsSol::CLinearSystem& ls = L(h);
  const int nrow = rows_of(ls, i);
  const int bs = len_of(ls, i) * len_of(ls, j);
  MatVec::CMat_BlkCrs& m = (i == j) ? (MatVec::CMat_BlkCrs&)ls.GetMatrix(i) : ls.GetMatrix(i, j);
  size_t pos = 0;
  for (int r = 0; r < nrow; r++) {
    unsigned int np = 0;
    double* v = m.GetPtrValPSuP(r, np);
    if (np) std::memcpy(v, &val[pos * bs], sizeof(double) * np * bs);
    pos += np;
  }
  if (i == j)
    for (int r = 0; r < nrow; r++) std::memcpy(ls.GetMatrix(i).GetPtrValDia(r), &dia[(size_t)r * bs], sizeof(double) * bs);
}

// iv: -1 residual, -2 update, >= 0 temporaries (grown as needed); v concatenated over the segments
void ref_ms_set_vector(void* h, int iv, const double* v) {
  LsSol::CLinearSystem& ls = L(h);
  if (iv >= 0 && (int)ls.GetTmpVectorArySize() <= iv) ls.ReSizeTmpVecSolver(iv + 1);
  size_t pos = 0;
  for (unsigned int i = 0; i < ls.GetNLinSysSeg(); i++) {
    MatVec::CVector_Blk& x = V(ls, iv, i);
    for (unsigned int b = 0; b < x.NBlk(); b++)
      for (int d = 0; d < x.Len(); d++) x.SetValue(b, d, v[pos++]);
  }
}

void ref_ms_get_vector(void* h, int iv, double* v) {
  LsSol::CLinearSystem& ls = L(h);
  size_t pos = 0;
  for (unsigned int i = 0; i < ls.GetNLinSysSeg(); i++) {
    const MatVec::CVector_Blk& x = V(ls, iv, i);
    for (unsigned int b = 0; b < x.NBlk(); b++)
      for (int d = 0; d < x.Len(); d++) v[pos++] = x.GetValue(b, d);
  }
}

// Dirichlet flags (concatenated) into the segments' CBCFlag, then FinalizeMarge; returns the residual norm.
double ref_ms_finalize(void* h, const int* flags) {
  LsSol::CLinearSystem& ls = L(h);
  ls.ClearFixedBoundaryCondition();
  size_t pos = 0;
  for (unsigned int i = 0; i < ls.GetNLinSysSeg(); i++) {
    MatVec::CBCFlag& bc = ls.GetBCFlag(i);
    const int n = rows_of(ls, i), len = len_of(ls, i);
    for (int b = 0; b < n; b++)
      for (int d = 0; d < len; d++)
        if (flags[pos++]) bc.SetBC(b, d);
  }
  return ls.FinalizeMarge();
}

// ILU(0) factor of the current matrices through CPreconditioner_ILU's steps (src/ls/preconditioner.cpp:91-107 and
// 215-297 with no fill level): val[i*nseg+j] receives the factor blocks of (i, j) in the matrix's pattern, dia[i]
// the inverted diagonal blocks of segment i. Returns 1, or 0 when a DoILUDecomp fails.
int ref_ms_factor(void* h, double** val, double** dia) {
  LsSol::CLinearSystem& ls = L(h);
  const unsigned int ns = ls.GetNLinSysSeg();
  std::vector<MatVec::CMatDiaFrac_BlkCrs*> D(ns, (MatVec::CMatDiaFrac_BlkCrs*)0);
  std::vector<MatVec::CMatFrac_BlkCrs*> F(ns * ns, (MatVec::CMatFrac_BlkCrs*)0);
  for (unsigned int i = 0; i < ns; i++) {
    D[i] = new MatVec::CMatDiaFrac_BlkCrs(0, ls.GetMatrix(i));
    for (unsigned int j = 0; j < ns; j++)
      if (i != j && ls.IsMatrix(i, j)) F[i * ns + j] = new MatVec::CMatFrac_BlkCrs(ls.GetMatrix(i, j));
  }
  for (unsigned int i = 0; i < ns; i++) {
    D[i]->SetValue_Initialize(ls.GetMatrix(i));
    for (unsigned int j = 0; j < ns; j++)
      if (F[i * ns + j]) F[i * ns + j]->SetValue_Initialize(ls.GetMatrix(i, j));
  }
  int ok = 1;
  for (unsigned int i = 0; i < ns && ok; i++) {
    for (unsigned int j = 0; j < ns && ok; j++) {
      if (i == j) {
        for (unsigned int k = 0; k < i; k++)
          if (F[i * ns + k] && F[k * ns + i] && !D[i]->DoILUDecompLowUp(*F[i * ns + k], *F[k * ns + i])) ok = 0;
        if (ok && !D[i]->DoILUDecomp()) ok = 0;
        continue;
      }
      MatVec::CMatFrac_BlkCrs* f = F[i * ns + j];
      if (!f) continue;
      const unsigned int kmax = i < j ? i : j;
      for (unsigned int k = 0; k < kmax + 1 && ok; k++) {
        if (k == i && k < j) { if (!f->DoILUDecompUp(*D[i])) ok = 0; continue; }
        if (k == j && k < i) { if (!f->DoILUDecompLow(*D[j])) ok = 0; continue; }
        if (!F[i * ns + k] || !F[k * ns + j]) continue;
        if (!f->DoILUDecompLowUp(*F[i * ns + k], *F[k * ns + j])) ok = 0;
      }
    }
  }
  for (unsigned int i = 0; i < ns; i++) {
    const int n = rows_of(ls, i), bs = len_of(ls, i) * len_of(ls, i);
    if (val[i * ns + i]) copy_factor(*D[i], n, bs, val[i * ns + i]);
    if (dia[i])
      for (int r = 0; r < n; r++) std::memcpy(&dia[i][(size_t)r * bs], D[i]->GetPtrValDia(r), sizeof(double) * bs);
    for (unsigned int j = 0; j < ns; j++)
      if (F[i * ns + j] && val[i * ns + j]) copy_factor(*F[i * ns + j], n, len_of(ls, i) * len_of(ls, j), val[i * ns + j]);
  }
  for (unsigned int i = 0; i < ns; i++) delete D[i];
  for (size_t k = 0; k < F.size(); k++) delete F[k];
  return ok;
}

// v := M^-1 v (handle iv) with the reference's CPreconditioner_ILU, ILU(0).
int ref_ms_precond(void* h, int iv) {
  LsSol::CLinearSystem& ls = L(h);
  LsSol::CPreconditioner_ILU prec;
  prec.SetFillInLevel(0);
  prec.SetLinearSystem(ls);
  if (!prec.SetValue(ls)) return 0;
  return prec.SolvePrecond(ls, iv) ? 1 : 0;
}

void ref_ms_matvec(void* h, double alpha, int iv1, double beta, int iv2) { L(h).MATVEC(alpha, iv1, beta, iv2); }

double ref_ms_dot(void* h, int iv1, int iv2) { return L(h).DOT(iv1, iv2); }

// method 0 Solve_PCG, 1 Solve_PBiCGSTAB with CPreconditioner_ILU (ILU(0)); the update stays in handle -2.
int ref_ms_solve(void* h, int method, double* conv_ratio, unsigned int* iter) {
  LsSol::CLinearSystem& ls = L(h);
  LsSol::CPreconditioner_ILU prec;
  prec.SetFillInLevel(0);
  prec.SetLinearSystem(ls);
  if (!prec.SetValue(ls)) return -1;
  LsSol::CLinearSystemPreconditioner lsp(ls, prec);
  const bool ok = method == 0 ? LsSol::Solve_PCG(*conv_ratio, *iter, lsp) : LsSol::Solve_PBiCGSTAB(*conv_ratio, *iter, lsp);
  return ok ? 1 : 0;
}

}  // extern "C"
