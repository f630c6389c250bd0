// dfm_delfem_adapter.hpp — the reference-side binding of libdfm_gpu (header only). A DelFEM build includes it with
// DelFEM's own include directory on the path; it is not part of the library's build.
//
// CLinearSystemGPU is the solver seam (LsSol::ILinearSystemPreconditioner_Sol,
// include/delfem/ls/linearsystem_interface_solver.h:39-82 in the reference) over a device system: the reference's
// own host loops (LsSol::Solve_PCG / Solve_PBiCGSTAB) run unchanged over device vectors, one C-ABI call per vector
// operation, and SolvePCG / SolvePBiCGSTAB run the device-resident loops.
//
// CLinearSystemGPU_Field is the assembly seam (Fem::Eqn::ILinearSystem_Eqn, include/delfem/
// linearsystem_interface_eqnsys.h:48-68): it is a CLinearSystem_Field, so the reference's element routines
// (AddLinSys_*) merge into it on the host as they always do; Upload installs its matrix pattern, values and
// residual in a device system (dfm_sys_create_pattern, dfm_sys_set_matrix), the device factors and solves, and
// DownloadUpdate puts the solution back in its update vector for UpdateValueOfField.
//
// CLinearSystemGPU_MultiSeg is the same seam for a linear system of several segments (LsSol::CLinearSystem,
// include/delfem/ls/linearsystem.h:51-173), e.g. the displacement (CORNER, BUBBLE) and lambda segments that
// CLinearSystem_Field::AddPattern_Field(id_disp) / AddPattern_Field(id_lambda, id_disp) make for a P1b/P1 element
// routine: the reference's element routines fill the host system, Upload mirrors every matrix, the Dirichlet flags
// and the residual into a device multi-segment system, the device factors (CPreconditioner_ILU, ILU(0)) and solves,
// DownloadUpdate writes the update back.
#ifndef DFM_DELFEM_ADAPTER_HPP
#define DFM_DELFEM_ADAPTER_HPP

#include <stdexcept>
#include <vector>

#include "delfem/femls/linearsystem_field.h"
#include "delfem/ls/linearsystem.h"
#include "delfem/ls/linearsystem_interface_solver.h"
#include "delfem/matvec/bcflag_blk.h"
#include "delfem/matvec/matdia_blkcrs.h"
#include "delfem/matvec/vector_blk.h"
#include "dfm_gpu.h"

namespace dfm {
namespace delfem {

inline void throw_if(int rc) {
  if (rc < 0) throw std::runtime_error(dfm_last_error());
}

class CLinearSystemGPU : public LsSol::ILinearSystemPreconditioner_Sol {
 public:
  CLinearSystemGPU(dfm_ctx ctx, int ndim, int nnode, const double* coords, int nelem, const int* lnods, int len) {
    throw_if(dfm_sys_create_mesh(ctx, ndim, nnode, coords, nelem, lnods, len, &sys_));
  }
  explicit CLinearSystemGPU(dfm_sys sys) : sys_(sys), owned_(false) {}
  ~CLinearSystemGPU() {
    if (owned_ && sys_) dfm_sys_destroy(sys_);
  }
  CLinearSystemGPU(const CLinearSystemGPU&) = delete;
  CLinearSystemGPU& operator=(const CLinearSystemGPU&) = delete;
  // ILinearSystem_Sol
  unsigned int GetTmpVectorArySize() const {
    int n = 0;
    dfm_sys_ntmp(sys_, &n);
    return (unsigned int)n;
  }
  bool ReSizeTmpVecSolver(unsigned int n) { return dfm_sys_resize_tmp(sys_, (int)n) == DFM_OK; }
  double DOT(int a, int b) {
    double d = 0.0;
    throw_if(dfm_sys_dot(sys_, a, b, &d));
    return d;
  }
  bool COPY(int a, int b) { return dfm_sys_copy(sys_, a, b) == DFM_OK; }
  bool SCAL(double s, int a) { return dfm_sys_scal(sys_, s, a) == DFM_OK; }
  bool AXPY(double s, int a, int b) { return dfm_sys_axpy(sys_, s, a, b) == DFM_OK; }
  bool MATVEC(double a, int x, double b, int y) { return dfm_sys_matvec(sys_, a, x, b, y) == DFM_OK; }
  // ILinearSystemPreconditioner_Sol
  bool SolvePrecond(int iv) { return dfm_prec_apply(sys_, iv) == DFM_OK; }
  // CPreconditioner_ILU(ls, lev) + SetValue
  void SetFillLevel(int lev) { throw_if(dfm_prec_set_fill_level(sys_, lev)); }
  bool SetValuePrecond() { return dfm_prec_factor(sys_) == DFM_OK; }
  // device-resident LsSol::Solve_PCG / Solve_PBiCGSTAB (same in/out meaning of conv_ratio and iteration)
  bool SolvePCG(double& conv_ratio, unsigned int& iteration, int precision = DFM_FP64) {
    const int rc = dfm_solve_pcg(sys_, precision, &conv_ratio, &iteration);
    throw_if(rc);
    return rc == 1;
  }
  bool SolvePBiCGSTAB(double& conv_ratio, unsigned int& iteration, int precision = DFM_FP64) {
    const int rc = dfm_solve_pbicgstab(sys_, precision, &conv_ratio, &iteration);
    throw_if(rc);
    return rc == 1;
  }
  dfm_sys handle() const { return sys_; }

 private:
  dfm_sys sys_ = nullptr;
  bool owned_ = true;
};

class CLinearSystemGPU_Field : public Fem::Ls::CLinearSystem_Field {
 public:
  explicit CLinearSystemGPU_Field(dfm_ctx ctx, int fill_level = 0) : ctx_(ctx), lev_(fill_level) {}
  ~CLinearSystemGPU_Field() {
    if (sys_) dfm_sys_destroy(sys_);
  }
  // After InitializeMarge, the AddLinSys_* element routines and FinalizeMarge on this object: the matrix of
  // id_field's corner segment and its residual into the device system (made at the first call)
  void Upload(unsigned int id_field, const Fem::Field::CFieldWorld& world) {
    MatVec::CMatDia_BlkCrs& a = this->GetMatrix(id_field, Fem::Field::CORNER, world);
    const unsigned int n = a.NBlkMatCol();
    const int len = a.LenBlkCol();
    const unsigned int bs = (unsigned int)(len * len);
    std::vector<double> crs, dia((size_t)n * bs);
    if (!sys_) {
      std::vector<unsigned> rp(n + 1, 0), ci;
      for (unsigned int i = 0; i < n; i++) {
        unsigned int np = 0;
        const unsigned int* ind = a.GetPtrIndPSuP(i, np);
        ci.insert(ci.end(), ind, ind + np);
        rp[i + 1] = rp[i] + np;
      }
      throw_if(dfm_sys_create_pattern(ctx_, (int)n, len, (int)n, rp.data(), ci.data(), &sys_));
      throw_if(dfm_prec_set_fill_level(sys_, lev_));
      len_ = len;
    }
    for (unsigned int i = 0; i < n; i++) {
      unsigned int np = 0;
      const double* val = a.GetPtrValPSuP(i, np);
      crs.insert(crs.end(), val, val + (size_t)np * bs);
      const double* d = a.GetPtrValDia(i);
      for (unsigned int t = 0; t < bs; t++) dia[(size_t)i * bs + t] = d[t];
    }
    throw_if(dfm_sys_set_matrix(sys_, crs.data(), dia.data()));
    MatVec::CVector_Blk& r = this->GetResidual(id_field, Fem::Field::CORNER, world);
    std::vector<double> rv((size_t)n * len);
    for (unsigned int i = 0; i < n; i++)
      for (int d = 0; d < len; d++) rv[(size_t)i * len + d] = r.GetValue(i, d);
    throw_if(dfm_sys_set_vector(sys_, -1, rv.data()));
  }
  // factor (CPreconditioner_ILU::SetValue) and the device-resident PCG
  bool SolvePCG(double& conv_ratio, unsigned int& iteration, int precision = DFM_FP64) {
    throw_if(dfm_prec_factor(sys_));
    const int rc = dfm_solve_pcg(sys_, precision, &conv_ratio, &iteration);
    throw_if(rc);
    return rc == 1;
  }
  // the device solution into this system's update vector (then UpdateValueOfField as usual)
  void DownloadUpdate(unsigned int id_field, const Fem::Field::CFieldWorld& world) {
    MatVec::CVector_Blk& u = this->GetUpdate(id_field, Fem::Field::CORNER, world);
    std::vector<double> x((size_t)u.NBlk() * len_);
    throw_if(dfm_sys_get_vector(sys_, -2, x.data()));
    for (unsigned int i = 0; i < u.NBlk(); i++)
      for (int d = 0; d < len_; d++) u.SetValue(i, d, x[(size_t)i * len_ + d]);
  }
  dfm_sys handle() const { return sys_; }

 private:
  dfm_ctx ctx_;
  int lev_;
  int len_ = 1;
  dfm_sys sys_ = nullptr;
};

class CLinearSystemGPU_MultiSeg {
 public:
  explicit CLinearSystemGPU_MultiSeg(dfm_ctx ctx) : ctx_(ctx) {}
  ~CLinearSystemGPU_MultiSeg() {
    if (ms_) dfm_msys_destroy(ms_);
  }
  CLinearSystemGPU_MultiSeg(const CLinearSystemGPU_MultiSeg&) = delete;
  CLinearSystemGPU_MultiSeg& operator=(const CLinearSystemGPU_MultiSeg&) = delete;
  // After InitializeMarge, the element routines and FinalizeMarge on ls: its matrices (patterns at the first call),
  // Dirichlet flags and residual into the device system
  void Upload(LsSol::CLinearSystem& ls) {
    const unsigned int ns = ls.GetNLinSysSeg();
    if (!ms_) {
      std::vector<int> nblk(ns), len(ns);
      for (unsigned int i = 0; i < ns; i++) {
        nblk[i] = (int)ls.GetVector(-1, i).NBlk();
        len[i] = ls.GetVector(-1, i).Len();
      }
      throw_if(dfm_msys_create(ctx_, (int)ns, nblk.data(), len.data(), &ms_));
      for (unsigned int i = 0; i < ns; i++)
        for (unsigned int j = 0; j < ns; j++) {
          if (!ls.IsMatrix(i, j)) continue;
          const MatVec::CMat_BlkCrs& a = matrix(ls, i, j);
          std::vector<unsigned> rp((size_t)nblk[i] + 1, 0), ci;
          for (int r = 0; r < nblk[i]; r++) {
            unsigned int np = 0;
            const unsigned int* ind = a.GetPtrIndPSuP(r, np);
            ci.insert(ci.end(), ind, ind + np);
            rp[r + 1] = rp[r] + np;
          }
          ci.push_back(0);  // a valid pointer also for an empty pattern
          throw_if(dfm_msys_add_pattern(ms_, (int)i, (int)j, rp.data(), ci.data()));
        }
      len_ = len;
    }
    std::vector<double> res, vals, dia;
    std::vector<int> flags;
    for (unsigned int i = 0; i < ns; i++) {
      const MatVec::CVector_Blk& r = ls.GetVector(-1, i);
      const MatVec::CBCFlag& bc = ls.GetBCFlag(i);
      for (unsigned int b = 0; b < r.NBlk(); b++)
        for (int d = 0; d < len_[i]; d++) {
          res.push_back(r.GetValue(b, d));
          flags.push_back(bc.GetBCFlag(b, d));
        }
      for (unsigned int j = 0; j < ns; j++) {
        if (!ls.IsMatrix(i, j)) continue;
        MatVec::CMat_BlkCrs& a = matrix(ls, i, j);
        const size_t bs = (size_t)len_[i] * len_[j];
        vals.clear();
        dia.clear();
        for (unsigned int b = 0; b < r.NBlk(); b++) {
          unsigned int np = 0;
          const double* v = a.GetPtrValPSuP(b, np);
          vals.insert(vals.end(), v, v + np * bs);
          if (i == j) {
            const double* d = ls.GetMatrix(i).GetPtrValDia(b);
            dia.insert(dia.end(), d, d + bs);
          }
        }
        vals.push_back(0.0);
        throw_if(dfm_msys_set_matrix(ms_, (int)i, (int)j, vals.data(), i == j ? dia.data() : nullptr));
      }
    }
    throw_if(dfm_msys_set_vector(ms_, -1, res.data()));
    // FinalizeMarge has run on the host: the device treatment of the same flags leaves the values as they are and
    // gives the device system its flags
    throw_if(dfm_msys_finalize(ms_, flags.data(), nullptr));
  }
  // CPreconditioner_ILU::SetValue and LsSol::Solve_PBiCGSTAB / Solve_PCG on the device (same in/out meaning)
  bool SolvePBiCGSTAB(double& conv_ratio, unsigned int& iteration) {
    throw_if(dfm_msys_factor(ms_));
    const int rc = dfm_msys_solve(ms_, DFM_METHOD_PBICGSTAB, &conv_ratio, &iteration);
    throw_if(rc);
    return rc == 1;
  }
  bool SolvePCG(double& conv_ratio, unsigned int& iteration) {
    throw_if(dfm_msys_factor(ms_));
    const int rc = dfm_msys_solve(ms_, DFM_METHOD_PCG, &conv_ratio, &iteration);
    throw_if(rc);
    return rc == 1;
  }
  // the device solution into ls's update vectors (then UpdateValueOfField as usual)
  void DownloadUpdate(LsSol::CLinearSystem& ls) {
    std::vector<double> x;
    size_t m = 0;
    for (unsigned int i = 0; i < ls.GetNLinSysSeg(); i++) m += (size_t)ls.GetVector(-2, i).NBlk() * len_[i];
    x.resize(m);
    throw_if(dfm_msys_get_vector(ms_, -2, x.data()));
    size_t pos = 0;
    for (unsigned int i = 0; i < ls.GetNLinSysSeg(); i++) {
      MatVec::CVector_Blk& u = ls.GetVector(-2, i);
      for (unsigned int b = 0; b < u.NBlk(); b++)
        for (int d = 0; d < len_[i]; d++) u.SetValue(b, d, x[pos++]);
    }
  }
  dfm_msys handle() const { return ms_; }

 private:
  static MatVec::CMat_BlkCrs& matrix(LsSol::CLinearSystem& ls, unsigned int i, unsigned int j) {
    if (i == j) return ls.GetMatrix(i);
    return ls.GetMatrix(i, j);
  }
  dfm_ctx ctx_;
  std::vector<int> len_;
  dfm_msys ms_ = nullptr;
};

}  // namespace delfem
}  // namespace dfm

#endif  // DFM_DELFEM_ADAPTER_HPP
