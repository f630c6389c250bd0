// Multi-segment linear systems: LsSol::CLinearSystem with several segments (src/ls/linearsystem.cpp:74-415) and
// CPreconditioner_ILU over them (src/ls/preconditioner.cpp:57-335), ILU(0), on the device.
//
// Each diagonal segment is a system made from its block pattern (a Sys in the natural numbering: the reference
// applies no ordering to multi-segment systems), which brings its matrix, vectors, level-scheduled ILU(0) factor and
// sweeps. The couplings (segment i rows x segment j columns, bi x bj blocks) are block-CRS matrices with their own
// factor copies (CMatFrac_BlkCrs). A vector handle of the system is the same handle of every segment.
//
// The factorisation follows CPreconditioner_ILU::SetValue's order of segment pairs; each step is one launch of a
// row-parallel kernel (k_msys.hip), except the diagonal segments' own factorisation (the lane-parallel ILU kernel of
// the single-segment path, in place after the coupling updates) and the upper couplings, which walk the diagonal
// segment's lower dependency levels in one workgroup. The Krylov loops are the reference's, host-driven over the
// handle operations (include/dfm_lssol.hpp).
#include <algorithm>
#include <cmath>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "dfm_gpu.h"
#include "dfm_internal.hpp"
#include "dfm_lssol.hpp"
#include "msys_kernels.hpp"

namespace dfm {

struct MBlk {
  int nrow = 0, ncol = 0, bi = 0, bj = 0;
  std::vector<unsigned> rowptr, colind;
  DBuf<unsigned> d_rowptr, d_colind;
  DBuf<double> d_val, d_fval;  // matrix values, ILU(0) factor values (same pattern)
  long long nnz() const { return rowptr.empty() ? 0 : (long long)rowptr.back(); }
  MsBlkDev dev() const { return MsBlkDev{nrow, bi, bj, d_rowptr.p, d_colind.p}; }
  ~MBlk() {
    d_rowptr.release(); d_colind.release(); d_val.release(); d_fval.release();
  }
};

struct MSys {
  dfm_ctx ctx = nullptr;
  int nseg = 0;
  std::vector<int> nblk, len;
  std::vector<long long> off;        // dof offsets of the segments in a concatenated vector
  std::vector<dfm_sys> seg;          // diagonal segments (nullptr until their pattern is added)
  std::vector<std::unique_ptr<MBlk>> od;  // [nseg*nseg] couplings, null where absent (and on the diagonal)
  std::vector<DBuf<int>> lptr, lrows;     // per segment: rows by dependency level of the lower factor
  std::vector<int> nlev;                  // per segment, -1 until built
  bool finalized = false, factored = false;
  ~MSys() {
    for (dfm_sys s : seg)
      if (s) dfm_sys_destroy(s);
    for (auto& b : lptr) b.release();
    for (auto& b : lrows) b.release();
  }
  MBlk* at(int i, int j) const { return od[(size_t)i * nseg + j].get(); }
};

namespace {

MsDiaDev dia_dev(Sys& s) {
  return MsDiaDev{s.n, s.len, s.d_frowptr.p, s.d_fcolind.p, s.d_fdiaind.p, s.d_fcrs.p, s.d_fdia.p};
}

// Rows of a diagonal segment grouped by dependency level of its lower factor (natural order): row r comes one level
// after the latest row it references below the diagonal.
int build_levels(MSys& m, int i) {
  if (m.nlev[i] >= 0) return DFM_OK;
  Sys& s = m.seg[i]->s;
  std::vector<int> lev(s.n, 0);
  int nl = 0;
  for (int r = 0; r < s.n; r++) {
    int l = 0;
    for (unsigned q = s.frowptr[r]; q < s.fdiaind[r]; q++) l = std::max(l, lev[s.fcolind[q]] + 1);
    lev[r] = l;
    nl = std::max(nl, l + 1);
  }
  std::vector<int> ptr(nl + 1, 0), rows(s.n);
  for (int r = 0; r < s.n; r++) ptr[lev[r] + 1]++;
  for (int l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int r = 0; r < s.n; r++) rows[fill[lev[r]]++] = r;
  hipStream_t st = s.ctx->stream;
  DFM_HIP(m.lptr[i].alloc(ptr.size()));
  DFM_HIP(m.lrows[i].alloc(std::max<size_t>(rows.size(), 1)));
  DFM_HIP(hipMemcpyAsync(m.lptr[i].p, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice, st));
  if (!rows.empty())
    DFM_HIP(hipMemcpyAsync(m.lrows[i].p, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice, st));
  DFM_HIP(hipStreamSynchronize(st));
  m.nlev[i] = nl;
  return DFM_OK;
}

int need_segments(const MSys& m, const char* who) {
  for (int i = 0; i < m.nseg; i++)
    DFM_CHECK(m.seg[i] != nullptr, DFM_ERR_STATE,
              std::string(who) + ": segment " + std::to_string(i) + " has no diagonal matrix (add its pattern first)");
  return DFM_OK;
}

}  // namespace

// The seam of the reference's Krylov loops over a multi-segment system (ILinearSystemPreconditioner_Sol).
class MultiSegLinearSystem : public lssol::ILinearSystemPreconditioner_Sol {
 public:
  explicit MultiSegLinearSystem(dfm_msys ms) : ms_(ms) {}
  unsigned int GetTmpVectorArySize() const override {
    int n = 0;
    check(dfm_msys_ntmp(ms_, &n));
    return (unsigned int)n;
  }
  bool ReSizeTmpVecSolver(unsigned int n) override { return dfm_msys_resize_tmp(ms_, (int)n) == DFM_OK; }
  double DOT(int a, int b) override {
    double d = 0.0;
    check(dfm_msys_dot(ms_, a, b, &d));
    return d;
  }
  bool COPY(int a, int b) override { return dfm_msys_copy(ms_, a, b) == DFM_OK; }
  bool SCAL(double s, int a) override { return dfm_msys_scal(ms_, s, a) == DFM_OK; }
  bool AXPY(double s, int a, int b) override { return dfm_msys_axpy(ms_, s, a, b) == DFM_OK; }
  bool MATVEC(double a, int x, double b, int y) override { return dfm_msys_matvec(ms_, a, x, b, y) == DFM_OK; }
  bool SolvePrecond(int iv) override { return dfm_msys_precond(ms_, iv) == DFM_OK; }

 private:
  static void check(int rc) {
    if (rc < 0) throw std::runtime_error(dfm_last_error());
  }
  dfm_msys ms_;
};

}  // namespace dfm

struct dfm_msys_s {
  dfm::MSys m;
};

using namespace dfm;

#define NEED(p, name) DFM_CHECK((p) != nullptr, DFM_ERR_ARG, std::string(name) + ": null argument")

extern "C" {

int dfm_msys_create(dfm_ctx ctx, int nseg, const int* nblk, const int* len, dfm_msys* out) {
  NEED(ctx, "dfm_msys_create");
  NEED(nblk, "dfm_msys_create");
  NEED(len, "dfm_msys_create");
  NEED(out, "dfm_msys_create");
  DFM_CHECK(nseg >= 1 && nseg <= DFM_MSYS_MAXSEG, DFM_ERR_ARG, "dfm_msys_create: 1 <= nseg <= DFM_MSYS_MAXSEG");
  for (int i = 0; i < nseg; i++) {
    DFM_CHECK(nblk[i] > 0, DFM_ERR_ARG, "dfm_msys_create: every segment needs at least one block");
    DFM_CHECK(len[i] >= 1 && len[i] <= 3, DFM_ERR_ARG, "dfm_msys_create: block length must be 1, 2 or 3");
  }
  dfm_msys h = new dfm_msys_s;
  MSys& m = h->m;
  m.ctx = ctx;
  m.nseg = nseg;
  m.nblk.assign(nblk, nblk + nseg);
  m.len.assign(len, len + nseg);
  m.off.assign(nseg + 1, 0);
  for (int i = 0; i < nseg; i++) m.off[i + 1] = m.off[i] + (long long)nblk[i] * len[i];
  m.seg.assign(nseg, nullptr);
  m.od.resize((size_t)nseg * nseg);
  m.lptr.resize(nseg);
  m.lrows.resize(nseg);
  m.nlev.assign(nseg, -1);
  ctx->c.refs++;
  *out = h;
  return DFM_OK;
}

int dfm_msys_destroy(dfm_msys ms) {
  if (!ms) return DFM_OK;
  dfm_ctx ctx = ms->m.ctx;
  DFM_HIP(hipStreamSynchronize(ctx->c.stream));
  {
    CacheRelease guard;
    delete ms;
  }
  ctx_release(ctx);
  return DFM_OK;
}

int dfm_msys_add_pattern(dfm_msys ms, int i, int j, const unsigned* rowptr, const unsigned* colind) {
  NEED(ms, "dfm_msys_add_pattern");
  NEED(rowptr, "dfm_msys_add_pattern");
  MSys& m = ms->m;
  DFM_CHECK(i >= 0 && i < m.nseg && j >= 0 && j < m.nseg, DFM_ERR_ARG, "dfm_msys_add_pattern: segment out of range");
  const int nrow = m.nblk[i], ncol = m.nblk[j];
  DFM_CHECK(rowptr[0] == 0, DFM_ERR_ARG, "dfm_msys_add_pattern: rowptr[0] must be 0");
  for (int r = 0; r < nrow; r++) {
    DFM_CHECK(rowptr[r + 1] >= rowptr[r], DFM_ERR_ARG, "dfm_msys_add_pattern: rowptr must not decrease");
    for (unsigned q = rowptr[r]; q < rowptr[r + 1]; q++) {
      DFM_CHECK(colind[q] < (unsigned)ncol, DFM_ERR_ARG, "dfm_msys_add_pattern: column out of range");
      DFM_CHECK(q == rowptr[r] || colind[q] > colind[q - 1], DFM_ERR_ARG,
                "dfm_msys_add_pattern: the columns of a row must ascend");
    }
  }
  const unsigned dummy = 0;
  const unsigned* ci = rowptr[nrow] ? colind : &dummy;
  if (rowptr[nrow]) NEED(colind, "dfm_msys_add_pattern");
  m.finalized = m.factored = false;
  if (i == j) {
    DFM_CHECK(m.seg[i] == nullptr, DFM_ERR_STATE, "dfm_msys_add_pattern: the segment's pattern is already set");
    dfm_sys s = nullptr;
    const int rc = dfm_sys_create_pattern(m.ctx, nrow, m.len[i], nrow, rowptr, ci, &s);
    if (rc != DFM_OK) return rc;
    m.seg[i] = s;
    m.nlev[i] = -1;
    return DFM_OK;
  }
  DFM_CHECK(m.at(i, j) == nullptr, DFM_ERR_STATE, "dfm_msys_add_pattern: the coupling's pattern is already set");
  auto b = std::make_unique<MBlk>();
  b->nrow = nrow; b->ncol = ncol; b->bi = m.len[i]; b->bj = m.len[j];
  b->rowptr.assign(rowptr, rowptr + nrow + 1);
  b->colind.assign(ci, ci + rowptr[nrow]);
  hipStream_t st = m.ctx->c.stream;
  const size_t nv = (size_t)std::max<long long>(b->nnz(), 1) * b->bi * b->bj;
  DFM_HIP(b->d_rowptr.alloc(b->rowptr.size()));
  DFM_HIP(b->d_colind.alloc(std::max<size_t>(b->colind.size(), 1)));
  DFM_HIP(b->d_val.alloc(nv));
  DFM_HIP(b->d_fval.alloc(nv));
  DFM_HIP(hipMemcpyAsync(b->d_rowptr.p, b->rowptr.data(), b->rowptr.size() * sizeof(unsigned),
                         hipMemcpyHostToDevice, st));
  if (!b->colind.empty())
    DFM_HIP(hipMemcpyAsync(b->d_colind.p, b->colind.data(), b->colind.size() * sizeof(unsigned),
                           hipMemcpyHostToDevice, st));
  DFM_HIP(hipMemsetAsync(b->d_val.p, 0, nv * sizeof(double), st));
  DFM_HIP(hipStreamSynchronize(st));
  m.od[(size_t)i * m.nseg + j] = std::move(b);
  return DFM_OK;
}

int dfm_msys_info(dfm_msys ms, int i, int j, long long* nnz) {
  NEED(ms, "dfm_msys_info");
  NEED(nnz, "dfm_msys_info");
  MSys& m = ms->m;
  DFM_CHECK(i >= 0 && i < m.nseg && j >= 0 && j < m.nseg, DFM_ERR_ARG, "dfm_msys_info: segment out of range");
  if (i == j) *nnz = m.seg[i] ? m.seg[i]->s.ncrs : -1;
  else *nnz = m.at(i, j) ? m.at(i, j)->nnz() : -1;
  return DFM_OK;
}

int dfm_msys_set_matrix(dfm_msys ms, int i, int j, const double* val, const double* dia) {
  NEED(ms, "dfm_msys_set_matrix");
  MSys& m = ms->m;
  DFM_CHECK(i >= 0 && i < m.nseg && j >= 0 && j < m.nseg, DFM_ERR_ARG, "dfm_msys_set_matrix: segment out of range");
  m.finalized = m.factored = false;
  if (i == j) {
    DFM_CHECK(m.seg[i], DFM_ERR_STATE, "dfm_msys_set_matrix: no pattern for the segment");
    return dfm_sys_set_matrix(m.seg[i], val, dia);
  }
  MBlk* b = m.at(i, j);
  DFM_CHECK(b, DFM_ERR_STATE, "dfm_msys_set_matrix: no pattern for the coupling");
  if (b->nnz() == 0) return DFM_OK;
  NEED(val, "dfm_msys_set_matrix");
  hipStream_t st = m.ctx->c.stream;
  DFM_HIP(hipMemcpyAsync(b->d_val.p, val, (size_t)b->nnz() * b->bi * b->bj * sizeof(double), hipMemcpyHostToDevice,
                         st));
  DFM_HIP(hipStreamSynchronize(st));
  return DFM_OK;
}

int dfm_msys_get_matrix(dfm_msys ms, int i, int j, double* val, double* dia) {
  NEED(ms, "dfm_msys_get_matrix");
  MSys& m = ms->m;
  DFM_CHECK(i >= 0 && i < m.nseg && j >= 0 && j < m.nseg, DFM_ERR_ARG, "dfm_msys_get_matrix: segment out of range");
  if (i == j) {
    DFM_CHECK(m.seg[i], DFM_ERR_STATE, "dfm_msys_get_matrix: no pattern for the segment");
    return dfm_sys_get_matrix(m.seg[i], val, dia);
  }
  MBlk* b = m.at(i, j);
  DFM_CHECK(b, DFM_ERR_STATE, "dfm_msys_get_matrix: no pattern for the coupling");
  hipStream_t st = m.ctx->c.stream;
  DFM_HIP(hipStreamSynchronize(st));
  if (b->nnz() == 0) return DFM_OK;
  NEED(val, "dfm_msys_get_matrix");
  DFM_HIP(hipMemcpy(val, b->d_val.p, (size_t)b->nnz() * b->bi * b->bj * sizeof(double), hipMemcpyDeviceToHost));
  return DFM_OK;
}

int dfm_msys_get_factor(dfm_msys ms, int i, int j, double* val, double* dia) {
  NEED(ms, "dfm_msys_get_factor");
  MSys& m = ms->m;
  DFM_CHECK(i >= 0 && i < m.nseg && j >= 0 && j < m.nseg, DFM_ERR_ARG, "dfm_msys_get_factor: segment out of range");
  DFM_CHECK(m.factored, DFM_ERR_STATE, "dfm_msys_get_factor: not factored");
  hipStream_t st = m.ctx->c.stream;
  DFM_HIP(hipStreamSynchronize(st));
  if (i == j) {
    Sys& s = m.seg[i]->s;
    if (val && s.ncrs) DFM_HIP(hipMemcpy(val, s.d_fcrs.p, (size_t)s.ncrs * s.bs * sizeof(double), hipMemcpyDeviceToHost));
    if (dia) DFM_HIP(hipMemcpy(dia, s.d_fdia.p, (size_t)s.n * s.bs * sizeof(double), hipMemcpyDeviceToHost));
    return DFM_OK;
  }
  MBlk* b = m.at(i, j);
  DFM_CHECK(b, DFM_ERR_STATE, "dfm_msys_get_factor: no coupling");
  if (b->nnz() && val)
    DFM_HIP(hipMemcpy(val, b->d_fval.p, (size_t)b->nnz() * b->bi * b->bj * sizeof(double), hipMemcpyDeviceToHost));
  return DFM_OK;
}

int dfm_msys_set_vector(dfm_msys ms, int iv, const double* v) {
  NEED(ms, "dfm_msys_set_vector");
  NEED(v, "dfm_msys_set_vector");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_set_vector");
  if (rc != DFM_OK) return rc;
  for (int i = 0; i < m.nseg; i++)
    if ((rc = dfm_sys_set_vector(m.seg[i], iv, v + m.off[i])) != DFM_OK) return rc;
  return DFM_OK;
}

int dfm_msys_get_vector(dfm_msys ms, int iv, double* v) {
  NEED(ms, "dfm_msys_get_vector");
  NEED(v, "dfm_msys_get_vector");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_get_vector");
  if (rc != DFM_OK) return rc;
  for (int i = 0; i < m.nseg; i++)
    if ((rc = dfm_sys_get_vector(m.seg[i], iv, v + m.off[i])) != DFM_OK) return rc;
  return DFM_OK;
}

// FinalizeMarge (src/ls/linearsystem.cpp:92-128)
int dfm_msys_finalize(dfm_msys ms, const int* flags, double* res_norm) {
  NEED(ms, "dfm_msys_finalize");
  NEED(flags, "dfm_msys_finalize");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_finalize");
  if (rc != DFM_OK) return rc;
  for (int i = 0; i < m.nseg; i++) {
    DFM_CHECK(m.seg[i]->s.assembled, DFM_ERR_STATE, "dfm_msys_finalize: set the matrices first");
    if ((rc = dfm_sys_set_bc(m.seg[i], flags + m.off[i])) != DFM_OK) return rc;
    if ((rc = sys_apply_bc(m.seg[i]->s)) != DFM_OK) return rc;
  }
  hipStream_t st = m.ctx->c.stream;
  for (int i = 0; i < m.nseg; i++)
    for (int j = 0; j < m.nseg; j++)
      if (MBlk* b = m.at(i, j))
        DFM_HIP(launch_ms_blk_bc(b->dev(), b->d_val.p, m.seg[i]->s.d_bc.p, m.seg[j]->s.d_bc.p, st));
  double sq = 0.0;
  for (int i = 0; i < m.nseg; i++) {
    double d = 0.0;
    if ((rc = dfm_sys_dot(m.seg[i], -1, -1, &d)) != DFM_OK) return rc;
    sq += d;
  }
  if (res_norm) *res_norm = std::sqrt(sq);
  m.finalized = true;
  m.factored = false;
  return DFM_OK;
}

// CPreconditioner_ILU::SetLinearSystem + SetValue with no fill level (src/ls/preconditioner.cpp:57-297)
int dfm_msys_factor(dfm_msys ms) {
  NEED(ms, "dfm_msys_factor");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_factor");
  if (rc != DFM_OK) return rc;
  for (int i = 0; i < m.nseg; i++)
    DFM_CHECK(m.seg[i]->s.assembled, DFM_ERR_STATE, "dfm_msys_factor: set the matrices first");
  DFM_CHECK(m.seg[0]->s.fill_level == 0, DFM_ERR_STATE, "dfm_msys_factor: ILU(0) only");
  hipStream_t st = m.ctx->c.stream;
  // SetValue_Initialize of every factor
  for (int i = 0; i < m.nseg; i++) {
    if ((rc = sys_factor_prefill(m.seg[i]->s)) != DFM_OK) return rc;
    for (int j = 0; j < m.nseg; j++)
      if (MBlk* b = m.at(i, j))
        if (b->nnz()) DFM_HIP(hipMemcpyAsync(b->d_fval.p, b->d_val.p, (size_t)b->nnz() * b->bi * b->bj * sizeof(double),
                                             hipMemcpyDeviceToDevice, st));
  }
  m.factored = false;
  for (int i = 0; i < m.nseg; i++) {
    for (int j = 0; j < m.nseg; j++) {
      if (i == j) {
        Sys& s = m.seg[i]->s;
        const MsDiaDev d = dia_dev(s);
        for (int k = 0; k < i; k++) {
          MBlk *low = m.at(i, k), *up = m.at(k, i);
          if (low && up) DFM_HIP(launch_ms_dia_lowup(d, low->dev(), low->d_fval.p, up->dev(), up->d_fval.p, st));
        }
        rc = sys_factor_run(s, true);
        if (rc == DFM_ERR_SINGULAR) set_error("dfm_msys_factor: more than 10 singular pivots in segment " +
                                              std::to_string(i) + " (DoILUDecomp returns false)");
        if (rc != DFM_OK) return rc;
        continue;
      }
      MBlk* f = m.at(i, j);
      if (!f) continue;
      const int kmax = std::min(i, j);
      for (int k = 0; k <= kmax; k++) {
        if (k == i && k < j) {
          if ((rc = build_levels(m, i)) != DFM_OK) return rc;
          DFM_HIP(launch_ms_od_up(m.nlev[i], m.lptr[i].p, m.lrows[i].p, f->dev(), f->d_fval.p, dia_dev(m.seg[i]->s),
                                  st));
          continue;
        }
        if (k == j && k < i) {
          DFM_HIP(launch_ms_od_low(f->dev(), f->d_fval.p, dia_dev(m.seg[j]->s), st));
          continue;
        }
        MBlk *low = m.at(i, k), *up = m.at(k, j);
        if (!low || !up) continue;
        DFM_HIP(launch_ms_od_lowup(f->dev(), f->d_fval.p, low->dev(), low->d_fval.p, up->dev(), up->d_fval.p, st));
      }
    }
  }
  DFM_HIP(hipStreamSynchronize(st));
  m.factored = true;
  return DFM_OK;
}

// CPreconditioner_ILU::SolvePrecond (src/ls/preconditioner.cpp:301-335)
int dfm_msys_precond(dfm_msys ms, int iv) {
  NEED(ms, "dfm_msys_precond");
  MSys& m = ms->m;
  DFM_CHECK(m.factored, DFM_ERR_STATE, "dfm_msys_precond: not factored");
  hipStream_t st = m.ctx->c.stream;
  std::vector<double*> v(m.nseg);
  for (int i = 0; i < m.nseg; i++) {
    v[i] = sys_vec_ptr(m.seg[i]->s, iv);
    DFM_CHECK(v[i], DFM_ERR_ARG, "dfm_msys_precond: bad vector handle");
  }
  for (int i = 0; i < m.nseg; i++) {
    for (int j = 0; j < i; j++)
      if (MBlk* f = m.at(i, j)) DFM_HIP(launch_ms_blk_matvec(f->dev(), -1.0, f->d_fval.p, v[j], 1.0, v[i], st));
    DFM_HIP(precond_apply64_dirs(m.seg[i]->s, v[i], 1, st));
  }
  for (int i = m.nseg - 1; i >= 0; i--) {
    for (int j = i + 1; j < m.nseg; j++)
      if (MBlk* f = m.at(i, j)) DFM_HIP(launch_ms_blk_matvec(f->dev(), -1.0, f->d_fval.p, v[j], 1.0, v[i], st));
    DFM_HIP(precond_apply64_dirs(m.seg[i]->s, v[i], 2, st));
  }
  return DFM_OK;
}

int dfm_msys_resize_tmp(dfm_msys ms, int ntmp) {
  NEED(ms, "dfm_msys_resize_tmp");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_resize_tmp");
  if (rc != DFM_OK) return rc;
  for (int i = 0; i < m.nseg; i++)
    if ((rc = dfm_sys_resize_tmp(m.seg[i], ntmp)) != DFM_OK) return rc;
  return DFM_OK;
}

int dfm_msys_ntmp(dfm_msys ms, int* ntmp) {
  NEED(ms, "dfm_msys_ntmp");
  NEED(ntmp, "dfm_msys_ntmp");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_ntmp");
  if (rc != DFM_OK) return rc;
  return dfm_sys_ntmp(m.seg[0], ntmp);
}

// CLinearSystem::DOT (src/ls/linearsystem.cpp:335-360): the segments' products summed in segment order
int dfm_msys_dot(dfm_msys ms, int iv1, int iv2, double* out) {
  NEED(ms, "dfm_msys_dot");
  NEED(out, "dfm_msys_dot");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_dot");
  if (rc != DFM_OK) return rc;
  double sum = 0.0;
  for (int i = 0; i < m.nseg; i++) {
    double d = 0.0;
    if ((rc = dfm_sys_dot(m.seg[i], iv1, iv2, &d)) != DFM_OK) return rc;
    sum += d;
  }
  *out = sum;
  return DFM_OK;
}

int dfm_msys_copy(dfm_msys ms, int iv1, int iv2) {
  NEED(ms, "dfm_msys_copy");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_copy");
  for (int i = 0; i < m.nseg && rc == DFM_OK; i++) rc = dfm_sys_copy(m.seg[i], iv1, iv2);
  return rc;
}

int dfm_msys_scal(dfm_msys ms, double alpha, int iv) {
  NEED(ms, "dfm_msys_scal");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_scal");
  for (int i = 0; i < m.nseg && rc == DFM_OK; i++) rc = dfm_sys_scal(m.seg[i], alpha, iv);
  return rc;
}

int dfm_msys_axpy(dfm_msys ms, double alpha, int iv1, int iv2) {
  NEED(ms, "dfm_msys_axpy");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_axpy");
  for (int i = 0; i < m.nseg && rc == DFM_OK; i++) rc = dfm_sys_axpy(m.seg[i], alpha, iv1, iv2);
  return rc;
}

// CLinearSystem::MATVEC (src/ls/linearsystem.cpp:250-291)
int dfm_msys_matvec(dfm_msys ms, double alpha, int iv1, double beta, int iv2) {
  NEED(ms, "dfm_msys_matvec");
  MSys& m = ms->m;
  int rc = need_segments(m, "dfm_msys_matvec");
  if (rc != DFM_OK) return rc;
  if (alpha == 0.0) return dfm_msys_scal(ms, beta, iv2);
  DFM_CHECK(iv1 != iv2, DFM_ERR_ARG, "dfm_msys_matvec: iv1 == iv2 is not supported (src/ls/linearsystem.cpp:275-278)");
  hipStream_t st = m.ctx->c.stream;
  for (int i = 0; i < m.nseg; i++) {
    if ((rc = dfm_sys_matvec(m.seg[i], alpha, iv1, beta, iv2)) != DFM_OK) return rc;
    double* y = sys_vec_ptr(m.seg[i]->s, iv2);
    for (int j = 0; j < m.nseg; j++)
      if (MBlk* b = m.at(i, j)) {
        const double* x = sys_vec_ptr(m.seg[j]->s, iv1);
        DFM_HIP(launch_ms_blk_matvec(b->dev(), alpha, b->d_val.p, x, 1.0, y, st));
      }
  }
  return DFM_OK;
}

// LsSol::Solve_PCG / Solve_PBiCGSTAB (src/ls/solver_ls_iter.cpp:100-172, 291-408) over the segments
int dfm_msys_solve(dfm_msys ms, int method, double* conv_ratio, unsigned* iteration) {
  NEED(ms, "dfm_msys_solve");
  NEED(conv_ratio, "dfm_msys_solve");
  NEED(iteration, "dfm_msys_solve");
  DFM_CHECK(method == DFM_METHOD_PCG || method == DFM_METHOD_PBICGSTAB, DFM_ERR_ARG, "dfm_msys_solve: unknown method");
  DFM_CHECK(ms->m.factored, DFM_ERR_STATE, "dfm_msys_solve: factor first");
  try {
    MultiSegLinearSystem ls(ms);
    const bool ok = method == DFM_METHOD_PCG ? lssol::Solve_PCG(*conv_ratio, *iteration, ls)
                                             : lssol::Solve_PBiCGSTAB(*conv_ratio, *iteration, ls);
    return ok ? 1 : 0;
  } catch (const std::exception& e) {
    set_error(std::string("dfm_msys_solve: ") + e.what());
    return DFM_ERR_HIP;
  }
}

}  // extern "C"
