/* Oracle: plain-C restatement of DelFEM's FEM hot path (test infrastructure only).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this. It is the checker the
 * HIP product is compared against, never part of the product. Each function follows the reference file
 * and lines cited beside it, in the same arithmetic order, so that on the same inputs it reproduces the
 * reference bit for bit (pinned against the reference itself, compiled from its sources into
 * oracle/_ref/, by tests/test_oracle_vs_reference.py).
 *
 * Storage is the reference's block-CRS (CMatDia_BlkCrs, include/delfem/matvec/mat_blkcrs.h:183-192 and
 * matdia_blkcrs.h:119): rowptr[n+1] (the reference's m_colInd_Blk), colind[ncrs] (its m_rowPtr_Blk),
 * crs[ncrs*b*b] off-diagonal blocks row-major, dia[n*b*b] diagonal blocks, the diagonal kept out of CRS.
 */
#ifndef DFM_ORACLE_CPU_REF_H
#define DFM_ORACLE_CPU_REF_H

#ifdef __cplusplus
extern "C" {
#endif

/* element kinds, shared with the product's C-ABI */
enum { CR_POISSON = 0, CR_SOLID2D = 1, CR_SOLID3D = 2, CR_STVK2D = 3, CR_STVK3D = 4, CR_DIFFUSION2D = 5 };

/* CElemAry::MakePattern_FEM: rowptr[n+1]; returns ncrs. */
long cr_pattern_count(int nnode, int nelem, int npoel, const int* lnods, unsigned* rowptr);
void cr_pattern_fill(int nnode, int nelem, int npoel, const int* lnods, const unsigned* rowptr,
                     unsigned* colind);

/* InitializeMarge + AddLinSys_* element loop + Mearge (values accumulate into crs/dia/res, which the
 * caller zeroes first). params: per element, npar doubles, stride pstride (0 = one set for all). */
int cr_assemble(int kind, int ndim, int nnode, const double* coords, int nelem, const int* lnods,
                const double* params, int npar, int pstride, const double* glob, const double* u,
                const double* v, int len, const unsigned* rowptr, const unsigned* colind, double* crs,
                double* dia, double* res);

/* res[node*len + d] += loads[k*len + d] for k in order (consistent traction loads; reference extension) */
void cr_add_nodal_loads(int nload, const int* node, const double* loads, int len, double* res);

/* P1 consistent surface loads: face triangles (3 nodes) with a traction vector each -> nodal loads,
 * t*A/3 per node, accumulated into res in face order. 3D only. */
void cr_traction_loads(int nface, const int* faces, const double* trac, const double* coords, double* res);

/* CMatDia_BlkCrs::SetBoundaryCondition + CBCFlag::SetZeroToBCDof; returns ||res|| (FinalizeMarge) */
double cr_set_bc(int n, int len, const unsigned* rowptr, const unsigned* colind, double* crs, double* dia,
                 const int* flags, double* res);

/* CMatDia_BlkCrs::MatVec: y = beta*y + alpha*A*x */
void cr_matvec(int n, int len, const unsigned* rowptr, const unsigned* colind, const double* crs,
               const double* dia, double alpha, const double* x, double beta, double* y);

/* ILU(0) symbolic (CMatDiaFrac_BlkCrs(0,mat[,order])): the factor pattern (permuted and sorted when
 * order != NULL; order[new] = old) and m_DiaInd. */
void cr_ilu0_pattern(int n, const unsigned* rowptr, const unsigned* colind, const int* order,
                     unsigned* frowptr, unsigned* fcolind, unsigned* diaind);
/* SetValue_Initialize: copy A into the factor pattern (permuted when order != NULL) */
/* Level-of-fill pattern (ILU(k) symbolic) of an ordered pattern: src/matvec/matdiafrac_blkcrs.cpp:1254-1381
   (AddFracPtn) with CListRLN (include/delfem/matvec/matdiafrac_blkcrs.h:185-301). lev_fill -1: no limit.
   Returns the entry count, or -(count needed) when cap is too small. flev may be NULL. */
/* COrdering_Blk's orderings of a block pattern (kind 3 MakeOrdering_RCM, 4 MakeOrdering_RCM2, 5 MakeOrdering_AMD as
 * compiled), written into order[n] as the reference's m_pOrder (read by NewToOld). Returns 0 where the reference
 * asserts (graph in several pieces), 1 otherwise. */
int cr_ordering(int kind, int n, const unsigned* rowptr, const unsigned* colind, int* order);
long cr_iluk_pattern(int n, const unsigned* rowptr, const unsigned* colind, int lev_fill, long cap, unsigned* frowptr,
                     unsigned* fcolind, unsigned* diaind, int* flev);
void cr_ilu0_setvalue(int n, int len, const unsigned* rowptr, const unsigned* colind, const double* crs,
                      const double* dia, const int* order, const unsigned* frowptr, const unsigned* fcolind,
                      double* fcrs, double* fdia);
/* DoILUDecomp: returns 1 on success, 0 when more than 10 singular pivots were met */
int cr_ilu0_factor(int n, int len, const unsigned* frowptr, const unsigned* fcolind, const unsigned* diaind,
                   double* fcrs, double* fdia);
/* CPreconditioner_ILU::SolvePrecond: ForwardSubstitution + BackwardSubstitution, in place, with the
 * ordering's permutation in and out (work: n*len doubles, used only with an ordering) */
void cr_ilu_solve(int n, int len, const unsigned* frowptr, const unsigned* fcolind, const unsigned* diaind,
                  const double* fcrs, const double* fdia, const int* order, double* vec, double* work);

/* A system with its ILU(0) preconditioner, for the Krylov loops. */
typedef struct {
  int n, len;
  const unsigned *rowptr, *colind;
  const double *crs, *dia;
  const unsigned *frowptr, *fcolind, *diaind;
  const double *fcrs, *fdia;
  const int* order;
} cr_system;

/* LsSol::Solve_PCG / Solve_PBiCGSTAB on handles -1 (res, overwritten) and -2 (upd).
 * conv_ratio and iter are in/out as in the reference. Returns the reference's bool. */
int cr_solve_pcg(const cr_system* s, double* res, double* upd, double* conv_ratio, unsigned* iter);
int cr_solve_pbicgstab(const cr_system* s, double* res, double* upd, double* conv_ratio, unsigned* iter);
/* unpreconditioned Solve_CG / Solve_BiCGSTAB (the factor fields of s are not used) */
int cr_solve_cg(const cr_system* s, double* res, double* upd, double* conv_ratio, unsigned* iter);
/* The matrix-level twins MatVec::Sol::Solve_CG / _PCG / _BiCGSTAB / _PBiCGSTAB (src/matvec/solver_mat_iter.cpp). */
int cr_solve_mat_cg(const cr_system* s, double* res, double* upd, double* conv_ratio, unsigned* iter);
int cr_solve_mat_pcg(const cr_system* s, double* res, double* upd, double* conv_ratio, unsigned* iter);
int cr_solve_mat_bicgstab(const cr_system* s, double* res, double* upd, double* conv_ratio, unsigned* iter);
int cr_solve_mat_pbicgstab(const cr_system* s, double* res, double* upd, double* conv_ratio, unsigned* iter);
int cr_solve_bicgstab(const cr_system* s, double* res, double* upd, double* conv_ratio, unsigned* iter);

/* Saved stiffness (CLinearSystem_Save, src/femls/linearsystem_fieldsave.cpp): the boundary matrix pattern
 * (CMat_BlkCrs::SetPatternBoundary; returns its block count, fills browptr/bcolind when non-NULL), the
 * *_SaveStiffMat element loops (matrix, boundary matrix, external force; kinds CR_POISSON on triangles, CR_SOLID2D,
 * CR_SOLID3D), the boundary matrix's Dirichlet treatment of FinalizeMarge, MakeResidual (returns ||res||) and
 * UpdateValueOfField. */
long cr_boundary_pattern(int n, int len, const unsigned* rowptr, const unsigned* colind, const int* flags,
                         unsigned* browptr, unsigned* bcolind);
int cr_assemble_save(int kind, int ndim, int nnode, const double* coords, int nelem, const int* lnods,
                     const double* params, int pstride, int len, const unsigned* rowptr, const unsigned* colind,
                     double* crs, double* dia, const unsigned* browptr, const unsigned* bcolind, double* bval,
                     double* force);
void cr_boundary_bc(int n, int len, const unsigned* browptr, const unsigned* bcolind, double* bval, const int* flags);
double cr_make_residual_save(int n, int len, const unsigned* browptr, const unsigned* bcolind, const double* bval,
                             const double* force, const double* u, double* res, double* upd);
void cr_update_field_save(int n, int len, const int* flags, double* upd, double* u);

/* Saved stiffness with a lumped mass (src/femls/linearsystem_fieldsave.cpp:363-870): CLinearSystem_SaveDiaM_Newmark
 * (kind CR_DIFFUSION2D, the element loop AddLinearSystem_Diffusion2D_P1 of src/femeqn/eqn_diffusion.cpp:565-650)
 * and CLinearSystem_SaveDiaM_NewmarkBeta (kind CR_SOLID2D, AddLinSys_LinearSolid2D_NonStatic_Save_NewmarkBeta_P1 of
 * src/femeqn/eqn_linear_solid2d.cpp:1051-1180). Boundary matrix pattern (CMat_BlkCrs::SetPatternDia: the whole
 * row with its diagonal; returns the block count, fills browptr/bcolind when non-NULL); the element loop (matrix,
 * boundary matrix, force; glob = (dt, gamma, beta)); the boundary matrix's fixed rows zeroed (FinalizeMarge);
 * MakeResidual (u, v, a: VALUE, VELOCITY, ACCELERATION; returns ||res||); UpdateValueOfField (VELOCITY for
 * CR_DIFFUSION2D, ACCELERATION for CR_SOLID2D). */
long cr_pattern_dia(int n, const unsigned* rowptr, const unsigned* colind, unsigned* browptr, unsigned* bcolind);
int cr_assemble_savedia(int kind, int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                        int pstride, const double* glob, int len, const unsigned* rowptr, const unsigned* colind,
                        double* crs, double* dia, const unsigned* browptr, const unsigned* bcolind, double* bval,
                        double* force);
void cr_boundary_rows(int n, int len, const unsigned* browptr, double* bval, const int* flags);
double cr_make_residual_savedia(int kind, int n, int len, const unsigned* browptr, const unsigned* bcolind,
                                const double* bval, const double* force, const double* glob, const int* flags,
                                const double* u, const double* v, const double* a, double* res, double* upd);
void cr_update_field_savedia(int kind, int n, int len, const double* glob, const int* flags, double* upd, double* u,
                             double* v, double* a);

/* Newmark-beta dynamic linear elasticity (2D and 3D P1): element loops into matrix and residual (u, v, a the
 * VALUE, VELOCITY and ACCELERATION fields; glob = (dt, gamma, beta, is_initial)), and UpdateValueOfField_NewmarkBeta
 * over m values. */
int cr_assemble_newmark(int ndim, int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                        int pstride, const double* glob, const double* u, const double* v, const double* a,
                        const unsigned* rowptr, const unsigned* colind, double* crs, double* dia, double* res);
void cr_update_field_newmark(long m, double gamma, double beta, double dt, int is_initial, const double* upd, double* u,
                             double* v, double* a);

/* Backward-Euler dynamic 2D elasticity: element loop (velo_pre the previous velocity) and
 * UpdateValueOfField_BackwardEular over m values. */
int cr_assemble_beuler(int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                       int pstride, double dt, int is_initial, const double* u, const double* v,
                       const double* velo_pre, const unsigned* rowptr, const unsigned* colind, double* crs,
                       double* dia, double* res);
void cr_update_field_beuler(long m, double dt, int is_initial, const double* upd, double* u, double* v);

/* Newmark-beta dynamic StVK 3D element loop; glob = (dt, gamma, beta, is_initial). */
int cr_assemble_stvk3d_newmark(int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                               int pstride, const double* glob, const double* u, const double* v, const double* a,
                               const unsigned* rowptr, const unsigned* colind, double* crs, double* dia, double* res);
int cr_assemble_stvk2d_newmark(int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                               int pstride, const double* glob, const double* u, const double* v, const double* a,
                               const unsigned* rowptr, const unsigned* colind, double* crs, double* dia, double* res);
int cr_assemble_stvk2d_beuler(int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                              int pstride, double dt, int is_initial, const double* u, const double* v,
                              const double* velo_pre, const unsigned* rowptr, const unsigned* colind, double* crs,
                              double* dia, double* res);

/* Multi-segment systems (LsSol::CLinearSystem with several segments, src/ls/linearsystem.cpp; CPreconditioner_ILU
 * over them, src/ls/preconditioner.cpp; the couplings' factors CMatFrac_BlkCrs, src/matvec/matfrac_blkcrs.cpp).
 * A block matrix of the system: nrow x ncol blocks of bi x bj, CRS rows (a diagonal segment: square, the
 * diagonal kept apart in dia, diaind = first upper entry of each row of its factor). */
typedef struct {
  int nrow, ncol, bi, bj;
  const unsigned *rowptr, *colind;
  double* val;
  double* dia;
  const unsigned* diaind;
} cr_blk;

#define CR_MAXSEG 8
typedef struct {
  int nseg;
  int nblk[CR_MAXSEG], len[CR_MAXSEG];
  long off[CR_MAXSEG + 1];           /* dof offset of each segment in a concatenated vector */
  cr_blk* a[CR_MAXSEG * CR_MAXSEG];  /* a[i*nseg + j]: matrix (i, j), NULL where absent; a[i*nseg+i] diagonal */
  cr_blk* f[CR_MAXSEG * CR_MAXSEG];  /* their ILU(0) factors (same patterns) */
} cr_msys;

/* CMat_BlkCrs::MatVec (src/matvec/mat_blkcrs.cpp:955-1175), not transposed */
void cr_blk_matvec(const cr_blk* A, double alpha, const double* x, double beta, double* y);
/* FinalizeMarge over the segments; flags/res concatenated. Returns the residual norm. */
double cr_ms_set_bc(cr_msys* s, const int* flags, double* res);
/* CLinearSystem::MATVEC over the segments */
void cr_ms_matvec(const cr_msys* s, double alpha, const double* x, double beta, double* y);
/* CPreconditioner_ILU::SetValue (ILU(0)): f from a, factored; 1 ok, 0 failed */
int cr_ms_ilu0(cr_msys* s);
/* CPreconditioner_ILU::SolvePrecond, in place */
void cr_ms_precond(const cr_msys* s, double* v);
/* Solve_PCG (method 0) / Solve_PBiCGSTAB (1) over the segments with the ILU(0) of f */
int cr_ms_solve(const cr_msys* s, int method, double* res, double* upd, double* conv_ratio, unsigned* iter);

#ifdef __cplusplus
}
#endif
#endif
