/* dfm_gpu.h — C-ABI of the MI355X-native DelFEM FEM hot path (libdfm_gpu.so).
 *
 * Plain pointers and sizes only. Every entry point returns an int status (DFM_OK = 0, negative on error;
 * dfm_last_error() gives the message). Device state lives in opaque handles; host arrays passed in are
 * copied, host arrays passed out are filled before the call returns.
 *
 * Numbering. Vectors and matrices are exchanged in the caller's node numbering ("original"). Internally a
 * system may run in a permuted "device" numbering chosen by dfm_sys_set_ordering (the ILU ordering of
 * LsSol::CPreconditioner_ILU::SetOrdering, include/delfem/ls/preconditioner.h:93-97 in the reference).
 * Handle operations (dot/copy/scal/axpy/matvec/precond) are permutation-invariant, so their meaning is
 * exactly the reference's.
 *
 * Vector handles follow the reference (src/ls/linearsystem.cpp:256-265): -1 = residual (right-hand side,
 * overwritten by the solvers), -2 = update (solution), 0.. = solver temporaries.
 */
#ifndef DFM_GPU_H
#define DFM_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFM_OK 0
#define DFM_ERR_ARG -1
#define DFM_ERR_HIP -2
#define DFM_ERR_STATE -3
#define DFM_ERR_NOMEM -4
#define DFM_ERR_SINGULAR -5

typedef struct dfm_ctx_s* dfm_ctx;
typedef struct dfm_sys_s* dfm_sys;

/* element kinds (Fem::Eqn::AddLinSys_* on P1 TRI/TET) and their per-element parameter tuples */
enum {
  DFM_POISSON = 0,     /* (alpha, source)                  src/femeqn/eqn_poisson.cpp:49-125, :249-323 */
  DFM_SOLID2D = 1,     /* (lambda, myu, rho, gx, gy)       src/femeqn/eqn_linear_solid2d.cpp:505-598   */
  DFM_SOLID3D = 2,     /* (lambda, myu, rho, gx, gy, gz)   src/femeqn/eqn_linear_solid3d.cpp:241-334   */
  DFM_STVK2D = 3,      /* (lambda, myu, rho, gx, gy)       src/femeqn/eqn_st_venant.cpp:47-272         */
  DFM_STVK3D = 4,      /* (lambda, myu, rho, gx, gy, gz)   src/femeqn/eqn_st_venant.cpp:724-939        */
  DFM_DIFFUSION2D = 5, /* (rho, alpha, source), glob=(dt, gamma)  src/femeqn/eqn_diffusion.cpp:240-345 */
  /* Newmark-beta dynamic linear elasticity (unknown: the acceleration increment; matrix M + beta dt^2 K):
   * AddLinSys_LinearSolid2D_NonStatic_NewmarkBeta_P1, src/femeqn/eqn_linear_solid2d.cpp:1182-1337,
   * (lambda, myu, rho, gx, gy), glob = (dt, gamma, beta, is_initial) */
  DFM_SOLID2D_NEWMARK = 6,
  /* AddLinSys_LinearSolid3D_NonStatic_NewmarkBeta_P1, src/femeqn/eqn_linear_solid3d.cpp:525-690 (its predictor
   * terms always on), (lambda, myu, rho, gx, gy, gz), glob = (dt, gamma, beta) */
  DFM_SOLID3D_NEWMARK = 7,
  /* backward-Euler dynamic 2D elasticity (unknown: the velocity increment; matrix M + dt^2 K):
   * AddLinSys_LinearSolid2D_NonStatic_BackwardEular_P1, src/femeqn/eqn_linear_solid2d.cpp:1338-1497,
   * (lambda, myu, rho, gx, gy), glob = (dt, is_initial); the previous velocity from dfm_sys_set_velocity_prev */
  DFM_SOLID2D_BEULER = 8,
  /* Newmark-beta dynamic St.Venant-Kirchhoff 3D (the Newton step of CEqn_Solid3D_Linear, geometrically nonlinear,
   * not stationary): AddLinSys_StVenant3D_NonStatic_NewmarkBeta_P1, src/femeqn/eqn_st_venant.cpp:1109-1242,
   * (lambda, myu, rho, gx, gy, gz), glob = (dt, gamma, beta, is_initial) */
  DFM_STVK3D_NEWMARK = 9,
  /* its 2D counterpart: AddLinSys_StVenant2D_NonStatic_NewmarkBeta_P1, src/femeqn/eqn_st_venant.cpp:314-460,
   * (lambda, myu, rho, gx, gy), glob = (dt, gamma, beta, is_initial) */
  DFM_STVK2D_NEWMARK = 10,
  /* backward-Euler dynamic St.Venant-Kirchhoff 2D: AddLinSys_StVenant2D_NonStatic_BackwardEular_P1,
   * src/femeqn/eqn_st_venant.cpp:509-664, (lambda, myu, rho, gx, gy), glob = (dt, is_initial); the previous velocity
   * from dfm_sys_set_velocity_prev, the update with dfm_sys_update_field_beuler */
  DFM_STVK2D_BEULER = 11
};

/* orderings for the ILU(0) preconditioner (dfm_sys_make_ordering); the last three are COrdering_Blk's, exactly
 * as the reference computes them (src/matvec/ordering_blk.cpp), for CPreconditioner_ILU::SetOrdering */
enum {
  DFM_ORDER_NATURAL = 0,     /* identity: the reference's default CPreconditioner_ILU                    */
  DFM_ORDER_MULTICOLOR = 1,  /* greedy node colouring, colour-major                                      */
  DFM_ORDER_BLOCKCOLOR = 2,  /* block multicolour: geometric tiles of <= 64 nodes, tiles coloured        */
  DFM_ORDER_RCM = 3,         /* MakeOrdering_RCM (ordering_blk.cpp:62-134): breadth-first levels from block 0,
                                each level by ascending degree, reversed; stored, like the reference, as the
                                old -> new positions in the new -> old slot (NewToOld returns them)        */
  DFM_ORDER_RCM2 = 4,        /* MakeOrdering_RCM2 (:169-262): level sets of the lower graph               */
  DFM_ORDER_AMD = 5          /* MakeOrdering_AMD (:320-485) as compiled: its element absorption calls are
                                commented out, so it takes blocks by ascending degree, lowest index first */
};

/* precision of the device Krylov solvers: DFM_FP64 the reference algorithm in FP64; DFM_MIXED FP32 matrix,
 * factor and Krylov vectors with FP64 dots, solution and true residual (reliable updates), continuing as
 * DFM_FP32_ILU from its FP64 solution when its reliable-update cycles make slow progress; DFM_FP32_ILU the FP64
 * loop with the FP32 copy of the factor as preconditioner (attainable accuracy of FP64, half the factor bytes);
 * DFM_FP64_TRUE the reference loop in FP64 whose true residual is checked when it stops, restarting from its
 * solution until the true residual meets the tolerance; DFM_MIXED_AUTO per system: DFM_MIXED for scalar systems
 * (one dof per node), DFM_FP64_TRUE for vector systems. The choice for vector systems is measured on the serving
 * stream's slender 2D elasticity: the FP32 matrix's reliable-update cycles need about 1.4x the FP64 iterations,
 * and the FP32 factor adds 10-30 % iterations while the batched sweeps of such systems, latency- rather than
 * bandwidth-bound, run no faster with it. */
enum { DFM_FP64 = 0, DFM_MIXED = 1, DFM_FP32_ILU = 2, DFM_MIXED_AUTO = 3, DFM_FP64_TRUE = 4 };

const char* dfm_last_error(void);
int dfm_device_count(int* count);

/* ---- context: one device, one stream ---------------------------------------------------------------- */
int dfm_ctx_create(int device, dfm_ctx* out);
int dfm_ctx_destroy(dfm_ctx ctx);
int dfm_ctx_synchronize(dfm_ctx ctx);

/* ---- seeded structured meshes (synthetic inputs; host-side) ------------------------------------------ */
/* 2D rectangle lx*ly, nx*ny quads split into 2 P1 triangles; interior nodes jittered by +-jitter*h.
 * coords[(nx+1)(ny+1)*2], tri[nx*ny*2*3]. Node order x-fastest (include/delfem/mesh_primitive.h:77). */
int dfm_mesh_rect_tri(int nx, int ny, double lx, double ly, unsigned long long seed, double jitter,
                      double* coords, int* tri);
/* 3D box, nx*ny*nz hexes split into 6 Kuhn tets each. coords[(nx+1)(ny+1)(nz+1)*3], tet[nx*ny*nz*6*4]. */
int dfm_mesh_box_tet(int nx, int ny, int nz, double lx, double ly, double lz, unsigned long long seed,
                     double jitter, double* coords, int* tet);
/* GiD mesh file (.msh) of P1 triangles or tetrahedra, as the reference's readers take it
 * (CMesher2D::ReadFromFile_GiDMsh, src/msh/mesher2d.cpp:2463-2620, for triangles: coplanar points of a
 * "dimension 3" file are projected onto the plane of their two widest axes; Msh::CMesh3D::ReadFromFile_GiDMsh,
 * src/msh/mesh3d.cpp:601-760, for tetrahedra), nodes and elements in file order, 0-based. Call with coords and
 * lnods NULL for the sizes (ndim, nnode, nelem, npoel = ndim + 1), then with arrays of those sizes.
 * DFM_ERR_ARG for other element types (quadrilaterals, hexahedra), a malformed file, or triangles that are not
 * coplanar; host only. */
int dfm_mesh_read_gid(const char* path, int* ndim, int* nnode, int* nelem, int* npoel, double* coords, int* lnods);
/* Counter-based SplitMix64 draws (uniform [0,1) / standard normal) for seeded materials and loads. */
int dfm_rng_uniform(unsigned long long seed, unsigned long long stream, size_t n, double* out);
int dfm_rng_normal(unsigned long long seed, unsigned long long stream, size_t n, double* out);

/* ---- system: CLinearSystem_Field + CMatDia_BlkCrs + CVector_Blk + CBCFlag on the device ------------- */
/* Pattern from the mesh (CLinearSystem_Field::AddPattern_Field, src/femls/linearsystem_field.cpp:196-280
 * -> CElemAry::MakePattern_FEM, src/femfield/elem_ary.cpp:225-245). len = dofs per node (1, 2 or 3). */
int dfm_sys_create_mesh(dfm_ctx ctx, int ndim, int nnode, const double* coords, int nelem, const int* lnods,
                        int len, dfm_sys* out);
int dfm_sys_destroy(dfm_sys sys);
/* A system from a host block pattern without a mesh: CMatDia_BlkCrs::AddPattern(const CIndexedArray&)
 * (src/matvec/matdia_blkcrs.cpp:346-505) on an nblk x nblk matrix of len x len blocks. rowptr/colind give the
 * off-diagonal blocks of rows 0..nrow-1 (nrow <= nblk; later rows empty), no diagonal, no repeated column. Its
 * values come from dfm_sys_set_matrix; orderings other than block colour, ILU(k), the solvers and the handle
 * operations apply as to a mesh system; element assembly does not. */
int dfm_sys_create_pattern(dfm_ctx ctx, int nblk, int len, int nrow, const unsigned* rowptr, const unsigned* colind,
                           dfm_sys* out);
/* A new system on the same mesh, ordering, ILU schedule and Dirichlet flags as src (no symbolic work; field,
 * parameters, loads and values are not copied). No counterpart in the reference: instances of one geometry
 * (Newton batches, repeated serving geometries). */
int dfm_sys_clone(dfm_sys src, dfm_sys* out);
int dfm_sys_info(dfm_sys sys, int* nblk, int* len, long long* ncrs);
/* The CRS pattern in the original numbering (rowptr[n+1], colind[ncrs]; CMat_BlkCrs::GetPtrIndPSuP). */
int dfm_sys_get_pattern(dfm_sys sys, unsigned* rowptr, unsigned* colind);

/* ILU ordering (CPreconditioner_ILU::SetOrdering): order[new] = old, NULL = natural. Call before assembly;
 * it re-labels the device numbering and rebuilds the schedule. */
int dfm_sys_set_ordering(dfm_sys sys, const int* order);
/* Build an ordering of the given kind (tile = nodes per tile edge for BLOCKCOLOR, 0 = default). */
int dfm_sys_make_ordering(dfm_sys sys, int kind, int tile, int* order_out, int* ncolors_out);
/* One of COrdering_Blk's orderings (DFM_ORDER_RCM, _RCM2, _AMD; NATURAL gives the identity) of a block pattern
 * (rowptr[n+1], colind: the off-diagonal blocks of each row, ascending, as CMatDia_BlkCrs::GetPtrIndPSuP gives
 * them) into order[n] (new -> old, NewToOld). Host only: needs no device. DFM_ERR_ARG where the reference would
 * stop on an assertion (RCM: a graph in several pieces; RCM2: blocks its level rule leaves unnumbered). */
int dfm_pattern_ordering(int kind, int n, const unsigned* rowptr, const unsigned* colind, int* order);
/* Sequential phases of the forward and backward triangular sweeps under the current ordering and fill level,
 * the most intra-group levels of a phase, and the sweep launches of one PCG iteration (one fewer than the phases
 * when the last forward and the first backward phase hold the same tiles and run as one kernel). No counterpart
 * in the reference. */
int dfm_sys_schedule_info(dfm_sys sys, int* nphase_fwd, int* nphase_bwd, int* max_group_levels,
                          int* pcg_sweep_launches);

/* Dirichlet flags, n*len ints (CLinearSystem_Field::SetFixedBoundaryCondition_Field -> CBCFlag::SetBC,
 * src/femls/linearsystem_field.cpp:618-710). */
int dfm_sys_set_bc(dfm_sys sys, const int* flags);
/* Nodal field values used by assembly (VALUE u and VELOCITY v; n*len each, v may be NULL). */
int dfm_sys_set_field(dfm_sys sys, const double* u, const double* v);
int dfm_sys_get_field(dfm_sys sys, double* u);
/* u_field += update (CLinearSystem_Field::UpdateValueOfField, src/femls/linearsystem_field.cpp:716-794) */
int dfm_sys_update_field(dfm_sys sys);
/* ACCELERATION field values (n*len; NULL = zero), and all three fields out (any pointer may be NULL). */
int dfm_sys_set_acceleration(dfm_sys sys, const double* a);
int dfm_sys_get_fields(dfm_sys sys, double* u, double* v, double* a);
/* CLinearSystem_Field::UpdateValueOfField_NewmarkBeta (src/femls/linearsystem_field.cpp:966-1059), the update
 * being the acceleration increment: [is_initial: u += dt v, u += dt^2/2 a], u += beta dt^2 du,
 * [is_initial: v += dt a], v += gamma dt du, a += du; in that order, as the reference's node-array passes. */
int dfm_sys_update_field_newmark(dfm_sys sys, double gamma, double beta, double dt, int is_initial);
/* The velo_pre vector of AddLinSys_LinearSolid2D_NonStatic_BackwardEular (n*len; NULL = zero), and
 * CLinearSystem_Field::UpdateValueOfField_BackwardEular (src/femls/linearsystem_field.cpp:1061-1101), the update
 * being the velocity increment: v += du, then u += dt v (is_initial) or u += dt du. */
int dfm_sys_set_velocity_prev(dfm_sys sys, const double* v0);
int dfm_sys_update_field_beuler(dfm_sys sys, double dt, int is_initial);
/* CLinearSystem_Field::UpdateValueOfField_Newmark with VELOCITY (src/femls/linearsystem_field.cpp:885-921), the
 * update being the velocity increment of the diffusion equation: [is_initial: u += dt v], u += (gamma dt) du,
 * v += du. */
int dfm_sys_update_field_newmark1(dfm_sys sys, double gamma, double dt, int is_initial);

/* InitializeMarge + Fem::Eqn::AddLinSys_<kind> + [surface loads] + FinalizeMarge in one call
 * (src/ls/linearsystem.cpp:74-128). params: npar doubles per element with stride pstride (0 = shared).
 * faces/trac: optional P1 surface loads (3D: triangles + traction vector each; a reference extension),
 * nface may be 0. Returns ||residual|| after Dirichlet zeroing in *res_norm (may be NULL). */
int dfm_sys_assemble(dfm_sys sys, int kind, const double* params, int npar, int pstride, const double* glob,
                     int nface, const int* faces, const double* trac, double* res_norm);
/* Resident inputs: upload element parameters (and optional surface loads) once, then re-assemble from HBM
 * (Newton steps, repeated solves). dfm_sys_assemble_resident = dfm_sys_assemble on the stored inputs. */
int dfm_sys_set_params(dfm_sys sys, const double* params, int npar, int pstride);
int dfm_sys_set_loads(dfm_sys sys, int nface, const int* faces, const double* trac);
int dfm_sys_assemble_resident(dfm_sys sys, int kind, const double* glob, double* res_norm);
/* Factor applied to the stored surface loads by later assemblies (load stepping; 1 = as given). */
int dfm_sys_set_load_scale(dfm_sys sys, double scale);
/* Matrix values in the original numbering: crs[ncrs*len*len], dia[n*len*len]. */
int dfm_sys_get_matrix(dfm_sys sys, double* crs, double* dia);
/* The matrix values from the host, in dfm_sys_get_matrix's layout (the caller's numbering and pattern order): the
 * values host-side element routines left in a CMatDia_BlkCrs (ILinearSystem_Eqn::GetMatrix,
 * include/delfem/linearsystem_interface_eqnsys.h:48-68). The system counts as assembled; factor it again. The
 * right-hand side goes in with dfm_sys_set_vector(-1). */
int dfm_sys_set_matrix(dfm_sys sys, const double* crs, const double* dia);

/* ---- Saved stiffness (CLinearSystem_Save, include/delfem/femls/linearsystem_fieldsave.h:59-106) -------- */
/* InitializeMarge + the *_SaveStiffMat element loop + FinalizeMarge of CLinearSystem_Save
 * (src/femls/linearsystem_fieldsave.cpp:169-204): K as dfm_sys_assemble builds it (Dirichlet rows and columns
 * applied), the external force f kept apart (GetForce, :158-167), and the boundary coupling K_b, the entries
 * of K in free rows and fixed columns (GetMatrix_Boundary, :144-156, pattern :334-362), kept from before the
 * Dirichlet zeroing. The residual handle is left zero. Kinds, as the reference has them:
 * DFM_POISSON on triangles (AddLinSys_Poisson on CLinearSystem_Save, src/femeqn/eqn_poisson.cpp:494-569;
 * the reference stops on tetrahedra), DFM_SOLID2D (AddLinSys_LinearSolid2D_Static_SaveStiffMat,
 * src/femeqn/eqn_linear_solid2d.cpp:861-960), DFM_SOLID3D (AddLinSys_LinearSolid3D_Static_SaveStiffMat,
 * src/femeqn/eqn_linear_solid3d.cpp:102-126, :425-521). Set the Dirichlet flags first. */
int dfm_sys_assemble_save(dfm_sys sys, int kind, const double* params, int npar, int pstride);
/* The saved systems with a lumped mass (src/femls/linearsystem_fieldsave.cpp:363-870), InitializeMarge + the
 * element loop + FinalizeMarge:
 *  kind DFM_DIFFUSION2D, CLinearSystem_SaveDiaM_Newmark (AddLinearSystem_Diffusion2D_P1,
 *    src/femeqn/eqn_diffusion.cpp:565-650), params (rho, alpha, source), glob = (dt, gamma): matrix
 *    gamma dt C + M (M lumped, rho area / 3), boundary matrix C, force the source term; the unknown is the new
 *    velocity;
 *  kind DFM_SOLID2D, CLinearSystem_SaveDiaM_NewmarkBeta (AddLinSys_LinearSolid2D_NonStatic_Save_NewmarkBeta_P1,
 *    src/femeqn/eqn_linear_solid2d.cpp:1051-1180), params (lambda, myu, rho, gx, gy), glob = (dt, gamma, beta):
 *    matrix M + beta dt^2 K, boundary matrix K, force the gravity (both of its components added to a node's first
 *    dof, as the reference adds them); the unknown is the new acceleration.
 * The boundary matrix is the whole matrix pattern with the diagonal (CMat_BlkCrs::SetPatternDia,
 * src/matvec/mat_blkcrs.cpp:1267-1319), its fixed rows zeroed. Triangles only. Set the Dirichlet flags first; the
 * residual handle is left zero. dfm_sys_make_residual and dfm_sys_update_field_save then follow these classes. */
int dfm_sys_assemble_save_dia(dfm_sys sys, int kind, const double* params, int npar, int pstride,
                              const double* glob);
/* MakeResidual of the saved systems: the update handle takes w, the residual handle r = f - K_b * w (the same
 * terms in the same order as the reference), and ||r|| is returned in *res_norm (may be NULL). w is the field
 * values for CLinearSystem_Save (src/femls/linearsystem_fieldsave.cpp:206-247); u + dt (1 - gamma) v for
 * SaveDiaM_Newmark (:425-472), with the residual zeroed at fixed dofs; u + dt v + dt^2 (0.5 - beta) a for
 * SaveDiaM_NewmarkBeta (:697-758), likewise. Re-solving after new Dirichlet values needs only this call. */
int dfm_sys_make_residual(dfm_sys sys, double* res_norm);
/* UpdateValueOfField of the saved systems: CLinearSystem_Save (VALUE, :249-332): the update takes the field
 * values at fixed dofs, then the field is set to the update (u = update, not u += update). SaveDiaM_Newmark
 * (VELOCITY, :474-520): u += dt (1 - gamma) v + dt gamma update at free dofs, then v = update (fixed dofs keep
 * their velocity). SaveDiaM_NewmarkBeta (ACCELERATION, :760-860): u += dt v + dt^2 (0.5 - beta) a +
 * dt^2 beta update, v += dt (1 - gamma) a + gamma dt update at free dofs, then a = update (fixed dofs keep
 * their acceleration). */
int dfm_sys_update_field_save(dfm_sys sys);
/* The saved external force f (original numbering, n*len) and the number of scalar terms of K_b. */
int dfm_sys_get_force(dfm_sys sys, double* f);
int dfm_sys_boundary_terms(dfm_sys sys, long long* nterm);

/* ---- ILinearSystemPreconditioner_Sol handle ops (include/delfem/ls/linearsystem_interface_solver.h:63-82) */
int dfm_sys_resize_tmp(dfm_sys sys, int ntmp);
int dfm_sys_ntmp(dfm_sys sys, int* ntmp);
int dfm_sys_dot(dfm_sys sys, int iv1, int iv2, double* out);
int dfm_sys_copy(dfm_sys sys, int iv1, int iv2);
int dfm_sys_scal(dfm_sys sys, double alpha, int iv);
int dfm_sys_axpy(dfm_sys sys, double alpha, int iv1, int iv2);
int dfm_sys_matvec(dfm_sys sys, double alpha, int iv1, double beta, int iv2);
int dfm_sys_get_vector(dfm_sys sys, int iv, double* out);
int dfm_sys_set_vector(dfm_sys sys, int iv, const double* in);

/* ---- ILU preconditioner (CPreconditioner_ILU, src/ls/preconditioner.cpp:57-335) ---------------------- */
/* Level of fill (CPreconditioner_ILU::SetFillInLevel, include/delfem/ls/preconditioner.h:77, symbolic part
 * SetLinearSystem, src/ls/preconditioner.cpp:57-211 -> CMatDiaFrac_BlkCrs::AddFracPtn,
 * src/matvec/matdiafrac_blkcrs.cpp:1254-1381): 0 ILU(0) (the default), k > 0 ILU(k), -1 no limit. The
 * level-of-fill pattern is built for the current ordering (the reference builds it in the natural order; under
 * an ordering it is that of the renumbered matrix); the factor must be computed again. */
int dfm_prec_set_fill_level(dfm_sys sys, int lev);
/* Number of off-diagonal blocks of the factor pattern (the matrix's for ILU(0)). */
int dfm_prec_factor_size(dfm_sys sys, long long* fncrs);
/* Numeric factorisation of the assembled matrix (SetValue -> DoILUDecomp, matdiafrac_blkcrs.cpp:774-1107).
 * Returns DFM_ERR_SINGULAR when more than 10 pivots were singular (the reference returns false). */
int dfm_prec_factor(dfm_sys sys);
/* In-place z = M^-1 z on handle iv (SolvePrecond). */
int dfm_prec_apply(dfm_sys sys, int iv);
/* The factor in the device numbering: frowptr[n+1], fcolind[fncrs], diaind[n], fcrs, fdia (L unscaled,
 * D^-1, D^-1 U), for comparison with CMatDiaFrac_BlkCrs under the same ordering. */
int dfm_prec_get_factor(dfm_sys sys, unsigned* frowptr, unsigned* fcolind, unsigned* diaind, double* fcrs,
                        double* fdia);

/* ---- device-resident Krylov loops (LsSol::Solve_PCG / Solve_PBiCGSTAB, src/ls/solver_ls_iter.cpp) ----
 * Same contract as the reference: conv_ratio in = tolerance, out = achieved ratio; iteration in = max,
 * out = the 0-based index at convergence for PCG (1-based loop index for PBiCGSTAB). Returns 1 when
 * converged, 0 when not, negative on error. Scalars stay on the device; the host polls a done flag.
 * DFM_FP64 runs the reference algorithm in FP64 on the recursive residual (its conv_ratio is that of the
 * reference; dfm_sys_true_residual gives the true one). DFM_MIXED keeps FP32 copies of the matrix and the ILU
 * factor and runs FP32 SpMV/ILU inside FP64 iterative refinement on the true residual r = b - A x (FP64 matrix,
 * FP64 dots and update). DFM_FP32_ILU, DFM_FP64_TRUE and a mixed solve that continued in DFM_FP32_ILU check the
 * true residual when their recursive one converged and restart from their solution (recursive target half the
 * tolerance) while it is above the tolerance and each restart lowers it by a fifth; where FP64 cannot resolve the
 * tolerance for the system the solve ends unconverged. In every mode but DFM_FP64 the returned conv_ratio is the
 * true relative residual ||b - A x|| / ||b|| (FP64 matrix). */
int dfm_solve_pcg(dfm_sys sys, int precision, double* conv_ratio, unsigned* iteration);
/* PBiCGSTAB: DFM_FP64 is the reference loop (it returns 1 also when not converged, with iteration = max and
 * conv_ratio untouched, as Solve_PBiCGSTAB returns true); DFM_FP32_ILU, and DFM_MIXED / DFM_MIXED_AUTO (which run
 * as DFM_FP32_ILU: the FP32-vector loop is PCG's), apply the FP32 copy of the factor in the FP64 loop; they and
 * DFM_FP64_TRUE check the true residual and restart on b - A x as PCG does, returning 0 when it cannot be met. */
int dfm_solve_pbicgstab(dfm_sys sys, int precision, double* conv_ratio, unsigned* iteration);
/* Host-driven loops over the handle operations (the C++ layer include/dfm_lssol.hpp), FP64, same contract as
 * the reference functions they follow: LsSol::Solve_CG and Solve_BiCGSTAB without preconditioner
 * (src/ls/solver_ls_iter.cpp:31-97, :177-286; both return 1 also when max_iter is reached), and Solve_PCG /
 * Solve_PBiCGSTAB with the ILU(0) preconditioner (:100-172, :291-408) as the reference drives them. */
int dfm_solve_cg(dfm_sys sys, double* conv_ratio, unsigned* iteration);
int dfm_solve_bicgstab(dfm_sys sys, double* conv_ratio, unsigned* iteration);
int dfm_solve_pcg_host(dfm_sys sys, double* conv_ratio, unsigned* iteration);
int dfm_solve_pbicgstab_host(dfm_sys sys, double* conv_ratio, unsigned* iteration);
/* The matrix-level twins MatVec::Sol::Solve_CG / Solve_PCG / Solve_BiCGSTAB / Solve_PBiCGSTAB
 * (src/matvec/solver_mat_iter.cpp:37-113, :115-184, :191-299, :305-423) on the system's matrix, the residual
 * (in: right-hand side, out: final residual) and the update as solution, the preconditioned ones with the
 * system's ILU factor; their own contracts: zero right-hand side below 1e-30 gives conv_ratio 0; Solve_CG
 * tests the residual of the previous step (one step late) and counts from 1; Solve_PCG returns 0 with
 * conv_ratio and iteration untouched when not converged; the others return 1. */
int dfm_solve_mat_cg(dfm_sys sys, double* conv_ratio, unsigned* iteration);
int dfm_solve_mat_pcg(dfm_sys sys, double* conv_ratio, unsigned* iteration);
int dfm_solve_mat_bicgstab(dfm_sys sys, double* conv_ratio, unsigned* iteration);
int dfm_solve_mat_pbicgstab(dfm_sys sys, double* conv_ratio, unsigned* iteration);
/* ||b - A x|| / ||b|| in FP64 for b = the residual as assembled and x = the update (after a solve). */
int dfm_sys_true_residual(dfm_sys sys, double* rel);

/* dfm_sys_true_residual without device allocation. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid it). Its
 * scratch, the partial sums of at most 2048 workgroups and the two totals, is a fixed array of 32,784 bytes on
 * each device the library runs on, held until the process ends; the result is bit-identical to
 * dfm_sys_true_residual's (same kernels). Systems of more than 2^31 - 1 dofs are refused (DFM_ERR_ARG). Calls that
 * overlap on one device share that array and race: serialise them. */
int dfm_sys_true_residual_noalloc(dfm_sys sys, double* rel);

/* ---- timing of the last device solve (HIP events on the system's stream) ----------------------------- */
int dfm_sys_last_solve_stats(dfm_sys sys, double* ms_total, double* ms_per_iter, unsigned* inner_iters);
/* Mixed precision: reliable updates of the last solve, and whether its FP64 true residual stalled (six
 * reliable updates in a row without halving it: cond(A) * eps32 too large for an FP32 matrix) so that the
 * solve was repeated in FP64. */
int dfm_sys_last_solve_info(dfm_sys sys, int* reliable_updates, int* fell_back);
/* Average device time of one launch of a hot kernel over `reps` back-to-back launches, timed with HIP events
 * on the system's own stream: which = 0 SpMV FP64 (y = A x), 1 SpMV FP32, 2 ILU apply FP64 (all phases of
 * the forward and backward sweeps), 3 ILU apply FP32, 4 / 5 the sweep phases of a PCG iteration as the mixed
 * (FP32) / FP64 loop runs them (forward phases with the r and x update, backward phases with (r, z)), 6 the
 * MakeResidual kernel of a saved system (it rewrites the residual and update handles as dfm_sys_make_residual
 * does), 7 the numeric ILU factorisation of the assembled matrix (its initial values and the factor kernels of
 * every phase; it rewrites the same factor), 8 the assembly kernels of the last dfm_sys_assemble[_resident] (the
 * system must be assembled again before it is solved). Also reports the algorithmic bytes of one launch (all
 * phases). No counterpart in the reference. */
int dfm_sys_bench_kernel(dfm_sys sys, int which, int reps, double* avg_ms, double* bytes);
/* ---- batches: many systems solved together (serving, instance batches) ------------------------------
 * A batch advances every system added to it with one launch per kernel for all running systems of one block
 * size; each system keeps its own Krylov scalars, done flag and iteration count, so its iterates are those of
 * LsSol::Solve_PCG (src/ls/solver_ls_iter.cpp:100-172) for DFM_FP64, or of the mixed-precision form of
 * dfm_solve_pcg for DFM_MIXED, and do not depend on which other systems share the batch. No counterpart in
 * the reference (single-threaded, one system at a time). Systems must be assembled and factored, live on the
 * batch's device (their own context may differ: a system prepared on another stream is synchronized when
 * added), and stay alive while in the batch (dfm_sys_destroy takes a system out first). */
typedef struct dfm_batch_s* dfm_batch;
int dfm_batch_create(dfm_ctx ctx, int precision, int max_systems, dfm_batch* out);
int dfm_batch_destroy(dfm_batch b);
/* Start solving sys (residual handle = right-hand side, solution into the update handle). */
int dfm_batch_add(dfm_batch b, dfm_sys sys, double tol, unsigned max_iter);
/* Enqueue `iterations` PCG iterations for all running systems, then collect the finished ones. */
int dfm_batch_step(dfm_batch b, int iterations, int* n_running);
/* state: 0 running, 1 converged, 2 stopped at max_iter; conv_ratio/iteration as dfm_solve_pcg; seconds from
 * dfm_batch_add to the step that found it finished (host clock). */
int dfm_batch_status(dfm_batch b, dfm_sys sys, int* state, double* conv_ratio, unsigned* iteration,
                     double* seconds);
/* Reliable updates and FP64 fallback of a system's solve (as dfm_sys_last_solve_info). */
int dfm_batch_info(dfm_batch b, dfm_sys sys, int* reliable_updates, int* fell_back);
int dfm_batch_remove(dfm_batch b, dfm_sys sys);

/* ---- continuous-batching solve server ------------------------------------------------------------------
 * No counterpart in the reference (single-threaded, one system at a time). Requests are admitted in arrival
 * order while fewer than max_active are in flight and their estimated device bytes fit budget_bytes; nworkers
 * threads (each with its own stream) build, order (block multicolour), assemble and factor admitted requests
 * while the running batch keeps iterating; a prepared request joins the batch at the next progress call and
 * leaves it when converged (PCG to tol on the true residual in DFM_MIXED). Each request's result is that of
 * the same system solved alone. Results stay on the device until dfm_server_release. */
typedef struct dfm_server_s* dfm_server;
typedef struct {
  int ndim, len, kind;          /* kind: DFM_POISSON, DFM_SOLID2D, ... (dfm_sys_assemble)            */
  int nnode, nelem;
  const double* coords;         /* [nnode*ndim]                                                        */
  const int* lnods;             /* [nelem*(ndim+1)]                                                    */
  const double* params;         /* npar per element (pstride) or shared (pstride = 0)                  */
  int npar, pstride;
  const double* glob;           /* the kind's scalars (see the kinds above) or NULL; exactly those are read */
  const int* bc;                /* [nnode*len] Dirichlet flags                                         */
  int nface;                    /* surface loads (3D), may be 0                                        */
  const int* faces;
  const double* trac;
  double tol;                   /* relative residual                                                   */
  unsigned max_iter;
  int method;                   /* DFM_METHOD_PCG (batched) or DFM_METHOD_PBICGSTAB (solved by a setup worker on
                                   its own stream beside the batch, dfm_solve_pbicgstab with the server's
                                   precision)                                                          */
} dfm_request;
enum { DFM_METHOD_PCG = 0, DFM_METHOD_PBICGSTAB = 1 };
enum { DFM_REQ_QUEUED = 0, DFM_REQ_SETUP = 1, DFM_REQ_SOLVING = 2, DFM_REQ_CONVERGED = 3, DFM_REQ_MAXITER = 4,
       DFM_REQ_FAILED = 5 };
int dfm_server_create(dfm_ctx ctx, int precision, long long budget_bytes, int max_active, int nworkers,
                      dfm_server* out);
int dfm_server_destroy(dfm_server s);
/* Copies the request; *ticket identifies it. */
int dfm_server_submit(dfm_server s, const dfm_request* req, long long* ticket);
/* Admit, join prepared requests, run `iterations` batch iterations, retire finished ones. With nothing to
 * iterate, waits up to wait_ms for a request to become ready. */
int dfm_server_progress(dfm_server s, int iterations, int wait_ms, int* n_active, int* n_queued);
/* latency_s: submission to completion; solve_s: admission to completion; fell_back: the mixed-precision solve
 * stalled and was repeated in FP64 (dfm_sys_last_solve_info). */
int dfm_server_poll(dfm_server s, long long ticket, int* state, double* latency_s, double* solve_s,
                    unsigned* iteration, double* conv_ratio, int* fell_back);
/* Host-time counters of the server, out[0..n): batch steps, seconds in them, seconds in batch table rebuilds,
 * rebuilds, seconds in flag polls (each waits for the device), restarts after a true-residual check, seconds
 * waiting with nothing to iterate, seconds of request setup summed over the workers, requests set up, seconds of
 * per-result true residuals, seconds starting the solves of joining requests, seconds in dfm_server_progress. */
int dfm_server_stats(dfm_server s, double* out, int n);
/* ||b - A x|| / ||b|| with the FP64 matrix of a finished request, computed when it finished (for DFM_FP64 requests
 * the conv_ratio of dfm_server_poll is the reference loop's recursive one). */
int dfm_server_true_residual(dfm_server s, long long ticket, double* rel);
/* The solution (update handle, caller's node numbering) of a finished request. */
int dfm_server_result(dfm_server s, long long ticket, double* u);
int dfm_server_release(dfm_server s, long long ticket);

/* ---- load-stepped Newton for many instances (CEqn_Solid3D_Linear::Solve, geometric-nonlinear branch,
 * src/femeqn/eqnsys_solid.cpp:174-209) ----------------------------------------------------------------------
 * Each system has its parameters, loads, Dirichlet flags and (zero) field set. For load step s = 1..nsteps
 * the surface loads are scaled by load_scale[i] * s / nsteps; then, per instance, up to max_newton iterations
 * of: assemble tangent + residual (kind), ||r|| (iteration 0: stop if < 1e-20, else ||r_0||), stop when
 * ||r|| < newton_tol * ||r_0||, ILU(0) factor, PCG (lin_tol, lin_maxit; the result is used whether or not it
 * converged, as the reference does), u += du. The PCG solves of one Newton iteration of all instances run as
 * one batch. Outputs (may be NULL): newton_iters[i*nsteps + s-1], total PCG iterations and unconverged PCG
 * solves per instance. */
int dfm_newton_batch(dfm_ctx ctx, int nsys, const dfm_sys* sys, int kind, const double* glob, int nsteps,
                     const double* load_scale, int max_newton, double newton_tol, int precision, double lin_tol,
                     unsigned lin_maxit, int* newton_iters, unsigned* lin_iters, int* lin_unconverged);

/* ---- equation systems: CEqnSystem_Solid2D, CEqn_Solid3D_Linear, CEqnSystem_Scalar2D, CEqn_Scalar3D
 * (include/delfem/eqnsys_solid.h, eqnsys_scalar.h, eqnsys.h) ----------------------------------------------------
 * The reference's classes and their Solve flows (src/femeqn/eqnsys_solid.cpp:175-247 and :671-761,
 * src/femeqn/eqnsys_scalar.cpp:294-359 and :554-584) over device linear systems: ILU(1) for solids, ILU(0) for
 * scalars, PCG to conv_ratio 1e-6 in at most 1000 iterations (GetAry_ItrNormRes gets each solve's iterations
 * and ratio), Newton under geometrical nonlinearity (10 steps in 2D, 40 in 3D, to 1e-6 of the first residual),
 * the saved stiffness matrix (CLinearSystem_Save) for stationary linear systems, Newmark-beta (solids) and
 * first-order Newmark (2D scalar) time stepping. The field world is replaced by the mesh: element groups are the
 * element arrays (one CEqn_* and material each; elements taken group by group), boundary element lists the fixed
 * element arrays (AddFixElemAry). dfm_eqn_create takes ngroup 1 for the 3D types. */
typedef struct dfm_eqn_s* dfm_eqn;
enum { DFM_EQN_SOLID2D = 0, DFM_EQN_SOLID3D = 1, DFM_EQN_SCALAR2D = 2, DFM_EQN_SCALAR3D = 3 };
/* dfm_eqn_set_flag: SetGeometricalNonlinear / SetGeometricalNonLinear (solids), SetStationary, SetSaveStiffMat */
enum { DFM_EQN_NONLINEAR = 0, DFM_EQN_STATIONARY = 1, DFM_EQN_SAVE_STIFF_MAT = 2 };
/* The equation system of `type` on nnode nodes and nelem P1 elements (lnods), element k in group elem_group[k]
 * (NULL: one group). Defaults are the reference constructors': 2D solid Young 1, Poisson 0, plane-stress flag
 * set, rho 1, stationary (the reference leaves that flag unset); 3D solid lambda = myu = 0, rho 1, not
 * stationary; scalars alpha 1, capacity 1, source 0, stationary. */
int dfm_eqn_create(dfm_ctx ctx, int type, int nnode, const double* coords, int nelem, const int* lnods,
                   const int* elem_group, int ngroup, dfm_eqn* out);
int dfm_eqn_destroy(dfm_eqn eq);
/* AddFixElemAry (eqnsys_solid.cpp:1341-1362): the nodes of nbnd boundary elements of nbno nodes each are fixed in
 * component idof (-1: all); every component of those nodes is set to zero. dfm_eqn_clear_fix: ClearFixElemAry(). */
int dfm_eqn_add_fix(dfm_eqn eq, int nbnd, int nbno, const int* bnd_lnods, int idof);
int dfm_eqn_clear_fix(dfm_eqn eq);
/* Material setters; group -1 the system's (CEqnSystem_Solid2D::SetYoungPoisson, :925-937, ...), else that
 * group's CEqn_* through SetEquation. SetYoungPoisson keeps the reference's flag: is_plane_stress != 0 leaves the
 * 3D lambda (plane strain), 0 reduces it to 2 lambda myu / (lambda + 2 myu) (plane stress), as
 * include/delfem/eqnsys_solid.h:190-196 does. set_rho: 2D solids; gz: 3D solids; set_scalar which: 0 alpha,
 * 1 capacity (2D), 2 source. */
int dfm_eqn_set_young_poisson(dfm_eqn eq, int group, double young, double poisson, int is_plane_stress);
int dfm_eqn_set_rho(dfm_eqn eq, int group, double rho);
int dfm_eqn_set_gravitation(dfm_eqn eq, int group, double gx, double gy, double gz);
int dfm_eqn_set_scalar(dfm_eqn eq, int which, int group, double value);
int dfm_eqn_set_flag(dfm_eqn eq, int which, int on);
/* SetTimeIntegrationParameter (include/delfem/eqnsys.h:88-93): beta < 0 gives 0.25 (gamma + 0.5)^2 */
int dfm_eqn_set_time(dfm_eqn eq, double dt, double gamma, double beta);
int dfm_eqn_solve(dfm_eqn eq);
/* GetAry_ItrNormRes (include/delfem/eqnsys.h:69): the (iterations, conv_ratio) of the last Solve's PCG solves */
int dfm_eqn_itr_norm_res(dfm_eqn eq, int* count, unsigned* iters, double* norms, int cap);
int dfm_eqn_get_field(dfm_eqn eq, double* value, double* velocity, double* acceleration);
int dfm_eqn_set_field(dfm_eqn eq, const double* value, const double* velocity, const double* acceleration);
/* what: 1 ClearValueLinearSystem, 2 ClearValuePreconditioner, 4 ClearLinearSystemPreconditioner */
int dfm_eqn_clear(dfm_eqn eq, int what);
/* The current linear system (NULL before the first Solve and after a clear), for the handle operations. */
int dfm_eqn_system(dfm_eqn eq, dfm_sys* sys);

/* ---- multi-segment systems (LsSol::CLinearSystem with several segments, src/ls/linearsystem.cpp:74-415) ----
 * Segments i = 0 .. nseg-1 of nblk[i] blocks of len[i] dofs (CLinearSystem::AddLinSysSeg, :138-167); matrix (i, i)
 * is a diagonal segment (CMatDia_BlkCrs: off-diagonal blocks in CRS, diagonal blocks apart), matrix (i, j), i != j,
 * a coupling of len[i] x len[j] blocks (CMat_BlkCrs). Vectors and Dirichlet flags are concatenated over the
 * segments, segment 0 first; handles as for dfm_sys (-1 residual, -2 update, >= 0 temporaries). The preconditioner
 * is CPreconditioner_ILU with no fill level (src/ls/preconditioner.cpp:57-335): ILU(0) of every diagonal segment
 * (CMatDiaFrac_BlkCrs) and of every coupling (CMatFrac_BlkCrs, src/matvec/matfrac_blkcrs.cpp), factored in the
 * reference's order of segment pairs; factor and preconditioner are the reference's bit for bit. The reference
 * applies no ordering to such systems, and neither does this (natural order, level-scheduled). Every diagonal
 * segment needs its pattern before vectors, Dirichlet flags or the factor are used. Values come from the host, as
 * the element routines leave them through ILinearSystem_Eqn::GetMatrix(id_field1, .., id_field2, ..)
 * (include/delfem/linearsystem_interface_eqnsys.h:48-68). */
#define DFM_MSYS_MAXSEG 8
typedef struct dfm_msys_s* dfm_msys;
int dfm_msys_create(dfm_ctx ctx, int nseg, const int* nblk, const int* len, dfm_msys* out);
int dfm_msys_destroy(dfm_msys ms);
/* CLinearSystem::AddMat_Dia (i == j, the off-diagonal pattern: no diagonal) / AddMat_NonDia (i != j)
 * (src/ls/linearsystem.cpp:419-500): CRS rowptr[nblk[i]+1], colind ascending in each row; once per matrix. */
int dfm_msys_add_pattern(dfm_msys ms, int i, int j, const unsigned* rowptr, const unsigned* colind);
/* Off-diagonal block count of matrix (i, j), -1 when it has no pattern. */
int dfm_msys_info(dfm_msys ms, int i, int j, long long* nnz);
/* Values of matrix (i, j): blocks row-major in pattern order; dia (i == j) the diagonal blocks. */
int dfm_msys_set_matrix(dfm_msys ms, int i, int j, const double* val, const double* dia);
int dfm_msys_get_matrix(dfm_msys ms, int i, int j, double* val, double* dia);
int dfm_msys_set_vector(dfm_msys ms, int iv, const double* v);
int dfm_msys_get_vector(dfm_msys ms, int iv, double* v);
/* FinalizeMarge (src/ls/linearsystem.cpp:92-128): the Dirichlet flags (concatenated), the diagonal segments'
 * SetBoundaryCondition, the couplings' fixed rows and columns zeroed, the fixed residual entries zeroed; the
 * residual norm (the segments' squared norms summed in segment order) in *res_norm when non-NULL. */
int dfm_msys_finalize(dfm_msys ms, const int* flags, double* res_norm);
/* CPreconditioner_ILU::SetLinearSystem + SetValue (ILU(0)); DFM_ERR_SINGULAR when a diagonal segment's
 * DoILUDecomp fails. dfm_msys_get_factor reads factor (i, j) in the layout of dfm_msys_get_matrix (dia: the
 * inverted diagonal blocks of segment i). */
int dfm_msys_factor(dfm_msys ms);
int dfm_msys_get_factor(dfm_msys ms, int i, int j, double* val, double* dia);
/* CPreconditioner_ILU::SolvePrecond (src/ls/preconditioner.cpp:301-335): v := M^-1 v in place. */
int dfm_msys_precond(dfm_msys ms, int iv);
/* Handle operations of CLinearSystem (src/ls/linearsystem.cpp:208-415); DOT sums the segments' products in
 * segment order, MATVEC adds each row's couplings after its diagonal segment. */
int dfm_msys_resize_tmp(dfm_msys ms, int ntmp);
int dfm_msys_ntmp(dfm_msys ms, int* ntmp);
int dfm_msys_dot(dfm_msys ms, int iv1, int iv2, double* out);
int dfm_msys_copy(dfm_msys ms, int iv1, int iv2);
int dfm_msys_scal(dfm_msys ms, double alpha, int iv);
int dfm_msys_axpy(dfm_msys ms, double alpha, int iv1, int iv2);
int dfm_msys_matvec(dfm_msys ms, double alpha, int iv1, double beta, int iv2);
/* LsSol::Solve_PCG (method DFM_METHOD_PCG) or Solve_PBiCGSTAB (DFM_METHOD_PBICGSTAB) with the factored ILU(0),
 * FP64, host-driven over the handle operations (src/ls/solver_ls_iter.cpp:100-172, 291-408): the reference's
 * contract (conv_ratio in = tolerance, out = ratio; iteration in = max, out = count; 1 converged, 0 not). */
int dfm_msys_solve(dfm_msys ms, int method, double* conv_ratio, unsigned* iteration);

/* Device memory currently allocated by the library (bytes). */
int dfm_device_bytes(long long* bytes);

#ifdef __cplusplus
}
#endif
#endif /* DFM_GPU_H */
