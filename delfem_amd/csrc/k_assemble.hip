// Element assembly on the device. Every matrix entry is the sum of its elements' contributions in ascending element
// index, the order in which the reference's element loop merges them (CMatDia_BlkCrs::Mearge,
// src/matvec/matdia_blkcrs.cpp:165-258, called per element from AddLinSys_*): the result is bit-identical,
// deterministic and needs no atomics. Three paths: (1) linear elasticity: one kernel, a group of lanes per node
// computes the node's block row of each incident element into LDS and each lane sums one entry of the row
// (k_asm_solid_rowgroups); (2) the other kinds, and elasticity rows too long for (1): the element store, one
// thread per element computes each of its block rows once and writes them (elasticity: only the element geometry);
// one thread per off-diagonal entry then sums its elements' blocks through a list built with the pattern, one thread
// per row its diagonal block and residual; in (1) and (2) every entry is written once, the matrix needs no zeroing.
// (3) where the store is not built: one thread per row recomputes the rows of its incident elements and adds them
// into the zeroed matrix.
//
// The element arithmetic follows the reference expression by expression; the library is compiled with
// -ffp-contract=off so that no product is fused into an FMA.
#include "dfm_internal.hpp"
#include "kernels.hpp"

namespace dfm {

// ---------------------------------------------------------------------------------------------------------
// P1 geometry: src/femeqn/ker_emat_tri.cpp:13-15 (TriArea), :36-52 (TriDlDx);
// include/delfem/femeqn/ker_emat_tet.h:32-39 (TetVolume, with its 0.1666666666666 literal), :42-77 (TetDlDx)

__device__ inline double tri_area(const double* p0, const double* p1, const double* p2) {
  return 0.5 * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
}

__device__ inline void tri_dldx(double d[3][2], const double co[4][3]) {
  const double* p0 = co[0];
  const double* p1 = co[1];
  const double* p2 = co[2];
  const double area = tri_area(p0, p1, p2);
  const double t = 0.5 / area;
  d[0][0] = t * (p1[1] - p2[1]);
  d[1][0] = t * (p2[1] - p0[1]);
  d[2][0] = t * (p0[1] - p1[1]);
  d[0][1] = t * (p2[0] - p1[0]);
  d[1][1] = t * (p0[0] - p2[0]);
  d[2][1] = t * (p1[0] - p0[0]);
}

__device__ inline double tet_volume(const double co[4][3]) {
  const double* p0 = co[0];
  const double* p1 = co[1];
  const double* p2 = co[2];
  const double* p3 = co[3];
  double vol = (p1[0] - p0[0]) * ((p2[1] - p0[1]) * (p3[2] - p0[2]) - (p3[1] - p0[1]) * (p2[2] - p0[2])) -
               (p1[1] - p0[1]) * ((p2[0] - p0[0]) * (p3[2] - p0[2]) - (p3[0] - p0[0]) * (p2[2] - p0[2])) +
               (p1[2] - p0[2]) * ((p2[0] - p0[0]) * (p3[1] - p0[1]) - (p3[0] - p0[0]) * (p2[1] - p0[1]));
  vol *= 0.1666666666666;
  return vol;
}

__device__ inline void tet_dldx(double d[4][3], const double co[4][3]) {
  const double* p0 = co[0];
  const double* p1 = co[1];
  const double* p2 = co[2];
  const double* p3 = co[3];
  const double vol = tet_volume(co);
  const double t = 1.0 / (vol * 6.0);
  d[0][0] = -t * ((p2[1] - p1[1]) * (p3[2] - p1[2]) - (p3[1] - p1[1]) * (p2[2] - p1[2]));
  d[0][1] = +t * ((p2[0] - p1[0]) * (p3[2] - p1[2]) - (p3[0] - p1[0]) * (p2[2] - p1[2]));
  d[0][2] = -t * ((p2[0] - p1[0]) * (p3[1] - p1[1]) - (p3[0] - p1[0]) * (p2[1] - p1[1]));
  d[1][0] = +t * ((p3[1] - p2[1]) * (p0[2] - p2[2]) - (p0[1] - p2[1]) * (p3[2] - p2[2]));
  d[1][1] = -t * ((p3[0] - p2[0]) * (p0[2] - p2[2]) - (p0[0] - p2[0]) * (p3[2] - p2[2]));
  d[1][2] = +t * ((p3[0] - p2[0]) * (p0[1] - p2[1]) - (p0[0] - p2[0]) * (p3[1] - p2[1]));
  d[2][0] = -t * ((p0[1] - p3[1]) * (p1[2] - p3[2]) - (p1[1] - p3[1]) * (p0[2] - p3[2]));
  d[2][1] = +t * ((p0[0] - p3[0]) * (p1[2] - p3[2]) - (p1[0] - p3[0]) * (p0[2] - p3[2]));
  d[2][2] = -t * ((p0[0] - p3[0]) * (p1[1] - p3[1]) - (p1[0] - p3[0]) * (p0[1] - p3[1]));
  d[3][0] = +t * ((p1[1] - p0[1]) * (p2[2] - p0[2]) - (p2[1] - p0[1]) * (p1[2] - p0[2]));
  d[3][1] = -t * ((p1[0] - p0[0]) * (p2[2] - p0[2]) - (p2[0] - p0[0]) * (p1[2] - p0[2]));
  d[3][2] = +t * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
}

// ---------------------------------------------------------------------------------------------------------
// Element rows. Each computes block row `a` of the element matrix, em[c*B*B + r*B + s] for block (a, c),
// and the element residual of node a, er[0..B).

// src/femeqn/eqn_poisson.cpp:49-125 (2D), :249-323 (3D). p = (alpha, source)
template <int NDIM>
__device__ inline void row_poisson(int a, const double co[4][3], const double* val, const double* p, double* em,
                                   double* er) {
  const double alpha = p[0], source = p[1];
  if (NDIM == 2) {
    const double area = tri_area(co[0], co[1], co[2]);
    double d[3][2];
    tri_dldx(d, co);
    for (int j = 0; j < 3; j++) em[j] = alpha * area * (d[a][0] * d[j][0] + d[a][1] * d[j][1]);
    double r = source * area * 0.33333333333333333;
    for (int j = 0; j < 3; j++) r -= em[j] * val[j];
    er[0] = r;
  } else {
    const double vol = tet_volume(co);
    double d[4][3];
    tet_dldx(d, co);
    for (int j = 0; j < 4; j++) em[j] = alpha * vol * (d[a][0] * d[j][0] + d[a][1] * d[j][1] + d[a][2] * d[j][2]);
    double r = source * vol * 0.25;
    for (int j = 0; j < 4; j++) r -= em[j] * val[j];
    er[0] = r;
  }
}

// src/femeqn/eqn_diffusion.cpp:240-345. p = (rho, alpha, source), glob = (dt, gamma)
__device__ inline void row_diffusion2d(int a, const double co[4][3], const double* val, const double* vval,
                                       const double* p, double dt, double gamma, double* em, double* er) {
  const double rho = p[0], alpha = p[1], source = p[2];
  const double area = tri_area(co[0], co[1], co[2]);
  double d[3][2];
  tri_dldx(d, co);
  double eC[3], eM[3];
  for (int j = 0; j < 3; j++) eC[j] = alpha * area * (d[a][0] * d[j][0] + d[a][1] * d[j][1]);
  const double m = rho * area * 0.08333333333333333;
  for (int j = 0; j < 3; j++) eM[j] = m;
  eM[a] += m;
  double r = source * area * 0.333333333333333333;
  const double gdt = gamma * dt;
  for (int j = 0; j < 3; j++) em[j] = eM[j] + gdt * eC[j];
  for (int j = 0; j < 3; j++) r -= eC[j] * (val[j] + dt * vval[j]) + eM[j] * vval[j];
  er[0] = r;
}

// src/femeqn/eqn_linear_solid2d.cpp:505-598. p = (lambda, myu, rho, gx, gy)
// Block (a, j) of the element matrix from the element's geometry (area, the gradients da = dldx[a], dj = dldx[j])
// and Lame parameters.
__device__ inline void solid2d_block(double area, const double* da, const double* dj, double lambda, double myu,
                                     double* e) {
  e[0] = area * (lambda + myu) * da[0] * dj[0];
  e[1] = area * (lambda * da[0] * dj[1] + myu * dj[0] * da[1]);
  e[2] = area * (lambda * da[1] * dj[0] + myu * dj[1] * da[0]);
  e[3] = area * (lambda + myu) * da[1] * dj[1];
  const double t = (da[1] * dj[1] + da[0] * dj[0]) * area * myu;
  e[0] += t;
  e[3] += t;
}

// Block row a and the residual of node a from the element's geometry.
__device__ inline void solid2d_rows(int a, double area, const double d[3][2], const double u[4][3], const double* p,
                                    double* em, double* er) {
  const double lambda = p[0], myu = p[1], rho = p[2], g_x = p[3], g_y = p[4];
  for (int j = 0; j < 3; j++) solid2d_block(area, d[a], d[j], lambda, myu, &em[j * 4]);
  double r0 = area * rho * g_x * 0.33333333333333333;
  double r1 = area * rho * g_y * 0.33333333333333333;
  for (int j = 0; j < 3; j++) {
    const double* e = &em[j * 4];
    r0 -= e[0] * u[j][0] + e[1] * u[j][1];
    r1 -= e[2] * u[j][0] + e[3] * u[j][1];
  }
  er[0] = r0;
  er[1] = r1;
}

// solid2d_rows handing each block to put(j, block) as it is made: the same values, each residual value taking the
// same operations in the same order.
template <class Put>
__device__ inline void solid2d_rows_put(int a, double area, const double d[3][2], const double u[4][3],
                                        const double* p, double* er, Put&& put) {
  const double lambda = p[0], myu = p[1], rho = p[2], g_x = p[3], g_y = p[4];
  double r0 = area * rho * g_x * 0.33333333333333333;
  double r1 = area * rho * g_y * 0.33333333333333333;
  for (int j = 0; j < 3; j++) {
    double e[4];
    solid2d_block(area, d[a], d[j], lambda, myu, e);
    put(j, e);
    r0 -= e[0] * u[j][0] + e[1] * u[j][1];
    r1 -= e[2] * u[j][0] + e[3] * u[j][1];
  }
  er[0] = r0;
  er[1] = r1;
}

__device__ inline void row_solid2d(int a, const double co[4][3], const double u[4][3], const double* p, double* em,
                                   double* er) {
  const double area = tri_area(co[0], co[1], co[2]);
  double d[3][2];
  tri_dldx(d, co);
  solid2d_rows(a, area, d, u, p, em, er);
}

// src/femeqn/eqn_linear_solid3d.cpp:241-334. p = (lambda, myu, rho, gx, gy, gz)
// Block (a, j) of the element matrix from the element's geometry (volume, the gradients da = dldx[a],
// dj = dldx[j]) and Lame parameters.
__device__ inline void solid3d_block(double vol, const double* da, const double* dj, double lambda, double myu,
                                     double* e) {
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) e[r * 3 + s] = vol * (lambda * da[r] * dj[s] + myu * dj[r] * da[s]);
  const double t = da[0] * dj[0] + da[1] * dj[1] + da[2] * dj[2];
  e[0] += vol * myu * t;
  e[4] += vol * myu * t;
  e[8] += vol * myu * t;
}

// Block row a and the residual of node a from the element's geometry.
__device__ inline void solid3d_rows(int a, double vol, const double d[4][3], const double u[4][3], const double* p,
                                    double* em, double* er) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  for (int j = 0; j < 4; j++) solid3d_block(vol, d[a], d[j], lambda, myu, &em[j * 9]);
  for (int r = 0; r < 3; r++) {
    double x = vol * rho * p[3 + r] * 0.25;
    for (int j = 0; j < 4; j++) {
      const double* e = &em[j * 9];
      x -= e[r * 3 + 0] * u[j][0] + e[r * 3 + 1] * u[j][1] + e[r * 3 + 2] * u[j][2];
    }
    er[r] = x;
  }
}

// solid3d_rows handing each block to put(j, block) as it is made: the same values, each residual value taking the
// same operations in the same order.
template <class Put>
__device__ inline void solid3d_rows_put(int a, double vol, const double d[4][3], const double u[4][3],
                                        const double* p, double* er, Put&& put) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  double x[3];
  for (int r = 0; r < 3; r++) x[r] = vol * rho * p[3 + r] * 0.25;
  for (int j = 0; j < 4; j++) {
    double e[9];
    solid3d_block(vol, d[a], d[j], lambda, myu, e);
    put(j, e);
    for (int r = 0; r < 3; r++) x[r] -= e[r * 3 + 0] * u[j][0] + e[r * 3 + 1] * u[j][1] + e[r * 3 + 2] * u[j][2];
  }
  for (int r = 0; r < 3; r++) er[r] = x[r];
}

__device__ inline void row_solid3d(int a, const double co[4][3], const double u[4][3], const double* p, double* em,
                                   double* er) {
  const double vol = tet_volume(co);
  double d[4][3];
  tet_dldx(d, co);
  solid3d_rows(a, vol, d, u, p, em, er);
}

// src/femeqn/eqn_st_venant.cpp:47-156 (AddElemMatFin_StVenant2D) and :157-272. p = (lambda, myu, rho, gx, gy)
__device__ inline void row_stvk2d(int a, const double co[4][3], const double u[4][3], const double* p, double* em,
                                  double* er) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  const double area = tri_area(co[0], co[1], co[2]);
  double dn[3][2];
  tri_dldx(dn, co);
  double du[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int n = 0; n < 3; n++)
    for (int r = 0; r < 2; r++)
      for (int s = 0; s < 2; s++) du[r][s] += u[n][r] * dn[n][s];
  const double w = area;
  double st[2][2];
  for (int r = 0; r < 2; r++)
    for (int s = 0; s < 2; s++) {
      st[r][s] = 0.5 * (du[r][s] + du[s][r]);
      for (int k = 0; k < 2; k++) st[r][s] += 0.5 * du[k][r] * du[k][s];
    }
  double z[2][2];
  for (int r = 0; r < 2; r++) {
    for (int s = 0; s < 2; s++) z[r][s] = du[r][s];
    z[r][r] += 1.0;
  }
  // d2s[k][kd][id][jd] = dn[k][jd] * z[kd][id]
  double sg[2][2];
  {
    double tr = 0.0;
    for (int r = 0; r < 2; r++) {
      for (int s = 0; s < 2; s++) sg[r][s] = 2.0 * myu * st[r][s];
      tr += st[r][r];
    }
    for (int r = 0; r < 2; r++) sg[r][r] += lambda * tr;
  }
  double fin[2] = {0.0, 0.0};
  for (int kd = 0; kd < 2; kd++)
    for (int r = 0; r < 2; r++)
      for (int s = 0; s < 2; s++) fin[kd] += w * sg[r][s] * (dn[a][s] * z[kd][r]);
  for (int j = 0; j < 3; j++) {
    double* e = &em[j * 4];
    for (int k = 0; k < 4; k++) e[k] = 0.0;
    for (int id = 0; id < 2; id++)
      for (int jd = 0; jd < 2; jd++) {
        {
          double t = 0.0;
          for (int gd = 0; gd < 2; gd++)
            for (int hd = 0; hd < 2; hd++) {
              const double di = dn[a][hd] * z[id][gd];
              t += di * (dn[j][hd] * z[jd][gd]) + di * (dn[j][gd] * z[jd][hd]);
            }
          e[id * 2 + jd] += w * myu * t;
        }
        {
          double t1 = 0.0, t2 = 0.0;
          for (int gd = 0; gd < 2; gd++) {
            t1 += dn[a][gd] * z[id][gd];
            t2 += dn[j][gd] * z[jd][gd];
          }
          e[id * 2 + jd] += w * lambda * t1 * t2;
        }
      }
    {
      double t = 0.0;
      for (int gd = 0; gd < 2; gd++)
        for (int hd = 0; hd < 2; hd++) t += sg[gd][hd] * dn[a][hd] * dn[j][gd];
      for (int id = 0; id < 2; id++) e[id * 2 + id] += w * t;
    }
  }
  const double g[2] = {p[3], p[4]};
  for (int r = 0; r < 2; r++) {
    double fex = 0.0;
    fex += area * rho * g[r] / 3.0;
    er[r] = fex - fin[r];
  }
}

// src/femeqn/eqn_st_venant.cpp:724-834 (SetElemMatFin_StVenant3D) and :838-939. p = (lambda, myu, rho, gx,gy,gz)
__device__ inline void row_stvk3d(int a, const double co[4][3], const double u[4][3], const double* p, double* em,
                                  double* er) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  const double vol = tet_volume(co);
  double dn[4][3];
  tet_dldx(dn, co);
  double du[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int n = 0; n < 4; n++)
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++) du[r][s] += u[n][r] * dn[n][s];
  const double w = vol;
  double sg[6];
  {
    double e[6];
    e[0] = du[0][0] + 0.5 * (du[0][0] * du[0][0] + du[1][0] * du[1][0] + du[2][0] * du[2][0]);
    e[1] = du[1][1] + 0.5 * (du[0][1] * du[0][1] + du[1][1] * du[1][1] + du[2][1] * du[2][1]);
    e[2] = du[2][2] + 0.5 * (du[0][2] * du[0][2] + du[1][2] * du[1][2] + du[2][2] * du[2][2]);
    e[3] = 0.5 * (du[0][1] + du[1][0] + du[0][0] * du[0][1] + du[1][0] * du[1][1] + du[2][0] * du[2][1]);
    e[4] = 0.5 * (du[1][2] + du[2][1] + du[0][1] * du[0][2] + du[1][1] * du[1][2] + du[2][1] * du[2][2]);
    e[5] = 0.5 * (du[2][0] + du[0][2] + du[0][2] * du[0][0] + du[1][2] * du[1][0] + du[2][2] * du[2][0]);
    const double tr = e[0] + e[1] + e[2];
    sg[0] = 2.0 * myu * e[0] + lambda * tr;
    sg[1] = 2.0 * myu * e[1] + lambda * tr;
    sg[2] = 2.0 * myu * e[2] + lambda * tr;
    sg[3] = 2.0 * myu * e[3];
    sg[4] = 2.0 * myu * e[4];
    sg[5] = 2.0 * myu * e[5];
  }
  const double z[3][3] = {{du[0][0] + 1, du[0][1], du[0][2]},
                          {du[1][0], du[1][1] + 1, du[1][2]},
                          {du[2][0], du[2][1], du[2][2] + 1}};
  double d2s[4][3][6];
  for (int k = 0; k < 4; k++)
    for (int kd = 0; kd < 3; kd++) {
      d2s[k][kd][0] = dn[k][0] * z[kd][0];
      d2s[k][kd][1] = dn[k][1] * z[kd][1];
      d2s[k][kd][2] = dn[k][2] * z[kd][2];
      d2s[k][kd][3] = (dn[k][0] * z[kd][1] + dn[k][1] * z[kd][0]);
      d2s[k][kd][4] = (dn[k][1] * z[kd][2] + dn[k][2] * z[kd][1]);
      d2s[k][kd][5] = (dn[k][2] * z[kd][0] + dn[k][0] * z[kd][2]);
    }
  for (int kd = 0; kd < 3; kd++) {
    double t = 0;
    t += sg[0] * d2s[a][kd][0];
    t += sg[1] * d2s[a][kd][1];
    t += sg[2] * d2s[a][kd][2];
    t += sg[3] * d2s[a][kd][3];
    t += sg[4] * d2s[a][kd][4];
    t += sg[5] * d2s[a][kd][5];
    er[kd] = t * w;  // internal force; the residual is formed below
  }
  // material part: the reference fills the pairs (ino > jno) and mirrors them, then the diagonal blocks
  for (int j = 0; j < 4; j++) {
    // (ino, jno) = (a, j) for j < a; (j, a) mirrored for j > a; diagonal formula for j == a
    const int ino = (j <= a) ? a : j;
    const int jno = (j <= a) ? j : a;
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++) {
        const int id = (j <= a) ? r : s;
        const int jd = (j <= a) ? s : r;
        double t = 0.0;
        t += d2s[ino][id][0] * d2s[jno][jd][0];
        t += d2s[ino][id][1] * d2s[jno][jd][1];
        t += d2s[ino][id][2] * d2s[jno][jd][2];
        t += d2s[ino][id][3] * d2s[jno][jd][3] * 0.5;
        t += d2s[ino][id][4] * d2s[jno][jd][4] * 0.5;
        t += d2s[ino][id][5] * d2s[jno][jd][5] * 0.5;
        const double t2 = d2s[ino][id][0] + d2s[ino][id][1] + d2s[ino][id][2];
        const double t3 = d2s[jno][jd][0] + d2s[jno][jd][1] + d2s[jno][jd][2];
        em[j * 9 + r * 3 + s] = w * (myu * 2 * t + lambda * t2 * t3);
      }
  }
  // geometric part
  for (int j = 0; j < 4; j++) {
    double t = 0.0;
    t += sg[0] * dn[a][0] * dn[j][0];
    t += sg[1] * dn[a][1] * dn[j][1];
    t += sg[2] * dn[a][2] * dn[j][2];
    t += sg[3] * (dn[a][0] * dn[j][1] + dn[a][1] * dn[j][0]);
    t += sg[4] * (dn[a][1] * dn[j][2] + dn[a][2] * dn[j][1]);
    t += sg[5] * (dn[a][2] * dn[j][0] + dn[a][0] * dn[j][2]);
    em[j * 9 + 0] += w * t;
    em[j * 9 + 4] += w * t;
    em[j * 9 + 8] += w * t;
  }
  for (int r = 0; r < 3; r++) {
    double fex = 0.0;
    fex += vol * rho * p[3 + r] / 4.0;
    er[r] = fex - er[r];
  }
}

// ---------------------------------------------------------------------------------------------------------
// Element rows of the saved-stiffness loops: the element matrix as those loops write it and the external
// force only (no internal force, the field values are not read).

// AddLinearSystem_Poisson2D_P1 on CLinearSystem_Save, src/femeqn/eqn_poisson.cpp:494-569. p = (alpha, source)
__device__ inline void row_poisson2d_save(int a, const double co[4][3], const double* p, double* em, double* er) {
  const double alpha = p[0], source = p[1];
  const double area = tri_area(co[0], co[1], co[2]);
  double d[3][2];
  tri_dldx(d, co);
  for (int j = 0; j < 3; j++) em[j] = alpha * area * (d[a][0] * d[j][0] + d[a][1] * d[j][1]);
  er[0] = source * area * 0.333333333333333333;
}

// AddLinSys_LinearSolid2D_Static_P1_SaveStiffMat, src/femeqn/eqn_linear_solid2d.cpp:861-960 (its element matrix
// is written in another order than the static loop's) and AddLinSys_LinearSolid3D_Static_P1 on
// CLinearSystem_Save, src/femeqn/eqn_linear_solid3d.cpp:425-521. p = (lambda, myu, rho, g...)
template <int NDIM>
__device__ inline void row_solid_save(int a, const double co[4][3], const double* p, double* em, double* er) {
  constexpr int NNO = NDIM + 1;
  constexpr int BS = NDIM * NDIM;
  const double lambda = p[0], myu = p[1], rho = p[2];
  double d[4][3];
  double w;
  if (NDIM == 2) {
    w = tri_area(co[0], co[1], co[2]);
    double d2[3][2];
    tri_dldx(d2, co);
    for (int k = 0; k < 3; k++) {
      d[k][0] = d2[k][0];
      d[k][1] = d2[k][1];
    }
  } else {
    w = tet_volume(co);
    tet_dldx(d, co);
  }
  for (int j = 0; j < NNO; j++) {
    double* e = &em[j * BS];
    double t = 0.0;
    for (int r = 0; r < NDIM; r++) {
      for (int s = 0; s < NDIM; s++) e[r * NDIM + s] = w * (lambda * d[a][r] * d[j][s] + myu * d[j][r] * d[a][s]);
      t += d[a][r] * d[j][r];
    }
    for (int r = 0; r < NDIM; r++) e[r * NDIM + r] += w * myu * t;
  }
  for (int r = 0; r < NDIM; r++) er[r] = (NDIM == 2) ? w * rho * p[3 + r] * 0.3333333333333333333 : w * rho * p[3 + r] * 0.25;
}

// Lumped-mass saved systems (CLinearSystem_SaveDiaM_Newmark / _NewmarkBeta): the rows of the system matrix; the
// boundary matrix is the stiffness alone, assembled with the saved kinds above.
// AddLinearSystem_Diffusion2D_P1 on CLinearSystem_SaveDiaM_Newmark, src/femeqn/eqn_diffusion.cpp:565-650: row a of
// (gamma dt) eC plus the lumped mass rho area / 3 on the diagonal; the source term with its 14-digit literal.
// p = (rho, alpha, source)
__device__ inline void row_diffusion2d_savedia(int a, const double co[4][3], const double* p, double dt, double gamma,
                                               double* em, double* er) {
  const double rho = p[0], alpha = p[1], source = p[2];
  const double area = tri_area(co[0], co[1], co[2]);
  double d[3][2];
  tri_dldx(d, co);
  const double gdt = gamma * dt;
  for (int j = 0; j < 3; j++) em[j] = gdt * (alpha * area * (d[a][0] * d[j][0] + d[a][1] * d[j][1]));
  em[a] += rho * area / 3.0;
  er[0] = source * area * 0.33333333333333;
}

// AddLinSys_LinearSolid2D_NonStatic_Save_NewmarkBeta_P1, src/femeqn/eqn_linear_solid2d.cpp:1051-1180: row a of
// M + (beta dt dt) K on the diagonal block (M: area rho / 3 on its diagonal, accumulated onto zero as written there)
// and (beta dt dt) K elsewhere, K as the saved static loop writes it; er = the node's gravity (x, y), which the
// reference adds to the node's first dof, x then y. p = (lambda, myu, rho, gx, gy)
__device__ inline void row_solid2d_savedia(int a, const double co[4][3], const double* p, double dt, double beta,
                                           double* em, double* er) {
  row_solid_save<2>(a, co, p, em, er);
  const double m = tri_area(co[0], co[1], co[2]) * p[2] * 0.33333333333333333333333;
  const double c = beta * dt * dt;
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 4; k++) {
      const double kv = c * em[j * 4 + k];
      em[j * 4 + k] = (j == a) ? ((k == 0 || k == 3) ? 0.0 + m : 0.0) + kv : kv;
    }
}

// Newmark-beta rows (AddLinSys_LinearSolid2D_NonStatic_NewmarkBeta_P1, src/femeqn/eqn_linear_solid2d.cpp:1182-1337;
// AddLinSys_LinearSolid3D_NonStatic_NewmarkBeta_P1, src/femeqn/eqn_linear_solid3d.cpp:525-690): stiffness K as
// the saved loops write it, consistent mass M (rho w / 12 in 2D, rho w / 20 in 3D, doubled on the diagonal
// blocks), element matrix M + beta dt^2 K, residual f - K u - M a [- dt K v - dt^2/2 K a: 2D when is_initial,
// 3D always]. p = (lambda, myu, rho, g...); dt, beta, initial from the launch.
template <int NDIM>
__device__ inline void row_solid_newmark(int a, const double co[4][3], const double u[4][3], const double vv[4][3],
                                         const double aa[4][3], const double* p, double dt, double beta, bool initial,
                                         double* em, double* er) {
  constexpr int NNO = NDIM + 1;
  constexpr int BS = NDIM * NDIM;
  const double lambda = p[0], myu = p[1], rho = p[2];
  double d[4][3];
  double w;
  if (NDIM == 2) {
    w = tri_area(co[0], co[1], co[2]);
    double d2[3][2];
    tri_dldx(d2, co);
    for (int k = 0; k < 3; k++) {
      d[k][0] = d2[k][0];
      d[k][1] = d2[k][1];
    }
  } else {
    w = tet_volume(co);
    tet_dldx(d, co);
  }
  // K is formed in em (the element matrix replaces it in place once the residual has used it); M is diagonal per
  // block and is written out as a function, zeros included, so that every product of the loops is kept
  double* K = em;
  for (int j = 0; j < NNO; j++) {
    double t = 0.0;
    for (int r = 0; r < NDIM; r++) {
      // 0.0 + x as the zero-initialised element matrix accumulates it (a -0.0 term becomes +0.0)
      for (int c = 0; c < NDIM; c++)
        K[j * BS + r * NDIM + c] = 0.0 + w * (lambda * d[a][r] * d[j][c] + myu * d[j][r] * d[a][c]);
      t += d[a][r] * d[j][r];
    }
    for (int r = 0; r < NDIM; r++) K[j * BS + r * NDIM + r] += w * myu * t;
  }
  const double mm = (NDIM == 2) ? w * rho * 0.0833333333333333333333333 : w * rho * 0.05;
  const double mm2 = mm + mm;
  auto Mv = [&](int j, int r, int c) { return c != r ? 0.0 : (j == a ? mm2 : mm); };
  auto kdot = [&](int j, int r, const double x[3]) {
    double t = K[j * BS + r * NDIM] * x[0];
    for (int c = 1; c < NDIM; c++) t = t + K[j * BS + r * NDIM + c] * x[c];
    return t;
  };
  auto mdot = [&](int j, int r, const double x[3]) {
    double t = Mv(j, r, 0) * x[0];
    for (int c = 1; c < NDIM; c++) t = t + Mv(j, r, c) * x[c];
    return t;
  };
  for (int r = 0; r < NDIM; r++)
    er[r] = (NDIM == 2) ? w * rho * p[3 + r] * 0.33333333333333333333333333 : w * rho * p[3 + r] * 0.25;
  for (int j = 0; j < NNO; j++)
    for (int r = 0; r < NDIM; r++) er[r] -= kdot(j, r, u[j]);
  for (int j = 0; j < NNO; j++)
    for (int r = 0; r < NDIM; r++) er[r] -= mdot(j, r, aa[j]);
  if (NDIM == 3 || initial) {
    for (int j = 0; j < NNO; j++)
      for (int r = 0; r < NDIM; r++) er[r] -= dt * kdot(j, r, vv[j]);
    for (int j = 0; j < NNO; j++)
      for (int r = 0; r < NDIM; r++) er[r] -= 0.5 * dt * dt * kdot(j, r, aa[j]);
  }
  const double bdt2 = beta * dt * dt;
  for (int j = 0; j < NNO; j++)
    for (int k = 0; k < BS; k++) em[j * BS + k] = Mv(j, k / NDIM, k % NDIM) + bdt2 * K[j * BS + k];
}

// StVK Newmark-beta row (AddLinSys_StVenant3D_NonStatic_NewmarkBeta_P1, src/femeqn/eqn_st_venant.cpp:1109-1242):
// tangent K and internal force from the static element (SetElemMatFin_StVenant3D; its row with density 0 gives
// -f_in exactly), mass rho vol / 20, element matrix M + beta dt^2 K, residual vol rho g / 4 - f_in - M a term by
// term, and when is_initial - (dt K v + dt^2/2 K a) term by term.
__device__ inline void row_stvk3d_newmark(int a, const double co[4][3], const double u[4][3], const double vv[4][3],
                                          const double aa[4][3], const double* p, double dt, double beta, bool initial,
                                          double* em, double* er) {
  const double p0[6] = {p[0], p[1], 0.0, p[3], p[4], p[5]};
  double fin[3];
  double* K = em;  // the tangent, replaced in place by M + beta dt^2 K once the residual has used it
  row_stvk3d(a, co, u, p0, K, fin);  // fin = -f_in
  const double rho = p[2];
  const double vol = tet_volume(co);
  const double mm = rho * vol * 0.05;
  const double mm2 = mm + mm;
  auto Mv = [&](int j, int r, int c) { return c != r ? 0.0 : (j == a ? mm2 : mm); };
  for (int r = 0; r < 3; r++) {
    double x = vol * rho * p[3 + r] * 0.25 + fin[r];
    for (int j = 0; j < 4; j++)
      for (int c = 0; c < 3; c++) x -= Mv(j, r, c) * aa[j][c];
    if (initial)
      for (int j = 0; j < 4; j++)
        for (int c = 0; c < 3; c++)
          x -= dt * K[j * 9 + r * 3 + c] * vv[j][c] + 0.5 * dt * dt * K[j * 9 + r * 3 + c] * aa[j][c];
    er[r] = x;
  }
  for (int j = 0; j < 4; j++)
    for (int k = 0; k < 9; k++) em[j * 9 + k] = Mv(j, k / 3, k % 3) + beta * dt * dt * K[j * 9 + k];
}

// StVK Newmark-beta row in 2D (AddLinSys_StVenant2D_NonStatic_NewmarkBeta_P1, src/femeqn/eqn_st_venant.cpp:
// 314-460): as in 3D, with mass rho area / 12 and the external force area rho g / 3 added to zero.
__device__ inline void row_stvk2d_newmark(int a, const double co[4][3], const double u[4][3], const double vv[4][3],
                                          const double aa[4][3], const double* p, double dt, double beta, bool initial,
                                          double* em, double* er) {
  const double p0[5] = {p[0], p[1], 0.0, p[3], p[4]};
  double K[3 * 4], fin[2];
  row_stvk2d(a, co, u, p0, K, fin);  // fin = -f_in
  const double rho = p[2];
  const double area = tri_area(co[0], co[1], co[2]);
  const double mm = rho * area / 12.0;
  double M[3][4];
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 4; k++) M[j][k] = (k % 3 == 0) ? mm : 0.0;
  for (int r = 0; r < 2; r++) M[a][r * 2 + r] += mm;
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 4; k++) em[j * 4 + k] = M[j][k] + beta * dt * dt * K[j * 4 + k];
  for (int r = 0; r < 2; r++) {
    double x = (0.0 + area * rho * p[3 + r] / 3.0) + fin[r];
    for (int j = 0; j < 3; j++)
      for (int c = 0; c < 2; c++) x -= M[j][r * 2 + c] * aa[j][c];
    if (initial)
      for (int j = 0; j < 3; j++)
        for (int c = 0; c < 2; c++)
          x -= dt * K[j * 4 + r * 2 + c] * vv[j][c] + 0.5 * dt * dt * K[j * 4 + r * 2 + c] * aa[j][c];
    er[r] = x;
  }
}

// StVK backward-Euler row (AddLinSys_StVenant2D_NonStatic_BackwardEular_P1, src/femeqn/eqn_st_venant.cpp:509-664):
// the static StVK tangent and internal force, mass rho area / 12, element matrix M + dt^2 K, residual
// (f_ex - f_in) dt, then - K v dt dt term by term (is_initial) or - M (v - v_prev) term by term.
__device__ inline void row_stvk2d_beuler(int a, const double co[4][3], const double u[4][3], const double vv[4][3],
                                         const double v0[4][3], const double* p, double dt, bool initial, double* em,
                                         double* er) {
  const double p0[5] = {p[0], p[1], 0.0, p[3], p[4]};
  double fin[2];
  double* K = em;
  row_stvk2d(a, co, u, p0, K, fin);  // fin = -f_in
  const double rho = p[2];
  const double area = tri_area(co[0], co[1], co[2]);
  const double mm = rho * area / 12.0;
  const double mm2 = mm + mm;
  auto Mv = [&](int j, int r, int c) { return c != r ? 0.0 : (j == a ? mm2 : mm); };
  for (int r = 0; r < 2; r++) {
    double x = ((0.0 + area * rho * p[3 + r] * 0.33333333333333333333) + fin[r]) * dt;
    if (initial) {
      for (int j = 0; j < 3; j++)
        for (int c = 0; c < 2; c++) x -= K[j * 4 + r * 2 + c] * vv[j][c] * dt * dt;
    } else {
      for (int j = 0; j < 3; j++)
        for (int c = 0; c < 2; c++) x -= Mv(j, r, c) * (vv[j][c] - v0[j][c]);
    }
    er[r] = x;
  }
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 4; k++) em[j * 4 + k] = Mv(j, k / 2, k % 2) + dt * dt * K[j * 4 + k];
}

// Backward-Euler row (AddLinSys_LinearSolid2D_NonStatic_BackwardEular_P1, src/femeqn/eqn_linear_solid2d.cpp:
// 1338-1497; unknown: the velocity increment): the Newmark loop's K and M, element matrix M + dt^2 K, residual
// dt f - dt K u, then - dt^2 K v when is_initial, else - M (v - v_prev). aa holds v_prev here.
__device__ inline void row_solid2d_beuler(int a, const double co[4][3], const double u[4][3], const double vv[4][3],
                                          const double v0[4][3], const double* p, double dt, bool initial,
                                          double* em, double* er) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  const double w = tri_area(co[0], co[1], co[2]);
  double d[3][2];
  tri_dldx(d, co);
  double K[3][4], M[3][4];
  for (int j = 0; j < 3; j++) {
    double t = 0.0;
    for (int r = 0; r < 2; r++) {
      for (int c = 0; c < 2; c++) K[j][r * 2 + c] = 0.0 + w * (lambda * d[a][r] * d[j][c] + myu * d[j][r] * d[a][c]);
      t += d[a][r] * d[j][r];
    }
    for (int r = 0; r < 2; r++) K[j][r * 2 + r] += w * myu * t;
  }
  const double mm = w * rho * 0.0833333333333333333333333;
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 4; k++) M[j][k] = (k % 3 == 0) ? mm : 0.0;
  for (int r = 0; r < 2; r++) M[a][r * 2 + r] += mm;
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 4; k++) em[j * 4 + k] = M[j][k] + dt * dt * K[j][k];
  for (int r = 0; r < 2; r++) er[r] = w * rho * p[3 + r] * 0.33333333333333333333333333 * dt;
  for (int j = 0; j < 3; j++)
    for (int r = 0; r < 2; r++) er[r] -= dt * (K[j][r * 2] * u[j][0] + K[j][r * 2 + 1] * u[j][1]);
  if (initial) {
    for (int j = 0; j < 3; j++)
      for (int r = 0; r < 2; r++) er[r] -= dt * dt * (K[j][r * 2] * vv[j][0] + K[j][r * 2 + 1] * vv[j][1]);
  } else {
    for (int j = 0; j < 3; j++)
      for (int r = 0; r < 2; r++)
        er[r] -= M[j][r * 2] * (vv[j][0] - v0[j][0]) + M[j][r * 2 + 1] * (vv[j][1] - v0[j][1]);
  }
}

// ---------------------------------------------------------------------------------------------------------

// The inputs of one element (nodes, coordinates, field values) and its block row a (em: NNO blocks, er: B values)
// by the kind's element routine.
template <int KIND, int NDIM, int B>
struct ElemIn {
  static constexpr int NNO = NDIM + 1;
  int no[NNO];
  double co[4][3], u[4][3], val[4], vval[4], vv[4][3], aa[4][3];
  const double* p;
};

// The node values of an element whose nodes are in E.no.
template <int KIND, int NDIM, int B>
__device__ inline void load_elem_nodes(const AsmArgs& A, ElemIn<KIND, NDIM, B>& E) {
  constexpr int NNO = NDIM + 1;
  constexpr bool NM = (KIND == DFM_SOLID2D_NEWMARK || KIND == DFM_SOLID3D_NEWMARK || KIND == DFM_SOLID2D_BEULER ||
                       KIND == DFM_STVK3D_NEWMARK || KIND == DFM_STVK2D_NEWMARK || KIND == DFM_STVK2D_BEULER);
  for (int c = 0; c < NNO; c++) {
    const long long nc = E.no[c];
    for (int d = 0; d < 3; d++) E.co[c][d] = (d < NDIM) ? A.coords[nc * NDIM + d] : 0.0;
    for (int d = 0; d < 3; d++) E.u[c][d] = (A.u && d < A.len) ? A.u[nc * A.len + d] : 0.0;
    E.val[c] = A.u ? A.u[nc * A.len] : 0.0;
    E.vval[c] = A.v ? A.v[nc * A.len] : 0.0;
    if (NM)
      for (int d = 0; d < 3; d++) {
        E.vv[c][d] = (d < B) ? A.v[nc * B + d] : 0.0;
        E.aa[c][d] = (d < B) ? A.acc[nc * B + d] : 0.0;
      }
  }
}

template <int KIND, int NDIM, int B>
__device__ inline void load_elem(const AsmArgs& A, int e, ElemIn<KIND, NDIM, B>& E) {
  constexpr int NNO = NDIM + 1;
  for (int c = 0; c < NNO; c++) E.no[c] = A.lnods[(long long)e * NNO + c];
  load_elem_nodes<KIND, NDIM, B>(A, E);
  E.p = &A.params[(long long)e * A.pstride];
}

template <int KIND, int NDIM, int B, bool SAVE>
__device__ inline void elem_row(const AsmArgs& A, const ElemIn<KIND, NDIM, B>& E, int a, double* em, double* er) {
  const double* p = E.p;
  if (SAVE) {
    if (KIND == DFM_POISSON) row_poisson2d_save(a, E.co, p, em, er);
    if (KIND == DFM_SOLID2D || KIND == DFM_SOLID3D) row_solid_save<NDIM>(a, E.co, p, em, er);
  } else {
    if (KIND == DFM_POISSON) row_poisson<NDIM>(a, E.co, E.val, p, em, er);
    if (KIND == DFM_DIFFUSION2D) row_diffusion2d(a, E.co, E.val, E.vval, p, A.glob0, A.glob1, em, er);
    if (KIND == DFM_SOLID2D) row_solid2d(a, E.co, E.u, p, em, er);
    if (KIND == DFM_SOLID3D) row_solid3d(a, E.co, E.u, p, em, er);
    if (KIND == DFM_STVK2D) row_stvk2d(a, E.co, E.u, p, em, er);
    if (KIND == DFM_STVK3D) row_stvk3d(a, E.co, E.u, p, em, er);
    if (KIND == DFM_SOLID2D_NEWMARK || KIND == DFM_SOLID3D_NEWMARK)
      row_solid_newmark<NDIM>(a, E.co, E.u, E.vv, E.aa, p, A.glob0, A.glob2, A.glob3 != 0.0, em, er);
    if (KIND == DFM_STVK3D_NEWMARK)
      row_stvk3d_newmark(a, E.co, E.u, E.vv, E.aa, p, A.glob0, A.glob2, A.glob3 != 0.0, em, er);
    if (KIND == DFM_STVK2D_NEWMARK)
      row_stvk2d_newmark(a, E.co, E.u, E.vv, E.aa, p, A.glob0, A.glob2, A.glob3 != 0.0, em, er);
    if (KIND == DFM_STVK2D_BEULER) row_stvk2d_beuler(a, E.co, E.u, E.vv, E.aa, p, A.glob0, A.glob1 != 0.0, em, er);
    if (KIND == DFM_SOLID2D_BEULER) row_solid2d_beuler(a, E.co, E.u, E.vv, E.aa, p, A.glob0, A.glob1 != 0.0, em, er);
    if (KIND == KIND_DIFFUSION2D_SAVEDIA) row_diffusion2d_savedia(a, E.co, p, A.glob0, A.glob1, em, er);
    if (KIND == KIND_SOLID2D_SAVEDIA) row_solid2d_savedia(a, E.co, p, A.glob0, A.glob2, em, er);
  }
}

// Per-row gather (the fallback where the element store is not built): one thread per device row recomputes the
// rows of its incident elements and adds them, in ascending element order, into the zeroed matrix and residual.
template <int KIND, int NDIM, int B, bool SAVE>
__global__ void __launch_bounds__(128) k_assemble(AsmArgs A) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 >= A.n) return;
  constexpr int NNO = NDIM + 1;
  constexpr int BS = B * B;
  const int i1 = A.order ? A.order[i0] : i0;
  const unsigned q0 = A.rowptr[i0], q1 = A.rowptr[i0 + 1];
  double* dia = &A.dia[(long long)i0 * BS];
  double* res = &A.res[(long long)i0 * B];
  double racc[B];
  for (int d = 0; d < B; d++) racc[d] = res[d];
  for (int k = A.elsup_ptr[i1]; k < A.elsup_ptr[i1 + 1]; k++) {
    const int e = A.elsup[k];
    ElemIn<KIND, NDIM, B> E;
    load_elem<KIND, NDIM, B>(A, e, E);
    int a = 0;
    for (int c = 0; c < NNO; c++)
      if (E.no[c] == i1) a = c;
    double em[NNO * BS], er[B];
    elem_row<KIND, NDIM, B, SAVE>(A, E, a, em, er);
    for (int c = 0; c < NNO; c++) {
      const double* blk = &em[c * BS];
      if (c == a) {
        for (int t = 0; t < BS; t++) dia[t] += blk[t];
      } else {
        const unsigned j0 = (unsigned)(A.inv ? A.inv[E.no[c]] : E.no[c]);
        for (unsigned q = q0; q < q1; q++) {
          if (A.colind[q] == j0) {
            double* o = &A.crs[(long long)q * BS];
            for (int t = 0; t < BS; t++) o[t] += blk[t];
            break;
          }
        }
      }
    }
    if (KIND == KIND_SOLID2D_SAVEDIA) {
      // both gravity components of the element go to the node's first dof, one after the other
      racc[0] += er[0];
      racc[0] += er[1];
    } else {
      for (int d = 0; d < B; d++) racc[d] += er[d];
    }
  }
  // surface loads (a reference extension): added after the element contributions, in face order
  if (A.load_ptr) {
    for (int k = A.load_ptr[i1]; k < A.load_ptr[i1 + 1]; k++)
      for (int d = 0; d < B; d++) racc[d] += A.load_scale * A.loads[(long long)k * B + d];
  }
  for (int d = 0; d < B; d++) res[d] = racc[d];
}

// Element store: one thread per element computes every block row of it once (the element routine is called per
// row a, as the row-gather kernel calls it); row a of element e (NNO blocks, then its B residual values) goes to
// ebuf[(a*nelem + e)*STR ...]. A workgroup of 64 elements stages each row a in LDS and writes the 64 rows, which
// are contiguous, with coalesced stores.
template <int KIND, int NDIM, int B, bool SAVE>
__global__ void __launch_bounds__(64) k_elem_store(AsmArgs A, int nelem, double* __restrict__ ebuf) {
  constexpr int NNO = NDIM + 1;
  constexpr int BS = B * B;
  constexpr int STR = NNO * BS + B;
  __shared__ double sh[64 * STR];
  const int e0 = (int)blockIdx.x * 64, e = e0 + (int)threadIdx.x;
  const int cnt = min(64, nelem - e0);
  ElemIn<KIND, NDIM, B> E;
  if (e < nelem) load_elem<KIND, NDIM, B>(A, e, E);
  for (int a = 0; a < NNO; a++) {
    if (e < nelem) {
      double em[NNO * BS], er[B];
      elem_row<KIND, NDIM, B, SAVE>(A, E, a, em, er);
      double* o = &sh[threadIdx.x * STR];
#pragma unroll
      for (int t = 0; t < NNO * BS; t++) o[t] = em[t];
#pragma unroll
      for (int d = 0; d < B; d++) o[NNO * BS + d] = er[d];
    }
    __syncthreads();
    double* out = &ebuf[((size_t)a * nelem + e0) * STR];
    for (int t = (int)threadIdx.x; t < cnt * STR; t += 64) out[t] = sh[t];
    __syncthreads();
  }
}

// Off-diagonal blocks from the element store: one thread per matrix entry q sums the blocks (a, c) of the elements
// holding both of its nodes, in ascending element order (the list eidx[eptr[q] .. eptr[q+1]) of (e*NNO + a)*NNO + c),
// and writes the block once: the same sums in the same order as the reference's element loop merges them.
template <int B, int NNO>
__global__ void __launch_bounds__(256) k_gather_blocks(long long ncrs, int nelem, const unsigned* __restrict__ eptr,
                                                      const unsigned* __restrict__ eidx,
                                                      const double* __restrict__ ebuf, double* __restrict__ crs) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= ncrs) return;
  constexpr int BS = B * B;
  constexpr int STR = NNO * BS + B;
  double acc[BS];
#pragma unroll
  for (int t = 0; t < BS; t++) acc[t] = 0.0;
  for (unsigned k = eptr[q]; k < eptr[q + 1]; k++) {
    const unsigned idx = eidx[k];
    const unsigned ea = idx / NNO, c = idx - ea * NNO, e = ea / NNO, a = ea - e * NNO;
    const double* src = &ebuf[((size_t)a * nelem + e) * STR + c * BS];
#pragma unroll
    for (int t = 0; t < BS; t++) acc[t] += src[t];
  }
#pragma unroll
  for (int t = 0; t < BS; t++) crs[q * BS + t] = acc[t];
}

// Diagonal blocks and residual from the element store: one thread per device row sums its incident elements' blocks
// (a, a) and residual rows in ascending element order (eloc: e*NNO + a per node-element incidence), then the surface
// loads in face order.
template <int B, int NNO>
__global__ void __launch_bounds__(256) k_gather_rows(int n, int nelem, const int* __restrict__ order,
                                                    const int* __restrict__ elsup_ptr, const unsigned* __restrict__ eloc,
                                                    const double* __restrict__ ebuf, const int* __restrict__ load_ptr,
                                                    const double* __restrict__ loads, double load_scale,
                                                    double* __restrict__ dia, double* __restrict__ res) {
  const int i0 = blockIdx.x * 256 + threadIdx.x;
  if (i0 >= n) return;
  constexpr int BS = B * B;
  constexpr int STR = NNO * BS + B;
  const int i1 = order ? order[i0] : i0;
  double d[BS], r[B];
#pragma unroll
  for (int t = 0; t < BS; t++) d[t] = 0.0;
#pragma unroll
  for (int t = 0; t < B; t++) r[t] = 0.0;
  for (int k = elsup_ptr[i1]; k < elsup_ptr[i1 + 1]; k++) {
    const unsigned ea = eloc[k], e = ea / NNO, a = ea - e * NNO;
    const double* src = &ebuf[((size_t)a * nelem + e) * STR];
#pragma unroll
    for (int t = 0; t < BS; t++) d[t] += src[a * BS + t];
#pragma unroll
    for (int t = 0; t < B; t++) r[t] += src[NNO * BS + t];
  }
  if (load_ptr)
    for (int k = load_ptr[i1]; k < load_ptr[i1 + 1]; k++)
#pragma unroll
      for (int t = 0; t < B; t++) r[t] += load_scale * loads[(long long)k * B + t];
#pragma unroll
  for (int t = 0; t < BS; t++) dia[(size_t)i0 * BS + t] = d[t];
#pragma unroll
  for (int t = 0; t < B; t++) res[(size_t)i0 * B + t] = r[t];
}

// Linear elasticity stores per element only its geometry, measure and shape-function gradients (1 + NNO*NDIM
// doubles), and the residual row of each of its nodes (NNO*NDIM doubles, from the element routine's full rows,
// computed once per element): 25 doubles in 3D against 156 for the element rows. The gathers compute each block
// from the geometry with the element routine's own block function, so the values are those of the full store, bit
// for bit.
template <int NDIM>
struct SolidStore {
  static constexpr int NNO = NDIM + 1, GS = 1 + NNO * NDIM, STR = GS + NNO * NDIM;
};

template <int NDIM>
__global__ void __launch_bounds__(64) k_geom_store(AsmArgs A, int nelem, double* __restrict__ gbuf) {
  constexpr int NNO = NDIM + 1, B = NDIM, BS = B * B, GS = SolidStore<NDIM>::GS, STR = SolidStore<NDIM>::STR;
  __shared__ double sh[64 * STR];
  const int e0 = (int)blockIdx.x * 64, e = e0 + (int)threadIdx.x;
  const int cnt = min(64, nelem - e0);
  if (e < nelem) {
    double co[4][3], u[4][3];
    for (int c = 0; c < NNO; c++) {
      const long long nc = A.lnods[(long long)e * NNO + c];
      for (int d = 0; d < 3; d++) co[c][d] = d < NDIM ? A.coords[nc * NDIM + d] : 0.0;
      for (int d = 0; d < 3; d++) u[c][d] = (A.u && d < A.len) ? A.u[nc * A.len + d] : 0.0;
    }
    const double* p = &A.params[(long long)e * A.pstride];
    double* o = &sh[threadIdx.x * STR];
    double em[NNO * BS], er[B];
    if (NDIM == 3) {
      double dd[4][3];
      const double vol = tet_volume(co);
      tet_dldx(dd, co);
      o[0] = vol;
      for (int c = 0; c < 4; c++)
        for (int k = 0; k < 3; k++) o[1 + c * 3 + k] = dd[c][k];
      for (int a = 0; a < 4; a++) {
        solid3d_rows(a, vol, dd, u, p, em, er);
        for (int k = 0; k < 3; k++) o[GS + a * 3 + k] = er[k];
      }
    } else {
      double dd[3][2];
      const double area = tri_area(co[0], co[1], co[2]);
      tri_dldx(dd, co);
      o[0] = area;
      for (int c = 0; c < 3; c++)
        for (int k = 0; k < 2; k++) o[1 + c * 2 + k] = dd[c][k];
      for (int a = 0; a < 3; a++) {
        solid2d_rows(a, area, dd, u, p, em, er);
        for (int k = 0; k < 2; k++) o[GS + a * 2 + k] = er[k];
      }
    }
  }
  __syncthreads();
  double* out = &gbuf[(size_t)e0 * STR];
  for (int t = (int)threadIdx.x; t < cnt * STR; t += 64) out[t] = sh[t];
}

// block (a, c) of element e from its stored geometry (only the gradients of nodes a and c are read)
template <int NDIM>
__device__ inline void solid_block_stored(const double* __restrict__ g, unsigned a, unsigned c, const double* p,
                                          double* blk) {
  double da[NDIM], dc[NDIM];
#pragma unroll
  for (int k = 0; k < NDIM; k++) {
    da[k] = g[1 + a * NDIM + k];
    dc[k] = g[1 + c * NDIM + k];
  }
  if constexpr (NDIM == 3) solid3d_block(g[0], da, dc, p[0], p[1], blk);
  else solid2d_block(g[0], da, dc, p[0], p[1], blk);
}

// The gathers load the data of up to four contributions before adding any (independent loads in flight together);
// the sums still run in list order.
template <int NDIM>
__global__ void __launch_bounds__(256) k_gather_blocks_solid(long long ncrs, const unsigned* __restrict__ eptr,
                                                            const unsigned* __restrict__ eidx,
                                                            const double* __restrict__ gbuf,
                                                            const double* __restrict__ params, int pstride,
                                                            double* __restrict__ crs) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= ncrs) return;
  constexpr int NNO = NDIM + 1, B = NDIM, BS = B * B, STR = SolidStore<NDIM>::STR;
  double acc[BS];
#pragma unroll
  for (int t = 0; t < BS; t++) acc[t] = 0.0;
  const unsigned k0 = eptr[q], k1 = eptr[q + 1];
  for (unsigned k = k0; k < k1; k += 4) {
    double blk[4][BS];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (k + u < k1) {
        const unsigned idx = eidx[k + u];
        const unsigned ea = idx / NNO, c = idx - ea * NNO, e = ea / NNO, a = ea - e * NNO;
        solid_block_stored<NDIM>(&gbuf[(size_t)e * STR], a, c, &params[(size_t)e * pstride], blk[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (k + u < k1)
#pragma unroll
        for (int t = 0; t < BS; t++) acc[t] += blk[u][t];
  }
#pragma unroll
  for (int t = 0; t < BS; t++) crs[q * BS + t] = acc[t];
}

template <int NDIM>
__global__ void __launch_bounds__(256) k_gather_rows_solid(AsmArgs A, const unsigned* __restrict__ eloc,
                                                          const double* __restrict__ gbuf) {
  const int i0 = blockIdx.x * 256 + threadIdx.x;
  if (i0 >= A.n) return;
  constexpr int NNO = NDIM + 1, B = NDIM, BS = B * B, GS = SolidStore<NDIM>::GS, STR = SolidStore<NDIM>::STR;
  const int i1 = A.order ? A.order[i0] : i0;
  double dg[BS], r[B];
#pragma unroll
  for (int t = 0; t < BS; t++) dg[t] = 0.0;
#pragma unroll
  for (int t = 0; t < B; t++) r[t] = 0.0;
  const int k0 = A.elsup_ptr[i1], k1 = A.elsup_ptr[i1 + 1];
  for (int k = k0; k < k1; k += 4) {
    double blk[4][BS], er[4][B];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (k + u < k1) {
        const unsigned ea = eloc[k + u], e = ea / NNO, a = ea - e * NNO;
        const double* g = &gbuf[(size_t)e * STR];
        solid_block_stored<NDIM>(g, a, a, &A.params[(long long)e * A.pstride], blk[u]);
#pragma unroll
        for (int t = 0; t < B; t++) er[u][t] = g[GS + a * B + t];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (k + u < k1) {
#pragma unroll
        for (int t = 0; t < BS; t++) dg[t] += blk[u][t];
#pragma unroll
        for (int t = 0; t < B; t++) r[t] += er[u][t];
      }
  }
  if (A.load_ptr)
    for (int k = A.load_ptr[i1]; k < A.load_ptr[i1 + 1]; k++)
#pragma unroll
      for (int t = 0; t < B; t++) r[t] += A.load_scale * A.loads[(long long)k * B + t];
#pragma unroll
  for (int t = 0; t < BS; t++) A.dia[(size_t)i0 * BS + t] = dg[t];
#pragma unroll
  for (int t = 0; t < B; t++) A.res[(size_t)i0 * B + t] = r[t];
}

// Linear elasticity without any store: one group of LPR lanes per node, nodes taken in their original numbering (so
// that the groups running together share elements in cache) and spread over the XCDs in contiguous ranges. Per chunk
// of up to CH incident elements (ascending index) each of CH lanes computes one element's block row and residual row
// for the node with the element routine, into LDS; then each lane, holding one entry of the node's block row (its
// off-diagonal entries in pattern order, then the diagonal block), runs through the chunk in order and adds the
// element's block for its column where the element holds that column. Every entry and the residual receive the
// element routine's values in ascending element order, as the reference's element loop merges them, and every entry
// of the matrix and residual is written once (entries no element touches become zero). Needs row length + 1 <= LPR;
// the workgroup is one wave, so its barriers only order the LDS traffic. CH is half of LPR from 16 lanes on, which
// keeps the LDS per wave near 11 KB (3D) so that the register budget, not the LDS, bounds the waves per CU.
// The field stride of the row-group records (at least nrec doubles): the one whose reads of one chunk position by
// all lanes (group g, block c of BS doubles) fall on the fewest equal banks of 32 four-byte banks.
constexpr int rowgroup_field_stride(int nrec, int BS, int NNO, int G) {
  int best = nrec, best_cost = 1 << 30;
  for (int f = nrec; f < nrec + 32; f++) {
    int hits[32] = {};
    for (int c = 0; c < NNO; c++)
      for (int g = 0; g < G; g++)
        for (int w = 0; w < 2; w++) hits[(c * BS * f * 2 + g * 2 + w) % 32]++;
    int cost = 0;
    for (int b = 0; b < 32; b++) cost = hits[b] > cost ? hits[b] : cost;
    if (cost < best_cost) {
      best_cost = cost;
      best = f;
    }
  }
  return best;
}

// KIND: linear elasticity (its rows made block by block, solid*_rows_put) or another vector kind with one degree of
// freedom per dimension (Newmark, backward Euler, StVK), whose rows come whole from elem_row.
template <int KIND, int NDIM, int LPR>
__global__ void __launch_bounds__(64) k_asm_solid_rowgroups(AsmArgs A) {
  constexpr bool kLinear = KIND == DFM_SOLID2D || KIND == DFM_SOLID3D;
  constexpr int NNO = NDIM + 1, B = NDIM, BS = B * B, G = 64 / LPR, CH = LPR >= 16 ? LPR / 2 : LPR;
  constexpr int FE = NNO * BS, NF = FE + B;  // the node's blocks (a, 0..NNO-1), then its residual row
  constexpr int FSTR = rowgroup_field_stride(CH * G, BS, NNO, G);
  __shared__ double rec[NF * FSTR];
  __shared__ int nod[NNO * CH * G];  // element nodes
  const unsigned per = gridDim.x / 8;     // the grid is a multiple of 8; workgroup w runs on XCD w % 8
  const unsigned lb = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const int g = (int)threadIdx.x / LPR, s = (int)threadIdx.x % LPR;
  const long long i1l = (long long)lb * G + g;
  const bool active = i1l < A.n;
  const int i1 = active ? (int)i1l : 0;
  // the row in the original pattern (column in original numbering, device entry), so that no load waits for the
  // ordering maps; the device row is needed only for the diagonal block and the residual
  int k0 = 0, k1 = 0, col = -1;
  unsigned q = 0;
  bool diag = false;
  if (active) {
    const unsigned q0 = A.rowptr0[i1], len = A.rowptr0[i1 + 1] - q0;
    if ((unsigned)s < len) {
      col = (int)A.colind0[q0 + s];
      q = A.permq[q0 + s];
    } else if ((unsigned)s == len) {
      col = i1;
      diag = true;
    }
    k0 = A.elsup_ptr[i1];
    k1 = A.elsup_ptr[i1 + 1];
  }
  // every lane of the wave runs as many chunks as its longest element list needs
  int nch = (k1 - k0 + CH - 1) / CH;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) nch = max(nch, __shfl_xor(nch, off));
  auto at = [&](int f, int m) { return f * FSTR + m * G + g; };
  auto ati = [&](int f, int m) { return (f * CH + m) * G + g; };
  double acc[BS], r[B];
#pragma unroll
  for (int t = 0; t < BS; t++) acc[t] = 0.0;
#pragma unroll
  for (int t = 0; t < B; t++) r[t] = 0.0;
  // a staging lane loads the element index and nodes (per incidence, elnod) of its next chunk one chunk ahead, so
  // a chunk waits only for coordinates, field values and parameters
  const int4* elnod = reinterpret_cast<const int4*>(A.elnod);
  int en = -1;
  int4 nn = make_int4(0, 0, 0, 0);
  if (s < CH && k0 + s < k1) {
    en = A.elsup[k0 + s];
    nn = elnod[k0 + s];
  }
  for (int ch = 0; ch < nch; ch++) {
    const int kb = k0 + ch * CH;
    const int e = en;
    const int no[4] = {nn.x, nn.y, nn.z, nn.w};
    en = -1;
    if (s < CH && kb + CH + s < k1) {
      en = A.elsup[kb + CH + s];
      nn = elnod[kb + CH + s];
    }
    __syncthreads();
    if (e >= 0 && !kLinear) {
      // the kind's element row as the element store computes it (same loads, same routine)
      ElemIn<KIND, NDIM, B> E;
      int a = 0;
#pragma unroll
      for (int c = 0; c < NNO; c++) {
        E.no[c] = no[c];
        nod[ati(c, s)] = no[c];
        if (no[c] == i1) a = c;
      }
      load_elem_nodes<KIND, NDIM, B>(A, E);
      E.p = &A.params[(long long)e * A.pstride];
      double em[NNO * BS], er[B];
      elem_row<KIND, NDIM, B, false>(A, E, a, em, er);
#pragma unroll
      for (int t = 0; t < NNO * BS; t++) rec[at(t, s)] = em[t];
#pragma unroll
      for (int t = 0; t < B; t++) rec[at(FE + t, s)] = er[t];
    }
    if (e >= 0 && kLinear) {
      double co[4][3], u[4][3];
      int a = 0;
#pragma unroll
      for (int c = 0; c < NNO; c++) {
        const int nc = no[c];
        nod[ati(c, s)] = nc;
        if (nc == i1) a = c;
#pragma unroll
        for (int d = 0; d < 3; d++) co[c][d] = d < NDIM ? A.coords[(long long)nc * NDIM + d] : 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) u[c][d] = (A.u && d < A.len) ? A.u[(long long)nc * A.len + d] : 0.0;
      }
      const double* p = &A.params[(long long)e * A.pstride];
      auto put = [&](int j, const double* blk) {
#pragma unroll
        for (int t = 0; t < BS; t++) rec[at(j * BS + t, s)] = blk[t];
      };
      double er[B];
      if constexpr (NDIM == 3) {
        double dd[4][3];
        const double vol = tet_volume(co);
        tet_dldx(dd, co);
        solid3d_rows_put(a, vol, dd, u, p, er, put);
      } else {
        double dd[3][2];
        const double area = tri_area(co[0], co[1], co[2]);
        tri_dldx(dd, co);
        solid2d_rows_put(a, area, dd, u, p, er, put);
      }
#pragma unroll
      for (int t = 0; t < B; t++) rec[at(FE + t, s)] = er[t];
    }
    __syncthreads();
    const int cnt = min(CH, k1 - kb);
    for (int m = 0; m < cnt; m++) {
      int c = -1;
#pragma unroll
      for (int t = 0; t < NNO; t++)
        if (nod[ati(t, m)] == col) c = t;
      if (c < 0) continue;
#pragma unroll
      for (int t = 0; t < BS; t++) acc[t] += rec[at(c * BS + t, m)];
      if (diag)
#pragma unroll
        for (int t = 0; t < B; t++) r[t] += rec[at(FE + t, m)];
    }
  }
  if (!active || col < 0) return;
  if (!diag) {
#pragma unroll
    for (int t = 0; t < BS; t++) A.crs[(size_t)q * BS + t] = acc[t];
    return;
  }
  if (A.load_ptr)
    for (int k = A.load_ptr[i1]; k < A.load_ptr[i1 + 1]; k++)
#pragma unroll
      for (int t = 0; t < B; t++) r[t] += A.load_scale * A.loads[(long long)k * B + t];
  const int i0 = A.inv ? A.inv[i1] : i1;
#pragma unroll
  for (int t = 0; t < BS; t++) A.dia[(size_t)i0 * BS + t] = acc[t];
#pragma unroll
  for (int t = 0; t < B; t++) A.res[(size_t)i0 * B + t] = r[t];
}

template <int KIND, int NDIM>
static hipError_t launch_solid_rowgroups(const AsmArgs& a, int lpr, hipStream_t st) {
  if (!a.rowptr0 || !a.colind0 || !a.permq || !a.elnod) return hipErrorInvalidValue;
  const long long groups = 64 / lpr;
  const long long nblk = (a.n + groups - 1) / groups;
  const dim3 grid((unsigned)((nblk + 7) / 8 * 8));
  if (lpr == 8) hipLaunchKernelGGL((k_asm_solid_rowgroups<KIND, NDIM, 8>), grid, dim3(64), 0, st, a);
  else if (lpr == 16) hipLaunchKernelGGL((k_asm_solid_rowgroups<KIND, NDIM, 16>), grid, dim3(64), 0, st, a);
  else if (lpr == 32) hipLaunchKernelGGL((k_asm_solid_rowgroups<KIND, NDIM, 32>), grid, dim3(64), 0, st, a);
  else if (lpr == 64) hipLaunchKernelGGL((k_asm_solid_rowgroups<KIND, NDIM, 64>), grid, dim3(64), 0, st, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int KIND, int NDIM, int B, bool SAVE>
static hipError_t assemble_kind(const AsmArgs& a, const ElemStore* es, hipStream_t st) {
  constexpr int NNO = NDIM + 1;
  // the lumped-mass saved kinds add two force values per element row to one dof: the per-row gather only
  if (!es || KIND == KIND_DIFFUSION2D_SAVEDIA || KIND == KIND_SOLID2D_SAVEDIA) {
    const int nt = 128;
    hipLaunchKernelGGL((k_assemble<KIND, NDIM, B, SAVE>), dim3((a.n + nt - 1) / nt), dim3(nt), 0, st, a);
    return hipGetLastError();
  }
  const dim3 ge((unsigned)((es->nelem + 63) / 64));
  const dim3 gq((unsigned)((es->ncrs + 255) / 256));
  if constexpr (!SAVE && B == NDIM && B > 1)
    if (es->lpr > 0) return launch_solid_rowgroups<KIND, NDIM>(a, es->lpr, st);
  if (!SAVE && (KIND == DFM_SOLID2D || KIND == DFM_SOLID3D)) {
    hipLaunchKernelGGL((k_geom_store<NDIM>), ge, dim3(64), 0, st, a, es->nelem, es->ebuf);
    if (es->ncrs > 0)
      hipLaunchKernelGGL((k_gather_blocks_solid<NDIM>), gq, dim3(256), 0, st, es->ncrs, es->eptr, es->eidx,
                         (const double*)es->ebuf, a.params, a.pstride, a.crs);
    hipLaunchKernelGGL((k_gather_rows_solid<NDIM>), dim3((a.n + 255) / 256), dim3(256), 0, st, a, es->eloc,
                       (const double*)es->ebuf);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_elem_store<KIND, NDIM, B, SAVE>), ge, dim3(64), 0, st, a, es->nelem, es->ebuf);
  if (es->ncrs > 0)
    hipLaunchKernelGGL((k_gather_blocks<B, NNO>), gq, dim3(256), 0, st, es->ncrs, es->nelem, es->eptr, es->eidx,
                       (const double*)es->ebuf, a.crs);
  hipLaunchKernelGGL((k_gather_rows<B, NNO>), dim3((a.n + 255) / 256), dim3(256), 0, st, a.n, es->nelem, a.order,
                     a.elsup_ptr, es->eloc, (const double*)es->ebuf, a.load_ptr, a.loads, a.load_scale, a.dia, a.res);
  return hipGetLastError();
}

hipError_t launch_assemble(const AsmArgs& a, const ElemStore* es, hipStream_t st) {
  if (a.n == 0) return hipSuccess;
#define DFM_ASM_SAVE(K, D, B) \
  if (a.kind == K && a.ndim == D && a.len == B) return assemble_kind<K, D, B, true>(a, es, st);
  if (a.save) {
    DFM_ASM_SAVE(DFM_POISSON, 2, 1)
    DFM_ASM_SAVE(DFM_SOLID2D, 2, 2)
    DFM_ASM_SAVE(DFM_SOLID3D, 3, 3)
    return hipErrorInvalidValue;
  }
#undef DFM_ASM_SAVE
#define DFM_ASM(K, D, B) \
  if (a.kind == K && a.ndim == D && a.len == B) return assemble_kind<K, D, B, false>(a, es, st);
  DFM_ASM(DFM_POISSON, 2, 1)
  DFM_ASM(DFM_POISSON, 3, 1)
  DFM_ASM(DFM_DIFFUSION2D, 2, 1)
  DFM_ASM(DFM_SOLID2D, 2, 2)
  DFM_ASM(DFM_SOLID3D, 3, 3)
  DFM_ASM(DFM_STVK2D, 2, 2)
  DFM_ASM(DFM_STVK3D, 3, 3)
  DFM_ASM(DFM_SOLID2D_NEWMARK, 2, 2)
  DFM_ASM(DFM_SOLID3D_NEWMARK, 3, 3)
  DFM_ASM(DFM_SOLID2D_BEULER, 2, 2)
  DFM_ASM(DFM_STVK3D_NEWMARK, 3, 3)
  DFM_ASM(DFM_STVK2D_NEWMARK, 2, 2)
  DFM_ASM(DFM_STVK2D_BEULER, 2, 2)
  DFM_ASM(KIND_DIFFUSION2D_SAVEDIA, 2, 1)
  DFM_ASM(KIND_SOLID2D_SAVEDIA, 2, 2)
#undef DFM_ASM
  return hipErrorInvalidValue;
}

// CMatDia_BlkCrs::SetBoundaryCondition (src/matvec/matdia_blkcrs.cpp:313-342) and CBCFlag::SetZeroToBCDof
// (include/delfem/matvec/bcflag_blk.h:102-129): rows and columns of fixed dofs are zeroed, their diagonal
// set to one, their residual zeroed. Only assignments of 0 and 1, so any traversal order gives the same
// matrix; one thread per row.
template <int B>
__global__ void k_set_bc(int n, const unsigned* rowptr, const unsigned* colind, double* crs, double* dia, double* res,
                         const int* bc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int BS = B * B;
  int fi[B];
  for (int d = 0; d < B; d++) fi[d] = bc[(long long)i * B + d];
  double* D = &dia[(long long)i * BS];
  for (int id = 0; id < B; id++) {
    if (!fi[id]) continue;
    for (int k = 0; k < B; k++) D[k * B + id] = 0.0;
    for (int j = 0; j < B; j++) D[id * B + j] = 0.0;
    D[id * B + id] = 1.0;
    res[(long long)i * B + id] = 0.0;
  }
  for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
    double* O = &crs[(long long)q * BS];
    const unsigned j = colind[q];
    for (int jd = 0; jd < B; jd++) {
      const bool fj = bc[(long long)j * B + jd] != 0;
      for (int id = 0; id < B; id++)
        if (fi[id] || fj) O[id * B + jd] = 0.0;
    }
  }
}

hipError_t launch_set_bc(int n, int len, const unsigned* rowptr, const unsigned* colind, double* crs, double* dia,
                         double* res, const int* bc, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const int nt = 256, nb = (n + nt - 1) / nt;
  if (len == 1) hipLaunchKernelGGL(k_set_bc<1>, dim3(nb), dim3(nt), 0, st, n, rowptr, colind, crs, dia, res, bc);
  else if (len == 2) hipLaunchKernelGGL(k_set_bc<2>, dim3(nb), dim3(nt), 0, st, n, rowptr, colind, crs, dia, res, bc);
  else if (len == 3) hipLaunchKernelGGL(k_set_bc<3>, dim3(nb), dim3(nt), 0, st, n, rowptr, colind, crs, dia, res, bc);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// u[old] += update[new] (CNodeAry::AddValueToNodeSegment with factor 1.0, src/femls/linearsystem_field.cpp:740)
__global__ void k_update_field(long long m, int len, const int* order, const double* upd, double* u) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const long long i0 = k / len;
  const int d = (int)(k - i0 * len);
  const long long i1 = order ? order[i0] : i0;
  u[i1 * len + d] += 1.0 * upd[k];
}

hipError_t launch_update_field(int n, int len, const int* order, const double* upd, double* u, hipStream_t st) {
  const long long m = (long long)n * len;
  if (m == 0) return hipSuccess;
  const int nt = 256;
  hipLaunchKernelGGL(k_update_field, dim3((unsigned)((m + nt - 1) / nt)), dim3(nt), 0, st, m, len, order, upd, u);
  return hipGetLastError();
}

// CLinearSystem_Field::UpdateValueOfField_NewmarkBeta (src/femls/linearsystem_field.cpp:966-1059): the reference
// makes whole-array passes (CNodeAry::AddValueToNodeSegment, src/femfield/node_ary.cpp:197-251, x += alpha * y),
// each dof reading only its own values, so one thread per dof runs the passes in their order.
__global__ void k_update_field_newmark(long long m, int len, const int* order, const double* upd, double gamma,
                                       double beta, double dt, int initial, double* u, double* v, double* a) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const long long i0 = k / len;
  const int d = (int)(k - i0 * len);
  const long long i1 = (order ? order[i0] : i0) * len + d;
  const double du = upd[k];
  double uu = u[i1], vv = v[i1], aa = a[i1];
  if (initial) {
    uu += dt * vv;
    uu += 0.5 * dt * dt * aa;
  }
  uu += beta * dt * dt * du;
  if (initial) vv += dt * aa;
  vv += gamma * dt * du;
  aa += 1.0 * du;
  u[i1] = uu;
  v[i1] = vv;
  a[i1] = aa;
}

hipError_t launch_update_field_newmark(int n, int len, const int* order, const double* upd, double gamma, double beta,
                                       double dt, int initial, double* u, double* v, double* a, hipStream_t st) {
  const long long m = (long long)n * len;
  if (m == 0) return hipSuccess;
  const int nt = 256;
  hipLaunchKernelGGL(k_update_field_newmark, dim3((unsigned)((m + nt - 1) / nt)), dim3(nt), 0, st, m, len, order, upd,
                     gamma, beta, dt, initial, u, v, a);
  return hipGetLastError();
}

// CLinearSystem_Field::UpdateValueOfField_BackwardEular (src/femls/linearsystem_field.cpp:1061-1101), the update
// being the velocity increment: v += du, then u += dt v (is_initial) or u += dt du.
__global__ void k_update_field_beuler(long long m, int len, const int* order, const double* upd, double dt,
                                      int initial, double* u, double* v) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const long long i0 = k / len;
  const int d = (int)(k - i0 * len);
  const long long i1 = (order ? order[i0] : i0) * len + d;
  const double du = upd[k];
  const double vv = v[i1] + 1.0 * du;
  v[i1] = vv;
  u[i1] = initial ? u[i1] + dt * vv : u[i1] + dt * du;
}

// CLinearSystem_Field::UpdateValueOfField_Newmark, VELOCITY (src/femls/linearsystem_field.cpp:885-921): on the
// first step the value takes dt times the velocity, then gamma dt (gdt, formed by the caller) times the update; the
// velocity takes the update.
__global__ void k_update_field_newmark1(long long m, int len, const int* order, const double* upd, double gdt,
                                       double dt, int initial, double* u, double* v) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const long long i0 = k / len;
  const int d = (int)(k - i0 * len);
  const long long i1 = (order ? order[i0] : i0) * len + d;
  const double du = upd[k];
  double uu = u[i1];
  if (initial) uu += dt * v[i1];
  u[i1] = uu + gdt * du;
  v[i1] = v[i1] + 1.0 * du;
}

hipError_t launch_update_field_newmark1(int n, int len, const int* order, const double* upd, double gdt, double dt,
                                        int initial, double* u, double* v, hipStream_t st) {
  const long long m = (long long)n * len;
  if (m == 0) return hipSuccess;
  const int nt = 256;
  hipLaunchKernelGGL(k_update_field_newmark1, dim3((unsigned)((m + nt - 1) / nt)), dim3(nt), 0, st, m, len, order,
                     upd, gdt, dt, initial, u, v);
  return hipGetLastError();
}

hipError_t launch_update_field_beuler(int n, int len, const int* order, const double* upd, double dt, int initial,
                                      double* u, double* v, hipStream_t st) {
  const long long m = (long long)n * len;
  if (m == 0) return hipSuccess;
  const int nt = 256;
  hipLaunchKernelGGL(k_update_field_beuler, dim3((unsigned)((m + nt - 1) / nt)), dim3(nt), 0, st, m, len, order, upd,
                     dt, initial, u, v);
  return hipGetLastError();
}

// K_b values, read from the assembled matrix before the Dirichlet zeroing. The boundary matrix of the reference
// receives the same element blocks in the same element order as the matrix (CMat_BlkCrs::Mearge,
// src/matvec/mat_blkcrs.cpp:754-823), so each of its entries is the matrix entry of the same position.
__global__ void k_gather_terms(long long nterm, const long long* src, long long ndia, const double* dia,
                               const double* crs, double* val) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nterm) return;
  const long long k = src[t];
  val[t] = (k < ndia) ? dia[k] : crs[k - ndia];
}

hipError_t launch_gather_terms(long long nterm, const long long* src, long long ndia, const double* dia,
                               const double* crs, double* val, hipStream_t st) {
  if (nterm == 0) return hipSuccess;
  const int nt = 256;
  hipLaunchKernelGGL(k_gather_terms, dim3((unsigned)((nterm + nt - 1) / nt)), dim3(nt), 0, st, nterm, src, ndia, dia,
                     crs, val);
  return hipGetLastError();
}

// The update vector w the saved systems' MakeResidual forms before its MatVec, dof j of the field (original
// numbering): mode SAVE_PLAIN u (CLinearSystem_Save); SAVE_DIA_NEWMARK u + v c0 with c0 = dt (1 - gamma)
// (CLinearSystem_SaveDiaM_Newmark, src/femls/linearsystem_fieldsave.cpp:433-457); SAVE_DIA_NEWMARKBETA
// u + v c0 + a c1 with c0 = dt, c1 = dt dt (0.5 - beta) (_NewmarkBeta, :705-737). Each term is added as
// CNodeAry::AddValueFromNodeSegment adds it (value * alpha, src/femfield/node_ary.cpp:302-325).
__device__ inline double save_w(int mode, const double* u, const double* v, const double* a, long long j, double c0,
                                double c1) {
  double w = u[j];
  if (mode != SAVE_PLAIN) w += v[j] * c0;
  if (mode == SAVE_DIA_NEWMARKBETA) w += a[j] * c1;
  return w;
}

// MakeResidual of the saved systems (CLinearSystem_Save, src/femls/linearsystem_fieldsave.cpp:206-247, and the
// lumped-mass ones above): update = w, residual = force, then CMat_BlkCrs::MatVec(-1, update, 1, residual) over
// the boundary matrix (src/matvec/mat_blkcrs.cpp:955-1170): residual[i] += (-1 * K_b[i][j]) * w[j] entry by entry,
// which is residual[i] - K_b[i][j] * w[j] exactly. Each dof runs over its kept terms in the reference's order. The
// lumped-mass systems then zero the residual at fixed dofs (bc non-null). One thread per scalar dof.
__global__ void k_make_residual(long long m, int len, int mode, double c0, double c1, const int* order,
                                const unsigned* bptr, const unsigned* bcol, const double* bval, const double* force,
                                const double* u, const double* v, const double* a, const int* bc, double* res,
                                double* upd) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const long long i0 = k / len;
  const int d = (int)(k - i0 * len);
  const long long i1 = order ? order[i0] : i0;
  double r = force[k];
  for (unsigned t = bptr[k]; t < bptr[k + 1]; t++) r = r - bval[t] * save_w(mode, u, v, a, bcol[t], c0, c1);
  if (bc && bc[k]) r = 0.0;
  res[k] = r;
  upd[k] = save_w(mode, u, v, a, i1 * len + d, c0, c1);
}

hipError_t launch_make_residual(int n, int len, int mode, double c0, double c1, const int* order,
                                const unsigned* bptr, const unsigned* bcol, const double* bval, const double* force,
                                const double* u, const double* v, const double* a, const int* bc, double* res,
                                double* upd, hipStream_t st) {
  const long long m = (long long)n * len;
  if (m == 0) return hipSuccess;
  const int nt = 256;
  hipLaunchKernelGGL(k_make_residual, dim3((unsigned)((m + nt - 1) / nt)), dim3(nt), 0, st, m, len, mode, c0, c1,
                     order, bptr, bcol, bval, force, u, v, a, bc, res, upd);
  return hipGetLastError();
}

// UpdateValueOfField of the lumped-mass saved systems, per dof (free: bc == 0). SAVE_DIA_NEWMARK, VELOCITY
// (src/femls/linearsystem_fieldsave.cpp:474-520): u += dt (1 - gamma) v0 + dt gamma v1 (v1 the update); a fixed dof
// keeps its velocity in the update; velocity = update. SAVE_DIA_NEWMARKBETA, ACCELERATION (:760-860):
// u += dt v0 + dt dt (0.5 - beta) a0 + dt dt beta a1, v += dt (1 - gamma) a0 + gamma dt a1 (a1 the update); a fixed
// dof keeps its acceleration in the update; acceleration = update. The expressions are written as there.
__global__ void k_update_field_savedia(long long m, int len, int mode, double dt, double gamma, double beta,
                                       const int* order, const int* bc, double* upd, double* u, double* v,
                                       double* a) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const long long i0 = k / len;
  const int d = (int)(k - i0 * len);
  const long long f = (order ? order[i0] : i0) * len + d;
  if (mode == SAVE_DIA_NEWMARK) {
    const double velo0 = v[f];
    if (!bc[k]) {
      const double velo1 = upd[k];
      u[f] += dt * (1 - gamma) * velo0 + dt * gamma * velo1;
    } else {
      upd[k] = velo0;
    }
    v[f] = upd[k];
  } else {
    const double velo0 = v[f], acc0 = a[f];
    if (!bc[k]) {
      const double acc1 = upd[k];
      u[f] += dt * velo0 + dt * dt * (0.5 - beta) * acc0 + dt * dt * beta * acc1;
      v[f] += dt * (1 - gamma) * acc0 + gamma * dt * acc1;
    } else {
      upd[k] = acc0;
    }
    a[f] = upd[k];
  }
}

hipError_t launch_update_field_savedia(int n, int len, int mode, double dt, double gamma, double beta,
                                       const int* order, const int* bc, double* upd, double* u, double* v, double* a,
                                       hipStream_t st) {
  const long long m = (long long)n * len;
  if (m == 0) return hipSuccess;
  const int nt = 256;
  hipLaunchKernelGGL(k_update_field_savedia, dim3((unsigned)((m + nt - 1) / nt)), dim3(nt), 0, st, m, len, mode,
                     dt, gamma, beta, order, bc, upd, u, v, a);
  return hipGetLastError();
}

// CLinearSystem_Save::UpdateValueOfField, VALUE (src/femls/linearsystem_fieldsave.cpp:249-296): the update takes
// the field values at fixed dofs, then the field takes the update (CNodeAry::SetValueToNodeSegment).
__global__ void k_update_field_save(long long m, int len, const int* order, const int* bc, double* upd, double* u) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const long long i0 = k / len;
  const int d = (int)(k - i0 * len);
  const long long i1 = order ? order[i0] : i0;
  double* f = &u[i1 * len + d];
  if (bc[k]) upd[k] = *f;
  *f = upd[k];
}

hipError_t launch_update_field_save(int n, int len, const int* order, const int* bc, double* upd, double* u,
                                    hipStream_t st) {
  const long long m = (long long)n * len;
  if (m == 0) return hipSuccess;
  const int nt = 256;
  hipLaunchKernelGGL(k_update_field_save, dim3((unsigned)((m + nt - 1) / nt)), dim3(nt), 0, st, m, len, order, bc,
                     upd, u);
  return hipGetLastError();
}

}  // namespace dfm
