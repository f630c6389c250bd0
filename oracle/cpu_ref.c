/* Oracle: plain-C restatement of DelFEM's FEM hot path. TEST INFRASTRUCTURE ONLY: the checker the HIP
 * product is compared against, never the product. Built with -ffp-contract=off so that every product
 * and sum is rounded as in the reference (g++ -O2 on x86-64 emits no FMA).
 *
 * Every function names the reference lines it restates. Parity pinning: tests/test_oracle_vs_reference.py
 * compares each function bit for bit with the reference compiled from its own sources (oracle/_ref), and
 * tests/test_oracle_math.py checks what the mathematics fixes (closed forms, rigid-body null space, row
 * sums, exact factorisation of banded matrices, CG termination on tiny SPD systems). The only function
 * without a reference counterpart is cr_traction_loads (the reference has no surface traction,
 * src/femeqn/eqn_linear_solid3d.cpp:312-314); it is pinned by closed form (sum of loads = t * area).
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------------------------------- */
/* pattern: src/femfield/elem_ary.cpp:675-724 (MakePointSurElem), :821-892 (MakePointSurPoint with        */
/* isnt_self), include/delfem/indexed_array.h:113-132 (Sort): per row, the distinct neighbours, ascending. */

static int* point_sur_elem(int nnode, int nelem, int npoel, const int* lnods, int** index_out) {
  int* index = (int*)calloc((size_t)nnode + 1, sizeof(int));
  for (long e = 0; e < nelem; e++)
    for (int k = 0; k < npoel; k++) index[lnods[e * npoel + k] + 1]++;
  for (int i = 0; i < nnode; i++) index[i + 1] += index[i];
  int* arr = (int*)malloc(sizeof(int) * (size_t)(index[nnode] > 0 ? index[nnode] : 1));
  int* fill = (int*)malloc(sizeof(int) * (size_t)(nnode + 1));
  memcpy(fill, index, sizeof(int) * (size_t)(nnode + 1));
  for (long e = 0; e < nelem; e++)
    for (int k = 0; k < npoel; k++) arr[fill[lnods[e * npoel + k]]++] = (int)e;
  free(fill);
  *index_out = index;
  return arr;
}

static int cmp_u(const void* a, const void* b) {
  const unsigned x = *(const unsigned*)a, y = *(const unsigned*)b;
  return (x > y) - (x < y);
}

long cr_pattern_count(int nnode, int nelem, int npoel, const int* lnods, unsigned* rowptr) {
  int* esi;
  int* es = point_sur_elem(nnode, nelem, npoel, lnods, &esi);
  int* mark = (int*)malloc(sizeof(int) * (size_t)(nnode > 0 ? nnode : 1));
  for (int i = 0; i < nnode; i++) mark[i] = -1;
  long cnt = 0;
  rowptr[0] = 0;
  for (int i = 0; i < nnode; i++) {
    mark[i] = i;
    for (int k = esi[i]; k < esi[i + 1]; k++) {
      const int e = es[k];
      for (int a = 0; a < npoel; a++) {
        const int j = lnods[(long)e * npoel + a];
        if (mark[j] != i) { mark[j] = i; cnt++; }
      }
    }
    rowptr[i + 1] = (unsigned)cnt;
  }
  free(mark); free(es); free(esi);
  return cnt;
}

void cr_pattern_fill(int nnode, int nelem, int npoel, const int* lnods, const unsigned* rowptr,
                     unsigned* colind) {
  int* esi;
  int* es = point_sur_elem(nnode, nelem, npoel, lnods, &esi);
  int* mark = (int*)malloc(sizeof(int) * (size_t)(nnode > 0 ? nnode : 1));
  for (int i = 0; i < nnode; i++) mark[i] = -1;
  for (int i = 0; i < nnode; i++) {
    unsigned pos = rowptr[i];
    mark[i] = i;
    for (int k = esi[i]; k < esi[i + 1]; k++) {
      const int e = es[k];
      for (int a = 0; a < npoel; a++) {
        const int j = lnods[(long)e * npoel + a];
        if (mark[j] != i) { mark[j] = i; colind[pos++] = (unsigned)j; }
      }
    }
    qsort(&colind[rowptr[i]], rowptr[i + 1] - rowptr[i], sizeof(unsigned), cmp_u);
  }
  free(mark); free(es); free(esi);
}

/* ---------------------------------------------------------------------------------------------------- */
/* Mearge: src/matvec/matdia_blkcrs.cpp:165-258. emat is [nno][nno][bs] (bs = b*b). */

static void merge(int nno, const int* no, int bs, const double* emat, const unsigned* rowptr,
                  const unsigned* colind, double* crs, double* dia) {
  for (int a = 0; a < nno; a++) {
    const int i = no[a];
    for (int c = 0; c < nno; c++) {
      const double* pin = &emat[(a * nno + c) * bs];
      if (a == c) {
        double* pout = &dia[(long)i * bs];
        for (int k = 0; k < bs; k++) pout[k] += pin[k];
      } else {
        const unsigned j = (unsigned)no[c];
        long pos = -1;
        for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++)
          if (colind[q] == j) { pos = q; break; }
        if (pos < 0) continue; /* pairs missing from the pattern are skipped (:242) */
        double* pout = &crs[pos * bs];
        for (int k = 0; k < bs; k++) pout[k] += pin[k];
      }
    }
  }
}

/* ---------------------------------------------------------------------------------------------------- */
/* P1 geometry: src/femeqn/ker_emat_tri.cpp:13-15 (TriArea), :36-52 (TriDlDx);                          */
/* include/delfem/femeqn/ker_emat_tet.h:32-39 (TetVolume, with its 0.1666666666666 literal), :42-77.     */

static double tri_area(const double* p0, const double* p1, const double* p2) {
  return 0.5 * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
}

static void tri_dldx(double dldx[3][2], const double* p0, const double* p1, const double* p2) {
  const double area = tri_area(p0, p1, p2);
  const double tmp1 = 0.5 / area;
  dldx[0][0] = tmp1 * (p1[1] - p2[1]);
  dldx[1][0] = tmp1 * (p2[1] - p0[1]);
  dldx[2][0] = tmp1 * (p0[1] - p1[1]);
  dldx[0][1] = tmp1 * (p2[0] - p1[0]);
  dldx[1][1] = tmp1 * (p0[0] - p2[0]);
  dldx[2][1] = tmp1 * (p1[0] - p0[0]);
}

static double tet_volume(const double* p0, const double* p1, const double* p2, const double* p3) {
  double vol = (p1[0] - p0[0]) * ((p2[1] - p0[1]) * (p3[2] - p0[2]) - (p3[1] - p0[1]) * (p2[2] - p0[2]))
             - (p1[1] - p0[1]) * ((p2[0] - p0[0]) * (p3[2] - p0[2]) - (p3[0] - p0[0]) * (p2[2] - p0[2]))
             + (p1[2] - p0[2]) * ((p2[0] - p0[0]) * (p3[1] - p0[1]) - (p3[0] - p0[0]) * (p2[1] - p0[1]));
  vol *= 0.1666666666666;
  return vol;
}

static void tet_dldx(double d[4][3], const double* p0, const double* p1, const double* p2, const double* p3) {
  const double vol = tet_volume(p0, p1, p2, p3);
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

/* ---------------------------------------------------------------------------------------------------- */
/* element kernels                                                                                       */

/* src/femeqn/eqn_poisson.cpp:49-125 (2D), :249-323 (3D). p = (alpha, source). */
static void elem_poisson(int ndim, const double co[][3], const double* val, const double* p, double* emat,
                         double* eres) {
  const double alpha = p[0], source = p[1];
  if (ndim == 2) {
    const double area = tri_area(co[0], co[1], co[2]);
    double dldx[3][2];
    tri_dldx(dldx, co[0], co[1], co[2]);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        emat[i * 3 + j] = alpha * area * (dldx[i][0] * dldx[j][0] + dldx[i][1] * dldx[j][1]);
    for (int i = 0; i < 3; i++) eres[i] = source * area * 0.33333333333333333;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) eres[i] -= emat[i * 3 + j] * val[j];
  } else {
    const double vol = tet_volume(co[0], co[1], co[2], co[3]);
    double dldx[4][3];
    tet_dldx(dldx, co[0], co[1], co[2], co[3]);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++)
        emat[i * 4 + j] =
            alpha * vol * (dldx[i][0] * dldx[j][0] + dldx[i][1] * dldx[j][1] + dldx[i][2] * dldx[j][2]);
    for (int i = 0; i < 4; i++) eres[i] = source * vol * 0.25;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) eres[i] -= emat[i * 4 + j] * val[j];
  }
}

/* src/femeqn/eqn_diffusion.cpp:240-345. p = (rho, alpha, source), glob = (dt, gamma). */
static void elem_diffusion2d(const double co[][3], const double* val, const double* vval, const double* p,
                             const double* glob, double* emat, double* eres) {
  const double rho = p[0], alpha = p[1], source = p[2], dt = glob[0], gamma = glob[1];
  double eC[3][3], eM[3][3];
  const double area = tri_area(co[0], co[1], co[2]);
  double dldx[3][2];
  tri_dldx(dldx, co[0], co[1], co[2]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) eC[i][j] = alpha * area * (dldx[i][0] * dldx[j][0] + dldx[i][1] * dldx[j][1]);
  {
    const double dtmp1 = rho * area * 0.08333333333333333;
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) eM[i][j] = dtmp1;
      eM[i][i] += dtmp1;
    }
  }
  for (int i = 0; i < 3; i++) eres[i] = source * area * 0.333333333333333333;
  {
    const double dtmp1 = gamma * dt;
    for (int i = 0; i < 9; i++) emat[i] = (&eM[0][0])[i] + dtmp1 * (&eC[0][0])[i];
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) eres[i] -= eC[i][j] * (val[j] + dt * vval[j]) + eM[i][j] * vval[j];
}

/* src/femeqn/eqn_linear_solid2d.cpp:505-598. p = (lambda, myu, rho, gx, gy). emat [3][3][2][2]. */
static void elem_solid2d(const double co[][3], const double disp[][3], const double* p, double* emat,
                         double* eres) {
  const double lambda = p[0], myu = p[1], rho = p[2], g_x = p[3], g_y = p[4];
  const double area = tri_area(co[0], co[1], co[2]);
  double dldx[3][2];
  tri_dldx(dldx, co[0], co[1], co[2]);
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) {
      double* e = &emat[(i * 3 + j) * 4];
      e[0] = area * (lambda + myu) * dldx[i][0] * dldx[j][0];
      e[1] = area * (lambda * dldx[i][0] * dldx[j][1] + myu * dldx[j][0] * dldx[i][1]);
      e[2] = area * (lambda * dldx[i][1] * dldx[j][0] + myu * dldx[j][1] * dldx[i][0]);
      e[3] = area * (lambda + myu) * dldx[i][1] * dldx[j][1];
      const double dtmp1 = (dldx[i][1] * dldx[j][1] + dldx[i][0] * dldx[j][0]) * area * myu;
      e[0] += dtmp1;
      e[3] += dtmp1;
    }
  }
  for (int i = 0; i < 3; i++) {
    eres[i * 2 + 0] = area * rho * g_x * 0.33333333333333333;
    eres[i * 2 + 1] = area * rho * g_y * 0.33333333333333333;
  }
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) {
      const double* e = &emat[(i * 3 + j) * 4];
      eres[i * 2 + 0] -= e[0] * disp[j][0] + e[1] * disp[j][1];
      eres[i * 2 + 1] -= e[2] * disp[j][0] + e[3] * disp[j][1];
    }
  }
}

/* src/femeqn/eqn_linear_solid3d.cpp:241-334. p = (lambda, myu, rho, gx, gy, gz). emat [4][4][3][3]. */
static void elem_solid3d(const double co[][3], const double disp[][3], const double* p, double* emat,
                         double* eres) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  const double g[3] = {p[3], p[4], p[5]};
  const double vol = tet_volume(co[0], co[1], co[2], co[3]);
  double d[4][3];
  tet_dldx(d, co[0], co[1], co[2], co[3]);
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) {
      double* e = &emat[(i * 4 + j) * 9];
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          e[a * 3 + b] = vol * (lambda * d[i][a] * d[j][b] + myu * d[j][a] * d[i][b]);
      const double dtmp1 = d[i][0] * d[j][0] + d[i][1] * d[j][1] + d[i][2] * d[j][2];
      e[0] += vol * myu * dtmp1;
      e[4] += vol * myu * dtmp1;
      e[8] += vol * myu * dtmp1;
    }
  }
  for (int i = 0; i < 4; i++) {
    eres[i * 3 + 0] = vol * rho * g[0] * 0.25;
    eres[i * 3 + 1] = vol * rho * g[1] * 0.25;
    eres[i * 3 + 2] = vol * rho * g[2] * 0.25;
    for (int j = 0; j < 4; j++) {
      const double* e = &emat[(i * 4 + j) * 9];
      eres[i * 3 + 0] -= e[0] * disp[j][0] + e[1] * disp[j][1] + e[2] * disp[j][2];
      eres[i * 3 + 1] -= e[3] * disp[j][0] + e[4] * disp[j][1] + e[5] * disp[j][2];
      eres[i * 3 + 2] -= e[6] * disp[j][0] + e[7] * disp[j][1] + e[8] * disp[j][2];
    }
  }
}

/* src/femeqn/eqn_st_venant.cpp:47-156 (AddElemMatFin_StVenant2D), :157-272 (element loop).
 * p = (lambda, myu, rho, gx, gy). */
static void elem_stvk2d(const double co[][3], const double disp[][3], const double* p, double* emat,
                        double* eres) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  const double g[2] = {p[3], p[4]};
  double efin[3][2] = {{0, 0}, {0, 0}, {0, 0}}, efex[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int i = 0; i < 36; i++) emat[i] = 0.0;
  const double area = tri_area(co[0], co[1], co[2]);
  double dndx[3][2];
  tri_dldx(dndx, co[0], co[1], co[2]);
  double dudx[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int n = 0; n < 3; n++)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) dudx[a][b] += disp[n][a] * dndx[n][b];
  const double detwei = area;
  double strain[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      strain[a][b] = 0.5 * (dudx[a][b] + dudx[b][a]);
      for (int k = 0; k < 2; k++) strain[a][b] += 0.5 * dudx[k][a] * dudx[k][b];
    }
  double d2s[3][2][2][2];
  {
    double z[2][2];
    for (int a = 0; a < 2; a++) {
      for (int b = 0; b < 2; b++) z[a][b] = dudx[a][b];
      z[a][a] += 1.0;
    }
    for (int k = 0; k < 3; k++)
      for (int jd = 0; jd < 2; jd++)
        for (int id = 0; id < 2; id++)
          for (int kd = 0; kd < 2; kd++) d2s[k][kd][id][jd] = dndx[k][jd] * z[kd][id];
  }
  double stress[2][2];
  {
    double dtmp1 = 0.0;
    for (int a = 0; a < 2; a++) {
      for (int b = 0; b < 2; b++) stress[a][b] = 2.0 * myu * strain[a][b];
      dtmp1 += strain[a][a];
    }
    for (int a = 0; a < 2; a++) stress[a][a] += lambda * dtmp1;
  }
  for (int k = 0; k < 3; k++)
    for (int kd = 0; kd < 2; kd++)
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) efin[k][kd] += detwei * stress[a][b] * d2s[k][kd][a][b];
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) {
      double* e = &emat[(i * 3 + j) * 4];
      for (int id = 0; id < 2; id++)
        for (int jd = 0; jd < 2; jd++) {
          {
            double dtmp1 = 0.0;
            for (int gd = 0; gd < 2; gd++)
              for (int hd = 0; hd < 2; hd++)
                dtmp1 += d2s[i][id][gd][hd] * d2s[j][jd][gd][hd] + d2s[i][id][gd][hd] * d2s[j][jd][hd][gd];
            e[id * 2 + jd] += detwei * myu * dtmp1;
          }
          {
            double dtmp1 = 0.0, dtmp2 = 0.0;
            for (int gd = 0; gd < 2; gd++) {
              dtmp1 += d2s[i][id][gd][gd];
              dtmp2 += d2s[j][jd][gd][gd];
            }
            e[id * 2 + jd] += detwei * lambda * dtmp1 * dtmp2;
          }
        }
      {
        double dtmp1 = 0.0;
        for (int gd = 0; gd < 2; gd++)
          for (int hd = 0; hd < 2; hd++) dtmp1 += stress[gd][hd] * dndx[i][hd] * dndx[j][gd];
        for (int id = 0; id < 2; id++) e[id * 2 + id] += detwei * dtmp1;
      }
    }
  }
  for (int n = 0; n < 3; n++)
    for (int a = 0; a < 2; a++) efex[n][a] += area * rho * g[a] / 3.0;
  for (int n = 0; n < 3; n++)
    for (int a = 0; a < 2; a++) eres[n * 2 + a] = efex[n][a] - efin[n][a];
}

/* src/femeqn/eqn_st_venant.cpp:724-834 (SetElemMatFin_StVenant3D), :838-939 (element loop).
 * p = (lambda, myu, rho, gx, gy, gz). */
static void elem_stvk3d(const double co[][3], const double disp[][3], const double* p, double* emat,
                        double* eres) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  const double g[3] = {p[3], p[4], p[5]};
  double efin[4][3], efex[4][3];
  for (int i = 0; i < 144; i++) emat[i] = 0.0;
  for (int i = 0; i < 12; i++) { (&efex[0][0])[i] = 0.0; (&efin[0][0])[i] = 0.0; }
  const double vol = tet_volume(co[0], co[1], co[2], co[3]);
  double dndx[4][3];
  tet_dldx(dndx, co[0], co[1], co[2], co[3]);
  double u[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int n = 0; n < 4; n++)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) u[a][b] += disp[n][a] * dndx[n][b];
  const double detwei = vol;
  double s[6];
  {
    double e[6];
    e[0] = u[0][0] + 0.5 * (u[0][0] * u[0][0] + u[1][0] * u[1][0] + u[2][0] * u[2][0]);
    e[1] = u[1][1] + 0.5 * (u[0][1] * u[0][1] + u[1][1] * u[1][1] + u[2][1] * u[2][1]);
    e[2] = u[2][2] + 0.5 * (u[0][2] * u[0][2] + u[1][2] * u[1][2] + u[2][2] * u[2][2]);
    e[3] = 0.5 * (u[0][1] + u[1][0] + u[0][0] * u[0][1] + u[1][0] * u[1][1] + u[2][0] * u[2][1]);
    e[4] = 0.5 * (u[1][2] + u[2][1] + u[0][1] * u[0][2] + u[1][1] * u[1][2] + u[2][1] * u[2][2]);
    e[5] = 0.5 * (u[2][0] + u[0][2] + u[0][2] * u[0][0] + u[1][2] * u[1][0] + u[2][2] * u[2][0]);
    const double dtmp1 = e[0] + e[1] + e[2];
    s[0] = 2.0 * myu * e[0] + lambda * dtmp1;
    s[1] = 2.0 * myu * e[1] + lambda * dtmp1;
    s[2] = 2.0 * myu * e[2] + lambda * dtmp1;
    s[3] = 2.0 * myu * e[3];
    s[4] = 2.0 * myu * e[4];
    s[5] = 2.0 * myu * e[5];
  }
  double d2s[4][3][6];
  {
    const double z[3][3] = {{u[0][0] + 1, u[0][1], u[0][2]},
                            {u[1][0], u[1][1] + 1, u[1][2]},
                            {u[2][0], u[2][1], u[2][2] + 1}};
    for (int k = 0; k < 4; k++)
      for (int kd = 0; kd < 3; kd++) {
        d2s[k][kd][0] = dndx[k][0] * z[kd][0];
        d2s[k][kd][1] = dndx[k][1] * z[kd][1];
        d2s[k][kd][2] = dndx[k][2] * z[kd][2];
        d2s[k][kd][3] = (dndx[k][0] * z[kd][1] + dndx[k][1] * z[kd][0]);
        d2s[k][kd][4] = (dndx[k][1] * z[kd][2] + dndx[k][2] * z[kd][1]);
        d2s[k][kd][5] = (dndx[k][2] * z[kd][0] + dndx[k][0] * z[kd][2]);
      }
  }
  for (int k = 0; k < 4; k++)
    for (int kd = 0; kd < 3; kd++) {
      double dtmp1 = 0;
      dtmp1 += s[0] * d2s[k][kd][0];
      dtmp1 += s[1] * d2s[k][kd][1];
      dtmp1 += s[2] * d2s[k][kd][2];
      dtmp1 += s[3] * d2s[k][kd][3];
      dtmp1 += s[4] * d2s[k][kd][4];
      dtmp1 += s[5] * d2s[k][kd][5];
      efin[k][kd] = dtmp1 * detwei;
    }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < i; j++)
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          double dtmp1 = 0.0;
          dtmp1 += d2s[i][a][0] * d2s[j][b][0];
          dtmp1 += d2s[i][a][1] * d2s[j][b][1];
          dtmp1 += d2s[i][a][2] * d2s[j][b][2];
          dtmp1 += d2s[i][a][3] * d2s[j][b][3] * 0.5;
          dtmp1 += d2s[i][a][4] * d2s[j][b][4] * 0.5;
          dtmp1 += d2s[i][a][5] * d2s[j][b][5] * 0.5;
          const double dtmp2 = d2s[i][a][0] + d2s[i][a][1] + d2s[i][a][2];
          const double dtmp3 = d2s[j][b][0] + d2s[j][b][1] + d2s[j][b][2];
          const double dtmp4 = detwei * (myu * 2 * dtmp1 + lambda * dtmp2 * dtmp3);
          emat[(i * 4 + j) * 9 + a * 3 + b] = dtmp4;
          emat[(j * 4 + i) * 9 + b * 3 + a] = dtmp4;
        }
  for (int i = 0; i < 4; i++)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double dtmp1 = 0.0;
        dtmp1 += d2s[i][a][0] * d2s[i][b][0];
        dtmp1 += d2s[i][a][1] * d2s[i][b][1];
        dtmp1 += d2s[i][a][2] * d2s[i][b][2];
        dtmp1 += d2s[i][a][3] * d2s[i][b][3] * 0.5;
        dtmp1 += d2s[i][a][4] * d2s[i][b][4] * 0.5;
        dtmp1 += d2s[i][a][5] * d2s[i][b][5] * 0.5;
        const double dtmp2 = d2s[i][a][0] + d2s[i][a][1] + d2s[i][a][2];
        const double dtmp3 = d2s[i][b][0] + d2s[i][b][1] + d2s[i][b][2];
        emat[(i * 4 + i) * 9 + a * 3 + b] = detwei * (myu * 2 * dtmp1 + lambda * dtmp2 * dtmp3);
      }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double dtmp1 = 0.0;
      dtmp1 += s[0] * dndx[i][0] * dndx[j][0];
      dtmp1 += s[1] * dndx[i][1] * dndx[j][1];
      dtmp1 += s[2] * dndx[i][2] * dndx[j][2];
      dtmp1 += s[3] * (dndx[i][0] * dndx[j][1] + dndx[i][1] * dndx[j][0]);
      dtmp1 += s[4] * (dndx[i][1] * dndx[j][2] + dndx[i][2] * dndx[j][1]);
      dtmp1 += s[5] * (dndx[i][2] * dndx[j][0] + dndx[i][0] * dndx[j][2]);
      emat[(i * 4 + j) * 9 + 0] += detwei * dtmp1;
      emat[(i * 4 + j) * 9 + 4] += detwei * dtmp1;
      emat[(i * 4 + j) * 9 + 8] += detwei * dtmp1;
    }
  for (int n = 0; n < 4; n++)
    for (int a = 0; a < 3; a++) efex[n][a] += vol * rho * g[a] / 4.0;
  for (int n = 0; n < 4; n++)
    for (int a = 0; a < 3; a++) eres[n * 3 + a] = efex[n][a] - efin[n][a];
}

int cr_assemble(int kind, int ndim, int nnode, const double* coords, int nelem, const int* lnods,
                const double* params, int npar, int pstride, const double* glob, const double* u,
                const double* v, int len, const unsigned* rowptr, const unsigned* colind, double* crs,
                double* dia, double* res) {
  (void)nnode;
  (void)npar;
  const int nno = (ndim == 2) ? 3 : 4;
  const int bs = len * len;
  double emat[144], eres[12], co[4][3], dsp[4][3], val[4], vval[4];
  for (long e = 0; e < nelem; e++) {
    const int* no = &lnods[e * nno];
    const double* p = &params[e * pstride];
    for (int a = 0; a < nno; a++) {
      for (int d = 0; d < 3; d++) co[a][d] = (d < ndim) ? coords[(long)no[a] * ndim + d] : 0.0;
      for (int d = 0; d < 3; d++) dsp[a][d] = (d < len && u) ? u[(long)no[a] * len + d] : 0.0;
      val[a] = u ? u[(long)no[a] * len] : 0.0;
      vval[a] = v ? v[(long)no[a] * len] : 0.0;
    }
    switch (kind) {
      case CR_POISSON: elem_poisson(ndim, co, val, p, emat, eres); break;
      case CR_DIFFUSION2D: elem_diffusion2d(co, val, vval, p, glob, emat, eres); break;
      case CR_SOLID2D: elem_solid2d(co, dsp, p, emat, eres); break;
      case CR_SOLID3D: elem_solid3d(co, dsp, p, emat, eres); break;
      case CR_STVK2D: elem_stvk2d(co, dsp, p, emat, eres); break;
      case CR_STVK3D: elem_stvk3d(co, dsp, p, emat, eres); break;
      default: return -1;
    }
    merge(nno, no, bs, emat, rowptr, colind, crs, dia);
    /* CVector_Blk::AddValue, include/delfem/matvec/vector_blk.h:170-181 */
    for (int a = 0; a < nno; a++)
      for (int d = 0; d < len; d++) res[(long)no[a] * len + d] += eres[a * len + d];
  }
  return 0;
}

void cr_add_nodal_loads(int nload, const int* node, const double* loads, int len, double* res) {
  for (int k = 0; k < nload; k++)
    for (int d = 0; d < len; d++) res[(long)node[k] * len + d] += loads[(long)k * len + d];
}

void cr_traction_loads(int nface, const int* faces, const double* trac, const double* coords, double* res) {
  for (int f = 0; f < nface; f++) {
    const double* p0 = &coords[(long)faces[f * 3 + 0] * 3];
    const double* p1 = &coords[(long)faces[f * 3 + 1] * 3];
    const double* p2 = &coords[(long)faces[f * 3 + 2] * 3];
    const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double area = 0.5 * sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int k = 0; k < 3; k++)
      for (int d = 0; d < 3; d++) res[(long)faces[f * 3 + k] * 3 + d] += trac[f * 3 + d] * area / 3.0;
  }
}

/* ---------------------------------------------------------------------------------------------------- */
/* src/matvec/matdia_blkcrs.cpp:261-343 (SetBoundaryCondition, fixed block size branch :313-342),        */
/* include/delfem/matvec/bcflag_blk.h:102-129 (SetZeroToBCDof), src/ls/linearsystem.cpp:92-128.          */
double cr_set_bc(int n, int len, const unsigned* rowptr, const unsigned* colind, double* crs, double* dia,
                 const int* flags, double* res) {
  const int bl = len, bs = len * len;
  for (long i = 0; i < n; i++) {
    for (int id = 0; id < bl; id++) {
      if (flags[i * bl + id] == 0) continue;
      for (int k = 0; k < bl; k++) dia[i * bs + k * bl + id] = 0.0;
      for (int j = 0; j < bl; j++) dia[i * bs + id * bl + j] = 0.0;
      dia[i * bs + id * bl + id] = 1.0;
      for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++)
        for (int j = 0; j < bl; j++) crs[(long)q * bs + id * bl + j] = 0.0;
    }
  }
  const unsigned ncrs = rowptr[n];
  for (unsigned q = 0; q < ncrs; q++) {
    const unsigned j1 = colind[q];
    for (int jd = 0; jd < bl; jd++) {
      if (flags[(long)j1 * bl + jd] == 0) continue;
      for (int id = 0; id < bl; id++) crs[(long)q * bs + id * bl + jd] = 0.0;
    }
  }
  double sq = 0.0;
  for (long i = 0; i < (long)n * len; i++) {
    if (flags[i] == 1) res[i] = 0;
  }
  for (long i = 0; i < (long)n * len; i++) sq += res[i] * res[i];
  return sqrt(sq);
}

/* ---------------------------------------------------------------------------------------------------- */
/* src/matvec/matdia_blkcrs.cpp:921-1094 (MatVec, b = 1, 2, 3 specialisations, general b)               */
void cr_matvec(int n, int len, const unsigned* rowptr, const unsigned* colind, const double* crs,
               const double* dia, double alpha, const double* x, double beta, double* y) {
  if (len == 1) {
    for (long i = 0; i < n; i++) {
      double* iy = y + i;
      (*iy) *= beta;
      for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) (*iy) += alpha * crs[q] * x[colind[q]];
      (*iy) += alpha * dia[i] * x[i];
    }
  } else if (len == 2) {
    for (long i = 0; i < n; i++) {
      double* iy = y + i * 2;
      iy[0] *= beta;
      iy[1] *= beta;
      for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
        const double* jx = x + (long)colind[q] * 2;
        iy[0] += alpha * (crs[q * 4] * jx[0] + crs[q * 4 + 1] * jx[1]);
        iy[1] += alpha * (crs[q * 4 + 2] * jx[0] + crs[q * 4 + 3] * jx[1]);
      }
      const double* ix = x + i * 2;
      iy[0] += alpha * (dia[i * 4] * ix[0] + dia[i * 4 + 1] * ix[1]);
      iy[1] += alpha * (dia[i * 4 + 2] * ix[0] + dia[i * 4 + 3] * ix[1]);
    }
  } else if (len == 3) {
    for (long i = 0; i < n; i++) {
      double* iy = y + i * 3;
      iy[0] *= beta;
      iy[1] *= beta;
      iy[2] *= beta;
      for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
        const double* jx = x + (long)colind[q] * 3;
        const double* a = crs + (long)q * 9;
        iy[0] += alpha * (a[0] * jx[0] + a[1] * jx[1] + a[2] * jx[2]);
        iy[1] += alpha * (a[3] * jx[0] + a[4] * jx[1] + a[5] * jx[2]);
        iy[2] += alpha * (a[6] * jx[0] + a[7] * jx[1] + a[8] * jx[2]);
      }
      const double* ix = x + i * 3;
      const double* a = dia + i * 9;
      iy[0] += alpha * (a[0] * ix[0] + a[1] * ix[1] + a[2] * ix[2]);
      iy[1] += alpha * (a[3] * ix[0] + a[4] * ix[1] + a[5] * ix[2]);
      iy[2] += alpha * (a[6] * ix[0] + a[7] * ix[1] + a[8] * ix[2]);
    }
  } else {
    const int bl = len, bs = len * len;
    for (long i = 0; i < n; i++) {
      double* iy = y + i * bl;
      for (int d = 0; d < bl; d++) iy[d] *= beta;
      for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
        const double* jx = x + (long)colind[q] * bl;
        for (int a = 0; a < bl; a++)
          for (int b = 0; b < bl; b++) iy[a] += alpha * crs[(long)q * bs + a * bl + b] * jx[b];
      }
      const double* ix = x + i * bl;
      for (int a = 0; a < bl; a++)
        for (int b = 0; b < bl; b++) iy[a] += alpha * dia[i * bs + a * bl + b] * ix[b];
    }
  }
}

/* ---------------------------------------------------------------------------------------------------- */
/* ILU(0). Symbolic: src/matvec/matdiafrac_blkcrs.cpp:227-287 (m_DiaInd), and with an ordering           */
/* src/matvec/mat_blkcrs.cpp:493-549 (permuted, std::sort per row). Values: matdia_blkcrs.cpp:673-760.   */
void cr_ilu0_pattern(int n, const unsigned* rowptr, const unsigned* colind, const int* order,
                     unsigned* frowptr, unsigned* fcolind, unsigned* diaind) {
  if (!order) {
    memcpy(frowptr, rowptr, sizeof(unsigned) * ((size_t)n + 1));
    memcpy(fcolind, colind, sizeof(unsigned) * rowptr[n]);
  } else {
    int* inv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) inv[order[i]] = i;
    frowptr[0] = 0;
    for (int i0 = 0; i0 < n; i0++) {
      const int i1 = order[i0];
      frowptr[i0 + 1] = frowptr[i0] + (rowptr[i1 + 1] - rowptr[i1]);
    }
    for (int i0 = 0; i0 < n; i0++) {
      const int i1 = order[i0];
      unsigned pos = frowptr[i0];
      for (unsigned q = rowptr[i1]; q < rowptr[i1 + 1]; q++) fcolind[pos++] = (unsigned)inv[colind[q]];
      qsort(&fcolind[frowptr[i0]], frowptr[i0 + 1] - frowptr[i0], sizeof(unsigned), cmp_u);
    }
    free(inv);
  }
  for (int i = 0; i < n; i++) {
    diaind[i] = frowptr[i + 1];
    for (unsigned q = frowptr[i]; q < frowptr[i + 1]; q++)
      if (fcolind[q] > (unsigned)i) { diaind[i] = q; break; }
  }
}

/* Row i of the new pattern is a sorted singly linked list over column numbers (nxt[], -1 ends it, head hd);  */
/* lv[] holds the entry levels. Rows before i are final in frowptr/fcolind/flev. Visiting the row's entries in */
/* ascending order while inserting behind the visited one is the traversal of AddFracPtn (:1330-1361): the   */
/* first insertion of a column fixes its level, max(lev_ik, lev_kj) + 1 (:1354-1356); entries at or above    */
/* lev_fill neither pivot nor spread.                                                                           */
long cr_iluk_pattern(int n, const unsigned* rowptr, const unsigned* colind, int lev_fill, long cap, unsigned* frowptr,
                     unsigned* fcolind, unsigned* diaind, int* flev) {
  int* nxt = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int* lv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int* tmplev = flev ? NULL : (int*)malloc(sizeof(int) * (size_t)(cap > 0 ? cap : 1));
  int* L = flev ? flev : tmplev;
  long pos = 0, need = 0;
  int overflow = 0;
  frowptr[0] = 0;
  for (int i = 0; i < n; i++) {
    int hd = -1, tail = -1;
    for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
      const int c = (int)colind[q];
      lv[c] = 0;
      nxt[c] = -1;
      if (tail < 0) hd = c; else nxt[tail] = c;
      tail = c;
    }
    if (lev_fill != 0 && !overflow) {
      for (int k = hd; k >= 0; k = nxt[k]) {
        const int lik = lv[k];
        if (lev_fill != -1 && lik + 1 > lev_fill) continue;
        if (k >= i) break;
        int at = k; /* insertion point: every new column of this pivot lies behind it, in ascending order */
        for (unsigned kj = diaind[k]; kj < frowptr[k + 1]; kj++) {
          const int lkj = L[kj];
          if (lev_fill != -1 && lkj + 1 > lev_fill) continue;
          const int j = (int)fcolind[kj];
          if (j == i) continue;
          while (nxt[at] >= 0 && nxt[at] < j) at = nxt[at];
          if (nxt[at] == j) continue;
          lv[j] = (lik > lkj ? lik : lkj) + 1;
          nxt[j] = nxt[at];
          nxt[at] = j;
          at = j;
        }
      }
    }
    long cnt = 0;
    for (int c = hd; c >= 0; c = nxt[c]) cnt++;
    need += cnt;
    if (overflow || pos + cnt > cap) {
      overflow = 1;
      continue;
    }
    const unsigned end = (unsigned)(pos + cnt);
    diaind[i] = end;
    for (int c = hd; c >= 0; c = nxt[c]) {
      if ((unsigned)c > (unsigned)i && diaind[i] == end) diaind[i] = (unsigned)pos;
      fcolind[pos] = (unsigned)c;
      L[pos] = lv[c];
      pos++;
    }
    frowptr[i + 1] = (unsigned)pos;
  }
  free(nxt);
  free(lv);
  free(tmplev);
  return overflow ? -need : pos;
}

/* src/matvec/matdiafrac_blkcrs.h:90-117 (SetValue_Initialize) -> CMat_BlkCrs::SetValue: the factor pattern  */
/* (possibly with fill) takes A's values where A has an entry, zero elsewhere; under an ordering new row i0   */
/* takes old row order[i0].                                                                                    */
void cr_ilu0_setvalue(int n, int len, const unsigned* rowptr, const unsigned* colind, const double* crs,
                      const double* dia, const int* order, const unsigned* frowptr, const unsigned* fcolind,
                      double* fcrs, double* fdia) {
  const int bs = len * len;
  memset(fcrs, 0, sizeof(double) * (size_t)frowptr[n] * bs);
  if (!order && frowptr[n] == rowptr[n]) {
    memcpy(fcrs, crs, sizeof(double) * (size_t)rowptr[n] * bs);
    memcpy(fdia, dia, sizeof(double) * (size_t)n * bs);
    return;
  }
  int* inv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) inv[order ? order[i] : i] = i;
  for (int i0 = 0; i0 < n; i0++) {
    const int i1 = order ? order[i0] : i0;
    for (unsigned q1 = rowptr[i1]; q1 < rowptr[i1 + 1]; q1++) {
      const unsigned j0 = (unsigned)inv[colind[q1]];
      for (unsigned q0 = frowptr[i0]; q0 < frowptr[i0 + 1]; q0++)
        if (fcolind[q0] == j0) {
          memcpy(&fcrs[(size_t)q0 * bs], &crs[(size_t)q1 * bs], sizeof(double) * bs);
          break;
        }
    }
    memcpy(&fdia[(size_t)i0 * bs], &dia[(size_t)i1 * bs], sizeof(double) * bs);
  }
  free(inv);
}

/* src/matvec/matdiafrac_blkcrs.cpp:774-1107 (DoILUDecomp), with the b = 1, 2, 3 forms and the general   */
/* form through CalcSubMatPr (:62-73), CalcInvMat (:107-136) and CalcMatPr (:45-60).                     */
static void inv_general(double* a, int n, int* info) {
  *info = 0;
  for (int i = 0; i < n; i++) {
    if (fabs(a[i * n + i]) < 1.0e-30) { *info = 1; return; }
    if (a[i * n + i] < 0.0) (*info)--;
    double tmp1 = 1.0 / a[i * n + i];
    a[i * n + i] = 1.0;
    for (int k = 0; k < n; k++) a[i * n + k] *= tmp1;
    for (int j = 0; j < n; j++) {
      if (j != i) {
        tmp1 = a[j * n + i];
        a[j * n + i] = 0.0;
        for (int k = 0; k < n; k++) a[j * n + k] -= tmp1 * a[i * n + k];
      }
    }
  }
}

int cr_ilu0_factor(int n, int len, const unsigned* frowptr, const unsigned* fcolind, const unsigned* diaind,
                   double* fcrs, double* fdia) {
  const int bl = len, bs = len * len;
  const unsigned nmax_sing = 10;
  unsigned icnt_sing = 0;
  int* row2crs = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; j++) row2crs[j] = -1;
  double tmp[64];
  int ok = 1;
  for (int i = 0; i < n && ok; i++) {
    for (unsigned q = frowptr[i]; q < frowptr[i + 1]; q++) row2crs[fcolind[q]] = (int)q;
    for (unsigned ik = frowptr[i]; ik < diaind[i]; ik++) {
      const unsigned k = fcolind[ik];
      const double* pik = &fcrs[(size_t)ik * bs];
      for (unsigned kj = diaind[k]; kj < frowptr[k + 1]; kj++) {
        const unsigned j = fcolind[kj];
        const double* pkj = &fcrs[(size_t)kj * bs];
        double* pij;
        if (j != (unsigned)i) {
          const int ij = row2crs[j];
          if (ij == -1) continue;
          pij = &fcrs[(size_t)ij * bs];
        } else {
          pij = &fdia[(size_t)i * bs];
        }
        if (bl == 1) {
          pij[0] -= pik[0] * pkj[0];
        } else if (bl == 2) {
          pij[0] -= pik[0] * pkj[0] + pik[1] * pkj[2];
          pij[1] -= pik[0] * pkj[1] + pik[1] * pkj[3];
          pij[2] -= pik[2] * pkj[0] + pik[3] * pkj[2];
          pij[3] -= pik[2] * pkj[1] + pik[3] * pkj[3];
        } else if (bl == 3) {
          for (int a = 0; a < 3; a++) {
            pij[a * 3 + 0] -= pik[a * 3 + 0] * pkj[0] + pik[a * 3 + 1] * pkj[3] + pik[a * 3 + 2] * pkj[6];
            pij[a * 3 + 1] -= pik[a * 3 + 0] * pkj[1] + pik[a * 3 + 1] * pkj[4] + pik[a * 3 + 2] * pkj[7];
            pij[a * 3 + 2] -= pik[a * 3 + 0] * pkj[2] + pik[a * 3 + 1] * pkj[5] + pik[a * 3 + 2] * pkj[8];
          }
        } else {
          for (int a = 0; a < bl; a++)
            for (int b = 0; b < bl; b++)
              for (int c = 0; c < bl; c++) pij[a * bl + b] -= pik[a * bl + c] * pkj[c * bl + b];
        }
      }
    }
    double* pii = &fdia[(size_t)i * bs];
    int singular = 0;
    if (bl == 1) {
      const double iiv = pii[0];
      if (fabs(iiv) > 1.0e-30) pii[0] = 1.0 / iiv;
      else singular = 1;
    } else if (bl == 2) {
      const double det = pii[0] * pii[3] - pii[1] * pii[2];
      if (fabs(det) > 1.0e-30) {
        const double inv_det = 1.0 / det;
        const double dtmp1 = pii[0];
        pii[0] = inv_det * pii[3];
        pii[1] = -inv_det * pii[1];
        pii[2] = -inv_det * pii[2];
        pii[3] = inv_det * dtmp1;
      } else singular = 1;
    } else if (bl == 3) {
      const double* a = pii;
      const double det = a[0] * a[4] * a[8] + a[3] * a[7] * a[2] + a[6] * a[1] * a[5] - a[0] * a[7] * a[5] -
                         a[6] * a[4] * a[2] - a[3] * a[1] * a[8];
      if (fabs(det) > 1.0e-30) {
        /* CalcInvMat3, :139-158 (it recomputes det) */
        const double det2 = a[0] * a[4] * a[8] + a[3] * a[7] * a[2] + a[6] * a[1] * a[5] -
                            a[0] * a[7] * a[5] - a[6] * a[4] * a[2] - a[3] * a[1] * a[8];
        const double inv_det = 1.0 / det2;
        double t[9];
        for (int k = 0; k < 9; k++) t[k] = pii[k];
        pii[0] = inv_det * (t[4] * t[8] - t[5] * t[7]);
        pii[1] = inv_det * (t[2] * t[7] - t[1] * t[8]);
        pii[2] = inv_det * (t[1] * t[5] - t[2] * t[4]);
        pii[3] = inv_det * (t[5] * t[6] - t[3] * t[8]);
        pii[4] = inv_det * (t[0] * t[8] - t[2] * t[6]);
        pii[5] = inv_det * (t[2] * t[3] - t[0] * t[5]);
        pii[6] = inv_det * (t[3] * t[7] - t[4] * t[6]);
        pii[7] = inv_det * (t[1] * t[6] - t[0] * t[7]);
        pii[8] = inv_det * (t[0] * t[4] - t[1] * t[3]);
      } else singular = 1;
    } else {
      int info = 0;
      inv_general(pii, bl, &info);
      if (info == 1) singular = 1;
    }
    if (singular) {
      icnt_sing++;
      if (icnt_sing > nmax_sing) { ok = 0; break; }
    }
    for (unsigned ij = diaind[i]; ij < frowptr[i + 1]; ij++) {
      double* pij = &fcrs[(size_t)ij * bs];
      if (bl == 1) {
        pij[0] = pij[0] * pii[0];
      } else if (bl == 2) {
        for (int k = 0; k < 4; k++) tmp[k] = pij[k];
        pij[0] = pii[0] * tmp[0] + pii[1] * tmp[2];
        pij[1] = pii[0] * tmp[1] + pii[1] * tmp[3];
        pij[2] = pii[2] * tmp[0] + pii[3] * tmp[2];
        pij[3] = pii[2] * tmp[1] + pii[3] * tmp[3];
      } else if (bl == 3) {
        for (int k = 0; k < 9; k++) tmp[k] = pij[k];
        for (int a = 0; a < 3; a++) {
          pij[a * 3 + 0] = pii[a * 3 + 0] * tmp[0] + pii[a * 3 + 1] * tmp[3] + pii[a * 3 + 2] * tmp[6];
          pij[a * 3 + 1] = pii[a * 3 + 0] * tmp[1] + pii[a * 3 + 1] * tmp[4] + pii[a * 3 + 2] * tmp[7];
          pij[a * 3 + 2] = pii[a * 3 + 0] * tmp[2] + pii[a * 3 + 1] * tmp[5] + pii[a * 3 + 2] * tmp[8];
        }
      } else {
        for (int a = 0; a < bl; a++)
          for (int b = 0; b < bl; b++) {
            tmp[a * bl + b] = 0.0;
            for (int c = 0; c < bl; c++) tmp[a * bl + b] += pii[a * bl + c] * pij[c * bl + b];
          }
        for (int k = 0; k < bs; k++) pij[k] = tmp[k];
      }
    }
    for (unsigned q = frowptr[i]; q < frowptr[i + 1]; q++) row2crs[fcolind[q]] = -1;
  }
  free(row2crs);
  return ok;
}

/* src/matvec/matdiafrac_blkcrs.cpp:332-507 (ForwardSubstitution), :509-669 (BackwardSubstitution),      */
/* src/ls/preconditioner.cpp:301-335 (SolvePrecond), src/matvec/ordering_blk.cpp:136-166 (permutation).  */
/* dirs: bit 0 the forward substitution, bit 1 the backward one */
static void fwd_bwd_dirs(int n, int bl, const unsigned* rp, const unsigned* ci, const unsigned* di,
                         const double* fcrs, const double* fdia, double* v, int dirs) {
  const int bs = bl * bl;
  double t[8];
  for (long i = 0; i < n && (dirs & 1); i++) {
    for (int d = 0; d < bl; d++) t[d] = v[i * bl + d];
    for (unsigned q = rp[i]; q < di[i]; q++) {
      const long j = ci[q];
      const double* a = &fcrs[(size_t)q * bs];
      if (bl == 1) {
        t[0] -= a[0] * v[j];
      } else if (bl == 2) {
        const double v0 = v[j * 2], v1 = v[j * 2 + 1];
        t[0] -= a[0] * v0 + a[1] * v1;
        t[1] -= a[2] * v0 + a[3] * v1;
      } else if (bl == 3) {
        const double v0 = v[j * 3], v1 = v[j * 3 + 1], v2 = v[j * 3 + 2];
        t[0] -= a[0] * v0 + a[1] * v1 + a[2] * v2;
        t[1] -= a[3] * v0 + a[4] * v1 + a[5] * v2;
        t[2] -= a[6] * v0 + a[7] * v1 + a[8] * v2;
      } else {
        for (int r = 0; r < bl; r++)
          for (int c = 0; c < bl; c++) t[r] -= a[r * bl + c] * v[j * bl + c];
      }
    }
    const double* a = &fdia[(size_t)i * bs];
    if (bl == 1) {
      v[i] = a[0] * t[0];
    } else if (bl == 2) {
      v[i * 2] = a[0] * t[0] + a[1] * t[1];
      v[i * 2 + 1] = a[2] * t[0] + a[3] * t[1];
    } else if (bl == 3) {
      v[i * 3] = a[0] * t[0] + a[1] * t[1] + a[2] * t[2];
      v[i * 3 + 1] = a[3] * t[0] + a[4] * t[1] + a[5] * t[2];
      v[i * 3 + 2] = a[6] * t[0] + a[7] * t[1] + a[8] * t[2];
    } else {
      for (int r = 0; r < bl; r++) {
        double s = 0.0;
        for (int c = 0; c < bl; c++) s += a[r * bl + c] * t[c];
        v[i * bl + r] = s;
      }
    }
  }
  for (long i = (long)n - 1; i >= 0 && (dirs & 2); i--) {
    for (int d = 0; d < bl; d++) t[d] = v[i * bl + d];
    for (unsigned q = di[i]; q < rp[i + 1]; q++) {
      const long j = ci[q];
      const double* a = &fcrs[(size_t)q * bs];
      if (bl == 1) {
        t[0] -= a[0] * v[j];
      } else if (bl == 2) {
        const double v0 = v[j * 2], v1 = v[j * 2 + 1];
        t[0] -= a[0] * v0 + a[1] * v1;
        t[1] -= a[2] * v0 + a[3] * v1;
      } else if (bl == 3) {
        const double v0 = v[j * 3], v1 = v[j * 3 + 1], v2 = v[j * 3 + 2];
        t[0] -= a[0] * v0 + a[1] * v1 + a[2] * v2;
        t[1] -= a[3] * v0 + a[4] * v1 + a[5] * v2;
        t[2] -= a[6] * v0 + a[7] * v1 + a[8] * v2;
      } else {
        for (int r = 0; r < bl; r++)
          for (int c = 0; c < bl; c++) t[r] -= a[r * bl + c] * v[j * bl + c];
      }
    }
    for (int d = 0; d < bl; d++) v[i * bl + d] = t[d];
  }
}

static void fwd_bwd(int n, int bl, const unsigned* rp, const unsigned* ci, const unsigned* di, const double* fcrs,
                    const double* fdia, double* v) {
  fwd_bwd_dirs(n, bl, rp, ci, di, fcrs, fdia, v, 3);
}

void cr_ilu_solve(int n, int len, const unsigned* frowptr, const unsigned* fcolind, const unsigned* diaind,
                  const double* fcrs, const double* fdia, const int* order, double* vec, double* work) {
  if (!order) {
    fwd_bwd(n, len, frowptr, fcolind, diaind, fcrs, fdia, vec);
    return;
  }
  for (long i = 0; i < n; i++)
    for (int d = 0; d < len; d++) work[i * len + d] = vec[(long)order[i] * len + d];
  fwd_bwd(n, len, frowptr, fcolind, diaind, fcrs, fdia, work);
  for (long i = 0; i < n; i++)
    for (int d = 0; d < len; d++) vec[(long)order[i] * len + d] = work[i * len + d];
}

/* ---------------------------------------------------------------------------------------------------- */
/* Multi-segment systems: LsSol::CLinearSystem with several segments (src/ls/linearsystem.cpp:74-415) and  */
/* CPreconditioner_ILU over them (src/ls/preconditioner.cpp:57-335), ILU(0): the diagonal segments are     */
/* CMatDiaFrac_BlkCrs, the couplings CMatFrac_BlkCrs (src/matvec/matfrac_blkcrs.cpp).                      */

/* CalcSubMatPr (src/matvec/matfrac_blkcrs.cpp:43-54): out -= a b, one product at a time */
static void sub_matpr(double* out, const double* a, const double* b, int ni, int nk, int nj) {
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < nj; j++)
      for (int k = 0; k < nk; k++) out[i * nj + j] -= a[i * nk + k] * b[k * nj + j];
}

/* CalcMatPr (:56-72): out := d out, d square ni x ni, each entry summed from 0.0 */
static void matpr(double* out, const double* d, int ni, int nj) {
  double tmp[64];
  for (int i = 0; i < ni; i++)
    for (int j = 0; j < nj; j++) {
      tmp[i * nj + j] = 0.0;
      for (int k = 0; k < ni; k++) tmp[i * nj + j] += d[i * ni + k] * out[k * nj + j];
    }
  for (int i = 0; i < ni * nj; i++) out[i] = tmp[i];
}

/* CMat_BlkCrs::MatVec(alpha, x, beta, y, isnt_trans = true), src/matvec/mat_blkcrs.cpp:955-1175: the shape
 * branches sum in different orders (1 x n: alpha (sum of the row's products); n x 1: products with alpha x;
 * otherwise one alpha a x term at a time). */
void cr_blk_matvec(const cr_blk* A, double alpha, const double* x, double beta, double* y) {
  const int bi = A->bi, bj = A->bj, bs = bi * bj;
  const unsigned *rp = A->rowptr, *ci = A->colind;
  const double* v = A->val;
  for (long i = 0; i < A->nrow; i++) {
    double* iy = y + i * bi;
    if (bi == 1 && bj == 1 && beta == 0.0) {
      iy[0] = 0.0;
    } else {
      for (int d = 0; d < bi; d++) iy[d] *= beta;
    }
    for (unsigned q = rp[i]; q < rp[i + 1]; q++) {
      const long j = ci[q];
      const double* a = v + (size_t)q * bs;
      const double* jx = x + j * bj;
      if (bi == 1 && bj == 1) {
        iy[0] += alpha * a[0] * jx[0];
      } else if (bi == 1 && bj == 2) {
        iy[0] += alpha * (a[0] * jx[0] + a[1] * jx[1]);
      } else if (bi == 1 && bj == 3) {
        iy[0] += alpha * (a[0] * jx[0] + a[1] * jx[1] + a[2] * jx[2]);
      } else if (bi == 1) {
        for (int c = 0; c < bj; c++) iy[0] += alpha * a[c] * jx[c];
      } else if (bj == 1) {
        const double ax = jx[0] * alpha;
        for (int r = 0; r < bi; r++) iy[r] += a[r] * ax;
      } else {
        for (int r = 0; r < bi; r++)
          for (int c = 0; c < bj; c++) iy[r] += alpha * a[r * bj + c] * jx[c];
      }
    }
  }
}

/* CLinearSystem::FinalizeMarge (src/ls/linearsystem.cpp:92-128): SetBoundaryCondition of every diagonal segment,
 * SetBoundaryCondition_Row of every coupling by its row segment's flags, then _Colum by its column segment's
 * (src/matvec/mat_blkcrs.cpp:829-874, 910-953), the fixed residual entries zeroed; the norm sums the segments'
 * squared norms in segment order. flags and res are concatenated over the segments. */
double cr_ms_set_bc(cr_msys* s, const int* flags, double* res) {
  const int ns = s->nseg;
  for (int i = 0; i < ns; i++) {
    cr_blk* A = s->a[i * ns + i];
    if (!A) continue;
    const int bl = A->bi, bs = bl * bl;
    const int* fl = flags + s->off[i];
    for (long r = 0; r < A->nrow; r++)
      for (int id = 0; id < bl; id++) {
        if (fl[r * bl + id] == 0) continue;
        for (int k = 0; k < bl; k++) A->dia[r * bs + k * bl + id] = 0.0;
        for (int j = 0; j < bl; j++) A->dia[r * bs + id * bl + j] = 0.0;
        A->dia[r * bs + id * bl + id] = 1.0;
        for (unsigned q = A->rowptr[r]; q < A->rowptr[r + 1]; q++)
          for (int j = 0; j < bl; j++) A->val[(size_t)q * bs + id * bl + j] = 0.0;
      }
    const unsigned ncrs = A->rowptr[A->nrow];
    for (unsigned q = 0; q < ncrs; q++) {
      const unsigned j1 = A->colind[q];
      for (int jd = 0; jd < bl; jd++) {
        if (fl[(long)j1 * bl + jd] == 0) continue;
        for (int id = 0; id < bl; id++) A->val[(size_t)q * bs + id * bl + jd] = 0.0;
      }
    }
  }
  for (int i = 0; i < ns; i++)
    for (int j = 0; j < ns; j++) {
      cr_blk* A = s->a[i * ns + j];
      if (!A || i == j) continue;
      const int bi = A->bi, bj = A->bj, bs = bi * bj;
      const int* fl = flags + s->off[i];
      for (long r = 0; r < A->nrow; r++)
        for (int id = 0; id < bi; id++) {
          if (fl[r * bi + id] == 0) continue;
          for (unsigned q = A->rowptr[r]; q < A->rowptr[r + 1]; q++)
            for (int jd = 0; jd < bj; jd++) A->val[(size_t)q * bs + id * bj + jd] = 0.0;
        }
    }
  for (int j = 0; j < ns; j++)
    for (int i = 0; i < ns; i++) {
      cr_blk* A = s->a[i * ns + j];
      if (!A || i == j) continue;
      const int bi = A->bi, bj = A->bj, bs = bi * bj;
      const int* fl = flags + s->off[j];
      const unsigned ncrs = A->rowptr[A->nrow];
      for (unsigned q = 0; q < ncrs; q++) {
        const unsigned c = A->colind[q];
        for (int jd = 0; jd < bj; jd++) {
          if (fl[(long)c * bj + jd] == 0) continue;
          for (int id = 0; id < bi; id++) A->val[(size_t)q * bs + id * bj + jd] = 0.0;
        }
      }
    }
  for (long k = 0; k < s->off[ns]; k++)
    if (flags[k] == 1) res[k] = 0.0;
  double sq = 0.0;
  for (int i = 0; i < ns; i++) {
    double si = 0.0;
    for (long k = s->off[i]; k < s->off[i + 1]; k++) si += res[k] * res[k];
    sq += si;
  }
  return sqrt(sq);
}

/* CLinearSystem::MATVEC (src/ls/linearsystem.cpp:250-291): per segment its diagonal matrix
 * (CMatDia_BlkCrs::MatVec) or y *= beta, then the couplings of its row in segment order with beta 1. */
void cr_ms_matvec(const cr_msys* s, double alpha, const double* x, double beta, double* y) {
  const int ns = s->nseg;
  if (alpha == 0.0) {
    for (long k = 0; k < s->off[ns]; k++) y[k] *= beta;
    return;
  }
  for (int i = 0; i < ns; i++) {
    const cr_blk* D = s->a[i * ns + i];
    double* yi = y + s->off[i];
    if (D) cr_matvec(D->nrow, D->bi, D->rowptr, D->colind, D->val, D->dia, alpha, x + s->off[i], beta, yi);
    else
      for (long k = s->off[i]; k < s->off[i + 1]; k++) y[k] *= beta;
    for (int j = 0; j < ns; j++) {
      const cr_blk* A = s->a[i * ns + j];
      if (!A || i == j) continue;
      cr_blk_matvec(A, alpha, x + s->off[j], 1.0, yi);
    }
  }
}

static double ms_dot(const cr_msys* s, const double* a, const double* b) {
  double d = 0.0;
  for (int i = 0; i < s->nseg; i++) {
    double di = 0.0;
    for (long k = s->off[i]; k < s->off[i + 1]; k++) di += a[k] * b[k];
    d += di;
  }
  return d;
}

/* CMatDiaFrac_BlkCrs::DoILUDecompLowUp (src/matvec/matdiafrac_blkcrs.cpp:671-772, uniform blocks):
 * D_ij -= L_ik U_kj over the coupling factors low (segment i x k) and up (k x i), rows in order, each row's
 * low entries ascending, entries outside the factor pattern dropped */
static void ilu_dia_lowup(cr_blk* D, const cr_blk* low, const cr_blk* up, int* row2crs) {
  const int bi = D->bi, bk = low->bj;
  for (long i = 0; i < D->nrow; i++) {
    for (unsigned q = D->rowptr[i]; q < D->rowptr[i + 1]; q++) row2crs[D->colind[q]] = (int)q;
    for (unsigned ik = low->rowptr[i]; ik < low->rowptr[i + 1]; ik++) {
      const unsigned k = low->colind[ik];
      const double* pik = low->val + (size_t)ik * bi * bk;
      for (unsigned kj = up->rowptr[k]; kj < up->rowptr[k + 1]; kj++) {
        const unsigned j = up->colind[kj];
        const double* pkj = up->val + (size_t)kj * bk * bi;
        if (j != (unsigned)i) {
          const int ij = row2crs[j];
          if (ij == -1) continue;
          sub_matpr(D->val + (size_t)ij * bi * bi, pik, pkj, bi, bk, bi);
        } else {
          sub_matpr(D->dia + (size_t)i * bi * bi, pik, pkj, bi, bk, bi);
        }
      }
    }
    for (unsigned q = D->rowptr[i]; q < D->rowptr[i + 1]; q++) row2crs[D->colind[q]] = -1;
  }
}

/* CMatFrac_BlkCrs::DoILUDecompLowUp (src/matvec/matfrac_blkcrs.cpp:540-599): F_ij -= L_ik U_kj */
static void ilu_od_lowup(cr_blk* F, const cr_blk* low, const cr_blk* up, int* row2crs) {
  const int bi = F->bi, bj = F->bj, bk = low->bj;
  for (long i = 0; i < F->nrow; i++) {
    for (unsigned q = F->rowptr[i]; q < F->rowptr[i + 1]; q++) row2crs[F->colind[q]] = (int)q;
    for (unsigned ik = low->rowptr[i]; ik < low->rowptr[i + 1]; ik++) {
      const unsigned k = low->colind[ik];
      const double* pik = low->val + (size_t)ik * bi * bk;
      for (unsigned kj = up->rowptr[k]; kj < up->rowptr[k + 1]; kj++) {
        const int ij = row2crs[up->colind[kj]];
        if (ij == -1) continue;
        sub_matpr(F->val + (size_t)ij * bi * bj, pik, up->val + (size_t)kj * bk * bj, bi, bk, bj);
      }
    }
    for (unsigned q = F->rowptr[i]; q < F->rowptr[i + 1]; q++) row2crs[F->colind[q]] = -1;
  }
}

/* CMatFrac_BlkCrs::DoILUDecompUp (:601-701, uniform blocks): the upper coupling F (segment i x j, i < j) is
 * eliminated with the lower part of the factored diagonal segment D = segment i (its rows in order, reading
 * the already scaled rows k < i of F), then each row is scaled by D's inverted diagonal block */
static void ilu_od_up(cr_blk* F, const cr_blk* D, int* row2crs) {
  const int bi = F->bi, bj = F->bj;
  for (long i = 0; i < F->nrow; i++) {
    for (unsigned q = F->rowptr[i]; q < F->rowptr[i + 1]; q++) row2crs[F->colind[q]] = (int)q;
    for (unsigned ik = D->rowptr[i]; ik < D->diaind[i]; ik++) {
      const unsigned k = D->colind[ik];
      const double* pik = D->val + (size_t)ik * bi * bi;
      for (unsigned kj = F->rowptr[k]; kj < F->rowptr[k + 1]; kj++) {
        const int ij = row2crs[F->colind[kj]];
        if (ij == -1) continue;
        sub_matpr(F->val + (size_t)ij * bi * bj, pik, F->val + (size_t)kj * bi * bj, bi, bi, bj);
      }
    }
    const double* dii = D->dia + (size_t)i * bi * bi;
    for (unsigned q = F->rowptr[i]; q < F->rowptr[i + 1]; q++) matpr(F->val + (size_t)q * bi * bj, dii, bi, bj);
    for (unsigned q = F->rowptr[i]; q < F->rowptr[i + 1]; q++) row2crs[F->colind[q]] = -1;
  }
}

/* CMatFrac_BlkCrs::DoILUDecompLow (:703-783, uniform blocks): the lower coupling F (segment i x j, i > j)
 * against the upper part of the factored diagonal segment D = segment j: F_ij -= F_ik U_kj, each row's
 * entries ascending */
static void ilu_od_low(cr_blk* F, const cr_blk* D, int* row2crs) {
  const int bi = F->bi, bj = F->bj;
  for (long i = 0; i < F->nrow; i++) {
    for (unsigned q = F->rowptr[i]; q < F->rowptr[i + 1]; q++) row2crs[F->colind[q]] = (int)q;
    for (unsigned ik = F->rowptr[i]; ik < F->rowptr[i + 1]; ik++) {
      const unsigned k = F->colind[ik];
      const double* pik = F->val + (size_t)ik * bi * bj;
      for (unsigned kj = D->diaind[k]; kj < D->rowptr[k + 1]; kj++) {
        const int ij = row2crs[D->colind[kj]];
        if (ij == -1) continue;
        sub_matpr(F->val + (size_t)ij * bi * bj, pik, D->val + (size_t)kj * bj * bj, bi, bj, bj);
      }
    }
    for (unsigned q = F->rowptr[i]; q < F->rowptr[i + 1]; q++) row2crs[F->colind[q]] = -1;
  }
}

/* CPreconditioner_ILU::SetValue (src/ls/preconditioner.cpp:215-297), ILU(0): the factors f take the matrices'
 * values (SetValue_Initialize), then segment row by segment row: the diagonal segment gets the LowUp updates of
 * the earlier segments and is factored (DoILUDecomp); a coupling of the row gets LowUp for the segments before
 * both, then Up (row segment before the column's) or Low (after) against the factored diagonal segment.
 * Returns 0 when a diagonal factorisation fails (more than 10 singular pivots), else 1. */
int cr_ms_ilu0(cr_msys* s) {
  const int ns = s->nseg;
  int nmax = 1;
  for (int i = 0; i < ns; i++) nmax = s->nblk[i] > nmax ? s->nblk[i] : nmax;
  int* row2crs = (int*)malloc(sizeof(int) * (size_t)nmax);
  for (int k = 0; k < nmax; k++) row2crs[k] = -1;
  for (int p = 0; p < ns * ns; p++) {
    const cr_blk* A = s->a[p];
    cr_blk* F = s->f[p];
    if (!A) continue;
    const size_t bs = (size_t)A->bi * A->bj;
    memcpy(F->val, A->val, sizeof(double) * bs * A->rowptr[A->nrow]);
    if (A->dia) memcpy(F->dia, A->dia, sizeof(double) * bs * A->nrow);
  }
  int ok = 1;
  for (int i = 0; i < ns && ok; i++) {
    for (int j = 0; j < ns && ok; j++) {
      if (i == j) {
        cr_blk* D = s->f[i * ns + i];
        for (int k = 0; k < i; k++) {
          const cr_blk *low = s->f[i * ns + k], *up = s->f[k * ns + i];
          if (low && up) ilu_dia_lowup(D, low, up, row2crs);
        }
        if (!cr_ilu0_factor(D->nrow, D->bi, D->rowptr, D->colind, D->diaind, D->val, D->dia)) ok = 0;
        continue;
      }
      cr_blk* F = s->f[i * ns + j];
      if (!F) continue;
      const int kmax = i < j ? i : j;
      for (int k = 0; k <= kmax; k++) {
        if (k == i && k < j) {
          ilu_od_up(F, s->f[i * ns + i], row2crs);
          continue;
        }
        if (k == j && k < i) {
          ilu_od_low(F, s->f[j * ns + j], row2crs);
          continue;
        }
        const cr_blk *low = s->f[i * ns + k], *up = s->f[k * ns + j];
        if (!low || !up) continue;
        ilu_od_lowup(F, low, up, row2crs);
      }
    }
  }
  free(row2crs);
  return ok;
}

/* CPreconditioner_ILU::SolvePrecond (src/ls/preconditioner.cpp:301-335): forward, segment by segment, the
 * lower couplings' MatVec(-1, v_j, 1, v_i) then the diagonal segment's forward substitution; backward from the
 * last segment, the upper couplings then the backward substitution. v is concatenated over the segments. */
void cr_ms_precond(const cr_msys* s, double* v) {
  const int ns = s->nseg;
  for (int i = 0; i < ns; i++) {
    for (int j = 0; j < i; j++) {
      const cr_blk* F = s->f[i * ns + j];
      if (F) cr_blk_matvec(F, -1.0, v + s->off[j], 1.0, v + s->off[i]);
    }
    const cr_blk* D = s->f[i * ns + i];
    fwd_bwd_dirs(D->nrow, D->bi, D->rowptr, D->colind, D->diaind, D->val, D->dia, v + s->off[i], 1);
  }
  for (int i = ns - 1; i >= 0; i--) {
    for (int j = i + 1; j < ns; j++) {
      const cr_blk* F = s->f[i * ns + j];
      if (F) cr_blk_matvec(F, -1.0, v + s->off[j], 1.0, v + s->off[i]);
    }
    const cr_blk* D = s->f[i * ns + i];
    fwd_bwd_dirs(D->nrow, D->bi, D->rowptr, D->colind, D->diaind, D->val, D->dia, v + s->off[i], 2);
  }
}

/* ---------------------------------------------------------------------------------------------------- */
/* Krylov loops on handle semantics: src/ls/solver_ls_iter.cpp:100-172 (Solve_PCG), :291-408             */
/* (Solve_PBiCGSTAB); vector ops src/matvec/vector_blk.cpp:37-112; MATVEC src/ls/linearsystem.cpp:250-291.*/
static double v_dot(long m, const double* a, const double* b) {
  double d = 0.0;
  for (long i = 0; i < m; i++) d += a[i] * b[i];
  return d;
}
static void v_axpy(long m, double alpha, const double* x, double* y) {
  if (alpha == 0) return; /* CLinearSystem::AXPY returns early on alpha == 0 (linearsystem.cpp:302) */
  for (long i = 0; i < m; i++) y[i] += alpha * x[i];
}
static void v_scal(long m, double alpha, double* x) {
  for (long i = 0; i < m; i++) x[i] *= alpha;
}
static void precond(const cr_system* s, double* v, double* work) {
  cr_ilu_solve(s->n, s->len, s->frowptr, s->fcolind, s->diaind, s->fcrs, s->fdia, s->order, v, work);
}
static void matvec(const cr_system* s, double alpha, const double* x, double beta, double* y) {
  if (alpha == 0.0) { v_scal((long)s->n * s->len, beta, y); return; }
  cr_matvec(s->n, s->len, s->rowptr, s->colind, s->crs, s->dia, alpha, x, beta, y);
}

/* The operator a preconditioned loop runs on: one segment (cr_system) or several (cr_msys, whose inner product
 * is the sum of the segments' sequential sums, src/ls/linearsystem.cpp:335-360). */
typedef struct {
  long m;
  const cr_system* s;
  const cr_msys* ms;
} kop;
static double k_dot(const kop* o, const double* a, const double* b) {
  if (o->s) return v_dot(o->m, a, b);
  return ms_dot(o->ms, a, b);
}
static void k_precond(const kop* o, double* v, double* work) {
  if (o->s) precond(o->s, v, work);
  else cr_ms_precond(o->ms, v);
}
static void k_matvec(const kop* o, double alpha, const double* x, double beta, double* y) {
  if (o->s) matvec(o->s, alpha, x, beta, y);
  else cr_ms_matvec(o->ms, alpha, x, beta, y);
}

static int pcg_loop(const kop* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  const double tol = *conv_ratio;
  const unsigned mx_iter = *iter;
  const long m = s->m;
  double* p = (double*)calloc((size_t)m + 1, sizeof(double));
  double* z = (double*)calloc((size_t)m + 1, sizeof(double));
  double* w = (double*)calloc((size_t)m + 1, sizeof(double));
  int ret = 0;
  v_scal(m, 0.0, x);
  double sq_inv_norm_res0;
  {
    const double sq0 = k_dot(s, r, r);
    if (sq0 < 1.0e-30) {
      *conv_ratio = 0.0;
      *iter = 0;
      ret = 1;
      goto done;
    }
    sq_inv_norm_res0 = 1.0 / sq0;
  }
  memcpy(z, r, sizeof(double) * m);
  k_precond(s, z, w);
  memcpy(p, z, sizeof(double) * m);
  {
    double inpro_rz = k_dot(s, r, z);
    for (unsigned it = 0; it < mx_iter; it++) {
      double alpha;
      {
        k_matvec(s, 1.0, p, 0.0, z);
        const double pap = k_dot(s, p, z);
        alpha = inpro_rz / pap;
      }
      v_axpy(m, -alpha, z, r);
      v_axpy(m, alpha, p, x);
      {
        const double sq = k_dot(s, r, r);
        if (sq * sq_inv_norm_res0 < tol * tol) {
          *conv_ratio = sqrt(sq * sq_inv_norm_res0);
          *iter = it;
          ret = 1;
          goto done;
        }
      }
      double beta;
      {
        memcpy(z, r, sizeof(double) * m);
        k_precond(s, z, w);
        const double rz_new = k_dot(s, r, z);
        beta = rz_new / inpro_rz;
        inpro_rz = rz_new;
      }
      v_scal(m, beta, p);
      v_axpy(m, 1.0, z, p);
    }
    {
      const double sq = k_dot(s, r, r);
      *conv_ratio = sqrt(sq * sq_inv_norm_res0);
      ret = 0;
    }
  }
done:
  free(p); free(z); free(w);
  return ret;
}

static int pbicgstab_loop(const kop* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  const double tol = *conv_ratio;
  const unsigned max_iter = *iter;
  const long m = s->m;
  double* buf = (double*)calloc((size_t)m * 8 + 1, sizeof(double));
  double *vs = buf, *vMs = buf + m, *vAMs = buf + 2 * m, *vp = buf + 3 * m, *vMp = buf + 4 * m,
         *vAMp = buf + 5 * m, *vr2 = buf + 6 * m, *w = buf + 7 * m;
  int ret = 1;
  double sq_inv_norm_res_ini;
  {
    const double sq0 = k_dot(s, r, r);
    if (sq0 < 1.0e-60) {
      *conv_ratio = sqrt(sq0);
      *iter = 0;
      goto done;
    }
    sq_inv_norm_res_ini = 1.0 / sq0;
  }
  v_scal(m, 0.0, x);
  memcpy(vr2, r, sizeof(double) * m);
  memcpy(vp, r, sizeof(double) * m);
  *iter = max_iter;
  for (unsigned it = 1; it < max_iter; it++) {
    memcpy(vMp, vp, sizeof(double) * m);
    k_precond(s, vMp, w);
    const double r_r2 = k_dot(s, r, vr2);
    k_matvec(s, 1.0, vMp, 0.0, vAMp);
    double alpha;
    {
      const double den = k_dot(s, vAMp, vr2);
      alpha = r_r2 / den;
    }
    memcpy(vs, r, sizeof(double) * m);
    v_axpy(m, -alpha, vAMp, vs);
    memcpy(vMs, vs, sizeof(double) * m);
    k_precond(s, vMs, w);
    k_matvec(s, 1.0, vMs, 0.0, vAMs);
    double omega;
    {
      const double den = k_dot(s, vAMs, vAMs);
      const double num = k_dot(s, vs, vAMs);
      omega = num / den;
    }
    v_axpy(m, alpha, vMp, x);
    v_axpy(m, omega, vMs, x);
    memcpy(r, vs, sizeof(double) * m);
    v_axpy(m, -omega, vAMs, r);
    {
      const double sq = k_dot(s, r, r);
      const double sqc = sq * sq_inv_norm_res_ini;
      if (sqc < tol * tol) {
        *conv_ratio = sqrt(sq * sq_inv_norm_res_ini);
        *iter = it;
        goto done;
      }
    }
    double beta;
    {
      const double tmp1 = k_dot(s, r, vr2);
      beta = tmp1 * alpha / (r_r2 * omega);
    }
    v_scal(m, beta, vp);
    v_axpy(m, 1.0, r, vp);
    v_axpy(m, -beta * omega, vAMp, vp);
  }
done:
  free(buf);
  return ret;
}

int cr_solve_pcg(const cr_system* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  const kop o = {(long)s->n * s->len, s, 0};
  return pcg_loop(&o, r, x, conv_ratio, iter);
}

int cr_solve_pbicgstab(const cr_system* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  const kop o = {(long)s->n * s->len, s, 0};
  return pbicgstab_loop(&o, r, x, conv_ratio, iter);
}

int cr_ms_solve(const cr_msys* ms, int method, double* r, double* x, double* conv_ratio, unsigned* iter) {
  const kop o = {ms->off[ms->nseg], 0, ms};
  return method == 0 ? pcg_loop(&o, r, x, conv_ratio, iter) : pbicgstab_loop(&o, r, x, conv_ratio, iter);
}

/* Solve_CG (src/ls/solver_ls_iter.cpp:31-97): no preconditioner; guard 1e-60 returns ||r0|| as conv_ratio;
 * x updated before r; returns 1 also when max_iter is reached (conv_ratio then keeps its input value). */
int cr_solve_cg(const cr_system* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  const unsigned max_iter = *iter;
  const double tol = *conv_ratio;
  const long m = (long)s->n * s->len;
  double* p = (double*)calloc((size_t)m + 1, sizeof(double));
  double* ap = (double*)calloc((size_t)m + 1, sizeof(double));
  v_scal(m, 0.0, x);
  double sq = v_dot(m, r, r);
  if (sq < 1.0e-60) {
    *iter = 0;
    *conv_ratio = sqrt(sq);
    free(p);
    free(ap);
    return 1;
  }
  const double sq_inv0 = 1.0 / sq;
  memcpy(p, r, sizeof(double) * m);
  unsigned it;
  for (it = 0; it < max_iter; it++) {
    double alpha;
    {
      matvec(s, 1.0, p, 0.0, ap);
      const double pap = v_dot(m, p, ap);
      alpha = sq / pap;
    }
    v_axpy(m, alpha, p, x);
    v_axpy(m, -alpha, ap, r);
    const double sq_new = v_dot(m, r, r);
    if (sq_new * sq_inv0 < tol * tol) {
      *conv_ratio = sqrt(sq_new * sq_inv0);
      break;
    }
    const double beta = sq_new / sq;
    sq = sq_new;
    v_scal(m, beta, p);
    v_axpy(m, 1.0, r, p);
  }
  *iter = it;
  free(p);
  free(ap);
  return 1;
}

/* Solve_BiCGSTAB (src/ls/solver_ls_iter.cpp:177-286): no preconditioner; loop from 1; iteration = max_iter
 * unless converged; returns 1 also when max_iter is reached. */
int cr_solve_bicgstab(const cr_system* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  const unsigned max_iter = *iter;
  const double tol = *conv_ratio;
  const long m = (long)s->n * s->len;
  double* buf = (double*)calloc((size_t)5 * m + 1, sizeof(double));
  double *vs = buf, *vas = buf + m, *vp = buf + 2 * m, *vap = buf + 3 * m, *vr2 = buf + 4 * m;
  v_scal(m, 0.0, x);
  double sq_inv0;
  {
    const double sq0 = v_dot(m, r, r);
    if (sq0 < 1.0e-60) {
      *conv_ratio = sqrt(sq0);
      *iter = 0;
      free(buf);
      return 1;
    }
    sq_inv0 = 1.0 / sq0;
  }
  memcpy(vr2, r, sizeof(double) * m);
  memcpy(vp, r, sizeof(double) * m);
  double r_r2 = v_dot(m, r, vr2);
  *iter = max_iter;
  for (unsigned it = 1; it < max_iter; it++) {
    matvec(s, 1.0, vp, 0.0, vap);
    const double alpha = r_r2 / v_dot(m, vap, vr2);
    memcpy(vs, r, sizeof(double) * m);
    v_axpy(m, -alpha, vap, vs);
    matvec(s, 1.0, vs, 0.0, vas);
    const double den = v_dot(m, vas, vas);
    const double num = v_dot(m, vas, vs);
    const double omega = num / den;
    v_axpy(m, alpha, vp, x);
    v_axpy(m, omega, vs, x);
    memcpy(r, vs, sizeof(double) * m);
    v_axpy(m, -omega, vas, r);
    {
      const double sq = v_dot(m, r, r);
      const double sqc = sq * sq_inv0;
      if (sqc < tol * tol) {
        *conv_ratio = sqrt(sqc);
        *iter = it;
        break;
      }
    }
    const double tmp1 = v_dot(m, r, vr2);
    const double beta = (tmp1 * alpha) / (r_r2 * omega);
    r_r2 = tmp1;
    v_scal(m, beta, vp);
    v_axpy(m, 1.0, r, vp);
    v_axpy(m, -beta * omega, vap, vp);
  }
  free(buf);
  return 1;
}

/* ---- COrdering_Blk (src/matvec/ordering_blk.cpp) ------------------------------------------------------- */

/* MakeOrdering_RCM (:62-134): level-by-level breadth-first search from block 0; within a level the blocks are taken
 * in ascending degree, ties in the order they were reached (:90-97, a std::multimap keyed by degree), each gets the
 * next number (:102-105) and appends its unreached neighbours in CRS order (:107-119); at the end every number is
 * reversed (:127-129). The numbers sit in the slot of their block (m_pOrder[iblk0] = iblk_cur, :103). */
static int cr_order_rcm(int n, const unsigned* rowptr, const unsigned* colind, int* order) {
  int *level = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1)), *cur = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1)),
      *nxt = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int ncur = 0, nnxt = 0, taken = 0, lev = 0, ok = 1;
  for (int i = 0; i < n; i++) { level[i] = -1; order[i] = -1; }
  if (n > 0) { level[0] = 0; nxt[nnxt++] = 0; }
  while (nnxt > 0) {
    /* stable insertion sort by degree */
    ncur = 0;
    for (int k = 0; k < nnxt; k++) {
      const int b = nxt[k];
      const unsigned d = rowptr[b + 1] - rowptr[b];
      int pos = ncur;
      while (pos > 0 && rowptr[cur[pos - 1] + 1] - rowptr[cur[pos - 1]] > d) { cur[pos] = cur[pos - 1]; pos--; }
      cur[pos] = b;
      ncur++;
    }
    nnxt = 0;
    for (int k = 0; k < ncur; k++) {
      const int b = cur[k];
      order[b] = taken++;
      for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++) {
        const int j = (int)colind[q];
        if (level[j] != -1) continue;
        level[j] = lev + 1;
        nxt[nnxt++] = j;
      }
    }
    lev++;
  }
  if (taken != n) ok = 0;
  else for (int b = 0; b < n; b++) order[b] = n - 1 - order[b];
  free(level); free(cur); free(nxt);
  return ok;
}

/* MakeOrdering_RCM2 (:169-262): level 0 = the blocks with no lower neighbour, in index order (:186-205), their
 * neighbours not yet in level 0 or 1 become candidates of level 1; a candidate is accepted unless one of its lower
 * neighbours is unnumbered or in the current level (then it drops back to unreached, :222-236); accepted blocks are
 * numbered in order and their unreached higher neighbours become candidates of the next level (:239-255).
 * m_pOrder[new] = old (:243). */
static int cr_order_rcm2(int n, const unsigned* rowptr, const unsigned* colind, int* order) {
  const size_t m = (size_t)(n > 0 ? n : 1);
  int *level = malloc(sizeof(int) * m), *cand = malloc(sizeof(int) * m * 4 + 4), *next = malloc(sizeof(int) * m);
  int ncand = 0, capc = (int)(m * 4 + 1), taken = 0, lev = 0;
  for (int i = 0; i < n; i++) { level[i] = -1; order[i] = -1; }
  for (int b = 0; b < n; b++) {
    int lower = 0;
    for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++) if ((int)colind[q] < b) { lower = 1; break; }
    if (lower) continue;
    level[b] = lev;
    order[taken++] = b;
    for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++) {
      const int j = (int)colind[q];
      if (level[j] == lev || level[j] == lev + 1) continue;
      if (ncand == capc) { capc *= 2; cand = realloc(cand, sizeof(int) * (size_t)capc); }
      cand[ncand++] = j;
      level[j] = lev + 1;
    }
  }
  lev++;
  while (ncand > 0) {
    int nnext = 0;
    for (int k = 0; k < ncand; k++) {
      const int b = cand[k];
      int blocked = 0;
      for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++) {
        const int j = (int)colind[q];
        if (j < b && (level[j] == lev || level[j] == -1)) { blocked = 1; break; }
      }
      if (blocked) { level[b] = -1; continue; }
      next[nnext++] = b;
    }
    ncand = 0;
    for (int k = 0; k < nnext; k++) {
      const int b = next[k];
      order[taken++] = b;
      for (unsigned q = rowptr[b]; q < rowptr[b + 1]; q++) {
        const int j = (int)colind[q];
        if (level[j] != -1 || j < b) continue;
        level[j] = lev + 1;
        if (ncand == capc) { capc *= 2; cand = realloc(cand, sizeof(int) * (size_t)capc); }
        cand[ncand++] = j;
      }
    }
    lev++;
  }
  free(level); free(cand); free(next);
  return taken == n;
}

/* MakeOrdering_AMD (:320-485) as compiled: each step takes the uneliminated block of least degree, the lowest index
 * among equals (:376-389), numbered in turn (m_pOrder[iblk_cur] = iblk_del, :393); the degree update loops run
 * over aEnode / aAroundEnode, which only the commented-out AroundEnode call fills (:414-419), so degrees never
 * change. Restated literally (a scan per step), which is quadratic: for test sizes. */
static int cr_order_amd(int n, const unsigned* rowptr, const unsigned* colind, int* order) {
  (void)colind;
  char* done = calloc((size_t)(n > 0 ? n : 1), 1);
  for (int k = 0; k < n; k++) {
    unsigned best = 0xffffffffu;
    int del = -1;
    for (int b = 0; b < n; b++)
      if (!done[b] && rowptr[b + 1] - rowptr[b] < best) { best = rowptr[b + 1] - rowptr[b]; del = b; }
    done[del] = 1;
    order[k] = del;
  }
  free(done);
  return 1;
}

int cr_ordering(int kind, int n, const unsigned* rowptr, const unsigned* colind, int* order) {
  if (kind == 3) return cr_order_rcm(n, rowptr, colind, order);
  if (kind == 4) return cr_order_rcm2(n, rowptr, colind, order);
  if (kind == 5) return cr_order_amd(n, rowptr, colind, order);
  return 0;
}

/* ---- MatVec::Sol twins of the Krylov loops (src/matvec/solver_mat_iter.cpp) ---------------------------- */

/* MatVec::Sol::Solve_CG (:37-113): x = 0; guard ||r0||^2 < 1e-30 gives conv_ratio 0, iteration 0 (:59-64); loop
 * from 1 (:74): alpha = (r,r)/(p,Ap), r -= alpha Ap, then the test on the residual norm of the iteration before
 * the update (sq_norm_res, not the one just formed: :86-93), x += alpha p only after the test, beta from the new
 * norm. Not converged: iteration = max, conv_ratio keeps its input; returns 1 either way. */
int cr_solve_mat_cg(const cr_system* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  const unsigned mx_iter = *iter;
  const double tol = *conv_ratio;
  const long m = (long)s->n * s->len;
  v_scal(m, 0.0, x);
  double sq = v_dot(m, r, r);
  if (sq < 1.0e-30) {
    *conv_ratio = 0.0;
    *iter = 0;
    return 1;
  }
  const double sq_inv = 1.0 / sq;
  double* p = (double*)calloc((size_t)m + 1, sizeof(double));
  double* ap = (double*)calloc((size_t)m + 1, sizeof(double));
  memcpy(p, r, sizeof(double) * m);
  *iter = mx_iter;
  for (unsigned it = 1; it < mx_iter; it++) {
    double alpha;
    {
      matvec(s, 1.0, p, 0.0, ap);
      const double pap = v_dot(m, p, ap);
      alpha = sq / pap;
    }
    v_axpy(m, -alpha, ap, r);
    const double sq1 = v_dot(m, r, r);
    if (sq * sq_inv < tol * tol) {
      *conv_ratio = sqrt(sq * sq_inv);
      *iter = it;
      break;
    }
    v_axpy(m, alpha, p, x);
    const double beta = sq1 / sq;
    sq = sq1;
    v_scal(m, beta, p);
    v_axpy(m, 1.0, r, p);
  }
  free(p);
  free(ap);
  return 1;
}

/* MatVec::Sol::Solve_PCG (:115-184): the steps of Solve_PCG (same guard, 0-based iteration, same arithmetic: its
 * (r, z) recomputed at the top of each iteration equals the carried one bit for bit); not converged: returns 0
 * and leaves conv_ratio and iteration as they came in. */
int cr_solve_mat_pcg(const cr_system* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  const double tol = *conv_ratio;
  const unsigned mx = *iter;
  const int ok = cr_solve_pcg(s, r, x, conv_ratio, iter);
  if (!ok) {
    *conv_ratio = tol;
    *iter = mx;
  }
  return ok;
}

/* MatVec::Sol::Solve_BiCGSTAB (:191-299) and Solve_PBiCGSTAB (:305-423): the steps of the LsSol loops except the
 * guard, ||r0||^2 < 1e-30 giving conv_ratio 0 (:219-224, :348-353); their (r, r0*) recomputed at the top of each
 * iteration equals the carried one bit for bit. */
static int mat_guard(const cr_system* s, const double* r, double* x, double* conv_ratio, unsigned* iter) {
  const long m = (long)s->n * s->len;
  if (v_dot(m, r, r) < 1.0e-30) {
    v_scal(m, 0.0, x);
    *conv_ratio = 0.0;
    *iter = 0;
    return 1;
  }
  return 0;
}
int cr_solve_mat_bicgstab(const cr_system* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  if (mat_guard(s, r, x, conv_ratio, iter)) return 1;
  return cr_solve_bicgstab(s, r, x, conv_ratio, iter);
}
int cr_solve_mat_pbicgstab(const cr_system* s, double* r, double* x, double* conv_ratio, unsigned* iter) {
  if (mat_guard(s, r, x, conv_ratio, iter)) return 1;
  cr_solve_pbicgstab(s, r, x, conv_ratio, iter);
  return 1;
}

/* ---------------------------------------------------------------------------------------------------- */
/* Saved stiffness: CLinearSystem_Save (src/femls/linearsystem_fieldsave.cpp) with its boundary matrix,  */
/* a CMat_BlkCrs (src/matvec/mat_blkcrs.cpp), restated step by step.                                     */

long cr_boundary_pattern(int n, int len, const unsigned* rowptr, const unsigned* colind, const int* flags,
                         unsigned* browptr, unsigned* bcolind) {
  /* SetPatternBoundary (:1177-1265): rows with a free dof; in them the diagonal block first when it has a
   * fixed dof, then the blocks of the row (pattern order) whose node has a fixed dof */
  long nb = 0;
  if (browptr) browptr[0] = 0;
  for (long i = 0; i < n; i++) {
    int has_free = 0, has_fixed = 0;
    for (int d = 0; d < len; d++) {
      if (flags[i * len + d] == 0) has_free = 1;
      else has_fixed = 1;
    }
    if (has_free) {
      if (has_fixed) {
        if (bcolind) bcolind[nb] = (unsigned)i;
        nb++;
      }
      for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
        const unsigned j = colind[q];
        int fj = 0;
        for (int d = 0; d < len; d++)
          if (flags[(long)j * len + d] != 0) fj = 1;
        if (!fj) continue;
        if (bcolind) bcolind[nb] = j;
        nb++;
      }
    }
    if (browptr) browptr[i + 1] = (unsigned)nb;
  }
  return nb;
}

/* AddLinearSystem_Poisson2D_P1 on CLinearSystem_Save, src/femeqn/eqn_poisson.cpp:494-569 */
static void elem_poisson2d_save(const double co[][3], const double* p, double* emat, double* eforce) {
  const double alpha = p[0], source = p[1];
  const double area = tri_area(co[0], co[1], co[2]);
  double dldx[3][2];
  tri_dldx(dldx, co[0], co[1], co[2]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      emat[i * 3 + j] = alpha * area * (dldx[i][0] * dldx[j][0] + dldx[i][1] * dldx[j][1]);
  for (int i = 0; i < 3; i++) eforce[i] = source * area * 0.333333333333333333;
}

/* AddLinSys_LinearSolid2D_Static_P1_SaveStiffMat, src/femeqn/eqn_linear_solid2d.cpp:861-960, and
 * AddLinSys_LinearSolid3D_Static_P1 on CLinearSystem_Save, src/femeqn/eqn_linear_solid3d.cpp:425-521:
 * emat[ino][jno][idim][jdim], zero-initialised and accumulated as written there */
static void elem_solid_save(int ndim, const double co[][3], const double* p, double* emat, double* eforce) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  const int nno = ndim + 1, bs = ndim * ndim;
  double dldx[4][3];
  double w;
  if (ndim == 2) {
    double d2[3][2];
    w = tri_area(co[0], co[1], co[2]);
    tri_dldx(d2, co[0], co[1], co[2]);
    for (int k = 0; k < 3; k++) { dldx[k][0] = d2[k][0]; dldx[k][1] = d2[k][1]; }
  } else {
    w = tet_volume(co[0], co[1], co[2], co[3]);
    tet_dldx(dldx, co[0], co[1], co[2], co[3]);
  }
  for (int i = 0; i < nno * nno * bs; i++) emat[i] = 0.0;
  for (int i = 0; i < nno * ndim; i++) eforce[i] = 0.0;
  for (int ino = 0; ino < nno; ino++) {
    for (int jno = 0; jno < nno; jno++) {
      double* e = &emat[(ino * nno + jno) * bs];
      double dtmp1 = 0.0;
      for (int idim = 0; idim < ndim; idim++) {
        for (int jdim = 0; jdim < ndim; jdim++)
          e[idim * ndim + jdim] += w * (lambda * dldx[ino][idim] * dldx[jno][jdim] + myu * dldx[jno][idim] * dldx[ino][jdim]);
        dtmp1 += dldx[ino][idim] * dldx[jno][idim];
      }
      for (int idim = 0; idim < ndim; idim++) e[idim * ndim + idim] += w * myu * dtmp1;
    }
  }
  for (int ino = 0; ino < nno; ino++)
    for (int d = 0; d < ndim; d++)
      eforce[ino * ndim + d] = (ndim == 2) ? w * rho * p[3 + d] * 0.3333333333333333333 : w * rho * p[3 + d] * 0.25;
}

/* CMat_BlkCrs::Mearge (src/matvec/mat_blkcrs.cpp:754-823): blocks absent from the pattern are skipped */
static void merge_crs(int nno, const int* no, int bs, const double* emat, const unsigned* browptr,
                      const unsigned* bcolind, double* bval) {
  for (int a = 0; a < nno; a++) {
    const int i = no[a];
    for (int c = 0; c < nno; c++) {
      const unsigned j = (unsigned)no[c];
      long pos = -1;
      for (unsigned q = browptr[i]; q < browptr[i + 1]; q++)
        if (bcolind[q] == j) { pos = q; break; }
      if (pos < 0) continue;
      const double* pin = &emat[(a * nno + c) * bs];
      double* pout = &bval[pos * bs];
      for (int k = 0; k < bs; k++) pout[k] += pin[k];
    }
  }
}

int cr_assemble_save(int kind, int ndim, int nnode, const double* coords, int nelem, const int* lnods,
                     const double* params, int pstride, int len, const unsigned* rowptr, const unsigned* colind,
                     double* crs, double* dia, const unsigned* browptr, const unsigned* bcolind, double* bval,
                     double* force) {
  (void)nnode;
  const int nno = (ndim == 2) ? 3 : 4;
  const int bs = len * len;
  double emat[144], eforce[12], co[4][3];
  if (kind == CR_POISSON && ndim != 2) return -1; /* TET11: "Not Implimented" there (eqn_poisson.cpp:704-707) */
  if (kind != CR_POISSON && kind != CR_SOLID2D && kind != CR_SOLID3D) return -1;
  for (long e = 0; e < nelem; e++) {
    const int* no = &lnods[e * nno];
    const double* p = &params[e * pstride];
    for (int a = 0; a < nno; a++)
      for (int d = 0; d < 3; d++) co[a][d] = (d < ndim) ? coords[(long)no[a] * ndim + d] : 0.0;
    if (kind == CR_POISSON) elem_poisson2d_save(co, p, emat, eforce);
    else elem_solid_save(ndim, co, p, emat, eforce);
    merge(nno, no, bs, emat, rowptr, colind, crs, dia);
    merge_crs(nno, no, bs, emat, browptr, bcolind, bval);
    for (int a = 0; a < nno; a++)
      for (int d = 0; d < len; d++) force[(long)no[a] * len + d] += eforce[a * len + d];
  }
  return 0;
}

void cr_boundary_bc(int n, int len, const unsigned* browptr, const unsigned* bcolind, double* bval, const int* flags) {
  const int bs = len * len;
  /* SetBoundaryConditionInverse_Colum (:876-900): columns of free dofs to zero */
  for (unsigned q = 0; q < browptr[n]; q++) {
    const unsigned j = bcolind[q];
    for (int jd = 0; jd < len; jd++) {
      if (flags[(long)j * len + jd] != 0) continue;
      for (int id = 0; id < len; id++) bval[(long)q * bs + id * len + jd] = 0.0;
    }
  }
  /* SetBoundaryCondition_Row (:910-953): rows of fixed dofs to zero */
  for (long i = 0; i < n; i++)
    for (int id = 0; id < len; id++) {
      if (flags[i * len + id] == 0) continue;
      for (unsigned q = browptr[i]; q < browptr[i + 1]; q++)
        for (int jd = 0; jd < len; jd++) bval[(long)q * bs + id * len + jd] = 0.0;
    }
}

/* CMat_BlkCrs::MatVec(-1, x, 1, res) (src/matvec/mat_blkcrs.cpp:955-1170) over a boundary matrix: 1x1 blocks
 * (:1007-1020), otherwise the general branch (:1150-1168) */
static void boundary_matvec(int n, int len, const unsigned* browptr, const unsigned* bcolind, const double* bval,
                            const double* x, double* res) {
  const double alpha = -1.0, beta = 1.0;
  const int bs = len * len;
  if (len == 1) {
    for (long i = 0; i < n; i++) {
      double* lval = &res[i];
      *lval *= beta;
      for (unsigned q = browptr[i]; q < browptr[i + 1]; q++) *lval += alpha * bval[q] * x[bcolind[q]];
    }
  } else {
    for (long i = 0; i < n; i++) {
      double* iy = &res[i * len];
      for (int id = 0; id < len; id++) iy[id] *= beta;
      for (unsigned q = browptr[i]; q < browptr[i + 1]; q++) {
        const double* jx = &x[(long)bcolind[q] * len];
        for (int id = 0; id < len; id++)
          for (int jd = 0; jd < len; jd++) iy[id] += alpha * bval[(long)q * bs + id * len + jd] * jx[jd];
      }
    }
  }
}

double cr_make_residual_save(int n, int len, const unsigned* browptr, const unsigned* bcolind, const double* bval,
                             const double* force, const double* u, double* res, double* upd) {
  const long m = (long)n * len;
  /* update = field values; residual = force (MakeResidual, src/femls/linearsystem_fieldsave.cpp:206-247) */
  for (long k = 0; k < m; k++) upd[k] = u[k];
  for (long k = 0; k < m; k++) res[k] = force[k];
  boundary_matvec(n, len, browptr, bcolind, bval, upd, res);
  /* sqrt(DOT(-1,-1)) */
  double sq = 0.0;
  for (long k = 0; k < m; k++) sq += res[k] * res[k];
  return sqrt(sq);
}

void cr_update_field_save(int n, int len, const int* flags, double* upd, double* u) {
  /* UpdateValueOfField of CLinearSystem_Save, VALUE (src/femls/linearsystem_fieldsave.cpp:249-296) */
  const long m = (long)n * len;
  for (long k = 0; k < m; k++)
    if (flags[k] != 0) upd[k] = u[k];
  for (long k = 0; k < m; k++) u[k] = upd[k];
}

/* ---------------------------------------------------------------------------------------------------- */
/* Saved stiffness with a lumped mass: CLinearSystem_SaveDiaM_Newmark / _NewmarkBeta                     */
/* (src/femls/linearsystem_fieldsave.cpp:363-870), restated step by step.                                */

long cr_pattern_dia(int n, const unsigned* rowptr, const unsigned* colind, unsigned* browptr, unsigned* bcolind) {
  /* CMat_BlkCrs::SetPatternDia (src/matvec/mat_blkcrs.cpp:1267-1319): the row of the matrix pattern with the
   * diagonal placed just before the first column below it; last when no column is below it */
  long nb = 0;
  if (browptr) browptr[0] = 0;
  for (long i = 0; i < n; i++) {
    int dia_done = 0;
    for (unsigned q = rowptr[i]; q < rowptr[i + 1]; q++) {
      const unsigned j = colind[q];
      if ((unsigned)i > j && !dia_done) {
        if (bcolind) bcolind[nb] = (unsigned)i;
        nb++;
        dia_done = 1;
      }
      if (bcolind) bcolind[nb] = j;
      nb++;
    }
    if (!dia_done) {
      if (bcolind) bcolind[nb] = (unsigned)i;
      nb++;
    }
    if (browptr) browptr[i + 1] = (unsigned)nb;
  }
  return nb;
}

/* AddLinearSystem_Diffusion2D_P1 on CLinearSystem_SaveDiaM_Newmark, src/femeqn/eqn_diffusion.cpp:565-650:
 * eC (the boundary matrix's element), emat = (gamma dt) eC + the lumped mass rho area / 3 on the diagonal, and the
 * source term with its 14-digit literal. p = (rho, alpha, source) */
static void elem_diffusion2d_savedia(const double co[][3], const double* p, double dt, double gamma, double* emat,
                                     double* ec, double* eforce) {
  const double rho = p[0], alpha = p[1], source = p[2];
  const double area = tri_area(co[0], co[1], co[2]);
  double dldx[3][2];
  tri_dldx(dldx, co[0], co[1], co[2]);
  double em[3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) ec[i * 3 + j] = alpha * area * (dldx[i][0] * dldx[j][0] + dldx[i][1] * dldx[j][1]);
  {
    const double dtmp1 = rho * area / 3.0;
    for (int i = 0; i < 3; i++) em[i] = dtmp1;
  }
  for (int i = 0; i < 3; i++) eforce[i] = source * area * 0.33333333333333;
  {
    const double dtmp1 = gamma * dt;
    for (int i = 0; i < 3; i++) {
      for (int j = 0; j < 3; j++) emat[i * 3 + j] = dtmp1 * ec[i * 3 + j];
      emat[i * 3 + i] += em[i];
    }
  }
}

/* AddLinSys_LinearSolid2D_NonStatic_Save_NewmarkBeta_P1, src/femeqn/eqn_linear_solid2d.cpp:1051-1180: eK (the
 * boundary matrix's element, written as the static saved loop writes it), the lumped mass area rho / 3 on the
 * diagonal of the diagonal blocks, emat = M + (beta dt dt) K on the diagonal blocks and (beta dt dt) K elsewhere,
 * and the gravity force of each node as two values (x, then y) that the reference adds to the node's first dof.
 * p = (lambda, myu, rho, gx, gy); eforce[2 a + d] is the value added d-th. */
static void elem_solid2d_savedia(const double co[][3], const double* p, double dt, double beta, double* emat,
                                 double* ek, double* eforce) {
  const double rho = p[2], g_x = p[3], g_y = p[4];
  double em[3][4], dummy[6];
  elem_solid_save(2, co, p, ek, dummy);
  const double area = tri_area(co[0], co[1], co[2]);
  for (int a = 0; a < 3; a++)
    for (int k = 0; k < 4; k++) em[a][k] = 0.0;
  {
    const double dtmp1 = area * rho * 0.33333333333333333333333;
    for (int a = 0; a < 3; a++) {
      em[a][0] += dtmp1;
      em[a][3] += dtmp1;
    }
  }
  for (int a = 0; a < 3; a++) {
    eforce[a * 2 + 0] = area * rho * g_x * 0.33333333333333333333333333;
    eforce[a * 2 + 1] = area * rho * g_y * 0.33333333333333333333333333;
  }
  const double dtmp1 = beta * dt * dt;
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++)
      for (int k = 0; k < 4; k++) {
        const double kv = ek[(a * 3 + c) * 4 + k];
        emat[(a * 3 + c) * 4 + k] = (a == c) ? em[a][k] + dtmp1 * kv : dtmp1 * kv;
      }
}

int cr_assemble_savedia(int kind, int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                        int pstride, const double* glob, int len, const unsigned* rowptr, const unsigned* colind,
                        double* crs, double* dia, const unsigned* browptr, const unsigned* bcolind, double* bval,
                        double* force) {
  (void)nnode;
  const double dt = glob[0], gamma = glob[1], beta = glob[2];
  const int bs = len * len;
  double emat[36], ek[36], eforce[6], co[4][3];
  if (!((kind == CR_DIFFUSION2D && len == 1) || (kind == CR_SOLID2D && len == 2))) return -1;
  for (long e = 0; e < nelem; e++) {
    const int* no = &lnods[e * 3];
    const double* p = &params[e * pstride];
    for (int a = 0; a < 3; a++)
      for (int d = 0; d < 3; d++) co[a][d] = (d < 2) ? coords[(long)no[a] * 2 + d] : 0.0;
    if (kind == CR_DIFFUSION2D) elem_diffusion2d_savedia(co, p, dt, gamma, emat, ek, eforce);
    else elem_solid2d_savedia(co, p, dt, beta, emat, ek, eforce);
    merge(3, no, bs, emat, rowptr, colind, crs, dia);
    merge_crs(3, no, bs, ek, browptr, bcolind, bval);
    if (kind == CR_DIFFUSION2D) {
      for (int a = 0; a < 3; a++) force[no[a]] += eforce[a];
    } else {
      /* force_c.AddValue(noes[ino], 0, eqf_out[ino][idim]) for idim = 0, 1 */
      for (int a = 0; a < 3; a++)
        for (int d = 0; d < 2; d++) force[(long)no[a] * 2] += eforce[a * 2 + d];
    }
  }
  return 0;
}

void cr_boundary_rows(int n, int len, const unsigned* browptr, double* bval, const int* flags) {
  /* SetBoundaryCondition_Row (src/matvec/mat_blkcrs.cpp:910-953): rows of fixed dofs to zero */
  const int bs = len * len;
  for (long i = 0; i < n; i++)
    for (int id = 0; id < len; id++) {
      if (flags[i * len + id] == 0) continue;
      for (unsigned q = browptr[i]; q < browptr[i + 1]; q++)
        for (int jd = 0; jd < len; jd++) bval[(long)q * bs + id * len + jd] = 0.0;
    }
}

double cr_make_residual_savedia(int kind, int n, int len, const unsigned* browptr, const unsigned* bcolind,
                                const double* bval, const double* force, const double* glob, const int* flags,
                                const double* u, const double* v, const double* a, double* res, double* upd) {
  const double dt = glob[0], gamma = glob[1], beta = glob[2];
  const long m = (long)n * len;
  /* residual = force; update = u + (dt (1 - gamma)) v (Newmark, :433-457), or u + dt v + (dt dt (0.5 - beta)) a
   * (NewmarkBeta, :705-737), each term added by CNodeAry::AddValueFromNodeSegment (value * alpha,
   * src/femfield/node_ary.cpp:302-325) */
  for (long k = 0; k < m; k++) res[k] = force[k];
  for (long k = 0; k < m; k++) upd[k] = u[k];
  if (kind == CR_DIFFUSION2D) {
    const double c1 = dt * (1 - gamma);
    for (long k = 0; k < m; k++) upd[k] += v[k] * c1;
  } else {
    const double c2 = dt * dt * (0.5 - beta);
    for (long k = 0; k < m; k++) upd[k] += v[k] * dt;
    for (long k = 0; k < m; k++) upd[k] += a[k] * c2;
  }
  /* the boundary matrix's MatVec(-1, update, 1, residual) (src/matvec/mat_blkcrs.cpp:955-1170) */
  boundary_matvec(n, len, browptr, bcolind, bval, upd, res);
  /* SetZeroToBCDof of the residual, then sqrt(DOT(-1,-1)) */
  for (long k = 0; k < m; k++)
    if (flags[k] != 0) res[k] = 0.0;
  double sq = 0.0;
  for (long k = 0; k < m; k++) sq += res[k] * res[k];
  return sqrt(sq);
}

void cr_update_field_savedia(int kind, int n, int len, const double* glob, const int* flags, double* upd, double* u,
                             double* v, double* a) {
  const double dt = glob[0], gamma = glob[1], beta = glob[2];
  const long m = (long)n * len;
  if (kind == CR_DIFFUSION2D) {
    /* CLinearSystem_SaveDiaM_Newmark::UpdateValueOfField (:474-520): free dofs u += dt (1 - gamma) v0 +
     * dt gamma v1 (v1: the update); fixed dofs keep their velocity in the update; then velocity = update */
    for (long k = 0; k < m; k++) {
      const double velo0 = v[k];
      if (flags[k] == 0) {
        const double velo1 = upd[k];
        u[k] += dt * (1 - gamma) * velo0 + dt * gamma * velo1;
      } else {
        upd[k] = velo0;
      }
    }
    for (long k = 0; k < m; k++) v[k] = upd[k];
  } else {
    /* CLinearSystem_SaveDiaM_NewmarkBeta::UpdateValueOfField (:760-860), ACCELERATION: free dofs
     * u += dt v0 + dt dt (0.5 - beta) a0 + dt dt beta a1, v += dt (1 - gamma) a0 + gamma dt a1 (a1: the update);
     * fixed dofs keep their acceleration in the update; then acceleration = update */
    for (long k = 0; k < m; k++) {
      const double velo0 = v[k], acc0 = a[k];
      if (flags[k] == 0) {
        const double acc1 = upd[k];
        u[k] += dt * velo0 + dt * dt * (0.5 - beta) * acc0 + dt * dt * beta * acc1;
        v[k] += dt * (1 - gamma) * acc0 + gamma * dt * acc1;
        a[k] += acc1;
      } else {
        upd[k] = acc0;
      }
    }
    for (long k = 0; k < m; k++) a[k] = upd[k];
  }
}

/* ---------------------------------------------------------------------------------------------------- */
/* Newmark-beta dynamic linear elasticity: AddLinSys_LinearSolid2D_NonStatic_NewmarkBeta_P1               */
/* (src/femeqn/eqn_linear_solid2d.cpp:1182-1337) and AddLinSys_LinearSolid3D_NonStatic_NewmarkBeta_P1      */
/* (src/femeqn/eqn_linear_solid3d.cpp:525-690), restated loop by loop. glob = (dt, gamma, beta, initial).  */

static void elem_solid_newmark(int ndim, const double co[][3], const double disp[][3], const double velo[][3],
                               const double acc[][3], const double* p, const double* glob, double* emat,
                               double* eres) {
  const double lambda = p[0], myu = p[1], rho = p[2];
  const double dt = glob[0], beta = glob[2];
  const int is_initial = glob[3] != 0.0;
  const int nno = ndim + 1, bs = ndim * ndim;
  double eK[4 * 4 * 9], eM[4 * 4 * 9], eqf[12];
  double dldx[4][3];
  double w;
  if (ndim == 2) {
    double d2[3][2];
    w = tri_area(co[0], co[1], co[2]);
    tri_dldx(d2, co[0], co[1], co[2]);
    for (int k = 0; k < 3; k++) { dldx[k][0] = d2[k][0]; dldx[k][1] = d2[k][1]; }
  } else {
    w = tet_volume(co[0], co[1], co[2], co[3]);
    tet_dldx(dldx, co[0], co[1], co[2], co[3]);
  }
  for (int i = 0; i < nno * nno * bs; i++) { emat[i] = 0.0; eK[i] = 0.0; eM[i] = 0.0; }
  for (int i = 0; i < nno * ndim; i++) eqf[i] = 0.0;
  for (int ino = 0; ino < nno; ino++)
    for (int jno = 0; jno < nno; jno++) {
      double* k = &eK[(ino * nno + jno) * bs];
      double dtmp1 = 0.0;
      for (int idim = 0; idim < ndim; idim++) {
        for (int jdim = 0; jdim < ndim; jdim++)
          k[idim * ndim + jdim] += w * (lambda * dldx[ino][idim] * dldx[jno][jdim] + myu * dldx[jno][idim] * dldx[ino][jdim]);
        dtmp1 += dldx[ino][idim] * dldx[jno][idim];
      }
      for (int idim = 0; idim < ndim; idim++) k[idim * ndim + idim] += w * myu * dtmp1;
    }
  {
    const double dtmp1 = (ndim == 2) ? w * rho * 0.0833333333333333333333333 : w * rho * 0.05;
    for (int ino = 0; ino < nno; ino++) {
      for (int jno = 0; jno < nno; jno++)
        for (int d = 0; d < ndim; d++) eM[(ino * nno + jno) * bs + d * ndim + d] += dtmp1;
      for (int d = 0; d < ndim; d++) eM[(ino * nno + ino) * bs + d * ndim + d] += dtmp1;
    }
  }
  for (int ino = 0; ino < nno; ino++)
    for (int d = 0; d < ndim; d++)
      eqf[ino * ndim + d] = (ndim == 2) ? w * rho * p[3 + d] * 0.33333333333333333333333333 : w * rho * p[3 + d] * 0.25;
  {
    const double dtmp1 = beta * dt * dt;
    for (int i = 0; i < nno * nno * bs; i++) emat[i] = eM[i] + dtmp1 * eK[i];
  }
  for (int ino = 0; ino < nno; ino++) {
    double* r = &eres[ino * ndim];
    for (int d = 0; d < ndim; d++) r[d] = eqf[ino * ndim + d];
    for (int jno = 0; jno < nno; jno++) {
      const double* k = &eK[(ino * nno + jno) * bs];
      for (int d = 0; d < ndim; d++) {
        double t = k[d * ndim] * disp[jno][0];
        for (int e = 1; e < ndim; e++) t = t + k[d * ndim + e] * disp[jno][e];
        r[d] -= t;
      }
    }
    for (int jno = 0; jno < nno; jno++) {
      const double* m = &eM[(ino * nno + jno) * bs];
      for (int d = 0; d < ndim; d++) {
        double t = m[d * ndim] * acc[jno][0];
        for (int e = 1; e < ndim; e++) t = t + m[d * ndim + e] * acc[jno][e];
        r[d] -= t;
      }
    }
    if (ndim == 3 || is_initial) {  /* the 3D loop has no is_initial switch */
      for (int jno = 0; jno < nno; jno++) {
        const double* k = &eK[(ino * nno + jno) * bs];
        for (int d = 0; d < ndim; d++) {
          double t = k[d * ndim] * velo[jno][0];
          for (int e = 1; e < ndim; e++) t = t + k[d * ndim + e] * velo[jno][e];
          r[d] -= dt * t;
        }
      }
      for (int jno = 0; jno < nno; jno++) {
        const double* k = &eK[(ino * nno + jno) * bs];
        for (int d = 0; d < ndim; d++) {
          double t = k[d * ndim] * acc[jno][0];
          for (int e = 1; e < ndim; e++) t = t + k[d * ndim + e] * acc[jno][e];
          r[d] -= 0.5 * dt * dt * t;
        }
      }
    }
  }
}

int cr_assemble_newmark(int ndim, int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                        int pstride, const double* glob, const double* u, const double* v, const double* a,
                        const unsigned* rowptr, const unsigned* colind, double* crs, double* dia, double* res) {
  (void)nnode;
  const int nno = ndim + 1, len = ndim, bs = ndim * ndim;
  double emat[144], eres[12], co[4][3], dsp[4][3], vel[4][3], acc[4][3];
  for (long e = 0; e < nelem; e++) {
    const int* no = &lnods[e * nno];
    for (int k = 0; k < nno; k++)
      for (int d = 0; d < 3; d++) {
        co[k][d] = (d < ndim) ? coords[(long)no[k] * ndim + d] : 0.0;
        dsp[k][d] = (d < len) ? u[(long)no[k] * len + d] : 0.0;
        vel[k][d] = (d < len) ? v[(long)no[k] * len + d] : 0.0;
        acc[k][d] = (d < len) ? a[(long)no[k] * len + d] : 0.0;
      }
    elem_solid_newmark(ndim, co, dsp, vel, acc, &params[e * pstride], glob, emat, eres);
    merge(nno, no, bs, emat, rowptr, colind, crs, dia);
    for (int k = 0; k < nno; k++)
      for (int d = 0; d < len; d++) res[(long)no[k] * len + d] += eres[k * len + d];
  }
  return 0;
}

void cr_update_field_newmark(long m, double gamma, double beta, double dt, int is_initial, const double* upd, double* u,
                             double* v, double* a) {
  /* UpdateValueOfField_NewmarkBeta (src/femls/linearsystem_field.cpp:966-1059): whole-array passes of
   * CNodeAry::AddValueToNodeSegment (src/femfield/node_ary.cpp:197-251) in this order */
  long k;
  if (is_initial) {
    const double c1 = dt, c2 = 0.5 * dt * dt;
    for (k = 0; k < m; k++) u[k] += c1 * v[k];
    for (k = 0; k < m; k++) u[k] += c2 * a[k];
  }
  { const double c = beta * dt * dt; for (k = 0; k < m; k++) u[k] += c * upd[k]; }
  if (is_initial) { for (k = 0; k < m; k++) v[k] += dt * a[k]; }
  { const double c = gamma * dt; for (k = 0; k < m; k++) v[k] += c * upd[k]; }
  for (k = 0; k < m; k++) a[k] += 1.0 * upd[k];
}

/* ---------------------------------------------------------------------------------------------------- */
/* Backward-Euler dynamic 2D elasticity: AddLinSys_LinearSolid2D_NonStatic_BackwardEular_P1               */
/* (src/femeqn/eqn_linear_solid2d.cpp:1338-1497) and UpdateValueOfField_BackwardEular                     */
/* (src/femls/linearsystem_field.cpp:1061-1101), restated loop by loop.                                   */

int cr_assemble_beuler(int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                       int pstride, double dt, int is_initial, const double* u, const double* v,
                       const double* velo_pre, const unsigned* rowptr, const unsigned* colind, double* crs,
                       double* dia, double* res) {
  (void)nnode;
  double eK[3][3][2][2], eM[3][3][2][2], emat[3][3][2][2], eqf[3][2], eres[3][2];
  double co[4][3], disp[3][2], velo[3][2], velo0[3][2];
  for (long e = 0; e < nelem; e++) {
    const int* no = &lnods[e * 3];
    const double* p = &params[e * pstride];
    const double lambda = p[0], myu = p[1], rho = p[2], g_x = p[3], g_y = p[4];
    for (int k = 0; k < 3; k++) {
      co[k][0] = coords[(long)no[k] * 2];
      co[k][1] = coords[(long)no[k] * 2 + 1];
      co[k][2] = 0.0;
      for (int d = 0; d < 2; d++) {
        disp[k][d] = u[(long)no[k] * 2 + d];
        velo[k][d] = v[(long)no[k] * 2 + d];
        velo0[k][d] = velo_pre[(long)no[k] * 2 + d];
      }
    }
    for (int i = 0; i < 36; i++) { (&emat[0][0][0][0])[i] = 0.0; (&eK[0][0][0][0])[i] = 0.0; (&eM[0][0][0][0])[i] = 0.0; }
    for (int i = 0; i < 6; i++) (&eqf[0][0])[i] = 0.0;
    const double area = tri_area(co[0], co[1], co[2]);
    double dldx[3][2];
    tri_dldx(dldx, co[0], co[1], co[2]);
    for (int ino = 0; ino < 3; ino++)
      for (int jno = 0; jno < 3; jno++) {
        double dtmp1 = 0.0;
        for (int idim = 0; idim < 2; idim++) {
          for (int jdim = 0; jdim < 2; jdim++)
            eK[ino][jno][idim][jdim] += area * (lambda * dldx[ino][idim] * dldx[jno][jdim] + myu * dldx[jno][idim] * dldx[ino][jdim]);
          dtmp1 += dldx[ino][idim] * dldx[jno][idim];
        }
        for (int idim = 0; idim < 2; idim++) eK[ino][jno][idim][idim] += area * myu * dtmp1;
      }
    {
      const double dtmp1 = area * rho * 0.0833333333333333333333333;
      for (int ino = 0; ino < 3; ino++) {
        for (int jno = 0; jno < 3; jno++) {
          eM[ino][jno][0][0] += dtmp1;
          eM[ino][jno][1][1] += dtmp1;
        }
        eM[ino][ino][0][0] += dtmp1;
        eM[ino][ino][1][1] += dtmp1;
      }
    }
    for (int ino = 0; ino < 3; ino++) {
      eqf[ino][0] = area * rho * g_x * 0.33333333333333333333333333;
      eqf[ino][1] = area * rho * g_y * 0.33333333333333333333333333;
    }
    for (int i = 0; i < 36; i++) (&emat[0][0][0][0])[i] = (&eM[0][0][0][0])[i] + dt * dt * (&eK[0][0][0][0])[i];
    for (int ino = 0; ino < 3; ino++) {
      eres[ino][0] = eqf[ino][0] * dt;
      eres[ino][1] = eqf[ino][1] * dt;
      for (int jno = 0; jno < 3; jno++) {
        eres[ino][0] -= dt * (eK[ino][jno][0][0] * disp[jno][0] + eK[ino][jno][0][1] * disp[jno][1]);
        eres[ino][1] -= dt * (eK[ino][jno][1][0] * disp[jno][0] + eK[ino][jno][1][1] * disp[jno][1]);
      }
    }
    if (is_initial) {
      for (int ino = 0; ino < 3; ino++)
        for (int jno = 0; jno < 3; jno++) {
          eres[ino][0] -= dt * dt * (eK[ino][jno][0][0] * velo[jno][0] + eK[ino][jno][0][1] * velo[jno][1]);
          eres[ino][1] -= dt * dt * (eK[ino][jno][1][0] * velo[jno][0] + eK[ino][jno][1][1] * velo[jno][1]);
        }
    } else {
      for (int ino = 0; ino < 3; ino++)
        for (int jno = 0; jno < 3; jno++) {
          eres[ino][0] -= eM[ino][jno][0][0] * (velo[jno][0] - velo0[jno][0]) + eM[ino][jno][0][1] * (velo[jno][1] - velo0[jno][1]);
          eres[ino][1] -= eM[ino][jno][1][0] * (velo[jno][0] - velo0[jno][0]) + eM[ino][jno][1][1] * (velo[jno][1] - velo0[jno][1]);
        }
    }
    merge(3, no, 4, &emat[0][0][0][0], rowptr, colind, crs, dia);
    for (int k = 0; k < 3; k++)
      for (int d = 0; d < 2; d++) res[(long)no[k] * 2 + d] += eres[k][d];
  }
  return 0;
}

void cr_update_field_beuler(long m, double dt, int is_initial, const double* upd, double* u, double* v) {
  long k;
  for (k = 0; k < m; k++) v[k] += 1.0 * upd[k];
  if (is_initial) { for (k = 0; k < m; k++) u[k] += dt * v[k]; }
  else { for (k = 0; k < m; k++) u[k] += dt * upd[k]; }
}

/* ---------------------------------------------------------------------------------------------------- */
/* AddLinSys_StVenant3D_NonStatic_NewmarkBeta_P1 (src/femeqn/eqn_st_venant.cpp:1109-1242): the tangent and */
/* internal force of SetElemMatFin_StVenant3D (the static element restated above, with density 0 its     */
/* residual is -f_in exactly), the mass, and the residual terms in the loop's order.                     */
int cr_assemble_stvk3d_newmark(int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                               int pstride, const double* glob, const double* u, const double* v, const double* a,
                               const unsigned* rowptr, const unsigned* colind, double* crs, double* dia, double* res) {
  (void)nnode;
  const double dt = glob[0], beta = glob[2];
  const int is_initial = glob[3] != 0.0;
  double eK[4][4][3][3], eM[4][4][3][3], emat[4][4][3][3], fin[4][3], eres[4][3];
  double co[4][3], disp[4][3], velo[4][3], acc[4][3];
  for (long e = 0; e < nelem; e++) {
    const int* no = &lnods[e * 4];
    const double* p = &params[e * pstride];
    const double rho = p[2];
    const double g[3] = {p[3], p[4], p[5]};
    const double p0[6] = {p[0], p[1], 0.0, p[3], p[4], p[5]};
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) {
        co[k][d] = coords[(long)no[k] * 3 + d];
        disp[k][d] = u[(long)no[k] * 3 + d];
        velo[k][d] = v[(long)no[k] * 3 + d];
        acc[k][d] = a[(long)no[k] * 3 + d];
      }
    elem_stvk3d(co, disp, p0, &eK[0][0][0][0], &fin[0][0]);
    const double vol = tet_volume(co[0], co[1], co[2], co[3]);
    for (int i = 0; i < 144; i++) (&eM[0][0][0][0])[i] = 0.0;
    {
      const double tmp1 = rho * vol * 0.05;
      for (int ino = 0; ino < 4; ino++)
        for (int jno = 0; jno < 4; jno++)
          for (int d = 0; d < 3; d++) eM[ino][jno][d][d] += tmp1;
      for (int ino = 0; ino < 4; ino++)
        for (int d = 0; d < 3; d++) eM[ino][ino][d][d] += tmp1;
    }
    for (int i = 0; i < 144; i++) (&emat[0][0][0][0])[i] = (&eM[0][0][0][0])[i] + beta * dt * dt * (&eK[0][0][0][0])[i];
    for (int ino = 0; ino < 4; ino++)
      for (int idim = 0; idim < 3; idim++) {
        eres[ino][idim] = vol * rho * g[idim] * 0.25 + fin[ino][idim];
        for (int jno = 0; jno < 4; jno++)
          for (int jdim = 0; jdim < 3; jdim++) eres[ino][idim] -= eM[ino][jno][idim][jdim] * acc[jno][jdim];
      }
    if (is_initial)
      for (int ino = 0; ino < 4; ino++)
        for (int idim = 0; idim < 3; idim++)
          for (int jno = 0; jno < 4; jno++)
            for (int jdim = 0; jdim < 3; jdim++)
              eres[ino][idim] -= dt * eK[ino][jno][idim][jdim] * velo[jno][jdim] +
                                 0.5 * dt * dt * eK[ino][jno][idim][jdim] * acc[jno][jdim];
    merge(4, no, 9, &emat[0][0][0][0], rowptr, colind, crs, dia);
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++) res[(long)no[k] * 3 + d] += eres[k][d];
  }
  return 0;
}

/* AddLinSys_StVenant2D_NonStatic_NewmarkBeta_P1 (src/femeqn/eqn_st_venant.cpp:314-460), as the 3D loop above with
 * mass rho area / 12 and external force area rho g / 3 accumulated from zero. */
int cr_assemble_stvk2d_newmark(int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                               int pstride, const double* glob, const double* u, const double* v, const double* a,
                               const unsigned* rowptr, const unsigned* colind, double* crs, double* dia, double* res) {
  (void)nnode;
  const double dt = glob[0], beta = glob[2];
  const int is_initial = glob[3] != 0.0;
  double eK[3][3][2][2], eM[3][3][2][2], emat[3][3][2][2], fin[3][2], fex[3][2], eres[3][2];
  double co[4][3], disp[4][3], velo[3][2], acc[3][2];
  for (long e = 0; e < nelem; e++) {
    const int* no = &lnods[e * 3];
    const double* p = &params[e * pstride];
    const double rho = p[2];
    const double g[2] = {p[3], p[4]};
    const double p0[5] = {p[0], p[1], 0.0, p[3], p[4]};
    for (int k = 0; k < 3; k++) {
      co[k][2] = 0.0;
      disp[k][2] = 0.0;
      for (int d = 0; d < 2; d++) {
        co[k][d] = coords[(long)no[k] * 2 + d];
        disp[k][d] = u[(long)no[k] * 2 + d];
        velo[k][d] = v[(long)no[k] * 2 + d];
        acc[k][d] = a[(long)no[k] * 2 + d];
      }
    }
    elem_stvk2d(co, disp, p0, &eK[0][0][0][0], &fin[0][0]);
    const double area = tri_area(co[0], co[1], co[2]);
    for (int i = 0; i < 36; i++) (&eM[0][0][0][0])[i] = 0.0;
    for (int i = 0; i < 6; i++) (&fex[0][0])[i] = 0.0;
    {
      const double tmp1 = rho * area / 12.0;
      for (int ino = 0; ino < 3; ino++) {
        for (int jno = 0; jno < 3; jno++)
          for (int d = 0; d < 2; d++) eM[ino][jno][d][d] += tmp1;
        for (int d = 0; d < 2; d++) eM[ino][ino][d][d] += tmp1;
      }
    }
    for (int ino = 0; ino < 3; ino++)
      for (int d = 0; d < 2; d++) fex[ino][d] += area * rho * g[d] / (double)3;
    for (int i = 0; i < 36; i++) (&emat[0][0][0][0])[i] = (&eM[0][0][0][0])[i] + beta * dt * dt * (&eK[0][0][0][0])[i];
    for (int ino = 0; ino < 3; ino++)
      for (int idim = 0; idim < 2; idim++) {
        eres[ino][idim] = fex[ino][idim] + fin[ino][idim];
        for (int jno = 0; jno < 3; jno++)
          for (int jdim = 0; jdim < 2; jdim++) eres[ino][idim] -= eM[ino][jno][idim][jdim] * acc[jno][jdim];
      }
    if (is_initial)
      for (int ino = 0; ino < 3; ino++)
        for (int idim = 0; idim < 2; idim++)
          for (int jno = 0; jno < 3; jno++)
            for (int jdim = 0; jdim < 2; jdim++)
              eres[ino][idim] -= dt * eK[ino][jno][idim][jdim] * velo[jno][jdim] +
                                 0.5 * dt * dt * eK[ino][jno][idim][jdim] * acc[jno][jdim];
    merge(3, no, 4, &emat[0][0][0][0], rowptr, colind, crs, dia);
    for (int k = 0; k < 3; k++)
      for (int d = 0; d < 2; d++) res[(long)no[k] * 2 + d] += eres[k][d];
  }
  return 0;
}

/* AddLinSys_StVenant2D_NonStatic_BackwardEular_P1 (src/femeqn/eqn_st_venant.cpp:509-664): the static StVK tangent and
 * internal force (density 0: -f_in), mass rho area / 12, M + dt^2 K, residual (f_ex - f_in) dt, then the
 * is_initial or previous-velocity terms one product at a time. */
int cr_assemble_stvk2d_beuler(int nnode, const double* coords, int nelem, const int* lnods, const double* params,
                              int pstride, double dt, int is_initial, const double* u, const double* v,
                              const double* velo_pre, const unsigned* rowptr, const unsigned* colind, double* crs,
                              double* dia, double* res) {
  (void)nnode;
  double eK[3][3][2][2], eM[3][3][2][2], emat[3][3][2][2], fin[3][2], fex[3][2], eres[3][2];
  double co[4][3], disp[4][3], velo[3][2], velo0[3][2];
  for (long e = 0; e < nelem; e++) {
    const int* no = &lnods[e * 3];
    const double* p = &params[e * pstride];
    const double rho = p[2];
    const double g[2] = {p[3], p[4]};
    const double p0[5] = {p[0], p[1], 0.0, p[3], p[4]};
    for (int k = 0; k < 3; k++) {
      co[k][2] = 0.0;
      disp[k][2] = 0.0;
      for (int d = 0; d < 2; d++) {
        co[k][d] = coords[(long)no[k] * 2 + d];
        disp[k][d] = u[(long)no[k] * 2 + d];
        velo[k][d] = v[(long)no[k] * 2 + d];
        velo0[k][d] = velo_pre[(long)no[k] * 2 + d];
      }
    }
    elem_stvk2d(co, disp, p0, &eK[0][0][0][0], &fin[0][0]);
    const double area = tri_area(co[0], co[1], co[2]);
    for (int i = 0; i < 36; i++) (&eM[0][0][0][0])[i] = 0.0;
    for (int i = 0; i < 6; i++) (&fex[0][0])[i] = 0.0;
    {
      const double tmp1 = rho * area / 12.0;
      for (int ino = 0; ino < 3; ino++) {
        for (int jno = 0; jno < 3; jno++)
          for (int d = 0; d < 2; d++) eM[ino][jno][d][d] += tmp1;
        for (int d = 0; d < 2; d++) eM[ino][ino][d][d] += tmp1;
      }
    }
    for (int ino = 0; ino < 3; ino++)
      for (int d = 0; d < 2; d++) fex[ino][d] += area * rho * g[d] * 0.33333333333333333333;
    for (int i = 0; i < 36; i++) (&emat[0][0][0][0])[i] = (&eM[0][0][0][0])[i] + dt * dt * (&eK[0][0][0][0])[i];
    for (int ino = 0; ino < 3; ino++)
      for (int idim = 0; idim < 2; idim++) eres[ino][idim] = (fex[ino][idim] + fin[ino][idim]) * dt;
    for (int ino = 0; ino < 3; ino++)
      for (int idim = 0; idim < 2; idim++)
        for (int jno = 0; jno < 3; jno++)
          for (int jdim = 0; jdim < 2; jdim++) {
            if (is_initial) eres[ino][idim] -= eK[ino][jno][idim][jdim] * velo[jno][jdim] * dt * dt;
            else eres[ino][idim] -= eM[ino][jno][idim][jdim] * (velo[jno][jdim] - velo0[jno][jdim]);
          }
    merge(3, no, 4, &emat[0][0][0][0], rowptr, colind, crs, dia);
    for (int k = 0; k < 3; k++)
      for (int d = 0; d < 2; d++) res[(long)no[k] * 2 + d] += eres[k][d];
  }
  return 0;
}
