/*
 * Oracle (test infrastructure only): restatement of the reference's TET4
 * P1-P1 VMS fluid element assembly.
 *
 *   construct_fluid  Code/Source/solver/fluid.cpp:464-708
 *   fluid_3d_m       Code/Source/solver/fluid.cpp:1697-2139
 *   fluid_3d_c       Code/Source/solver/fluid.cpp:1389-1689
 *   get_viscosity    Code/Source/solver/fluid.cpp:2142-2162 (viscType_Const)
 *   gnn (insd=3)     Code/Source/solver/nn.cpp:503-540
 *   gn_nxx           Code/Source/solver/nn.cpp:853-923 — for TET4 the second
 *                    parent derivatives Nxi2 are never initialised (lShpF is
 *                    true, fs.cpp:307-312), so the right-hand side B is zero
 *                    and the 6x6 solve returns Nwxx = 0; restated as 0.
 *   TET4 tables      nn_elem_gip.h:501-517, nn_elem_gnn.h:547-569
 *   do_assem         Code/Source/solver/lhsa.cpp:97-141
 *   is_zero          Code/Source/solver/utils.cpp:170-189
 *
 * Every arithmetic expression keeps the reference's association order. The
 * second-derivative (uxx, es_x, mu_x, d2u2) and shear-rate-derivative terms
 * are kept even though they vanish for TET4 with constant viscosity, so that
 * the HIP kernel's removal of them is checked against this code.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

static char g_err[256];
const char* orc_last_error(void) { return g_err; }

static int is_zero1(double value1)
{
  double a = fabs(value1), b = 0.0;
  if (b > a) { double t = a; a = b; b = t; }
  double eps = DBL_EPSILON;
  double nrm = fmax(a, eps);
  return ((a - b) / nrm < 10.0 * eps);
}

void orc_tet4_shape(double qmTET4, double w[4], double N[4][4], double Nxi[3][4])
{
  /* N[a][g], Nxi[i][a] (TET4 parent derivatives are constant). */
  double s = qmTET4;
  double t = (1 - s) / 3.0;
  double xi[4][3] = {{s, t, t}, {t, s, t}, {t, t, s}, {t, t, t}};
  for (int g = 0; g < 4; g++) {
    w[g] = 1.0 / 24.0;
    N[0][g] = xi[g][0];
    N[1][g] = xi[g][1];
    N[2][g] = xi[g][2];
    N[3][g] = 1.0 - xi[g][0] - xi[g][1] - xi[g][2];
  }
  static const double NxiT[3][4] = {{1.0, 0.0, 0.0, -1.0}, {0.0, 1.0, 0.0, -1.0}, {0.0, 0.0, 1.0, -1.0}};
  memcpy(Nxi, NxiT, sizeof NxiT);
}

int orc_gnn_tet4(const double Nxi[3][4], const double x[3][4], double Nx[3][4], double* Jac_out,
                 double ks[3][3])
{
  double xXi[3][3] = {{0}};
  double xiX[3][3];
  for (int a = 0; a < 4; a++) {
    for (int i = 0; i < 3; i++) {
      xXi[i][0] = xXi[i][0] + x[i][a] * Nxi[0][a];
      xXi[i][1] = xXi[i][1] + x[i][a] * Nxi[1][a];
      xXi[i][2] = xXi[i][2] + x[i][a] * Nxi[2][a];
    }
  }
  double Jac = xXi[0][0] * xXi[1][1] * xXi[2][2] + xXi[0][1] * xXi[1][2] * xXi[2][0] +
               xXi[0][2] * xXi[1][0] * xXi[2][1] - xXi[0][0] * xXi[1][2] * xXi[2][1] -
               xXi[0][1] * xXi[1][0] * xXi[2][2] - xXi[0][2] * xXi[1][1] * xXi[2][0];

  xiX[0][0] = (xXi[1][1] * xXi[2][2] - xXi[1][2] * xXi[2][1]) / Jac;
  xiX[0][1] = (xXi[2][1] * xXi[0][2] - xXi[2][2] * xXi[0][1]) / Jac;
  xiX[0][2] = (xXi[0][1] * xXi[1][2] - xXi[0][2] * xXi[1][1]) / Jac;
  xiX[1][0] = (xXi[1][2] * xXi[2][0] - xXi[1][0] * xXi[2][2]) / Jac;
  xiX[1][1] = (xXi[2][2] * xXi[0][0] - xXi[2][0] * xXi[0][2]) / Jac;
  xiX[1][2] = (xXi[0][2] * xXi[1][0] - xXi[0][0] * xXi[1][2]) / Jac;
  xiX[2][0] = (xXi[1][0] * xXi[2][1] - xXi[1][1] * xXi[2][0]) / Jac;
  xiX[2][1] = (xXi[2][0] * xXi[0][1] - xXi[2][1] * xXi[0][0]) / Jac;
  xiX[2][2] = (xXi[0][0] * xXi[1][1] - xXi[0][1] * xXi[1][0]) / Jac;

  ks[0][0] = xiX[0][0] * xiX[0][0] + xiX[1][0] * xiX[1][0] + xiX[2][0] * xiX[2][0];
  ks[0][1] = xiX[0][1] * xiX[0][0] + xiX[1][1] * xiX[1][0] + xiX[2][1] * xiX[2][0];
  ks[0][2] = xiX[0][2] * xiX[0][0] + xiX[1][2] * xiX[1][0] + xiX[2][2] * xiX[2][0];
  ks[1][1] = xiX[0][1] * xiX[0][1] + xiX[1][1] * xiX[1][1] + xiX[2][1] * xiX[2][1];
  ks[1][2] = xiX[0][1] * xiX[0][2] + xiX[1][1] * xiX[1][2] + xiX[2][1] * xiX[2][2];
  ks[2][2] = xiX[0][2] * xiX[0][2] + xiX[1][2] * xiX[1][2] + xiX[2][2] * xiX[2][2];
  ks[1][0] = ks[0][1];
  ks[2][0] = ks[0][2];
  ks[2][1] = ks[1][2];

  for (int a = 0; a < 4; a++) {
    Nx[0][a] = 0.0 + Nxi[0][a] * xiX[0][0] + Nxi[1][a] * xiX[1][0] + Nxi[2][a] * xiX[2][0];
    Nx[1][a] = 0.0 + Nxi[0][a] * xiX[0][1] + Nxi[1][a] * xiX[1][1] + Nxi[2][a] * xiX[2][1];
    Nx[2][a] = 0.0 + Nxi[0][a] * xiX[0][2] + Nxi[1][a] * xiX[1][2] + Nxi[2][a] * xiX[2][2];
  }
  *Jac_out = Jac;
  return is_zero1(Jac) ? 1 : 0;
}

/* lK index helper: lK(k,a,b) of Array3<double>(16,4,4). */
#define LK(k, a, b) lK[(k) + 16 * ((a) + 4 * (b))]
#define LR(i, a) lR[(i) + 4 * (a)]

/* fluid_3d_m, fluid.cpp:1697-2139 (vmsFlag = 1, eNoNw = eNoNq = 4, mvMsh = 0). */
static void fluid_3d_m(const orc_fluid_params* P, double w, double Kxi[3][3], const double* Nw,
                       const double* Nq, double Nwx[3][4], double Nqx[3][4], double Nwxx[6][4],
                       const double al[4][4], const double yl[4][4], const double bfl[3][4],
                       double* lR, double* lK)
{
  const double K = P->K_darcy;
  const double dt = P->dt;
  const double ctM = 1.0, ctC = 36.0;
  const double rho = P->rho;
  const double f[3] = {P->f[0], P->f[1], P->f[2]};
  double T1 = P->af * P->gam * dt;
  const double amd = P->am / T1;
  const double wl = w * T1;
  const double wr = w * rho;

  double ud[3] = {-f[0], -f[1], -f[2]};
  double u[3] = {0, 0, 0};
  double ux[3][3] = {{0}};
  double uxx[3][3][3] = {{{0}}};
  for (int a = 0; a < 4; a++) {
    ud[0] = ud[0] + Nw[a] * (al[0][a] - bfl[0][a]);
    ud[1] = ud[1] + Nw[a] * (al[1][a] - bfl[1][a]);
    ud[2] = ud[2] + Nw[a] * (al[2][a] - bfl[2][a]);
    u[0] = u[0] + Nw[a] * yl[0][a];
    u[1] = u[1] + Nw[a] * yl[1][a];
    u[2] = u[2] + Nw[a] * yl[2][a];
    ux[0][0] += Nwx[0][a] * yl[0][a];
    ux[1][0] += Nwx[1][a] * yl[0][a];
    ux[2][0] += Nwx[2][a] * yl[0][a];
    ux[0][1] += Nwx[0][a] * yl[1][a];
    ux[1][1] += Nwx[1][a] * yl[1][a];
    ux[2][1] += Nwx[2][a] * yl[1][a];
    ux[0][2] += Nwx[0][a] * yl[2][a];
    ux[1][2] += Nwx[1][a] * yl[2][a];
    ux[2][2] += Nwx[2][a] * yl[2][a];
    for (int j = 0; j < 3; j++) {
      uxx[0][j][0] += Nwxx[0][a] * yl[j][a];
      uxx[1][j][1] += Nwxx[1][a] * yl[j][a];
      uxx[2][j][2] += Nwxx[2][a] * yl[j][a];
      uxx[1][j][0] += Nwxx[3][a] * yl[j][a];
      uxx[2][j][1] += Nwxx[4][a] * yl[j][a];
      uxx[0][j][2] += Nwxx[5][a] * yl[j][a];
    }
  }
  double divU = ux[0][0] + ux[1][1] + ux[2][2];
  for (int j = 0; j < 3; j++) {
    uxx[0][j][1] = uxx[1][j][0];
    uxx[1][j][2] = uxx[2][j][1];
    uxx[2][j][0] = uxx[0][j][2];
  }
  double d2u2[3];
  for (int j = 0; j < 3; j++) d2u2[j] = uxx[0][j][0] + uxx[1][j][1] + uxx[2][j][2];

  double p = 0.0, px[3] = {0, 0, 0};
  for (int a = 0; a < 4; a++) {
    p = p + Nq[a] * yl[3][a];
    px[0] = px[0] + Nqx[0][a] * yl[3][a];
    px[1] = px[1] + Nqx[1][a] * yl[3][a];
    px[2] = px[2] + Nqx[2][a] * yl[3][a];
  }

  double es[3][3];
  es[0][0] = ux[0][0] + ux[0][0];
  es[1][1] = ux[1][1] + ux[1][1];
  es[2][2] = ux[2][2] + ux[2][2];
  es[1][0] = ux[1][0] + ux[0][1];
  es[2][1] = ux[2][1] + ux[1][2];
  es[0][2] = ux[0][2] + ux[2][0];
  es[0][1] = es[1][0];
  es[1][2] = es[2][1];
  es[2][0] = es[0][2];

  double esNx[3][4];
  for (int a = 0; a < 4; a++) {
    esNx[0][a] = es[0][0] * Nwx[0][a] + es[1][0] * Nwx[1][a] + es[2][0] * Nwx[2][a];
    esNx[1][a] = es[0][1] * Nwx[0][a] + es[1][1] * Nwx[1][a] + es[2][1] * Nwx[2][a];
    esNx[2][a] = es[0][2] * Nwx[0][a] + es[1][2] * Nwx[1][a] + es[2][2] * Nwx[2][a];
  }
  double es_x[3][3][3];
  for (int k = 0; k < 3; k++) {
    es_x[0][0][k] = uxx[0][0][k] + uxx[0][0][k];
    es_x[1][1][k] = uxx[1][1][k] + uxx[1][1][k];
    es_x[2][2][k] = uxx[2][2][k] + uxx[2][2][k];
    es_x[1][0][k] = uxx[1][0][k] + uxx[0][1][k];
    es_x[2][1][k] = uxx[2][1][k] + uxx[1][2][k];
    es_x[0][2][k] = uxx[0][2][k] + uxx[2][0][k];
    es_x[0][1][k] = es_x[1][0][k];
    es_x[1][2][k] = es_x[2][1][k];
    es_x[2][0][k] = es_x[0][2][k];
  }
  double mu_x[3];
  for (int k = 0; k < 3; k++) {
    mu_x[k] = (es_x[0][0][k] * es[0][0] + es_x[1][1][k] * es[1][1] + es_x[2][2][k] * es[2][2]) * 0.5 +
              es_x[1][0][k] * es[1][0] + es_x[2][1][k] * es[2][1] + es_x[0][2][k] * es[0][2];
  }
  double gam = es[0][0] * es[0][0] + es[1][0] * es[1][0] + es[2][0] * es[2][0] + es[0][1] * es[0][1] +
               es[1][1] * es[1][1] + es[2][1] * es[2][1] + es[0][2] * es[0][2] + es[1][2] * es[1][2] +
               es[2][2] * es[2][2];
  gam = sqrt(0.5 * gam);
  /* get_viscosity, viscType_Const (fluid.cpp:2158-2162) */
  double mu = P->mu, mu_g = 0.0;
  if (is_zero1(gam)) mu_g = 0.0; else mu_g = mu_g / gam;
  for (int k = 0; k < 3; k++) mu_x[k] = mu_g * mu_x[k];

  double kT = 4.0 * pow(ctM / dt, 2.0);
  kT = kT + pow(K * mu / rho, 2.0);
  double kU = u[0] * u[0] * Kxi[0][0] + u[1] * u[0] * Kxi[1][0] + u[2] * u[0] * Kxi[2][0] +
              u[0] * u[1] * Kxi[0][1] + u[1] * u[1] * Kxi[1][1] + u[2] * u[1] * Kxi[2][1] +
              u[0] * u[2] * Kxi[0][2] + u[1] * u[2] * Kxi[1][2] + u[2] * u[2] * Kxi[2][2];
  double kS = Kxi[0][0] * Kxi[0][0] + Kxi[1][0] * Kxi[1][0] + Kxi[2][0] * Kxi[2][0] +
              Kxi[0][1] * Kxi[0][1] + Kxi[1][1] * Kxi[1][1] + Kxi[2][1] * Kxi[2][1] +
              Kxi[0][2] * Kxi[0][2] + Kxi[1][2] * Kxi[1][2] + Kxi[2][2] * Kxi[2][2];
  kS = ctC * kS * pow(mu / rho, 2.0);
  double tauM = 1.0 / (rho * sqrt(kT + kU + kS));

  double rV[3], rS[3], up[3];
  for (int j = 0; j < 3; j++) rV[j] = ud[j] + u[0] * ux[0][j] + u[1] * ux[1][j] + u[2] * ux[2][j];
  for (int j = 0; j < 3; j++) rS[j] = mu_x[0] * es[0][j] + mu_x[1] * es[1][j] + mu_x[2] * es[2][j] + mu * d2u2[j];
  for (int j = 0; j < 3; j++) up[j] = -tauM * (rho * rV[j] + px[j] - rS[j] + mu * K * u[j]);

  double eps = DBL_EPSILON;
  double tauC = 1.0 / (tauM * (Kxi[0][0] + Kxi[1][1] + Kxi[2][2]));
  double tauB = up[0] * up[0] * Kxi[0][0] + up[1] * up[0] * Kxi[1][0] + up[2] * up[0] * Kxi[2][0] +
                up[0] * up[1] * Kxi[0][1] + up[1] * up[1] * Kxi[1][1] + up[2] * up[1] * Kxi[2][1] +
                up[0] * up[2] * Kxi[0][2] + up[1] * up[2] * Kxi[1][2] + up[2] * up[2] * Kxi[2][2];
  if (is_zero1(tauB)) tauB = eps;
  tauB = rho / sqrt(tauB);
  double ua[3] = {u[0] + up[0], u[1] + up[1], u[2] + up[2]};
  double pa = p - tauC * divU;

  for (int j = 0; j < 3; j++) rV[j] = tauB * (up[0] * ux[0][j] + up[1] * ux[1][j] + up[2] * ux[2][j]);

  double rM[3][3];
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) {
      rM[i][j] = mu * es[i][j] - rho * up[j] * ua[i] + rV[j] * up[i];
      if (i == j) rM[i][j] = rM[i][j] - pa;
    }
  for (int j = 0; j < 3; j++) rV[j] = ud[j] + ua[0] * ux[0][j] + ua[1] * ux[1][j] + ua[2] * ux[2][j];

  double updu[3][3][4];
  double uNx[4], upNx[4], uaNx[4];
  for (int a = 0; a < 4; a++) {
    for (int i = 0; i < 3; i++)
      LR(i, a) = LR(i, a) + wr * Nw[a] * rV[i] +
                 w * (Nwx[0][a] * rM[0][i] + Nwx[1][a] * rM[1][i] + Nwx[2][a] * rM[2][i]);
    uNx[a] = u[0] * Nwx[0][a] + u[1] * Nwx[1][a] + u[2] * Nwx[2][a];
    upNx[a] = up[0] * Nwx[0][a] + up[1] * Nwx[1][a] + up[2] * Nwx[2][a];
    uaNx[a] = uNx[a] + upNx[a];
    T1 = -rho * uNx[a] + mu * (Nwxx[0][a] + Nwxx[1][a] + Nwxx[2][a]) + mu_x[0] * Nwx[0][a] +
         mu_x[1] * Nwx[1][a] + mu_x[2] * Nwx[2][a] - mu * K * Nw[a];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        updu[i][j][a] = mu_x[i] * Nwx[j][a] + d2u2[i] * mu_g * esNx[j][a];
        if (i == j) updu[i][j][a] = updu[i][j][a] + T1;
      }
  }

  for (int b = 0; b < 4; b++) {
    for (int a = 0; a < 4; a++) {
      double r[3][3];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r[i][j] = Nwx[i][a] * Nwx[j][b];
      double NxNx = Nwx[0][a] * Nwx[0][b] + Nwx[1][a] * Nwx[1][b] + Nwx[2][a] * Nwx[2][b];
      T1 = mu * NxNx + rho * amd * Nw[b] * (Nw[a] + rho * tauM * uaNx[a]) + rho * Nw[a] * (uNx[b] + upNx[b]) +
           tauB * upNx[a] * upNx[b];
      /* row i of the momentum block, column j: lK(4i+j) */
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double T2;
          if (i == j) {
            T2 = (mu + tauC) * r[i][i] + esNx[i][a] * mu_g * esNx[i][b] - rho * tauM * uaNx[a] * updu[i][i][b];
            LK(4 * i + j, a, b) = LK(4 * i + j, a, b) + wl * (T2 + T1);
            LK(4 * i + j, a, b) = LK(4 * i + j, a, b) + mu * K * wl * Nw[b] * Nw[a];
          } else {
            T2 = mu * r[j][i] + tauC * r[i][j] + esNx[i][a] * mu_g * esNx[j][b] -
                 rho * tauM * uaNx[a] * updu[j][i][b];
            LK(4 * i + j, a, b) = LK(4 * i + j, a, b) + wl * (T2);
          }
        }
    }
  }
  for (int b = 0; b < 4; b++) {
    for (int a = 0; a < 4; a++) {
      T1 = rho * tauM * uaNx[a];
      LK(3, a, b) = LK(3, a, b) - wl * (Nwx[0][a] * Nq[b] - Nqx[0][b] * T1);
      LK(7, a, b) = LK(7, a, b) - wl * (Nwx[1][a] * Nq[b] - Nqx[1][b] * T1);
      LK(11, a, b) = LK(11, a, b) - wl * (Nwx[2][a] * Nq[b] - Nqx[2][b] * T1);
    }
  }
  for (int a = 0; a < 4; a++) {
    LR(0, a) = LR(0, a) + mu * K * w * Nw[a] * (u[0] + up[0]);
    LR(1, a) = LR(1, a) + mu * K * w * Nw[a] * (u[1] + up[1]);
    LR(2, a) = LR(2, a) + mu * K * w * Nw[a] * (u[2] + up[2]);
  }
}

/* fluid_3d_c, fluid.cpp:1389-1689 (vmsFlag = 1). */
static void fluid_3d_c(const orc_fluid_params* P, double w, double Kxi[3][3], const double* Nw,
                       const double* Nq, double Nwx[3][4], double Nqx[3][4], double Nwxx[6][4],
                       const double al[4][4], const double yl[4][4], const double bfl[3][4],
                       double* lR, double* lK)
{
  const double K = P->K_darcy;
  const double dt = P->dt;
  const double ctM = 1.0, ctC = 36.0;
  const double rho = P->rho;
  const double f[3] = {P->f[0], P->f[1], P->f[2]};
  double T1 = P->af * P->gam * dt;
  const double amd = P->am / T1;
  const double wl = w * T1;

  double ud[3] = {-f[0], -f[1], -f[2]};
  double u[3] = {0, 0, 0};
  double ux[3][3] = {{0}};
  double uxx[3][3][3] = {{{0}}};
  for (int a = 0; a < 4; a++) {
    ud[0] = ud[0] + Nw[a] * (al[0][a] - bfl[0][a]);
    ud[1] = ud[1] + Nw[a] * (al[1][a] - bfl[1][a]);
    ud[2] = ud[2] + Nw[a] * (al[2][a] - bfl[2][a]);
    u[0] = u[0] + Nw[a] * yl[0][a];
    u[1] = u[1] + Nw[a] * yl[1][a];
    u[2] = u[2] + Nw[a] * yl[2][a];
    ux[0][0] = ux[0][0] + Nwx[0][a] * yl[0][a];
    ux[1][0] = ux[1][0] + Nwx[1][a] * yl[0][a];
    ux[2][0] = ux[2][0] + Nwx[2][a] * yl[0][a];
    ux[0][1] = ux[0][1] + Nwx[0][a] * yl[1][a];
    ux[1][1] = ux[1][1] + Nwx[1][a] * yl[1][a];
    ux[2][1] = ux[2][1] + Nwx[2][a] * yl[1][a];
    ux[0][2] = ux[0][2] + Nwx[0][a] * yl[2][a];
    ux[1][2] = ux[1][2] + Nwx[1][a] * yl[2][a];
    ux[2][2] = ux[2][2] + Nwx[2][a] * yl[2][a];
    for (int j = 0; j < 3; j++) {
      uxx[0][j][0] += Nwxx[0][a] * yl[j][a];
      uxx[1][j][1] += Nwxx[1][a] * yl[j][a];
      uxx[2][j][2] += Nwxx[2][a] * yl[j][a];
      uxx[1][j][0] += Nwxx[3][a] * yl[j][a];
      uxx[2][j][1] += Nwxx[4][a] * yl[j][a];
      uxx[0][j][2] += Nwxx[5][a] * yl[j][a];
    }
  }
  double divU = ux[0][0] + ux[1][1] + ux[2][2];
  for (int j = 0; j < 3; j++) {
    uxx[0][j][1] = uxx[1][j][0];
    uxx[1][j][2] = uxx[2][j][1];
    uxx[2][j][0] = uxx[0][j][2];
  }
  double d2u2[3];
  for (int j = 0; j < 3; j++) d2u2[j] = uxx[0][j][0] + uxx[1][j][1] + uxx[2][j][2];
  double px[3] = {0, 0, 0};
  for (int a = 0; a < 4; a++) {
    px[0] = px[0] + Nqx[0][a] * yl[3][a];
    px[1] = px[1] + Nqx[1][a] * yl[3][a];
    px[2] = px[2] + Nqx[2][a] * yl[3][a];
  }
  double es[3][3];
  es[0][0] = ux[0][0] + ux[0][0];
  es[1][1] = ux[1][1] + ux[1][1];
  es[2][2] = ux[2][2] + ux[2][2];
  es[1][0] = ux[1][0] + ux[0][1];
  es[2][1] = ux[2][1] + ux[1][2];
  es[0][2] = ux[0][2] + ux[2][0];
  es[0][1] = es[1][0];
  es[1][2] = es[2][1];
  es[2][0] = es[0][2];
  double esNx[3][4];
  for (int a = 0; a < 4; a++) {
    esNx[0][a] = es[0][0] * Nwx[0][a] + es[1][0] * Nwx[1][a] + es[2][0] * Nwx[2][a];
    esNx[1][a] = es[0][1] * Nwx[0][a] + es[1][1] * Nwx[1][a] + es[2][1] * Nwx[2][a];
    esNx[2][a] = es[0][2] * Nwx[0][a] + es[1][2] * Nwx[1][a] + es[2][2] * Nwx[2][a];
  }
  double es_x[3][3][3];
  for (int k = 0; k < 3; k++) {
    es_x[0][0][k] = uxx[0][0][k] + uxx[0][0][k];
    es_x[1][1][k] = uxx[1][1][k] + uxx[1][1][k];
    es_x[2][2][k] = uxx[2][2][k] + uxx[2][2][k];
    es_x[1][0][k] = uxx[1][0][k] + uxx[0][1][k];
    es_x[2][1][k] = uxx[2][1][k] + uxx[1][2][k];
    es_x[0][2][k] = uxx[0][2][k] + uxx[2][0][k];
    es_x[0][1][k] = es_x[1][0][k];
    es_x[1][2][k] = es_x[2][1][k];
    es_x[2][0][k] = es_x[0][2][k];
  }
  double mu_x[3];
  for (int k = 0; k < 3; k++) {
    mu_x[k] = (es_x[0][0][k] * es[0][0] + es_x[1][1][k] * es[1][1] + es_x[2][2][k] * es[2][2]) * 0.5 +
              es_x[1][0][k] * es[1][0] + es_x[2][1][k] * es[2][1] + es_x[0][2][k] * es[0][2];
  }
  double gam = es[0][0] * es[0][0] + es[1][0] * es[1][0] + es[2][0] * es[2][0] + es[0][1] * es[0][1] +
               es[1][1] * es[1][1] + es[2][1] * es[2][1] + es[0][2] * es[0][2] + es[1][2] * es[1][2] +
               es[2][2] * es[2][2];
  gam = sqrt(0.5 * gam);
  double mu = P->mu, mu_g = 0.0;
  if (is_zero1(gam)) mu_g = 0.0; else mu_g = mu_g / gam;
  for (int i = 0; i < 3; i++) mu_x[i] = mu_g * mu_x[i];

  double up[3];
  double updu[3][3][4];
  double kT = 4.0 * pow(ctM / dt, 2.0);
  kT = kT + pow(K * mu / rho, 2.0);
  double kU = u[0] * u[0] * Kxi[0][0] + u[1] * u[0] * Kxi[1][0] + u[2] * u[0] * Kxi[2][0] +
              u[0] * u[1] * Kxi[0][1] + u[1] * u[1] * Kxi[1][1] + u[2] * u[1] * Kxi[2][1] +
              u[0] * u[2] * Kxi[0][2] + u[1] * u[2] * Kxi[1][2] + u[2] * u[2] * Kxi[2][2];
  double kS = Kxi[0][0] * Kxi[0][0] + Kxi[1][0] * Kxi[1][0] + Kxi[2][0] * Kxi[2][0] +
              Kxi[0][1] * Kxi[0][1] + Kxi[1][1] * Kxi[1][1] + Kxi[2][1] * Kxi[2][1] +
              Kxi[0][2] * Kxi[0][2] + Kxi[1][2] * Kxi[1][2] + Kxi[2][2] * Kxi[2][2];
  kS = ctC * kS * pow(mu / rho, 2.0);
  double tauM = 1.0 / (rho * sqrt(kT + kU + kS));
  double rV[3], rS[3];
  for (int j = 0; j < 3; j++) rV[j] = ud[j] + u[0] * ux[0][j] + u[1] * ux[1][j] + u[2] * ux[2][j];
  for (int j = 0; j < 3; j++) rS[j] = mu_x[0] * es[0][j] + mu_x[1] * es[1][j] + mu_x[2] * es[2][j] + mu * d2u2[j];
  for (int j = 0; j < 3; j++) up[j] = -tauM * (rho * rV[j] + px[j] - rS[j] + mu * K * u[j]);
  for (int a = 0; a < 4; a++) {
    double uNx = u[0] * Nwx[0][a] + u[1] * Nwx[1][a] + u[2] * Nwx[2][a];
    T1 = -rho * uNx + mu * (Nwxx[0][a] + Nwxx[1][a] + Nwxx[2][a]) + mu_x[0] * Nwx[0][a] + mu_x[1] * Nwx[1][a] +
         mu_x[2] * Nwx[2][a] - mu * K * Nw[a];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        updu[i][j][a] = mu_x[i] * Nwx[j][a] + d2u2[i] * mu_g * esNx[j][a];
        if (i == j) updu[i][j][a] = updu[i][j][a] + T1;
      }
  }

  for (int a = 0; a < 4; a++) {
    double upNx = up[0] * Nqx[0][a] + up[1] * Nqx[1][a] + up[2] * Nqx[2][a];
    LR(3, a) = LR(3, a) + w * (Nq[a] * divU - upNx);
  }
  for (int b = 0; b < 4; b++) {
    T1 = rho * amd * Nw[b];
    for (int a = 0; a < 4; a++) {
      double T2 = Nqx[0][a] * (updu[0][0][b] - T1) + Nqx[1][a] * updu[0][1][b] + Nqx[2][a] * updu[0][2][b];
      LK(12, a, b) = LK(12, a, b) + wl * (Nq[a] * Nwx[0][b] - tauM * T2);
      T2 = Nqx[0][a] * updu[1][0][b] + Nqx[1][a] * (updu[1][1][b] - T1) + Nqx[2][a] * updu[1][2][b];
      LK(13, a, b) = LK(13, a, b) + wl * (Nq[a] * Nwx[1][b] - tauM * T2);
      T2 = Nqx[0][a] * updu[2][0][b] + Nqx[1][a] * updu[2][1][b] + Nqx[2][a] * (updu[2][2][b] - T1);
      LK(14, a, b) = LK(14, a, b) + wl * (Nq[a] * Nwx[2][b] - tauM * T2);
    }
  }
  for (int b = 0; b < 4; b++)
    for (int a = 0; a < 4; a++) {
      double NxNx = Nqx[0][a] * Nqx[0][b] + Nqx[1][a] * Nqx[1][b] + Nqx[2][a] * Nqx[2][b];
      LK(15, a, b) = LK(15, a, b) + wl * tauM * NxNx;
    }
}

int orc_fluid_element(const double xl[3][4], const double al[4][4], const double yl[4][4],
                      const double bfl[3][4], const orc_fluid_params* p, double lR[16], double lK[256])
{
  double w[4], N[4][4], Nxi[3][4];
  orc_tet4_shape(p->qmTET4, w, N, Nxi);
  memset(lR, 0, 16 * sizeof(double));
  memset(lK, 0, 256 * sizeof(double));

  double Nwx[3][4], Nqx[3][4], Nwxx[6][4];
  double ksix[3][3], Jac = 0.0;
  memset(Nwxx, 0, sizeof Nwxx);

  /* Gauss integration 1 (fluid.cpp:605-648); TET4 lShpF: geometry at g = 0 only. */
  for (int g = 0; g < 4; g++) {
    if (g == 0) {
      if (orc_gnn_tet4(Nxi, xl, Nqx, &Jac, ksix)) return 1;
      if (orc_gnn_tet4(Nxi, xl, Nwx, &Jac, ksix)) return 1;
      memset(Nwxx, 0, sizeof Nwxx); /* gn_nxx with Nxi2 = 0 */
    }
    double wg = w[g] * Jac;
    double Ng[4] = {N[0][g], N[1][g], N[2][g], N[3][g]};
    fluid_3d_m(p, wg, ksix, Ng, Ng, Nwx, Nqx, Nwxx, al, yl, bfl, lR, lK);
  }
  /* Gauss integration 2 (fluid.cpp:665-698). */
  for (int g = 0; g < 4; g++) {
    if (g == 0) {
      if (orc_gnn_tet4(Nxi, xl, Nwx, &Jac, ksix)) return 1;
      if (orc_gnn_tet4(Nxi, xl, Nqx, &Jac, ksix)) return 1;
    }
    double wg = w[g] * Jac;
    double Ng[4] = {N[0][g], N[1][g], N[2][g], N[3][g]};
    fluid_3d_c(p, wg, ksix, Ng, Ng, Nwx, Nqx, Nwxx, al, yl, bfl, lR, lK);
  }
  return 0;
}

int orc_construct_fluid(int nEl, const int* IEN, const double* x, const double* Bf, const double* Ag,
                        const double* Yg, const orc_fluid_params* p, const int* rowPtr, const int* colPtr,
                        double* R, double* Val)
{
  double xl[3][4], al[4][4], yl[4][4], bfl[3][4];
  double lR[16], lK[256];
  int ptr[4];
  for (int e = 0; e < nEl; e++) {
    for (int a = 0; a < 4; a++) {
      int Ac = IEN[a + 4 * e];
      ptr[a] = Ac;
      for (int i = 0; i < 3; i++) {
        xl[i][a] = x[i + 3 * Ac];
        bfl[i][a] = Bf ? Bf[i + 3 * Ac] : 0.0;
      }
      for (int i = 0; i < 4; i++) {
        al[i][a] = Ag[i + 4 * Ac];
        yl[i][a] = Yg[i + 4 * Ac];
      }
    }
    if (orc_fluid_element(xl, al, yl, bfl, p, lR, lK)) {
      snprintf(g_err, sizeof g_err, "[construct_fluid] Jacobian for element %d is < 0.", e);
      return 1;
    }
    /* do_assem, lhsa.cpp:97-141 */
    for (int a = 0; a < 4; a++) {
      int rowN = ptr[a];
      for (int i = 0; i < 4; i++) R[i + 4 * rowN] = R[i + 4 * rowN] + lR[i + 4 * a];
      for (int b = 0; b < 4; b++) {
        int colN = ptr[b];
        int left = rowPtr[rowN];
        int right = rowPtr[rowN + 1];
        int q = (right + left) / 2;
        while (colN != colPtr[q]) {
          if (colN > colPtr[q]) left = q; else right = q;
          q = (right + left) / 2;
        }
        for (int i = 0; i < 16; i++) Val[i + 16 * q] = Val[i + 16 * q] + lK[i + 16 * (a + 4 * b)];
      }
    }
  }
  return 0;
}

/* Neumann-face assembly of the fluid equation on TRI3 faces of a TET4 mesh:
 *   b_assem_neu_bc   Code/Source/solver/eq_assem.cpp:58-172 (one face, fluid physics)
 *   gnnb             Code/Source/solver/nn.cpp:552-760 (non-shell branch: area-weighted
 *                    normal (x0 - x2) x (x1 - x2), flipped to point from the interior node
 *                    of the parent element towards face node 0)
 *   TRI3 face rule   nn_elem_gip.h:745-762 (w = 1/6, s = qmTRI3, t = (1 - s)/2),
 *                    shape functions nn_elem_gnn.h:1561-1572
 *   b_fluid          Code/Source/solver/fluid.cpp:46-133 (fixed mesh, nsd = 3)
 *   do_assem         Code/Source/solver/lhsa.cpp:97-141
 * faceIEN(3,nFe); inner(nFe) = the parent tet's node off the face; hg(nNo) the
 * Neumann value per node (NULL: zero traction). */
int orc_b_assem_neu_fluid(int nFe, const int* faceIEN, const int* inner, const double* x, const double* hg,
                          const double* Yg, const orc_fluid_params* p, double beta, double qmTRI3,
                          const int* rowPtr, const int* colPtr, double* R, double* Val)
{
  const double s = qmTRI3, t = -0.5 * s + 0.5;
  const double xi[3][2] = {{t, t}, {s, t}, {t, s}};
  for (int e = 0; e < nFe; e++) {
    int ptr[3];
    double xl[3][3], yl[3][3], hl[3];
    for (int a = 0; a < 3; a++) {
      int Ac = faceIEN[a + 3 * e];
      ptr[a] = Ac;
      hl[a] = hg ? hg[Ac] : 0.0;
      for (int i = 0; i < 3; i++) {
        xl[i][a] = x[i + 3 * Ac];
        yl[i][a] = Yg[i + 4 * Ac];
      }
    }
    double lR[3][3] = {{0}}, lKd[3][3] = {{0}};  /* lK(0,a,b) = lK(5,a,b) = lK(10,a,b) */
    for (int g = 0; g < 3; g++) {
      double N[3] = {xi[g][0], xi[g][1], 1.0 - xi[g][0] - xi[g][1]};
      /* gnnb: xXi(j,i) = sum_a Nx(i,a) x(j,a) with Nx = [1 0 -1; 0 1 -1] */
      double X0[3], X1[3], n[3];
      for (int j = 0; j < 3; j++) {
        X0[j] = 0.0 + 1.0 * xl[j][0] + 0.0 * xl[j][1] + (-1.0) * xl[j][2];
        X1[j] = 0.0 + 0.0 * xl[j][0] + 1.0 * xl[j][1] + (-1.0) * xl[j][2];
      }
      n[0] = X0[1] * X1[2] - X0[2] * X1[1];
      n[1] = X0[2] * X1[0] - X0[0] * X1[2];
      n[2] = X0[0] * X1[1] - X0[1] * X1[0];
      int in = inner[e];
      double v[3], nv = 0.0;
      for (int i = 0; i < 3; i++) v[i] = xl[i][0] - x[i + 3 * in];
      for (int i = 0; i < 3; i++) nv += n[i] * v[i];
      if (nv < 0.0)
        for (int i = 0; i < 3; i++) n[i] = -n[i];
      double nn = 0.0;
      for (int i = 0; i < 3; i++) nn += n[i] * n[i];
      double Jac = sqrt(nn);
      for (int i = 0; i < 3; i++) n[i] = n[i] / Jac;
      double w = (1.0 / 6.0) * Jac;
      double h = 0.0, y[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < 3; a++) {
        h = h + N[a] * hl[a];
        for (int i = 0; i < 3; i++) y[i] = y[i] + N[a] * yl[i][a];
      }
      /* b_fluid */
      double wl = w * p->af * p->gam * p->dt;
      double udn = 0.0;
      for (int i = 0; i < 3; i++) udn = udn + y[i] * n[i];
      udn = 0.50 * beta * p->rho * (udn - fabs(udn));
      double hc[3];
      for (int i = 0; i < 3; i++) hc[i] = h * n[i] + udn * y[i];
      for (int a = 0; a < 3; a++) {
        for (int i = 0; i < 3; i++) lR[i][a] = lR[i][a] - w * N[a] * hc[i];
        for (int b = 0; b < 3; b++) {
          double T1 = wl * N[a] * N[b] * udn;
          lKd[a][b] = lKd[a][b] - T1;
        }
      }
    }
    /* do_assem */
    for (int a = 0; a < 3; a++) {
      int rowN = ptr[a];
      for (int i = 0; i < 3; i++) R[i + 4 * rowN] = R[i + 4 * rowN] + lR[i][a];
      R[3 + 4 * rowN] = R[3 + 4 * rowN] + 0.0;
      for (int b = 0; b < 3; b++) {
        int colN = ptr[b];
        int left = rowPtr[rowN];
        int right = rowPtr[rowN + 1];
        int q = (right + left) / 2;
        while (colN != colPtr[q]) {
          if (colN > colPtr[q]) left = q; else right = q;
          q = (right + left) / 2;
        }
        for (int k = 0; k < 16; k++) {
          double lk = (k == 0 || k == 5 || k == 10) ? lKd[a][b] : 0.0;
          Val[k + 16 * q] = Val[k + 16 * q] + lk;
        }
      }
    }
  }
  return 0;
}
