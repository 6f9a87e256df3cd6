/*
 * Oracle (test infrastructure only): the generalized-alpha pieces of one
 * Newton iteration for a single fluid equation (s = 0, e = 3, tDof = dof = 4).
 *
 *   pici   Code/Source/solver/pic.cpp:486-574   Ag = (1-am) Ao + am An, Yg = (1-af) Yo + af Yn
 *   picc   Code/Source/solver/pic.cpp:155-167   An -= R, Yn -= gam dt R
 *          (the Dn update of the same loop is not part of a fluid-only run:
 *           dFlag is false for fluid, initialize.cpp:420-421)
 *   picp   Code/Source/solver/pic.cpp:676-693   An = Ao (gam-1)/gam, Yn = Yo (rows s..e of the
 *          equation; Dn = Do is not carried)
 *   set_bc_dir  Code/Source/solver/set_bc.cpp:794-1063 for one strong Dirichlet face:
 *          set_bc_dir_l's non-generic branch (tmpA = dirA gx nV or dirA gx, tmpY the
 *          same with dirY; here gxn = gx nV or gx is given) and the scatter without
 *          impD, over all lDof rows (eDir all false) or the eDrn directions
 */
#include "oracle.h"

void orc_pici(int nNo, double am, double af, const double* Ao, const double* An, const double* Yo,
              const double* Yn, double* Ag, double* Yg)
{
  double coef0 = 1.0 - am, coef1 = am, coef2 = 1.0 - af, coef3 = af;
  for (int a = 0; a < nNo; a++)
    for (int j = 0; j < 4; j++) {
      Ag[j + 4 * a] = Ao[j + 4 * a] * coef0 + An[j + 4 * a] * coef1;
      Yg[j + 4 * a] = Yo[j + 4 * a] * coef2 + Yn[j + 4 * a] * coef3;
    }
}

void orc_picc_update(int nNo, double gam, double dt, const double* R, double* An, double* Yn)
{
  double coef0 = gam * dt;
  for (int a = 0; a < nNo; a++)
    for (int i = 0; i < 4; i++) {
      An[i + 4 * a] = An[i + 4 * a] - R[i + 4 * a];
      Yn[i + 4 * a] = Yn[i + 4 * a] - R[i + 4 * a] * coef0;
    }
}

void orc_picp(int nNo, double gam, const double* Ao, const double* Yo, double* An, double* Yn)
{
  double coef = (gam - 1.0) / gam;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < nNo; j++) {
      An[i + 4 * j] = Ao[i + 4 * j] * coef;
      Yn[i + 4 * j] = Yo[i + 4 * j];
    }
}

void orc_set_bc_dir(int nNo, const int* gN, int tDof, int s, int lDof, unsigned eDir, const double* gxn,
                    double dirA, double dirY, double* A, double* Y)
{
  /* set_bc_dir_l: tmpA(i,a) = dirA * gxn(i,a), tmpY(i,a) = dirY * gxn(i,a) */
  for (int a = 0; a < nNo; a++) {
    int Ac = gN[a];
    if (eDir) {
      int l = 0;
      for (int i = 0; i < 32; i++)
        if ((eDir >> i) & 1u) {
          A[(s + i) + tDof * Ac] = dirA * gxn[l + lDof * a];
          Y[(s + i) + tDof * Ac] = dirY * gxn[l + lDof * a];
          l = l + 1;
        }
    } else {
      for (int i = 0; i < lDof; i++) {
        A[(i + s) + tDof * Ac] = dirA * gxn[i + lDof * a];
        Y[(i + s) + tDof * Ac] = dirY * gxn[i + lDof * a];
      }
    }
  }
}
