/*
 * Oracle (test infrastructure only): the generalized-alpha pieces of one
 * Newton iteration for a single fluid equation (s = 0, e = 3, tDof = dof = 4).
 *
 *   pici   Code/Source/solver/pic.cpp:486-574   Ag = (1-am) Ao + am An, Yg = (1-af) Yo + af Yn
 *   picc   Code/Source/solver/pic.cpp:155-167   An -= R, Yn -= gam dt R
 *          (the Dn update of the same loop is not part of a fluid-only run:
 *           dFlag is false for fluid, initialize.cpp:420-421)
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
