/* CPU restatement of the reference TTM2.1-F/TTM3-F force path and ring-polymer
 * propagation.  TEST INFRASTRUCTURE ONLY: the oracle that the HIP library is
 * checked against, and the "port" CPU baseline.  Never linked into the product.
 *
 * Array layouts follow the reference's Fortran column-major arrays read as C
 * row-major: RRt(3,Natoms,P) -> [P][Natoms][3], virt(3,3) -> [col][row].
 */
#pragma once
#include <stddef.h>

typedef struct {
  int pot_model;       /* 2 = TTM2.1-F, 3 = TTM3-F */
  int nwaters;
  int nbeads;
  int polar_maxiter;
  int guess_initdip;
  double box[3];
  double rc, rc1, eps_ewald, polar_sor, polar_eps;
} orc_params;

typedef struct orc_sys orc_sys;

orc_sys *orc_create(const orc_params *p);
void orc_destroy(orc_sys *s);
void orc_get_ttm_params(const orc_sys *s, double *theta16);
void orc_set_ttm_params(orc_sys *s, const double *theta16);
double orc_get_gamma_m(const orc_sys *s);
void orc_set_gamma_m(orc_sys *s, double gamma_m);
int orc_nkvec(const orc_sys *s);
/* aewald, Uvdw_lrc, shiftA, shiftB, const_ts1..3, polfac[4] (10 doubles) */
void orc_consts(const orc_sys *s, double *out);
void orc_kvecs(const orc_sys *s, double *kvec, double *Ak, int *hkl);

/* One pot_ttm call per bead.  hist: per-bead predictor history
 * [P][4][ndip][3] (read and written; may be NULL when guess_initdip == 0).
 * terms: [P][5] = Umon, Uvdw, Uelec, Uind, Uvdw_lrc. */
int orc_forces(orc_sys *s, const double *RRt, const double *RRc, int t,
               double *En, double *terms, double *dRRt, double *virt,
               double *virialc, double *dipI, double *dipE, double *chg,
               int *iters, double *hist);

/* Scalar kernels for unit checks. */
double orc_erfc(double x);
double orc_gammq23(double x);
void orc_smear(const orc_sys *s, double drsq, double apol, double *out13);
void orc_monomer(const double *r1, double *e1, double *dr1, double *q3,
                 double *dq3, const double *box, double dms1, double dms2,
                 double dms3);
/* PES with its parameters as inputs (for finite differences of a fit). */
void orc_pes_default(double *theta);
void orc_pes_monomer(const double *r1, const double *box, const double *theta,
                     double *e1, double *dr1);

/* Ring polymer (NormalModes.f90, Langevin.f90, Nose_Hoover.f90). */
void orc_nm_init(int P, double omegan, double delta, double setnm, double *C,
                 double *A11, double *A12, double *A21, double *A22,
                 double *omegalist, double *msf);
void orc_nm_evolve(int P, int natoms, const double *C, const double *A11,
                   const double *A12, const double *A21, const double *A22,
                   const double *msf, const double *masses, double *RRt,
                   double *PPt);
void orc_langevin_init(int P, double delt2, int centroid, double tau_bead,
                       double temp, const double *omegalist, double *c1,
                       double *c2);
void orc_langevin(int P, int natoms, const double *C, const double *msf,
                  const double *c1, const double *c2, double massO,
                  double massH, const double *noise, double *PPt);
void orc_nose_hoover(double *s, double *uk, double *vxi, double tau,
                     double delt2, int L, double temp);
