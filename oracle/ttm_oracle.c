/* CPU restatement of the reference TTM force path (TEST INFRASTRUCTURE ONLY).
 *
 * Each function states the reference lines whose algorithm it follows.  The
 * statement order and the association of floating-point sums mirror the
 * reference so that results agree to rounding with the Fortran build in
 * oracle/_ref.  Sites are indexed type-major as in the reference
 * (pot_ttm.f90:53-80): site(type, water) = type*N + water with types
 * O=0, H1=1, H2=2, M=3.
 */
#include "ttm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../pimd-f90_amd/csrc/ps_tables.h"

/* consts.f90:22-42 */
#define CHARGECON 18.22234397655801030455
#define DEBYE 4.8033324
#define PI_ 3.14159265358979323844
#define TWOPI_ 6.28318530717958647688
#define FOURPI_ 12.56637061435917295376
#define TSP_ 1.12837916709551257390
#define SQPI_ 1.77245385090551602729
#define KB_AMU 0.831447148

struct orc_sys {
  orc_params p;
  int N, natoms, ns;
  double boxi[3], vol, rc2;
  double vdwA, vdwB, vdwC, vdwD, vdwE, aDD, aCC, polarO, polarH, polarM, gammaM;
  double dms1, dms2, dms3;
  double polfac[4], shiftA, shiftB, Uvdw_lrc, aewald, cts1, cts2, cts3;
  double polfac_in[3]; /* polfacO, polfacH, polfacM as given (potential.f90:88,99) */
  /* q-SPC/Fw and SPC/F (pot_mod.f90:31-44): charges, bond and angle terms */
  double qO, qH, roh0, th0, harmkb, harmka, spcA, spcB, spcC, spcD;
  int fd, ld, fdI, ldI; /* dipole site range [fd, ld] and type range */
  int hmax, kmax, lmax, nk;
  double *kvec, *Ak;
  int *hkl;
};

static inline double anint(double x) { return round(x); }
static inline void minimg(const orc_sys *s, double *r) {
  for (int c = 0; c < 3; ++c) r[c] = r[c] - s->p.box[c] * anint(r[c] * s->boxi[c]);
}

/* ---------------- special functions (math.f90:188-314) ---------------- */

/* A&S 7.1.26, math.f90:287-314 */
double orc_erfc(double x) {
  const double A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741,
               A4 = -1.453152027, A5 = 1.061405429, P = 0.3275911;
  double T = 1.0 / (1.0 + P * x);
  double XSQ = x * x;
  double TP = T * (A1 + T * (A2 + T * (A3 + T * (A4 + T * A5))));
  return TP * exp(-XSQ);
}

/* gser2_3, math.f90:211-247 */
static double gser23(double x) {
  const double a = 2.0 / 3.0, gln = 0.303150275147547, EPS = 3e-7;
  if (x <= 0.0) return 0.0;
  double ap = a, sum = 1.0 / a, del = sum;
  for (int n = 1; n <= 800; ++n) {
    ap = ap + 1.0;
    del = del * x / ap;
    sum = sum + del;
    if (fabs(del) < fabs(sum) * EPS) return sum * exp(-x + a * log(x) - gln);
  }
  return NAN;
}

/* gcf2_3, math.f90:250-285 */
static double gcf23(double x) {
  const double a = 2.0 / 3.0, gln = 0.303150275147547, EPS = 3e-7, FPMIN = 1e-30;
  double b = x + 1.0 - a, c = 1.0 / FPMIN, d = 1.0 / b, h = d;
  for (int i = 1; i <= 100; ++i) {
    double an = -((double)i * ((double)i - a));
    b = b + 2.0;
    d = an * d + b;
    if (fabs(d) < FPMIN) d = FPMIN;
    c = b + an / c;
    if (fabs(c) < FPMIN) c = FPMIN;
    d = 1.0 / d;
    double del = d * c;
    h = h * del;
    if (fabs(del - 1.0) < EPS) return exp(-x + a * log(x) - gln) * h;
  }
  return NAN;
}

/* gammq2_3, math.f90:188-208: regularized upper incomplete gamma Q(2/3, x) */
double orc_gammq23(double x) {
  const double a = 2.0 / 3.0;
  if (x < a + 1.0) return 1.0 - gser23(x);
  return gcf23(x);
}

/* ---------------- smeared interaction tensors (smear.f90) ---------------- */

/* PBCsmear01, smear.f90:2-34 */
static void smear01(const orc_sys *s, double drsq, double apol, double *ts0, double *ts1) {
  const double g23 = 1.3541179394264;
  double dd = sqrt(drsq), dri = 1.0 / dd, drsqi = dri * dri;
  double expon = exp(-(s->aewald * s->aewald * drsq));
  double er = orc_erfc(s->aewald * dd);
  double t0 = er * dri;
  double t1 = (t0 + s->cts1 * expon) * drsqi;
  if (drsq < 25.0) {
    double rA3 = dd * drsq * apol;
    double ra = pow(apol, 1.0 / 3.0);
    double exp1 = exp(-rA3);
    double ch1 = -1.0 * dri * drsqi;
    double sr1 = (1.0 - exp1) * dri * drsqi;
    t1 = t1 + ch1 + sr1;
    t0 = t0 - exp1 * dri + ra * g23 * orc_gammq23(rA3);
  }
  *ts0 = t0;
  *ts1 = t1;
}

/* PBCsmear2, smear.f90:36-69 */
static void smear2(const orc_sys *s, double drsq, double apol, double *ts1, double *ts2) {
  double dd = sqrt(drsq), dri = 1.0 / dd, drsqi = dri * dri;
  double expon = exp(-(s->aewald * s->aewald * drsq));
  double er = orc_erfc(s->aewald * dd);
  double t1 = (er * dri + s->cts1 * expon) * drsqi;
  double t2 = (t1 + s->cts2 * expon) * drsqi;
  if (drsq < 25.0) {
    double rA3 = (dd * drsq) * apol;
    double exp1 = exp(-rA3);
    double ch1 = -1.0 * dri * drsqi, ch2 = ch1 * drsqi;
    double sr1 = (1.0 - exp1) * dri * drsqi;
    double sr2 = (sr1 - exp1 * apol) * drsqi;
    t1 = t1 + ch1 + sr1;
    t2 = t2 + ch2 + sr2;
  }
  *ts1 = t1;
  *ts2 = t2;
}

/* PBCsmear3, smear.f90:71-107 */
static void smear3(const orc_sys *s, double drsq, double apol, double *ts1, double *ts2,
                   double *ts3) {
  double dd = sqrt(drsq), dri = 1.0 / dd, drsqi = dri * dri;
  double expon = exp(-(s->aewald * s->aewald * drsq));
  double er = orc_erfc(s->aewald * dd);
  double t1 = (er * dri + s->cts1 * expon) * drsqi;
  double t2 = (t1 + s->cts2 * expon) * drsqi;
  double t3 = (t2 + s->cts3 * expon) * drsqi;
  if (drsq < 25.0) {
    double rA3 = (dd * drsq) * apol;
    double exp1 = exp(-rA3);
    double ch1 = -1.0 * dri * drsqi, ch2 = ch1 * drsqi, ch3 = ch2 * drsqi;
    double sr1 = (1.0 - exp1) * dri * drsqi;
    double sr2 = (sr1 - exp1 * apol) * drsqi;
    double sr3 = (sr2 - 0.6 * exp1 * dd * apol * apol) * drsqi;
    t1 = t1 + ch1 + sr1;
    t2 = t2 + ch2 + sr2;
    t3 = t3 + ch3 + sr3;
  }
  *ts1 = t1;
  *ts2 = t2;
  *ts3 = t3;
}

/* self1, smear.f90:110-131 */
static void self1(const orc_sys *s, double drsq, double *ts0, double *ts1) {
  double dd = sqrt(drsq), drsqi = 1.0 / drsq, dri = sqrt(drsqi);
  double expon = exp(-(s->aewald * s->aewald * drsq));
  double kr0 = orc_erfc(s->aewald * dd) * dri;
  double kr1 = (kr0 + s->cts1 * expon) * drsqi;
  double ch0 = -dri, ch1 = ch0 * drsqi;
  *ts0 = ch0 + kr0;
  *ts1 = kr1 + ch1;
}

/* self2, smear.f90:133-156 */
static void self2(const orc_sys *s, double drsq, double *ts1, double *ts2) {
  double dd = sqrt(drsq), drsqi = 1.0 / drsq, dri = sqrt(drsqi);
  double expon = exp(-(s->aewald * s->aewald * drsq));
  double kr0 = orc_erfc(s->aewald * dd) * dri;
  double kr1 = (kr0 + s->cts1 * expon) * drsqi;
  double kr2 = (kr1 + s->cts2 * expon) * drsqi;
  double ch0 = -dri, ch1 = ch0 * drsqi, ch2 = ch1 * drsqi;
  *ts1 = kr1 + ch1;
  *ts2 = kr2 + ch2;
}

void orc_smear(const orc_sys *s, double drsq, double apol, double *o) {
  smear01(s, drsq, apol, &o[0], &o[1]);
  smear2(s, drsq, apol, &o[2], &o[3]);
  smear3(s, drsq, apol, &o[4], &o[5], &o[6]);
  self1(s, drsq, &o[7], &o[8]);
  self2(s, drsq, &o[9], &o[10]);
}

/* get_dd3, pot_ttm.f90:633-647; d[a][b] symmetric */
static void get_dd3(const double *R, double ts1, double ts2, double d[3][3]) {
  d[0][0] = 3.0 * ts2 * R[0] * R[0] - ts1;
  d[1][1] = 3.0 * ts2 * R[1] * R[1] - ts1;
  d[2][2] = 3.0 * ts2 * R[2] * R[2] - ts1;
  d[0][1] = 3.0 * ts2 * R[0] * R[1];
  d[0][2] = 3.0 * ts2 * R[0] * R[2];
  d[1][2] = 3.0 * ts2 * R[1] * R[2];
  d[1][0] = d[0][1];
  d[2][0] = d[0][2];
  d[2][1] = d[1][2];
}

/* ---------------- Partridge-Schwenke monomer (nasa.f90) ---------------- */

static void fmat_fill(double f[16][3], double x1, double x2, double x3) {
  for (int c = 0; c < 3; ++c) {
    f[0][c] = 0.0;
    f[1][c] = 1.0;
  }
  for (int j = 2; j <= 15; ++j) {
    f[j][0] = f[j - 1][0] * x1;
    f[j][1] = f[j - 1][1] * x2;
    f[j][2] = f[j - 1][2] * x3;
  }
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* The reference PES parameters: the combined c5z (nasa.f90:45), then deohA,
 * alphaoh, roh, phh1A, phh2. */
void orc_pes_default(double *theta) {
  for (int j = 0; j < PS_NTERM; ++j)
    theta[j] = PS_F5Z * PS_C5ZA[j] + PS_FBASIS * PS_CBASIS[j] + PS_FCORE * PS_CCORE[j] +
               PS_FREST * PS_CREST[j];
  theta[PS_NTERM + 0] = PS_DEOHA;
  theta[PS_NTERM + 1] = PS_ALPHAOH;
  theta[PS_NTERM + 2] = PS_ROH;
  theta[PS_NTERM + 3] = PS_PHH1A;
  theta[PS_NTERM + 4] = PS_PHH2;
}

/* pot_nasa, nasa.f90:2-103.  r1[atom][xyz] (O,H1,H2) -> e1, dr1[atom][xyz];
 * theta (NULL: the reference's values) as orc_pes_default. */
static void pot_nasa_theta(const double r1[3][3], double dr1[3][3], double *e1, const double *box,
                           const double *boxi, const double *theta) {
  double dflt[PS_NTERM + 5];
  if (!theta) {
    orc_pes_default(dflt);
    theta = dflt;
  }
  const double *c5z = theta;
  const double deohA = theta[PS_NTERM], alphaoh = theta[PS_NTERM + 1], roh = theta[PS_NTERM + 2],
               phh1A = theta[PS_NTERM + 3], phh2 = theta[PS_NTERM + 4];
  double ROH1[3], ROH2[3], RHH[3];
  for (int c = 0; c < 3; ++c) {
    ROH1[c] = r1[1][c] - r1[0][c];
    ROH1[c] = ROH1[c] - box[c] * anint(ROH1[c] * boxi[c]);
    ROH2[c] = r1[2][c] - r1[0][c];
    ROH2[c] = ROH2[c] - box[c] * anint(ROH2[c] * boxi[c]);
    RHH[c] = r1[1][c] - r1[2][c];
    RHH[c] = RHH[c] - box[c] * anint(RHH[c] * boxi[c]);
  }
  double dROH1 = sqrt(dot3(ROH1, ROH1)), dROH2 = sqrt(dot3(ROH2, ROH2));
  double dRHH = sqrt(dot3(RHH, RHH));
  double costh = (ROH1[0] * ROH2[0] + ROH1[1] * ROH2[1] + ROH1[2] * ROH2[2]) / (dROH1 * dROH2);
  double deoh = PS_F5Z * deohA;
  double phh1 = PS_F5Z * phh1A;
  phh1 = phh1 * exp(phh2);
  const double costhe = -.24780227221366464506;
  double exp1 = exp(-alphaoh * (dROH1 - roh));
  double exp2 = exp(-alphaoh * (dROH2 - roh));
  double Va = deoh * (exp1 * (exp1 - 2.0) + exp2 * (exp2 - 2.0));
  double Vb = phh1 * exp(-phh2 * dRHH);
  double dVa1 = 2.0 * alphaoh * deoh * exp1 * (1.0 - exp1) / dROH1;
  double dVa2 = 2.0 * alphaoh * deoh * exp2 * (1.0 - exp2) / dROH2;
  double dVb = -phh2 * Vb / dRHH;
  double x1 = (dROH1 - PS_REOH) / PS_REOH, x2 = (dROH2 - PS_REOH) / PS_REOH, x3 = costh - costhe;
  double f[16][3];
  fmat_fill(f, x1, x2, x3);
  double efac = exp(-PS_B1 * ((dROH1 - PS_REOH) * (dROH1 - PS_REOH) +
                              (dROH2 - PS_REOH) * (dROH2 - PS_REOH)));
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int j = 1; j < PS_NTERM; ++j) {
    int I = PS_IDX1[j], J = PS_IDX2[j], K = PS_IDX3[j];
    s0 = s0 + c5z[j] * (f[I][0] * f[J][1] + f[J][0] * f[I][1]) * f[K][2];
    s1 = s1 + c5z[j] * ((double)(I - 1) * f[I - 1][0] * f[J][1] +
                        (double)(J - 1) * f[J - 1][0] * f[I][1]) * f[K][2];
    s2 = s2 + c5z[j] * ((double)(J - 1) * f[I][0] * f[J - 1][1] +
                        (double)(I - 1) * f[J][0] * f[I - 1][1]) * f[K][2];
    s3 = s3 + c5z[j] * (f[I][0] * f[J][1] + f[J][0] * f[I][1]) * (double)(K - 1) * f[K - 1][2];
  }
  double Vc = 2.0 * c5z[0] + efac * s0;
  double e = Va + Vb + Vc;
  e = e + 0.44739574026257;
  *e1 = e * 0.00285914375100642899;
  double dVcdr1 = (-2.0 * PS_B1 * efac * (dROH1 - PS_REOH) * s0 + efac * s1 / PS_REOH) / dROH1;
  double dVcdr2 = (-2.0 * PS_B1 * efac * (dROH2 - PS_REOH) * s0 + efac * s2 / PS_REOH) / dROH2;
  double dVcdcth = efac * s3;
  for (int c = 0; c < 3; ++c) {
    dr1[1][c] = dVa1 * ROH1[c] + dVb * RHH[c] + dVcdr1 * ROH1[c] +
                dVcdcth * (ROH2[c] / (dROH1 * dROH2) - costh * ROH1[c] / (dROH1 * dROH1));
    dr1[2][c] = dVa2 * ROH2[c] - dVb * RHH[c] + dVcdr2 * ROH2[c] +
                dVcdcth * (ROH1[c] / (dROH1 * dROH2) - costh * ROH2[c] / (dROH2 * dROH2));
    dr1[0][c] = -(dr1[1][c] + dr1[2][c]);
  }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) dr1[a][c] = dr1[a][c] * .00285914375100642899;
}

static void pot_nasa(const double r1[3][3], double dr1[3][3], double *e1, const double *box,
                     const double *boxi) {
  pot_nasa_theta(r1, dr1, e1, box, boxi, NULL);
}

/* pot_nasa with the PES parameters theta (orc_pes_default order). */
void orc_pes_monomer(const double *r1, const double *box, const double *theta, double *e1,
                     double *dr1) {
  double boxi[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
  double r[3][3], d[3][3];
  memcpy(r, r1, sizeof r);
  pot_nasa_theta((const double(*)[3])r, d, e1, box, boxi, theta);
  memcpy(dr1, d, sizeof d);
}

/* dms_nasa, nasa.f90:105-293.  q3 = (qO, qH1, qH2); gq[disp][charge][xyz] with
 * disp = (H1, H2, O) and charge = (H1, H2, O) as in the reference's gradq. */
static void dms_nasa(const double r1[3][3], double q3[3], double gq[3][3][3], const double *box,
                     const double *boxi, double dms1, double dms2, double dms3) {
  const double ath0 = 1.82400520401572996557, costhe = -.24780227221366464506;
  const double B2A = 0.529177249;
  const double bfac = 1.0 / (double)0.529177249f; /* real(4) literal, nasa.f90:218 */
  double ROH1[3], ROH2[3], RHH[3];
  for (int c = 0; c < 3; ++c) {
    ROH1[c] = r1[1][c] - r1[0][c];
    ROH1[c] = ROH1[c] - box[c] * anint(ROH1[c] * boxi[c]);
    ROH2[c] = r1[2][c] - r1[0][c];
    ROH2[c] = ROH2[c] - box[c] * anint(ROH2[c] * boxi[c]);
    RHH[c] = r1[1][c] - r1[2][c];
    RHH[c] = RHH[c] - box[c] * anint(RHH[c] * boxi[c]);
  }
  double dROH1 = sqrt(dot3(ROH1, ROH1)), dROH2 = sqrt(dot3(ROH2, ROH2));
  double costh = dot3(ROH1, ROH2) / (dROH1 * dROH2);
  double efac = exp(-DMS_B1D * ((dROH1 - PS_REOH) * (dROH1 - PS_REOH) +
                                (dROH2 - PS_REOH) * (dROH2 - PS_REOH)));
  double x1 = (dROH1 - PS_REOH) / PS_REOH, x2 = (dROH2 - PS_REOH) / PS_REOH, x3 = costh - costhe;
  double f[16][3];
  fmat_fill(f, x1, x2, x3);
  double P1 = 0, P2 = 0, PL1 = costh, PL2 = 0.5 * (3.0 * PL1 * PL1 - 1.0);
  double dp1dr1 = 0, dp1dr2 = 0, dp1dc = 0, dp2dr1 = 0, dp2dr2 = 0, dp2dc = 0;
  for (int j = 1; j < PS_NDMS; ++j) {
    int I = PS_IDXD1[j], J = PS_IDXD2[j], K = PS_IDXD3[j];
    double cj = PS_COEFD[j];
    P1 = P1 + cj * f[I][0] * f[J][1] * f[K][2];
    P2 = P2 + cj * f[J][0] * f[I][1] * f[K][2];
    dp1dr1 = dp1dr1 + cj * ((double)(I - 1) * f[I - 1][0] * f[J][1] * f[K][2]);
    dp1dr2 = dp1dr2 + cj * ((double)(J - 1) * f[I][0] * f[J - 1][1] * f[K][2]);
    dp1dc = dp1dc + cj * ((double)(K - 1) * f[I][0] * f[J][1] * f[K - 1][2]);
    dp2dr1 = dp2dr1 + cj * ((double)(J - 1) * f[J - 1][0] * f[I][1] * f[K][2]);
    dp2dr2 = dp2dr2 + cj * ((double)(I - 1) * f[J][0] * f[I - 1][1] * f[K][2]);
    dp2dc = dp2dc + cj * ((double)(K - 1) * f[J][0] * f[I][1] * f[K - 1][2]);
  }
  dp1dr1 = dp1dr1 / PS_REOH * B2A;
  dp1dr2 = dp1dr2 / PS_REOH * B2A;
  dp2dr1 = dp2dr1 / PS_REOH * B2A;
  dp2dr2 = dp2dr2 / PS_REOH * B2A;
  double PC0 = DMS_A * (pow(dROH1, DMS_B) + pow(dROH2, DMS_B)) * (DMS_C0 + PL1 * DMS_C1 + PL2 * DMS_C2);
  double dpc0dr1 = DMS_A * (DMS_B * pow(dROH1, DMS_B - 1)) * (DMS_C0 + PL1 * DMS_C1 + PL2 * DMS_C2) * B2A * B2A;
  double dpc0dr2 = DMS_A * (DMS_B * pow(dROH2, DMS_B - 1)) * (DMS_C0 + PL1 * DMS_C1 + PL2 * DMS_C2) * B2A * B2A;
  double dpc0dc = DMS_A * (pow(dROH1, DMS_B) + pow(dROH2, DMS_B)) * (DMS_C1 + 0.5 * (6.0 * PL1) * DMS_C2) * B2A;
  double defacdr1 = -2.0 * DMS_B1D * (dROH1 - PS_REOH) * efac * B2A;
  double defacdr2 = -2.0 * DMS_B1D * (dROH2 - PS_REOH) * efac * B2A;
  dp1dr1 = dp1dr1 * efac + P1 * defacdr1 + dpc0dr1;
  dp1dr2 = dp1dr2 * efac + P1 * defacdr2 + dpc0dr2;
  dp1dc = dp1dc * efac + dpc0dc;
  dp2dr1 = dp2dr1 * efac + P2 * defacdr1 + dpc0dr1;
  dp2dr2 = dp2dr2 * efac + P2 * defacdr2 + dpc0dr2;
  dp2dc = dp2dc * efac + dpc0dc;
  P1 = PS_COEFD[0] + P1 * efac + PC0 * B2A;
  P2 = PS_COEFD[0] + P2 * efac + PC0 * B2A;
  q3[0] = -(P1 + P2);
  q3[1] = P1;
  q3[2] = P2;
  dp1dr1 = dp1dr1 * bfac;
  dp1dr2 = dp1dr2 * bfac;
  dp2dr1 = dp2dr1 * bfac;
  dp2dr2 = dp2dr2 * bfac;
  double AxB[3];
  AxB[0] = ROH1[1] * ROH2[2] - ROH1[2] * ROH2[1];
  AxB[1] = -ROH1[0] * ROH2[2] + ROH1[2] * ROH2[0];
  AxB[2] = ROH1[0] * ROH2[1] - ROH1[1] * ROH2[0];
  double dAxB = sqrt(dot3(AxB, AxB));
  double sinth = dAxB / (dROH1 * dROH2);
  double ang = atan2(sinth, costh);
  P1 = dms1 * (dROH1 - dms2) + dms3 * (ang - ath0);
  P2 = dms1 * (dROH2 - dms2) + dms3 * (ang - ath0);
  q3[0] = q3[0] - (P1 + P2);
  q3[1] = q3[1] + P1;
  q3[2] = q3[2] + P2;
  dp1dr1 = dp1dr1 + dms1;
  dp2dr2 = dp2dr2 + dms1;
  dp1dc = dp1dc - dms3 / sinth;
  dp2dc = dp2dc - dms3 / sinth;
  double f1q1r13 = (dp1dr1 - (dp1dc * costh / dROH1)) / dROH1;
  double f1q1r23 = dp1dc / (dROH1 * dROH2);
  double f2q1r23 = (dp1dr2 - (dp1dc * costh / dROH2)) / dROH2;
  double f2q1r13 = dp1dc / (dROH2 * dROH1);
  double f1q2r13 = (dp2dr1 - (dp2dc * costh / dROH1)) / dROH1;
  double f1q2r23 = dp2dc / (dROH1 * dROH2);
  double f2q2r23 = (dp2dr2 - (dp2dc * costh / dROH2)) / dROH2;
  double f2q2r13 = dp2dc / (dROH2 * dROH1);
  for (int c = 0; c < 3; ++c) {
    gq[0][0][c] = f1q1r13 * ROH1[c] + f1q1r23 * ROH2[c];
    gq[1][0][c] = f2q1r13 * ROH1[c] + f2q1r23 * ROH2[c];
    gq[2][0][c] = -(gq[0][0][c] + gq[1][0][c]);
    gq[0][1][c] = f1q2r13 * ROH1[c] + f1q2r23 * ROH2[c];
    gq[1][1][c] = f2q2r13 * ROH1[c] + f2q2r23 * ROH2[c];
    gq[2][1][c] = -(gq[0][1][c] + gq[1][1][c]);
    gq[0][2][c] = -(gq[0][0][c] + gq[0][1][c]);
    gq[1][2][c] = -(gq[1][0][c] + gq[1][1][c]);
    gq[2][2][c] = -(gq[2][0][c] + gq[2][1][c]);
  }
}

void orc_monomer(const double *r1, double *e1, double *dr1, double *q3, double *dq3,
                 const double *box, double dms1, double dms2, double dms3) {
  double boxi[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
  double r[3][3], d[3][3], g[3][3][3];
  memcpy(r, r1, sizeof r);
  pot_nasa((const double(*)[3])r, d, e1, box, boxi);
  memcpy(dr1, d, sizeof d);
  dms_nasa((const double(*)[3])r, q3, g, box, boxi, dms1, dms2, dms3);
  memcpy(dq3, g, sizeof g);
}

/* ---------------- set-up (potential.f90:43-199, ewald.f90:293-378) ---------------- */

/* The vdW long-range correction and polfac = (polfac^(1/6))^-3
 * (potential.f90:152-160) from the model parameters. */
static void derive_params(orc_sys *s) {
  const double Rc = s->p.rc;
  s->Uvdw_lrc = TWOPI_ * (double)(s->N * s->N) / s->vol *
                (s->vdwA / (9.0 * pow(Rc, 9)) + s->vdwB / (7.0 * pow(Rc, 7)) +
                 s->vdwC / (3.0 * (Rc * Rc * Rc)));
  double pf[4] = {pow(s->polfac_in[0], 1.0 / 6.0), pow(s->polfac_in[1], 1.0 / 6.0),
                  pow(s->polfac_in[1], 1.0 / 6.0), pow(s->polfac_in[2], 1.0 / 6.0)};
  for (int i = 0; i < 4; ++i) s->polfac[i] = 1.0 / (pf[i] * pf[i] * pf[i]);
}

/* Model parameters in the order vdwA..vdwE, polarO, polarH, polarM, polfacO,
 * polfacH, polfacM, aDD, aCCaCD, dms_param1..3 (the values of pot_mod.f90:7-26
 * unless set), for finite-difference checks of parameter derivatives. */
void orc_get_ttm_params(const orc_sys *s, double *th) {
  const double v[16] = {s->vdwA, s->vdwB, s->vdwC, s->vdwD, s->vdwE, s->polarO, s->polarH,
                        s->polarM, s->polfac_in[0], s->polfac_in[1], s->polfac_in[2], s->aDD,
                        s->aCC, s->dms1, s->dms2, s->dms3};
  memcpy(th, v, sizeof v);
}

void orc_set_ttm_params(orc_sys *s, const double *th) {
  s->vdwA = th[0]; s->vdwB = th[1]; s->vdwC = th[2]; s->vdwD = th[3]; s->vdwE = th[4];
  s->polarO = th[5]; s->polarH = th[6]; s->polarM = th[7];
  s->polfac_in[0] = th[8]; s->polfac_in[1] = th[9]; s->polfac_in[2] = th[10];
  s->aDD = th[11]; s->aCC = th[12];
  s->dms1 = th[13]; s->dms2 = th[14]; s->dms3 = th[15];
  derive_params(s);
}

/* The M-site fraction gammaM (pot_mod.f90:12,24), for finite-difference
 * checks of its derivative. */
double orc_get_gamma_m(const orc_sys *s) { return s->gammaM; }
void orc_set_gamma_m(orc_sys *s, double g) { s->gammaM = g; }

orc_sys *orc_create(const orc_params *p) {
  orc_sys *s = calloc(1, sizeof *s);
  s->p = *p;
  s->N = p->nwaters;
  s->natoms = 3 * s->N;
  s->ns = 4 * s->N;
  for (int c = 0; c < 3; ++c) s->boxi[c] = 1.0 / p->box[c];
  s->vol = p->box[0] * p->box[1] * p->box[2];
  s->rc2 = p->rc * p->rc;
  double polfacO, polfacH, polfacM;
  const double Rc = p->rc, rc1 = p->rc1;
  if (p->pot_model == 4 || p->pot_model == 5) {
    /* q-SPC/Fw (pot_mod.f90:31-35, potential.f90:121-127) and SPC/F
     * (pot_mod.f90:40-44, potential.f90:140-145): three charged sites per
     * water (potential.f90:65-71), no dipoles, no vdW shift */
    s->ns = 3 * s->N;
    s->fd = 0; s->ld = -1; s->fdI = 0; s->ldI = -1;
    s->vdwB = 0.0; s->vdwD = 0.0; s->vdwE = 0.0;
    if (p->pot_model == 4) {
      s->vdwA = 629326.57249877361701443513; s->vdwC = -625.50206521141523463400;
      s->qO = -0.84; s->qH = 0.42; s->roh0 = 1.012; s->th0 = 1.95476876223364912614;
      s->harmkb = 529.581; s->harmka = 37.95;
    } else {
      s->vdwA = 628213.72817108343399678768; s->vdwC = -624.97858519194071637260;
      s->qO = -0.82; s->qH = 0.41; s->roh0 = 1.0; s->th0 = 1.91061193215819258784;
      s->spcA = 671.288262912; s->spcB = 164.376; s->spcC = -211.536; s->spcD = 111.744;
    }
    polfacO = polfacH = polfacM = 1.0; /* not used by these models */
    s->shiftA = 0.0;
    s->shiftB = 0.0;
  } else if (p->pot_model == 2) { /* TTM2.1-F, pot_mod.f90:7-14, potential.f90:84-94 */
    s->fd = 0;
    s->ld = 3 * s->N - 1;
    s->vdwA = -1.329565985e+6; s->vdwB = 3.632560798e+5; s->vdwC = -2.147141323e+3;
    s->vdwD = 1e13; s->vdwE = 13.2;
    polfacO = 0.837; polfacH = 0.496; polfacM = 0.837;
    s->polarO = 0.837; s->polarH = 0.496; s->polarM = 0.837;
    s->gammaM = 0.426706882; s->aDD = 0.3; s->aCC = 0.2;
    s->dms1 = 0.0; s->dms2 = 0.0; s->dms3 = 0.0;
    s->fdI = 0; s->ldI = 2;
    s->shiftA = 0.0; s->shiftB = 0.0;
  } else { /* TTM3-F, pot_mod.f90:19-26, potential.f90:95-113 */
    s->fd = 3 * s->N;
    s->ld = 4 * s->N - 1;
    s->vdwA = 0.0; s->vdwB = 0.0; s->vdwC = -0.72298855e+03; s->vdwD = 0.10211829e+06;
    s->vdwE = 0.37170376e+01;
    polfacO = 0.837; polfacH = 0.496; polfacM = 0.837;
    s->polarO = 0.0; s->polarH = 0.0; s->polarM = 1.444;
    s->gammaM = 0.46; s->aDD = 0.175; s->aCC = 0.175;
    s->dms1 = 0.5; s->dms2 = 0.9578; s->dms3 = 0.012;
    s->fdI = 3; s->ldI = 3;
    if (Rc == rc1) {
      s->shiftA = 0.0;
      s->shiftB = 0.0;
    } else {
      double Rc8 = pow(Rc, 8);
      s->shiftA = -(10 * Rc - 7 * rc1) / (Rc8 * ((Rc - rc1) * (Rc - rc1)));
      s->shiftB = (9 * Rc - 7 * rc1) / (Rc8 * ((Rc - rc1) * (Rc - rc1) * (Rc - rc1)));
    }
  }
  s->polfac_in[0] = polfacO;
  s->polfac_in[1] = polfacH;
  s->polfac_in[2] = polfacM;
  derive_params(s);
  /* ewald_set(.true.), ewald.f90:308-377 */
  double PP = -log(p->eps_ewald);
  s->aewald = sqrt(PP) / Rc;
  double kcut = 2.0 * s->aewald * sqrt(PP), kcut2 = kcut * kcut;
  s->hmax = (int)floor(kcut / TWOPI_ * p->box[0]);
  s->kmax = (int)floor(kcut / TWOPI_ * p->box[1]);
  s->lmax = (int)floor(kcut / TWOPI_ * p->box[2]);
  double fac = -0.25 / (s->aewald * s->aewald);
  int nmax = 4 * (s->hmax + 1) * (s->kmax + 1) * (s->lmax + 1);
  s->kvec = malloc(sizeof(double) * 3 * nmax);
  s->Ak = malloc(sizeof(double) * nmax);
  s->hkl = malloc(sizeof(int) * 3 * nmax);
  s->nk = 0;
  for (int kx = 0; kx <= s->hmax; ++kx) {
    double kv0 = TWOPI_ * s->boxi[0] * (double)kx;
    int ky1 = kx == 0 ? 0 : -s->kmax;
    for (int ky = ky1; ky <= s->kmax; ++ky) {
      double kv1 = TWOPI_ * s->boxi[1] * (double)ky;
      int kz1 = (kx == 0 && ky == 0) ? 1 : -s->lmax;
      double kvxy = kv0 * kv0 + kv1 * kv1;
      for (int kz = kz1; kz <= s->lmax; ++kz) {
        double kv2 = TWOPI_ * s->boxi[2] * (double)kz;
        double kvsq = kvxy + kv2 * kv2;
        if (kvsq < kcut2 && kvsq > 1e-16) {
          int n = s->nk++;
          s->hkl[3 * n] = kx; s->hkl[3 * n + 1] = ky; s->hkl[3 * n + 2] = kz;
          s->kvec[3 * n] = kv0; s->kvec[3 * n + 1] = kv1; s->kvec[3 * n + 2] = kv2;
          s->Ak[n] = 2.0 * exp(fac * kvsq) / kvsq;
        }
      }
    }
  }
  s->cts1 = 2.0 * s->aewald / SQPI_;
  s->cts2 = 4.0 * s->aewald * s->aewald * s->aewald / (3.0 * SQPI_);
  s->cts3 = 8.0 * pow(s->aewald, 5) / (15.0 * SQPI_);
  return s;
}

void orc_destroy(orc_sys *s) {
  if (!s) return;
  free(s->kvec);
  free(s->Ak);
  free(s->hkl);
  free(s);
}

int orc_nkvec(const orc_sys *s) { return s->nk; }

void orc_consts(const orc_sys *s, double *o) {
  o[0] = s->aewald; o[1] = s->Uvdw_lrc; o[2] = s->shiftA; o[3] = s->shiftB;
  o[4] = s->cts1; o[5] = s->cts2; o[6] = s->cts3;
  for (int i = 0; i < 4; ++i) o[7 + i] = s->polfac[i];
}

void orc_kvecs(const orc_sys *s, double *kvec, double *Ak, int *hkl) {
  memcpy(kvec, s->kvec, sizeof(double) * 3 * s->nk);
  memcpy(Ak, s->Ak, sizeof(double) * s->nk);
  memcpy(hkl, s->hkl, sizeof(int) * 3 * s->nk);
}

/* ---------------- one bead: pot_ttm (pot_ttm.f90:3-630) ---------------- */

typedef struct {
  int n;
  int *j;
  double *Rij; /* [n][3] */
  double *R2;
} neigh_t;

/* find_neigh, find_neigh.f90:14-120: O-O partners j > i inside Rc (strict) */
static void find_neigh(const orc_sys *s, int iO, double (*R)[3], neigh_t *nb) {
  int n = 0;
  for (int j = iO + 1; j < s->N; ++j) {
    double r[3];
    for (int c = 0; c < 3; ++c) r[c] = R[iO][c] - R[j][c];
    minimg(s, r);
    double R2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    if (R2 < s->rc2) {
      nb->j[n] = j;
      memcpy(&nb->Rij[3 * n], r, sizeof r);
      nb->R2[n] = R2;
      ++n;
    }
  }
  nb->n = n;
}

typedef struct {
  double (*R)[3], (*Rc)[3], (*dR)[3], (*Efq)[3], (*Efd)[3], (*dip)[3], (*olddip)[3];
  double *charge, *phi, *grdq; /* grdq[w][disp][chg][xyz] */
  double *cx, *sx, *cy, *sy, *cz, *sz; /* [site][h], y/z offset by kmax/lmax */
  double *cxy, *sxy, *ckr, *skr, *dKk, *dKkX, *dKkXY;
  double *rQq, *iQq, *rQm, *iQm;
  neigh_t nb;
} scratch_t;

static void scratch_alloc(const orc_sys *s, scratch_t *w) {
  int ns = s->ns, nk = s->nk;
  w->R = calloc(ns, sizeof *w->R); w->Rc = calloc(ns, sizeof *w->Rc);
  w->dR = calloc(ns, sizeof *w->dR); w->Efq = calloc(ns, sizeof *w->Efq);
  w->Efd = calloc(ns, sizeof *w->Efd); w->dip = calloc(ns, sizeof *w->dip);
  w->olddip = calloc(ns, sizeof *w->olddip);
  w->charge = calloc(ns, sizeof(double)); w->phi = calloc(ns, sizeof(double));
  w->grdq = calloc((size_t)s->N * 27, sizeof(double));
  w->cx = calloc((size_t)ns * (s->hmax + 1), sizeof(double));
  w->sx = calloc((size_t)ns * (s->hmax + 1), sizeof(double));
  w->cy = calloc((size_t)ns * (2 * s->kmax + 1), sizeof(double));
  w->sy = calloc((size_t)ns * (2 * s->kmax + 1), sizeof(double));
  w->cz = calloc((size_t)ns * (2 * s->lmax + 1), sizeof(double));
  w->sz = calloc((size_t)ns * (2 * s->lmax + 1), sizeof(double));
  w->cxy = calloc(ns, sizeof(double)); w->sxy = calloc(ns, sizeof(double));
  w->ckr = calloc(ns, sizeof(double)); w->skr = calloc(ns, sizeof(double));
  w->dKk = calloc(ns, sizeof(double)); w->dKkX = calloc(ns, sizeof(double));
  w->dKkXY = calloc(ns, sizeof(double));
  w->rQq = calloc(nk, sizeof(double)); w->iQq = calloc(nk, sizeof(double));
  w->rQm = calloc(nk, sizeof(double)); w->iQm = calloc(nk, sizeof(double));
  w->nb.j = calloc(s->N, sizeof(int));
  w->nb.Rij = calloc((size_t)s->N * 3, sizeof(double));
  w->nb.R2 = calloc(s->N, sizeof(double));
}

static void scratch_free(scratch_t *w) {
  void *ptrs[] = {w->R, w->Rc, w->dR, w->Efq, w->Efd, w->dip, w->olddip, w->charge, w->phi,
                  w->grdq, w->cx, w->sx, w->cy, w->sy, w->cz, w->sz, w->cxy, w->sxy, w->ckr,
                  w->skr, w->dKk, w->dKkX, w->dKkXY, w->rQq, w->iQq, w->rQm, w->iQm,
                  w->nb.j, w->nb.Rij, w->nb.R2};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
}

#define CX(i, h) w->cx[(size_t)(i) * (s->hmax + 1) + (h)]
#define SX(i, h) w->sx[(size_t)(i) * (s->hmax + 1) + (h)]
#define CY(i, h) w->cy[(size_t)(i) * (2 * s->kmax + 1) + (h) + s->kmax]
#define SY(i, h) w->sy[(size_t)(i) * (2 * s->kmax + 1) + (h) + s->kmax]
#define CZ(i, h) w->cz[(size_t)(i) * (2 * s->lmax + 1) + (h) + s->lmax]
#define SZ(i, h) w->sz[(size_t)(i) * (2 * s->lmax + 1) + (h) + s->lmax]

/* ssincos, ewald.f90:403-437 */
static void ssincos(const orc_sys *s, scratch_t *w) {
  for (int i = 0; i < s->ns; ++i) {
    CX(i, 0) = 1.0; CY(i, 0) = 1.0; CZ(i, 0) = 1.0;
    SX(i, 0) = 0.0; SY(i, 0) = 0.0; SZ(i, 0) = 0.0;
    CX(i, 1) = cos(TWOPI_ * s->boxi[0] * w->R[i][0]);
    SX(i, 1) = sin(TWOPI_ * s->boxi[0] * w->R[i][0]);
    CY(i, 1) = cos(TWOPI_ * s->boxi[1] * w->R[i][1]); CY(i, -1) = CY(i, 1);
    SY(i, 1) = sin(TWOPI_ * s->boxi[1] * w->R[i][1]); SY(i, -1) = -SY(i, 1);
    CZ(i, 1) = cos(TWOPI_ * s->boxi[2] * w->R[i][2]); CZ(i, -1) = CZ(i, 1);
    SZ(i, 1) = sin(TWOPI_ * s->boxi[2] * w->R[i][2]); SZ(i, -1) = -SZ(i, 1);
    for (int k = 2; k <= s->hmax; ++k) {
      CX(i, k) = CX(i, k - 1) * CX(i, 1) - SX(i, k - 1) * SX(i, 1);
      SX(i, k) = CX(i, k - 1) * SX(i, 1) + SX(i, k - 1) * CX(i, 1);
    }
    for (int k = 2; k <= s->kmax; ++k) {
      CY(i, k) = CY(i, k - 1) * CY(i, 1) - SY(i, k - 1) * SY(i, 1);
      SY(i, k) = CY(i, k - 1) * SY(i, 1) + SY(i, k - 1) * CY(i, 1);
      CY(i, -k) = CY(i, k);
      SY(i, -k) = -SY(i, k);
    }
    for (int k = 2; k <= s->lmax; ++k) {
      CZ(i, k) = CZ(i, k - 1) * CZ(i, 1) - SZ(i, k - 1) * SZ(i, 1);
      SZ(i, k) = CZ(i, k - 1) * SZ(i, 1) + SZ(i, k - 1) * CZ(i, 1);
      CZ(i, -k) = CZ(i, k);
      SZ(i, -k) = -SZ(i, k);
    }
  }
}

/* ewald_std_qq, ewald.f90:1-52 */
static void ewald_qq(const orc_sys *s, scratch_t *w) {
  int N = s->N, fH = N, lM = 4 * N - 1;
  ssincos(s, w);
  int lastkx = -1000, lastky = -1000;
  for (int k = 0; k < s->nk; ++k) {
    int kx = s->hkl[3 * k], ky = s->hkl[3 * k + 1], kz = s->hkl[3 * k + 2];
    const double *kv = &s->kvec[3 * k];
    double p4vAkk = FOURPI_ * s->Ak[k] / s->vol;
    if (!(kx == lastkx && ky == lastky)) {
      for (int i = 0; i <= lM; ++i) {
        double tmp = SX(i, kx) * CY(i, ky) + CX(i, kx) * SY(i, ky);
        w->cxy[i] = CX(i, kx) * CY(i, ky) - SX(i, kx) * SY(i, ky);
        w->sxy[i] = tmp;
      }
    }
    lastkx = kx;
    lastky = ky;
    for (int i = 0; i <= lM; ++i) {
      double tmp = w->sxy[i] * CZ(i, kz) + w->cxy[i] * SZ(i, kz);
      w->ckr[i] = w->cxy[i] * CZ(i, kz) - w->sxy[i] * SZ(i, kz);
      w->skr[i] = tmp;
    }
    double rQ = 0, iQ = 0;
    for (int i = fH; i <= lM; ++i) rQ += w->charge[i] * w->ckr[i];
    for (int i = fH; i <= lM; ++i) iQ += w->charge[i] * w->skr[i];
    w->rQq[k] = rQ;
    w->iQq[k] = iQ;
    for (int i = 0; i <= lM; ++i) {
      double tmp = p4vAkk * (w->ckr[i] * iQ - w->skr[i] * rQ);
      for (int c = 0; c < 3; ++c) w->Efq[i][c] = w->Efq[i][c] - tmp * kv[c];
    }
    for (int i = fH; i <= lM; ++i)
      w->phi[i] = w->phi[i] + p4vAkk * (w->ckr[i] * rQ + w->skr[i] * iQ);
  }
}

/* ewald_std_dd, ewald.f90:56-119 */
static void ewald_dd(const orc_sys *s, scratch_t *w) {
  int fd = s->fd, ld = s->ld;
  int lastkx = -1000, lastky = -1000;
  for (int k = 0; k < s->nk; ++k) {
    int kx = s->hkl[3 * k], ky = s->hkl[3 * k + 1], kz = s->hkl[3 * k + 2];
    const double *kv = &s->kvec[3 * k];
    double p4vAkk = FOURPI_ * s->Ak[k] / s->vol;
    if (!(kx == lastkx))
      for (int i = fd; i <= ld; ++i) w->dKkX[i] = kv[0] * w->dip[i][0];
    if (!(kx == lastkx && ky == lastky)) {
      for (int i = fd; i <= ld; ++i) w->dKkXY[i] = w->dKkX[i] + kv[1] * w->dip[i][1];
      for (int i = fd; i <= ld; ++i) {
        double tmp = SX(i, kx) * CY(i, ky) + CX(i, kx) * SY(i, ky);
        w->cxy[i] = CX(i, kx) * CY(i, ky) - SX(i, kx) * SY(i, ky);
        w->sxy[i] = tmp;
      }
    }
    for (int i = fd; i <= ld; ++i) w->dKk[i] = w->dKkXY[i] + kv[2] * w->dip[i][2];
    double rQ = 0, iQ = 0;
    for (int i = fd; i <= ld; ++i) {
      double tmp = w->sxy[i] * CZ(i, kz) + w->cxy[i] * SZ(i, kz);
      w->ckr[i] = w->cxy[i] * CZ(i, kz) - w->sxy[i] * SZ(i, kz);
      w->skr[i] = tmp;
      rQ = rQ - w->dKk[i] * w->skr[i];
      iQ = iQ + w->dKk[i] * w->ckr[i];
    }
    w->rQm[k] = rQ;
    w->iQm[k] = iQ;
    lastkx = kx;
    lastky = ky;
    for (int i = fd; i <= ld; ++i) {
      double tmp = p4vAkk * (w->ckr[i] * iQ - w->skr[i] * rQ);
      for (int c = 0; c < 3; ++c) w->Efd[i][c] = w->Efd[i][c] - kv[c] * tmp;
    }
  }
}

/* ewald_std_qd, ewald.f90:123-226; returns vir_rec[col][row] */
static void ewald_qd(const orc_sys *s, scratch_t *w, double vir_rec[3][3]) {
  int N = s->N, fH = N, lM = 4 * N - 1, fd = s->fd, ld = s->ld;
  int lastkx = -1000, lastky = -1000;
  double vt[3][3] = {{0}}, vt1[3][3] = {{0}}; /* [row][col] */
  for (int k = 0; k < s->nk; ++k) {
    int kx = s->hkl[3 * k], ky = s->hkl[3 * k + 1], kz = s->hkl[3 * k + 2];
    const double *kv = &s->kvec[3 * k];
    double Akk = s->Ak[k];
    double p4vAkk = FOURPI_ * Akk / s->vol;
    if (!(kx == lastkx))
      for (int i = fd; i <= ld; ++i) w->dKkX[i] = kv[0] * w->dip[i][0];
    if (!(kx == lastkx && ky == lastky)) {
      for (int i = fd; i <= ld; ++i) w->dKkXY[i] = w->dKkX[i] + kv[1] * w->dip[i][1];
      for (int i = 0; i <= lM; ++i) {
        double tmp = SX(i, kx) * CY(i, ky) + CX(i, kx) * SY(i, ky);
        w->cxy[i] = CX(i, kx) * CY(i, ky) - SX(i, kx) * SY(i, ky);
        w->sxy[i] = tmp;
      }
    }
    for (int i = fd; i <= ld; ++i) w->dKk[i] = w->dKkXY[i] + kv[2] * w->dip[i][2];
    for (int i = 0; i <= lM; ++i) {
      double tmp = w->sxy[i] * CZ(i, kz) + w->cxy[i] * SZ(i, kz);
      w->ckr[i] = w->cxy[i] * CZ(i, kz) - w->sxy[i] * SZ(i, kz);
      w->skr[i] = tmp;
    }
    lastkx = kx;
    lastky = ky;
    double rQmk = w->rQm[k], iQmk = w->iQm[k];
    double rQqmk = w->rQq[k] + w->rQm[k], iQqmk = w->iQq[k] + w->iQm[k];
    for (int i = fH; i <= lM; ++i) {
      double tmp = w->charge[i] * p4vAkk * (w->ckr[i] * iQmk - w->skr[i] * rQmk);
      w->phi[i] = w->phi[i] + p4vAkk * (w->ckr[i] * rQmk + w->skr[i] * iQmk);
      for (int c = 0; c < 3; ++c) w->dR[i][c] = w->dR[i][c] + tmp * kv[c];
    }
    for (int i = fd; i <= ld; ++i) {
      double tmp = w->dKk[i] * p4vAkk * (w->ckr[i] * rQqmk + w->skr[i] * iQqmk);
      for (int c = 0; c < 3; ++c) w->dR[i][c] = w->dR[i][c] - tmp * kv[c];
    }
    double ksq = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    double coeff2 = 2.0 / ksq + 0.5 / (s->aewald * s->aewald);
    double Ureck = Akk * (rQqmk * rQqmk + iQqmk * iQqmk);
    vt[0][0] = vt[0][0] + Ureck - Ureck * coeff2 * kv[0] * kv[0];
    vt[1][0] = vt[1][0] - Ureck * coeff2 * kv[0] * kv[1];
    vt[2][0] = vt[2][0] - Ureck * coeff2 * kv[0] * kv[2];
    vt[1][1] = vt[1][1] + Ureck - Ureck * coeff2 * kv[1] * kv[1];
    vt[2][1] = vt[2][1] - Ureck * coeff2 * kv[1] * kv[2];
    vt[2][2] = vt[2][2] + Ureck - Ureck * coeff2 * kv[2] * kv[2];
    vt[0][1] = vt[0][1] - Ureck * coeff2 * kv[0] * kv[1];
    vt[0][2] = vt[0][2] - Ureck * coeff2 * kv[0] * kv[2];
    vt[1][2] = vt[1][2] - Ureck * coeff2 * kv[1] * kv[2];
    double rdipQ[3], idipQ[3];
    for (int c = 0; c < 3; ++c) {
      double a = 0, b = 0;
      for (int i = fd; i <= ld; ++i) a += w->skr[i] * w->dip[i][c];
      for (int i = fd; i <= ld; ++i) b += w->ckr[i] * w->dip[i][c];
      rdipQ[c] = -a;
      idipQ[c] = b;
    }
    for (int ix = 0; ix < 3; ++ix) {
      double tmp = 2.0 * Akk * kv[ix];
      for (int iy = 0; iy < 3; ++iy)
        vt1[ix][iy] = vt1[ix][iy] + tmp * (rQqmk * rdipQ[iy] + iQqmk * idipQ[iy]);
    }
  }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double a = -0.5 * vt[r][c] * FOURPI_ / s->vol;
      double b = -0.5 * vt1[r][c] * FOURPI_ / s->vol;
      vir_rec[c][r] = a + b;
    }
}

/* calc_dpcf, pot_ttm.f90:649-675 */
static void calc_dpcf(int t, double d[4]) {
  static const double tab[4][4] = {{-1, 4, -6, 4}, {4, -1, 4, -6}, {-6, 4, -1, 4}, {4, -6, 4, -1}};
  if (t < 4) {
    double a[3][4] = {{1, 0, 0, 0}, {-1, 2, 0, 0}, {1, -3, 3, 0}};
    memcpy(d, a[t - 1], sizeof(double) * 4);
  } else {
    memcpy(d, tab[t % 4], sizeof(double) * 4);
  }
}

#define VADD(v, c, r, x) (v)[(c) * 3 + (r)] += (x) /* virt(row, col) stored [col][row] */

static int pot_ttm_bead(const orc_sys *s, scratch_t *w, const double *RR, const double *RRc,
                        int t, double *En, double *terms, double *dRR, double *virt,
                        double *virialc_out, double *dipI, double *dipE, double *chg,
                        double *hist) {
  const int N = s->N, ns = s->ns, fd = s->fd, ld = s->ld, fdI = s->fdI, ldI = s->ldI;
  const double *box = s->p.box, *boxi = s->boxi;
  double virialc = 0.0, vir_rec[3][3];
  double Umon = 0, Uvdw = 0, Uelec = 0, Uind = 0;
  memset(virt, 0, 9 * sizeof(double));
  memset(w->phi, 0, ns * sizeof(double));
  memset(w->Efq, 0, ns * sizeof *w->Efq);
  memset(w->dip, 0, ns * sizeof *w->dip);
  memset(w->olddip, 0, ns * sizeof *w->olddip);
  memset(w->dR, 0, ns * sizeof *w->dR);
  const double gM = s->gammaM;
  /* OHH -> site-major and M-sites, pot_ttm.f90:56-80 */
  for (int iw = 0; iw < N; ++iw) {
    int iO = iw, iH1 = iw + N, iH2 = iw + 2 * N, iM = iw + 3 * N;
    for (int c = 0; c < 3; ++c) {
      w->R[iO][c] = RR[9 * iw + c];
      w->R[iH1][c] = RR[9 * iw + 3 + c];
      w->R[iH2][c] = RR[9 * iw + 6 + c];
      w->Rc[iO][c] = RRc[9 * iw + c];
      w->Rc[iH1][c] = RRc[9 * iw + 3 + c];
      w->Rc[iH2][c] = RRc[9 * iw + 6 + c];
    }
    double roh1[3], roh2[3];
    for (int c = 0; c < 3; ++c) {
      roh1[c] = w->R[iH1][c] - w->R[iO][c];
      roh1[c] = roh1[c] - box[c] * anint(roh1[c] * boxi[c]);
      roh2[c] = w->R[iH2][c] - w->R[iO][c];
      roh2[c] = roh2[c] - box[c] * anint(roh2[c] * boxi[c]);
    }
    for (int c = 0; c < 3; ++c) w->R[iM][c] = 0.5 * gM * (roh1[c] + roh2[c]) + w->R[iO][c];
    for (int c = 0; c < 3; ++c) {
      roh1[c] = w->Rc[iH1][c] - w->Rc[iO][c];
      roh1[c] = roh1[c] - box[c] * anint(roh1[c] * boxi[c]);
      roh2[c] = w->Rc[iH2][c] - w->Rc[iO][c];
      roh2[c] = roh2[c] - box[c] * anint(roh2[c] * boxi[c]);
    }
    for (int c = 0; c < 3; ++c) w->Rc[iM][c] = 0.5 * gM * (roh1[c] + roh2[c]) + w->Rc[iO][c];
  }
  /* monomer energy, DMS charges, intramolecular virial: pot_ttm.f90:84-151 */
  double tmpq = 0.5 * gM / (1.0 - gM);
  memset(w->charge, 0, ns * sizeof(double));
  for (int iw = 0; iw < N; ++iw) {
    int iO = iw, iH1 = iw + N, iH2 = iw + 2 * N, iM = iw + 3 * N;
    double r1[3][3], dr1[3][3], e1, q3[3], dq3[3][3][3];
    for (int c = 0; c < 3; ++c) {
      r1[0][c] = w->R[iO][c]; r1[1][c] = w->R[iH1][c]; r1[2][c] = w->R[iH2][c];
    }
    pot_nasa((const double(*)[3])r1, dr1, &e1, box, boxi);
    Umon = Umon + e1;
    for (int c = 0; c < 3; ++c) {
      w->dR[iO][c] = dr1[0][c]; w->dR[iH1][c] = dr1[1][c]; w->dR[iH2][c] = dr1[2][c];
    }
    dms_nasa((const double(*)[3])r1, q3, dq3, box, boxi, s->dms1, s->dms2, s->dms3);
    for (int a = 0; a < 3; ++a) q3[a] = q3[a] * CHARGECON;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) dq3[a][b][c] = dq3[a][b][c] * CHARGECON;
    for (int a = 0; a < 3; ++a) chg[3 * iw + a] = q3[a];
    double roh1[3], roh2[3];
    for (int c = 0; c < 3; ++c) {
      roh1[c] = r1[1][c] - r1[0][c];
      roh1[c] = roh1[c] - box[c] * anint(roh1[c] * boxi[c]);
      roh2[c] = r1[2][c] - r1[0][c];
      roh2[c] = roh2[c] - box[c] * anint(roh2[c] * boxi[c]);
    }
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row < 3; ++row)
        VADD(virt, col, row, roh1[row] * dr1[1][col] + roh2[row] * dr1[2][col]);
    double rc1_[3], rc2_[3];
    for (int c = 0; c < 3; ++c) {
      rc1_[c] = w->Rc[iH1][c] - w->Rc[iO][c];
      rc1_[c] = rc1_[c] - box[c] * anint(rc1_[c] * boxi[c]);
      rc2_[c] = w->Rc[iH2][c] - w->Rc[iO][c];
      rc2_[c] = rc2_[c] - box[c] * anint(rc2_[c] * boxi[c]);
    }
    virialc = virialc + dot3(rc1_, dr1[1]);
    virialc = virialc + dot3(rc2_, dr1[2]);
    w->charge[iO] = 0.0;
    w->charge[iH1] = q3[1] + tmpq * (q3[1] + q3[2]);
    w->charge[iH2] = q3[2] + tmpq * (q3[1] + q3[2]);
    w->charge[iM] = q3[0] / (1.0 - gM);
    double *g = &w->grdq[(size_t)iw * 27]; /* g[disp][chg][xyz] */
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        g[a * 9 + 0 * 3 + c] = dq3[a][0][c] + tmpq * (dq3[a][0][c] + dq3[a][1][c]);
        g[a * 9 + 1 * 3 + c] = dq3[a][1][c] + tmpq * (dq3[a][0][c] + dq3[a][1][c]);
        g[a * 9 + 2 * 3 + c] = dq3[a][2][c] - 2.0 * tmpq * (dq3[a][0][c] + dq3[a][1][c]);
      }
  }
  /* real-space vdW + charge-charge: pot_ttm.f90:159-228 */
  for (int iw = 0; iw < N; ++iw) {
    int iO = iw;
    find_neigh(s, iO, w->R, &w->nb);
    for (int jn = 0; jn < w->nb.n; ++jn) {
      int jO = w->nb.j[jn];
      double Rij[3] = {w->nb.Rij[3 * jn], w->nb.Rij[3 * jn + 1], w->nb.Rij[3 * jn + 2]};
      double dRsq = w->nb.R2[jn];
      double dRij = sqrt(dRsq);
      double R2i = 1.0 / dRsq, R4i = R2i * R2i, R6i = R2i * R2i * R2i, R8i = R6i * R2i;
      double tmp = s->vdwD * exp(-s->vdwE * dRij);
      double uij = R6i * ((s->vdwA * R6i + s->vdwB * R4i) + s->vdwC) + tmp;
      double x = dRij - s->p.rc1;
      double fsc = -(R8i * 6.0 * s->vdwC + tmp * s->vdwE / dRij + s->shiftA * (x * x) +
                     s->shiftB * (x * x * x));
      double dij[3] = {fsc * Rij[0], fsc * Rij[1], fsc * Rij[2]};
      Uvdw = Uvdw + uij;
      for (int c = 0; c < 3; ++c) {
        w->dR[iO][c] = w->dR[iO][c] + dij[c];
        w->dR[jO][c] = w->dR[jO][c] - dij[c];
      }
      double Rcij[3];
      for (int c = 0; c < 3; ++c) Rcij[c] = w->Rc[iO][c] - w->Rc[jO][c];
      minimg(s, Rcij);
      virialc = virialc + dot3(Rcij, dij);
      for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) VADD(virt, col, row, dij[col] * Rij[row]);
      double Rij0[3];
      for (int c = 0; c < 3; ++c) Rij0[c] = Rij[c] - w->R[iO][c] + w->R[jO][c];
      for (int isp = 0; isp < 4; ++isp) {
        int iat = iw + isp * N;
        double Ri[3];
        for (int c = 0; c < 3; ++c) Ri[c] = w->R[iat][c] + Rij0[c];
        double qi = w->charge[iat];
        double polfacI = s->aCC * s->polfac[isp];
        int jsp0 = isp == 0 ? 1 : 0;
        for (int jsp = jsp0; jsp < 4; ++jsp) {
          int jat = jO + jsp * N;
          double qj = w->charge[jat], r[3], ts0, ts1;
          for (int c = 0; c < 3; ++c) r[c] = Ri[c] - w->R[jat][c];
          minimg(s, r);
          double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          smear01(s, d2, polfacI * s->polfac[jsp], &ts0, &ts1);
          w->phi[iat] = w->phi[iat] + ts0 * qj;
          w->phi[jat] = w->phi[jat] + ts0 * qi;
          for (int c = 0; c < 3; ++c) {
            w->Efq[iat][c] = w->Efq[iat][c] + ts1 * qj * r[c];
            w->Efq[jat][c] = w->Efq[jat][c] - ts1 * qi * r[c];
          }
          double d[3];
          for (int c = 0; c < 3; ++c) d[c] = -(ts1 * qi * qj * r[c]);
          double rc[3];
          for (int c = 0; c < 3; ++c) rc[c] = w->Rc[iat][c] - w->Rc[jat][c] + Rij0[c];
          virialc = virialc + dot3(rc, d);
          for (int col = 0; col < 3; ++col)
            for (int row = 0; row < 3; ++row) VADD(virt, col, row, d[col] * r[row]);
        }
      }
    }
  }
  /* Ewald self energy and intramolecular exclusions: pot_ttm.f90:232-261 */
  for (int iat = N; iat < ns; ++iat) w->phi[iat] = w->phi[iat] - s->aewald * TSP_ * w->charge[iat];
  double ts1_stale = 0.0;
  for (int iw = 0; iw < N; ++iw)
    for (int isp = 0; isp < 3; ++isp) {
      int iat = iw + isp * N;
      for (int jsp = isp + 1; jsp < 4; ++jsp) {
        int jat = iw + jsp * N;
        double r[3], ts0, ts1;
        for (int c = 0; c < 3; ++c) r[c] = w->R[iat][c] - w->R[jat][c];
        minimg(s, r);
        double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        self1(s, d2, &ts0, &ts1);
        ts1_stale = ts1;
        w->phi[iat] = w->phi[iat] + ts0 * w->charge[jat];
        w->phi[jat] = w->phi[jat] + ts0 * w->charge[iat];
        for (int c = 0; c < 3; ++c) {
          w->Efq[iat][c] = w->Efq[iat][c] + ts1 * w->charge[jat] * r[c];
          w->Efq[jat][c] = w->Efq[jat][c] - ts1 * w->charge[iat] * r[c];
        }
        double d[3], rc[3];
        for (int c = 0; c < 3; ++c) d[c] = -(ts1 * w->charge[iat] * w->charge[jat] * r[c]);
        for (int c = 0; c < 3; ++c) rc[c] = w->Rc[iat][c] - w->Rc[jat][c];
        minimg(s, rc);
        virialc = virialc + dot3(rc, d);
        for (int col = 0; col < 3; ++col)
          for (int row = 0; row < 3; ++row) VADD(virt, col, row, d[col] * r[row]);
      }
    }
  /* reciprocal space, charges: pot_ttm.f90:269-277 */
  ewald_qq(s, w);
  for (int i = N; i < ns; ++i) Uelec += w->charge[i] * w->phi[i];
  Uelec = 0.5 * Uelec;
  for (int i = N; i < ns; ++i)
    for (int c = 0; c < 3; ++c) w->dR[i][c] = w->dR[i][c] - w->charge[i] * w->Efq[i][c];
  /* initial dipoles: pot_ttm.f90:281-297 */
  int ndip = ld - fd + 1;
  if (t > 4 && s->p.guess_initdip) {
    double dp[4];
    calc_dpcf(t, dp);
    for (int i = fd; i <= ld; ++i)
      for (int c = 0; c < 3; ++c) {
        double v = 0;
        for (int m = 0; m < 4; ++m) {
          double h = hist[((size_t)m * ndip + (i - fd)) * 3 + c];
          v = m == 0 ? dp[0] * h : v + dp[m] * h;
        }
        w->dip[i][c] = v;
      }
  } else if (s->p.pot_model == 2) {
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < 3; ++c) w->dip[i][c] = s->polarO * w->Efq[i][c];
    for (int i = N; i < 3 * N; ++i)
      for (int c = 0; c < 3; ++c) w->dip[i][c] = s->polarH * w->Efq[i][c];
  } else {
    for (int i = 3 * N; i < 4 * N; ++i)
      for (int c = 0; c < 3; ++c) w->dip[i][c] = s->polarM * w->Efq[i][c];
  }
  /* induced-dipole SOR iterations: pot_ttm.f90:306-392 */
  double stath = DEBYE / CHARGECON / sqrt((double)s->natoms);
  double factor = 4.0 * (s->aewald * s->aewald * s->aewald) / (3.0 * sqrt(PI_));
  int iter;
  for (iter = 1; iter <= s->p.polar_maxiter; ++iter) {
    memset(w->Efd, 0, ns * sizeof *w->Efd);
    for (int i = 0; i < N; ++i) {
      int iO = i;
      find_neigh(s, iO, w->R, &w->nb);
      for (int jn = 0; jn < w->nb.n; ++jn) {
        int jO = w->nb.j[jn];
        double Rij0[3];
        for (int c = 0; c < 3; ++c) {
          Rij0[c] = w->nb.Rij[3 * jn + c] - w->R[iO][c] + w->R[jO][c];
          Rij0[c] = Rij0[c] - box[c] * anint(Rij0[c] * boxi[c]);
        }
        for (int isp = fdI; isp <= ldI; ++isp) {
          int iat = iO + isp * N;
          double polfacI = s->aDD * s->polfac[isp];
          double di[3] = {w->dip[iat][0], w->dip[iat][1], w->dip[iat][2]}, Ri[3];
          for (int c = 0; c < 3; ++c) Ri[c] = w->R[iat][c] + Rij0[c];
          for (int jsp = fdI; jsp <= ldI; ++jsp) {
            int jat = jO + jsp * N;
            double dj[3] = {w->dip[jat][0], w->dip[jat][1], w->dip[jat][2]}, r[3], t1, t2, dd[3][3];
            for (int c = 0; c < 3; ++c) r[c] = Ri[c] - w->R[jat][c];
            minimg(s, r);
            double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            smear2(s, d2, polfacI * s->polfac[jsp], &t1, &t2);
            get_dd3(r, t1, t2, dd);
            for (int c = 0; c < 3; ++c)
              w->Efd[iat][c] = w->Efd[iat][c] + dd[0][c] * dj[0] + dd[1][c] * dj[1] + dd[2][c] * dj[2];
            for (int c = 0; c < 3; ++c)
              w->Efd[jat][c] = w->Efd[jat][c] + dd[0][c] * di[0] + dd[1][c] * di[1] + dd[2][c] * di[2];
          }
        }
      }
    }
    for (int iw = 0; iw < N; ++iw)
      for (int isp = fdI; isp <= ldI - 1; ++isp) {
        int iat = iw + isp * N;
        double di[3] = {w->dip[iat][0], w->dip[iat][1], w->dip[iat][2]};
        double polfacI = s->aDD * s->polfac[isp];
        for (int jsp = isp + 1; jsp <= ldI; ++jsp) {
          int jat = iw + jsp * N;
          double dj[3] = {w->dip[jat][0], w->dip[jat][1], w->dip[jat][2]}, r[3], t1, t2, dd[3][3];
          for (int c = 0; c < 3; ++c) r[c] = w->R[iat][c] - w->R[jat][c];
          minimg(s, r);
          double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          smear2(s, d2, polfacI * s->polfac[jsp], &t1, &t2);
          get_dd3(r, t1, t2, dd);
          for (int c = 0; c < 3; ++c)
            w->Efd[iat][c] = w->Efd[iat][c] + dd[0][c] * dj[0] + dd[1][c] * dj[1] + dd[2][c] * dj[2];
          for (int c = 0; c < 3; ++c)
            w->Efd[jat][c] = w->Efd[jat][c] + dd[0][c] * di[0] + dd[1][c] * di[1] + dd[2][c] * di[2];
        }
      }
    ewald_dd(s, w);
    for (int i = fd; i <= ld; ++i)
      for (int c = 0; c < 3; ++c) w->Efd[i][c] = w->Efd[i][c] + w->Efq[i][c] + factor * w->dip[i][c];
    double om = s->p.polar_sor;
    if (s->p.pot_model == 2) {
      for (int i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c)
          w->dip[i][c] = (1.0 - om) * w->dip[i][c] + om * s->polarO * w->Efd[i][c];
      for (int i = N; i < 3 * N; ++i)
        for (int c = 0; c < 3; ++c)
          w->dip[i][c] = (1.0 - om) * w->dip[i][c] + om * s->polarH * w->Efd[i][c];
    } else {
      for (int i = 3 * N; i < 4 * N; ++i)
        for (int c = 0; c < 3; ++c)
          w->dip[i][c] = (1.0 - om) * w->dip[i][c] + om * s->polarM * w->Efd[i][c];
    }
    double ss = 0;
    for (int i = fd; i <= ld; ++i)
      for (int c = 0; c < 3; ++c) {
        double d = w->dip[i][c] - w->olddip[i][c];
        ss += d * d;
      }
    double deltadip = stath * sqrt(ss);
    if (deltadip < s->p.polar_eps) break;
    memcpy(w->olddip, w->dip, ns * sizeof *w->dip);
  }
  for (int i = fd; i <= ld; ++i)
    for (int c = 0; c < 3; ++c) Uind += w->dip[i][c] * w->Efq[i][c];
  Uind = -0.5 * Uind;
  if (s->p.guess_initdip && hist) {
    int slot = t % 4;
    for (int i = fd; i <= ld; ++i)
      for (int c = 0; c < 3; ++c) hist[((size_t)slot * ndip + (i - fd)) * 3 + c] = w->dip[i][c];
  }
  /* dipole forces, real space: pot_ttm.f90:399-478 */
  double ts1 = ts1_stale, ts2 = 0.0;
  for (int iw = 0; iw < N; ++iw) {
    int iO = iw;
    find_neigh(s, iO, w->R, &w->nb);
    for (int jn = 0; jn < w->nb.n; ++jn) {
      int jO = w->nb.j[jn];
      double Rij0[3];
      for (int c = 0; c < 3; ++c) {
        Rij0[c] = w->nb.Rij[3 * jn + c] - w->R[iO][c] + w->R[jO][c];
        Rij0[c] = Rij0[c] - box[c] * anint(Rij0[c] * boxi[c]);
      }
      for (int isp = 0; isp < 4; ++isp) {
        int iat = iO + isp * N;
        double Ri[3], di[3] = {0, 0, 0};
        for (int c = 0; c < 3; ++c) Ri[c] = w->R[iat][c] + Rij0[c];
        double qi = w->charge[iat];
        int idip = isp >= fdI && isp <= ldI;
        if (idip)
          for (int c = 0; c < 3; ++c) di[c] = w->dip[iat][c];
        double polfacI = s->polfac[isp];
        for (int jsp = 0; jsp < 4; ++jsp) {
          int jat = jO + jsp * N;
          double r[3], dj[3] = {0, 0, 0}, dij[3] = {0, 0, 0}, qdqd[3], dd[3][3];
          for (int c = 0; c < 3; ++c) r[c] = Ri[c] - w->R[jat][c];
          minimg(s, r);
          double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          double qj = w->charge[jat];
          int jdip = jsp >= fdI && jsp <= ldI;
          if (jdip)
            for (int c = 0; c < 3; ++c) dj[c] = w->dip[jat][c];
          for (int c = 0; c < 3; ++c) qdqd[c] = qj * di[c] - qi * dj[c];
          double qsq = qdqd[0] * qdqd[0] + qdqd[1] * qdqd[1] + qdqd[2] * qdqd[2];
          if (idip && jdip) {
            double t1d, t2d, t3d, T[3][3][3], a33[3][3];
            smear3(s, d2, s->aDD * polfacI * s->polfac[jsp], &t1d, &t2d, &t3d);
            for (int ii = 0; ii < 3; ++ii)
              for (int jj = 0; jj < 3; ++jj)
                for (int kk = 0; kk < 3; ++kk) T[ii][jj][kk] = 15.0 * r[ii] * r[jj] * r[kk] * t3d;
            for (int jj = 0; jj < 3; ++jj) {
              for (int a = 0; a < 3; ++a) T[a][jj][jj] = T[a][jj][jj] - 3.0 * t2d * r[a];
              for (int a = 0; a < 3; ++a) T[jj][a][jj] = T[jj][a][jj] - 3.0 * t2d * r[a];
              for (int a = 0; a < 3; ++a) T[jj][jj][a] = T[jj][jj][a] - 3.0 * t2d * r[a];
            }
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b)
                a33[a][b] = T[a][b][0] * dj[0] + T[a][b][1] * dj[1] + T[a][b][2] * dj[2];
            for (int c = 0; c < 3; ++c)
              dij[c] = dij[c] + di[0] * a33[0][c] + di[1] * a33[1][c] + di[2] * a33[2][c];
            if (qsq > 1e-9) {
              if (fabs(s->aCC - s->aDD) < 1e-12) {
                ts1 = t1d;
                ts2 = t2d;
              } else {
                smear2(s, d2, s->aCC * polfacI * s->polfac[jsp], &ts1, &ts2);
              }
              get_dd3(r, ts1, ts2, dd);
              for (int c = 0; c < 3; ++c)
                dij[c] = dij[c] + dd[0][c] * qdqd[0] + dd[1][c] * qdqd[1] + dd[2][c] * qdqd[2];
            }
          } else if (qsq > 1e-9) {
            smear2(s, d2, s->aCC * polfacI * s->polfac[jsp], &ts1, &ts2);
            get_dd3(r, ts1, ts2, dd);
            for (int c = 0; c < 3; ++c)
              dij[c] = dij[c] + dd[0][c] * qdqd[0] + dd[1][c] * qdqd[1] + dd[2][c] * qdqd[2];
          }
          w->phi[iat] = w->phi[iat] + ts1 * (r[0] * dj[0] + r[1] * dj[1] + r[2] * dj[2]);
          w->phi[jat] = w->phi[jat] - ts1 * (r[0] * di[0] + r[1] * di[1] + r[2] * di[2]);
          for (int c = 0; c < 3; ++c) {
            w->dR[iat][c] = w->dR[iat][c] + dij[c];
            w->dR[jat][c] = w->dR[jat][c] - dij[c];
          }
          double rc[3];
          for (int c = 0; c < 3; ++c) rc[c] = w->Rc[iat][c] - w->Rc[jat][c] + Rij0[c];
          virialc = virialc + dot3(rc, dij);
          for (int col = 0; col < 3; ++col)
            for (int row = 0; row < 3; ++row) VADD(virt, col, row, dij[row] * r[col]);
        }
      }
    }
  }
  /* intramolecular dipole terms: pot_ttm.f90:482-542 */
  for (int iw = 0; iw < N; ++iw)
    for (int isp = 0; isp < 3; ++isp) {
      int iat = iw + isp * N;
      double qi = w->charge[iat], di[3] = {0, 0, 0};
      int idip = isp >= fdI && isp <= ldI;
      if (idip)
        for (int c = 0; c < 3; ++c) di[c] = w->dip[iat][c];
      for (int jsp = isp + 1; jsp < 4; ++jsp) {
        int jat = iw + jsp * N;
        double qj = w->charge[jat], dj[3] = {0, 0, 0};
        int jdip = jsp >= fdI && jsp <= ldI;
        if (jdip)
          for (int c = 0; c < 3; ++c) dj[c] = w->dip[jat][c];
        double r[3], rc[3], dij[3], dd[3][3], qdqd[3];
        for (int c = 0; c < 3; ++c) r[c] = w->R[iat][c] - w->R[jat][c];
        minimg(s, r);
        for (int c = 0; c < 3; ++c) rc[c] = w->Rc[iat][c] - w->Rc[jat][c];
        minimg(s, rc);
        double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        if (idip && jdip) {
          double t3, T[3][3][3], a33[3][3];
          smear3(s, d2, s->aDD * s->polfac[isp] * s->polfac[jsp], &ts1, &ts2, &t3);
          for (int ii = 0; ii < 3; ++ii)
            for (int jj = 0; jj < 3; ++jj)
              for (int kk = 0; kk < 3; ++kk) {
                double tm = 0.0;
                if (jj == kk) tm = tm + r[ii];
                if (ii == kk) tm = tm + r[jj];
                if (ii == jj) tm = tm + r[kk];
                T[ii][jj][kk] = 15.0 * r[ii] * r[jj] * r[kk] * t3 - 3.0 * tm * ts2;
              }
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              a33[a][b] = T[a][b][0] * dj[0] + T[a][b][1] * dj[1] + T[a][b][2] * dj[2];
          for (int c = 0; c < 3; ++c) dij[c] = di[0] * a33[0][c] + di[1] * a33[1][c] + di[2] * a33[2][c];
          for (int c = 0; c < 3; ++c) {
            w->dR[iat][c] = w->dR[iat][c] + dij[c];
            w->dR[jat][c] = w->dR[jat][c] - dij[c];
          }
          virialc = virialc + dot3(rc, dij);
          for (int col = 0; col < 3; ++col)
            for (int row = 0; row < 3; ++row) VADD(virt, col, row, dij[row] * r[col]);
        }
        for (int c = 0; c < 3; ++c) qdqd[c] = qj * di[c] - qi * dj[c];
        self2(s, d2, &ts1, &ts2);
        get_dd3(r, ts1, ts2, dd);
        for (int c = 0; c < 3; ++c) dij[c] = dd[0][c] * qdqd[0] + dd[1][c] * qdqd[1] + dd[2][c] * qdqd[2];
        w->phi[iat] = w->phi[iat] + ts1 * (r[0] * dj[0] + r[1] * dj[1] + r[2] * dj[2]);
        w->phi[jat] = w->phi[jat] - ts1 * (r[0] * di[0] + r[1] * di[1] + r[2] * di[2]);
        for (int c = 0; c < 3; ++c) {
          w->dR[iat][c] = w->dR[iat][c] + dij[c];
          w->dR[jat][c] = w->dR[jat][c] - dij[c];
        }
        virialc = virialc + dot3(rc, dij);
        for (int col = 0; col < 3; ++col)
          for (int row = 0; row < 3; ++row) VADD(virt, col, row, dij[row] * r[col]);
      }
    }
  /* reciprocal space, dipoles: pot_ttm.f90:551-553 */
  ewald_qd(s, w, vir_rec);
  for (int i = 0; i < 9; ++i) virt[i] = virt[i] + ((double *)vir_rec)[i];
  virialc = virialc + vir_rec[0][0] + vir_rec[1][1] + vir_rec[2][2];
  /* charge-derivative chain rule: pot_ttm.f90:556-584 */
  for (int iw = 0; iw < N; ++iw) {
    int iO = iw, iH1 = iw + N, iH2 = iw + 2 * N, iM = iw + 3 * N;
    const double *g = &w->grdq[(size_t)iw * 27];
    double dg[3][3]; /* dg[atom: O,H1,H2][xyz] */
    for (int c = 0; c < 3; ++c) {
      dg[1][c] = g[0 * 9 + 0 * 3 + c] * w->phi[iH1] + g[0 * 9 + 1 * 3 + c] * w->phi[iH2] +
                 g[0 * 9 + 2 * 3 + c] * w->phi[iM];
      dg[2][c] = g[1 * 9 + 0 * 3 + c] * w->phi[iH1] + g[1 * 9 + 1 * 3 + c] * w->phi[iH2] +
                 g[1 * 9 + 2 * 3 + c] * w->phi[iM];
      dg[0][c] = g[2 * 9 + 0 * 3 + c] * w->phi[iH1] + g[2 * 9 + 1 * 3 + c] * w->phi[iH2] +
                 g[2 * 9 + 2 * 3 + c] * w->phi[iM];
    }
    for (int c = 0; c < 3; ++c) {
      w->dR[iH1][c] = w->dR[iH1][c] + dg[1][c];
      w->dR[iH2][c] = w->dR[iH2][c] + dg[2][c];
      w->dR[iO][c] = w->dR[iO][c] + dg[0][c];
    }
    double r1[3], r2[3];
    for (int c = 0; c < 3; ++c) {
      r1[c] = w->R[iH1][c] - w->R[iO][c];
      r1[c] = r1[c] - box[c] * anint(r1[c] * boxi[c]);
      r2[c] = w->R[iH2][c] - w->R[iO][c];
      r2[c] = r2[c] - box[c] * anint(r2[c] * boxi[c]);
    }
    for (int ix = 0; ix < 3; ++ix)
      for (int iy = 0; iy < 3; ++iy) VADD(virt, iy, ix, r1[ix] * dg[1][iy] + r2[ix] * dg[2][iy]);
    for (int c = 0; c < 3; ++c) {
      r1[c] = w->Rc[iH1][c] - w->Rc[iO][c];
      r1[c] = r1[c] - box[c] * anint(r1[c] * boxi[c]);
      r2[c] = w->Rc[iH2][c] - w->Rc[iO][c];
      r2[c] = r2[c] - box[c] * anint(r2[c] * boxi[c]);
    }
    for (int c = 0; c < 3; ++c) virialc = virialc + r1[c] * dg[1][c] + r2[c] * dg[2][c];
  }
  /* long-range vdW correction: pot_ttm.f90:588-591 */
  virt[0] -= s->Uvdw_lrc;
  virt[4] -= s->Uvdw_lrc;
  virt[8] -= s->Uvdw_lrc;
  virialc = virialc - 3 * s->Uvdw_lrc;
  /* M-site force redistribution: pot_ttm.f90:597-602 */
  for (int iw = 0; iw < N; ++iw) {
    int iO = iw, iH1 = iw + N, iH2 = iw + 2 * N, iM = iw + 3 * N;
    for (int c = 0; c < 3; ++c) {
      dRR[9 * iw + 3 + c] = w->dR[iH1][c] + 0.5 * gM * w->dR[iM][c];
      dRR[9 * iw + 6 + c] = w->dR[iH2][c] + 0.5 * gM * w->dR[iM][c];
      dRR[9 * iw + c] = w->dR[iO][c] + (1.0 - gM) * w->dR[iM][c];
    }
  }
  /* molecular dipoles: pot_ttm.f90:605-627 */
  for (int iw = 0; iw < N; ++iw) {
    int iH1 = iw + N, iH2 = iw + 2 * N, iM = iw + 3 * N;
    double a[3], b[3];
    for (int c = 0; c < 3; ++c) {
      a[c] = w->R[iH1][c] - w->R[iM][c];
      a[c] = a[c] - box[c] * anint(a[c] * boxi[c]);
      b[c] = w->R[iH2][c] - w->R[iM][c];
      b[c] = b[c] - box[c] * anint(b[c] * boxi[c]);
    }
    for (int c = 0; c < 3; ++c) dipI[3 * iw + c] = w->charge[iH1] * a[c] + w->charge[iH2] * b[c];
    for (int isp = fdI; isp <= ldI; ++isp) {
      int iat = iw + isp * N;
      for (int c = 0; c < 3; ++c) {
        dipI[3 * iw + c] = w->dip[iat][c] + dipI[3 * iw + c];
        dipE[3 * iw + c] = w->dip[iat][c];
      }
    }
  }
  *En = Umon + Uvdw + Uelec + Uind + s->Uvdw_lrc;
  *virialc_out = virialc;
  terms[0] = Umon; terms[1] = Uvdw; terms[2] = Uelec; terms[3] = Uind; terms[4] = s->Uvdw_lrc;
  return iter;
}

/* ---------------- one bead: pot_spc (pot_ttm/pot_spc.f90:1-277) ---------------- */

/* Minimum image with a division, as pot_spc.f90 writes it for intramolecular
 * vectors (r - box*anint(r/box)). */
static void minimg_div(const orc_sys *s, double *r) {
  for (int c = 0; c < 3; ++c) r[c] = r[c] - s->p.box[c] * anint(r[c] / s->p.box[c]);
}

/* ewald_pw_std_qq, ewald.f90:233-289: reciprocal-space potential and field of
 * the charges of all sites, and the reciprocal virial vir[col][row]. */
static void ewald_pw_qq(const orc_sys *s, scratch_t *w, double *vir) {
  const int ns = s->ns;
  memset(vir, 0, 9 * sizeof(double));
  ssincos(s, w);
  int lastkx = -1000, lastky = -1000;
  for (int k = 0; k < s->nk; ++k) {
    int kx = s->hkl[3 * k], ky = s->hkl[3 * k + 1], kz = s->hkl[3 * k + 2];
    const double *kv = &s->kvec[3 * k];
    double p4vAkk = FOURPI_ * s->Ak[k] / s->vol;
    if (!(kx == lastkx && ky == lastky)) {
      for (int i = 0; i < ns; ++i) {
        double tmp = SX(i, kx) * CY(i, ky) + CX(i, kx) * SY(i, ky);
        w->cxy[i] = CX(i, kx) * CY(i, ky) - SX(i, kx) * SY(i, ky);
        w->sxy[i] = tmp;
      }
    }
    lastkx = kx;
    lastky = ky;
    for (int i = 0; i < ns; ++i) {
      double tmp = w->sxy[i] * CZ(i, kz) + w->cxy[i] * SZ(i, kz);
      w->ckr[i] = w->cxy[i] * CZ(i, kz) - w->sxy[i] * SZ(i, kz);
      w->skr[i] = tmp;
    }
    double rQ = 0.0, iQ = 0.0;
    for (int i = 0; i < ns; ++i) rQ += w->charge[i] * w->ckr[i];
    for (int i = 0; i < ns; ++i) iQ += w->charge[i] * w->skr[i];
    for (int i = 0; i < ns; ++i) {
      double tmp = p4vAkk * (w->ckr[i] * iQ - w->skr[i] * rQ);
      for (int c = 0; c < 3; ++c) w->Efq[i][c] = w->Efq[i][c] - tmp * kv[c];
    }
    for (int i = 0; i < ns; ++i) w->phi[i] = w->phi[i] + p4vAkk * (w->ckr[i] * rQ + w->skr[i] * iQ);
    double ksq = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    double coeff2 = 2.0 / ksq + 0.5 / (s->aewald * s->aewald);
    double U = -0.5 * p4vAkk * (rQ * rQ + iQ * iQ);
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r)
        vir[c * 3 + r] = vir[c * 3 + r] + (r == c ? U : 0.0) - U * coeff2 * kv[r] * kv[c];
  }
}

static void pot_spc_bead(const orc_sys *s, scratch_t *w, const double *RR, double *En,
                         double *terms, double *dRR, double *virt, double *dipI, double *chg) {
  const int N = s->N, ns = s->ns;
  const double *box = s->p.box, *boxi = s->boxi;
  double Umon = 0.0, Uvdw = 0.0, Uelec = 0.0;
  memset(virt, 0, 9 * sizeof(double));
  memset(w->phi, 0, ns * sizeof(double));
  memset(w->Efq, 0, ns * sizeof *w->Efq);
  memset(w->dR, 0, ns * sizeof *w->dR);
  /* OHH -> type-major sites (pot_spc.f90:55-60); charges of init_pot
   * (potential.f90:186-192) */
  for (int iw = 0; iw < N; ++iw) {
    int iO = iw, iH1 = iw + N, iH2 = iw + 2 * N;
    for (int c = 0; c < 3; ++c) {
      w->R[iO][c] = RR[9 * iw + c];
      w->R[iH1][c] = RR[9 * iw + 3 + c];
      w->R[iH2][c] = RR[9 * iw + 6 + c];
    }
    w->charge[iO] = s->qO * CHARGECON;
    w->charge[iH1] = s->qH * CHARGECON;
    w->charge[iH2] = s->qH * CHARGECON;
    chg[3 * iw] = w->charge[iO];
    chg[3 * iw + 1] = w->charge[iH1];
    chg[3 * iw + 2] = w->charge[iH2];
  }
  if (s->p.pot_model == 4) {
    /* harmonic bonds and angle (pot_spc.f90:69-106) */
    for (int iw = 0; iw < N; ++iw) {
      int iO = iw, iH1 = iw + N, iH2 = iw + 2 * N;
      double roh1[3], roh2[3], AxB[3], A1[3], A2[3], dh1[3], dh2[3];
      for (int c = 0; c < 3; ++c) {
        roh1[c] = w->R[iH1][c] - w->R[iO][c];
        roh2[c] = w->R[iH2][c] - w->R[iO][c];
      }
      minimg_div(s, roh1);
      minimg_div(s, roh2);
      double d1 = sqrt(dot3(roh1, roh1)), d2 = sqrt(dot3(roh2, roh2));
      AxB[0] = roh1[1] * roh2[2] - roh1[2] * roh2[1];
      AxB[1] = -roh1[0] * roh2[2] + roh1[2] * roh2[0];
      AxB[2] = roh1[0] * roh2[1] - roh1[1] * roh2[0];
      double dAxB = sqrt(dot3(AxB, AxB));
      double costh = dot3(roh1, roh2) / (d1 * d2);
      double sinth = dAxB / (d1 * d2);
      double ang = atan2(sinth, costh);
      double e1 = s->harmkb * ((d1 - s->roh0) * (d1 - s->roh0) + (d2 - s->roh0) * (d2 - s->roh0)) +
                  s->harmka * ((ang - s->th0) * (ang - s->th0));
      double tmp1 = 2.0 * s->harmkb * (d1 - s->roh0) / d1;
      double tmp2 = 2.0 * s->harmkb * (d2 - s->roh0) / d2;
      double tmp = 2.0 * s->harmka * (ang - s->th0);
      for (int c = 0; c < 3; ++c) {
        A1[c] = -(roh2[c] / (d1 * d2) - costh * roh1[c] / (d1 * d1)) / sinth;
        A2[c] = -(roh1[c] / (d1 * d2) - costh * roh2[c] / (d2 * d2)) / sinth;
      }
      for (int c = 0; c < 3; ++c) {
        dh1[c] = tmp1 * roh1[c] + tmp * A1[c];
        dh2[c] = tmp2 * roh2[c] + tmp * A2[c];
      }
      for (int c = 0; c < 3; ++c) {
        w->dR[iH1][c] = dh1[c];
        w->dR[iH2][c] = dh2[c];
        w->dR[iO][c] = -dh1[c] - dh2[c];
      }
      Umon = Umon + e1;
      for (int iy = 0; iy < 3; ++iy)
        for (int ix = 0; ix < 3; ++ix)
          virt[iy * 3 + ix] = virt[iy * 3 + ix] + roh1[ix] * dh1[iy] + roh2[ix] * dh2[iy];
    }
  } else {
    /* SPC/F bond, bond and H-H terms (pot_spc.f90:107-137) */
    double rhh0 = 2.0 * sin(0.5 * s->th0) * s->roh0;
    for (int iw = 0; iw < N; ++iw) {
      int iO = iw, iH1 = iw + N, iH2 = iw + 2 * N;
      double roh1[3], roh2[3], rhh[3], dr1[3][3];
      for (int c = 0; c < 3; ++c) {
        roh1[c] = w->R[iO][c] - w->R[iH1][c];
        roh2[c] = w->R[iO][c] - w->R[iH2][c];
        rhh[c] = w->R[iH1][c] - w->R[iH2][c];
      }
      minimg_div(s, roh1);
      minimg_div(s, roh2);
      minimg_div(s, rhh);
      double d1 = sqrt(roh1[0] * roh1[0] + roh1[1] * roh1[1] + roh1[2] * roh1[2]);
      double d2 = sqrt(roh2[0] * roh2[0] + roh2[1] * roh2[1] + roh2[2] * roh2[2]);
      double d3 = sqrt(rhh[0] * rhh[0] + rhh[1] * rhh[1] + rhh[2] * rhh[2]);
      double dd1 = d1 - s->roh0, dd2 = d2 - s->roh0, dd3 = d3 - rhh0;
      double e1 = s->spcA * (dd1 * dd1 + dd2 * dd2) + s->spcB * dd3 * dd3 +
                  s->spcC * (dd1 + dd2) * dd3 + s->spcD * dd1 * dd2;
      double v1 = 2.0 * s->spcA * dd1 + s->spcC * dd3 + s->spcD * dd2;
      double v2 = 2.0 * s->spcA * dd2 + s->spcC * dd3 + s->spcD * dd1;
      double v3 = 2.0 * s->spcB * dd3 + s->spcC * (dd1 + dd2);
      for (int c = 0; c < 3; ++c) {
        dr1[0][c] = v1 * roh1[c] / d1 + v2 * roh2[c] / d2;
        dr1[1][c] = -v1 * roh1[c] / d1 + v3 * rhh[c] / d3;
        dr1[2][c] = -v2 * roh2[c] / d2 - v3 * rhh[c] / d3;
      }
      Umon = Umon + e1;
      for (int c = 0; c < 3; ++c) {
        w->dR[iO][c] = dr1[0][c];
        w->dR[iH1][c] = dr1[1][c];
        w->dR[iH2][c] = dr1[2][c];
      }
      for (int iy = 0; iy < 3; ++iy)
        for (int ix = 0; ix < 3; ++ix)
          virt[iy * 3 + ix] = virt[iy * 3 + ix] - roh1[ix] * dr1[1][iy] - roh2[ix] * dr1[2][iy];
    }
  }
  /* vdW and real-space charge pairs over O-O neighbours (pot_spc.f90:142-191) */
  for (int i = 0; i < N; ++i) {
    find_neigh(s, i, w->R, &w->nb);
    for (int n = 0; n < w->nb.n; ++n) {
      int jO = w->nb.j[n];
      double Rij[3], dij[3], rij0[3];
      memcpy(Rij, &w->nb.Rij[3 * n], sizeof Rij);
      double dRsq = w->nb.R2[n];
      double R2i = 1.0 / dRsq, R6i = R2i * R2i * R2i, R8i = R6i * R2i;
      double uij = R6i * (s->vdwA * R6i + s->vdwC);
      for (int c = 0; c < 3; ++c) dij[c] = -R8i * (12.0 * s->vdwA * R6i + 6.0 * s->vdwC) * Rij[c];
      Uvdw = Uvdw + uij;
      for (int c = 0; c < 3; ++c) {
        w->dR[i][c] = w->dR[i][c] + dij[c];
        w->dR[jO][c] = w->dR[jO][c] - dij[c];
      }
      for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) VADD(virt, col, row, dij[col] * Rij[row]);
      for (int c = 0; c < 3; ++c) rij0[c] = Rij[c] - w->R[i][c] + w->R[jO][c];
      for (int isp = 0; isp < 3; ++isp) {
        int iat = i + isp * N;
        double Ri[3], qi = w->charge[iat];
        for (int c = 0; c < 3; ++c) Ri[c] = w->R[iat][c] + rij0[c];
        for (int jsp = 0; jsp < 3; ++jsp) {
          int jat = jO + jsp * N;
          double qj = w->charge[jat], r[3];
          for (int c = 0; c < 3; ++c) r[c] = Ri[c] - w->R[jat][c];
          for (int c = 0; c < 3; ++c) r[c] = r[c] - box[c] * anint(r[c] * boxi[c]);
          double drsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          double dd = sqrt(drsq);
          double expon = exp(-(s->aewald * s->aewald * drsq));
          double ts0 = orc_erfc(s->aewald * dd) / dd;
          double ts1 = (ts0 + s->cts1 * expon) / drsq;
          w->phi[iat] = w->phi[iat] + ts0 * qj;
          w->phi[jat] = w->phi[jat] + ts0 * qi;
          Uelec = Uelec + ts0 * qi * qj;
          for (int c = 0; c < 3; ++c) {
            w->Efq[iat][c] = w->Efq[iat][c] + ts1 * qj * r[c];
            w->Efq[jat][c] = w->Efq[jat][c] - ts1 * qi * r[c];
            dij[c] = -ts1 * qi * qj * r[c];
          }
          for (int col = 0; col < 3; ++col)
            for (int row = 0; row < 3; ++row) VADD(virt, col, row, dij[col] * r[row]);
        }
      }
    }
  }
  /* Ewald self energy and the screened intramolecular pairs (pot_spc.f90:202-233) */
  for (int iat = 0; iat < ns; ++iat) w->phi[iat] = w->phi[iat] - s->aewald * TSP_ * w->charge[iat];
  for (int iw = 0; iw < N; ++iw)
    for (int isp = 0; isp < 2; ++isp) {
      int iat = iw + isp * N;
      for (int jsp = isp + 1; jsp < 3; ++jsp) {
        int jat = iw + jsp * N;
        double r[3], dij[3];
        for (int c = 0; c < 3; ++c) r[c] = w->R[iat][c] - w->R[jat][c];
        minimg_div(s, r);
        double drsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        double dd = sqrt(drsq);
        double expon = exp(-(s->aewald * s->aewald * drsq));
        double kr0 = orc_erfc(s->aewald * dd) / dd;
        double kr1 = (kr0 + s->cts1 * expon) / drsq;
        double ch0 = -1.0 / dd, ch1 = ch0 / drsq;
        double ts0 = ch0 + kr0, ts1 = ch1 + kr1;
        w->phi[iat] = w->phi[iat] + ts0 * w->charge[jat];
        w->phi[jat] = w->phi[jat] + ts0 * w->charge[iat];
        for (int c = 0; c < 3; ++c) {
          w->Efq[iat][c] = w->Efq[iat][c] + ts1 * w->charge[jat] * r[c];
          w->Efq[jat][c] = w->Efq[jat][c] - ts1 * w->charge[iat] * r[c];
          dij[c] = -ts1 * w->charge[iat] * w->charge[jat] * r[c];
        }
        for (int col = 0; col < 3; ++col)
          for (int row = 0; row < 3; ++row) VADD(virt, col, row, dij[col] * r[row]);
      }
    }
  /* reciprocal space (pot_spc.f90:239-246) */
  double vir_rec[9];
  ewald_pw_qq(s, w, vir_rec);
  for (int i = 0; i < 9; ++i) virt[i] = virt[i] + vir_rec[i];
  Uelec = 0.0;
  for (int iat = 0; iat < ns; ++iat) Uelec += w->charge[iat] * w->phi[iat];
  Uelec = 0.5 * Uelec;
  for (int iat = 0; iat < ns; ++iat)
    for (int c = 0; c < 3; ++c) w->dR[iat][c] = w->dR[iat][c] - w->charge[iat] * w->Efq[iat][c];
  virt[0] -= s->Uvdw_lrc;
  virt[4] -= s->Uvdw_lrc;
  virt[8] -= s->Uvdw_lrc;
  /* molecular dipoles (pot_spc.f90:257-268) and the gradient in OHH order */
  for (int iw = 0; iw < N; ++iw) {
    int iO = iw, iH1 = iw + N, iH2 = iw + 2 * N;
    double a[3], b[3];
    for (int c = 0; c < 3; ++c) {
      a[c] = w->R[iH1][c] - w->R[iO][c];
      b[c] = w->R[iH2][c] - w->R[iO][c];
    }
    minimg_div(s, a);
    minimg_div(s, b);
    for (int c = 0; c < 3; ++c) {
      dipI[3 * iw + c] = w->charge[iH1] * a[c] + w->charge[iH2] * b[c];
      dRR[9 * iw + c] = w->dR[iO][c];
      dRR[9 * iw + 3 + c] = w->dR[iH1][c];
      dRR[9 * iw + 6 + c] = w->dR[iH2][c];
    }
  }
  *En = Umon + Uvdw + Uelec + s->Uvdw_lrc;
  terms[0] = Umon; terms[1] = Uvdw; terms[2] = Uelec; terms[3] = 0.0; terms[4] = s->Uvdw_lrc;
}

int orc_forces(orc_sys *s, const double *RRt, const double *RRc, int t, double *En,
               double *terms, double *dRRt, double *virt, double *virialc, double *dipI,
               double *dipE, double *chg, int *iters, double *hist) {
  scratch_t w;
  scratch_alloc(s, &w);
  int na = s->natoms, N = s->N, ndip = s->ld - s->fd + 1;
  for (int b = 0; b < s->p.nbeads; ++b) {
    if (s->p.pot_model >= 4) {
      /* pot_spc sets neither virialc nor Edip_mom (pot_spc.f90:1-13): zero here */
      pot_spc_bead(s, &w, RRt + (size_t)b * na * 3, &En[b], &terms[5 * b],
                   dRRt + (size_t)b * na * 3, virt + 9 * b, dipI + (size_t)b * N * 3,
                   chg + (size_t)b * na);
      virialc[b] = 0.0;
      memset(dipE + (size_t)b * N * 3, 0, sizeof(double) * N * 3);
      iters[b] = 0;
      continue;
    }
    iters[b] = pot_ttm_bead(s, &w, RRt + (size_t)b * na * 3, RRc, t, &En[b], &terms[5 * b],
                            dRRt + (size_t)b * na * 3, virt + 9 * b, &virialc[b],
                            dipI + (size_t)b * N * 3, dipE + (size_t)b * N * 3,
                            chg + (size_t)b * na, hist ? hist + (size_t)b * 4 * ndip * 3 : NULL);
  }
  scratch_free(&w);
  return 0;
}

/* ---------------- ring polymer ---------------- */

/* InitNormalModes, NormalModes.f90:68-218; C[mode][bead] */
void orc_nm_init(int P, double omegan, double delta, double setnm, double *C, double *A11,
                 double *A12, double *A21, double *A22, double *om, double *msf) {
  int l = (P % 2 == 0) ? (P - 2) / 2 : (P - 1) / 2;
  for (int i = 1; i <= P; ++i) {
    C[0 * P + (i - 1)] = sqrt(1.0 / P);
    for (int k = 1; k <= l; ++k) {
      C[k * P + (i - 1)] = sqrt(2.0 / P) * cos((double)(k * i) * 2.0 * PI_ / P);
      C[(k + l) * P + (i - 1)] = sqrt(2.0 / P) * sin((double)(k * i) * 2.0 * PI_ / P);
    }
    if (P % 2 == 0) C[(2 * l + 1) * P + (i - 1)] = sqrt(1.0 / P) * ((i % 2) ? -1.0 : 1.0);
  }
  om[0] = 0.0; A11[0] = 1.0; A12[0] = delta; A21[0] = 0.0; A22[0] = 1.0;
  for (int k = 1; k <= l; ++k) {
    double w = 2 * omegan * sin(PI_ * (double)((float)k / (float)P));
    for (int m = 0; m < 2; ++m) {
      int j = m == 0 ? k : k + l;
      om[j] = w;
      A11[j] = cos(w * delta);
      A12[j] = (1 / w) * sin(w * delta);
      A21[j] = -w * sin(w * delta);
      A22[j] = cos(w * delta);
    }
  }
  if (P % 2 == 0) {
    int k = l + 1, j = 2 * l + 1;
    double w = 2 * omegan * sin(PI_ * (double)((float)k / (float)P));
    om[j] = w;
    A11[j] = cos(w * delta);
    A12[j] = (1 / w) * sin(w * delta);
    A21[j] = -w * sin(w * delta);
    A22[j] = cos(w * delta);
  }
  for (int i = 0; i < P; ++i) msf[i] = 1.0;
  if (setnm != 0.0) {
    om[0] = 0.0; A11[0] = 1.0; A12[0] = delta; A21[0] = 0.0; A22[0] = 1.0;
    double w = (2 * PI_) * setnm / 33.33333;
    for (int i = 0; i < P; ++i) msf[i] = om[i] == 0 ? 1.0 : (om[i] / w) * (om[i] / w);
    int last = (P % 2 == 0) ? 2 * l + 1 : 2 * l;
    for (int j = 1; j <= last; ++j) {
      om[j] = w;
      A11[j] = cos(w * delta);
      A12[j] = (1 / w) * sin(w * delta);
      A21[j] = -w * sin(w * delta);
      A22[j] = cos(w * delta);
    }
  }
}

/* real2NM / NM2real, NormalModes.f90:308-341, on a strided (3, P) section */
static void real2nm(int P, const double *C, const double *a, size_t stride, double *out) {
  for (int j = 0; j < P; ++j)
    for (int c = 0; c < 3; ++c) {
      double v = 0;
      for (int k = 0; k < P; ++k) v = v + C[j * P + k] * a[k * stride + c];
      out[j * 3 + c] = v;
    }
}
static void nm2real(int P, const double *C, const double *a, double *out, size_t stride) {
  for (int j = 0; j < P; ++j)
    for (int c = 0; c < 3; ++c) {
      double v = 0;
      for (int k = 0; k < P; ++k) v = v + C[k * P + j] * a[k * 3 + c];
      out[j * stride + c] = v;
    }
}

/* EvolveRing for every atom, NormalModes.f90:34-63 */
void orc_nm_evolve(int P, int natoms, const double *C, const double *A11, const double *A12,
                   const double *A21, const double *A22, const double *msf, const double *masses,
                   double *RRt, double *PPt) {
  double *rt = malloc(sizeof(double) * 3 * P), *pt = malloc(sizeof(double) * 3 * P);
  size_t stride = (size_t)natoms * 3;
  for (int a = 0; a < natoms; ++a) {
    double m = masses[a];
    real2nm(P, C, RRt + 3 * a, stride, rt);
    real2nm(P, C, PPt + 3 * a, stride, pt);
    for (int j = 0; j < P; ++j)
      for (int c = 0; c < 3; ++c) {
        double t2 = A11[j] * rt[3 * j + c] + A12[j] * pt[3 * j + c] / (m * msf[j]);
        pt[3 * j + c] = A21[j] * m * msf[j] * rt[3 * j + c] + A22[j] * pt[3 * j + c];
        rt[3 * j + c] = t2;
      }
    nm2real(P, C, rt, RRt + 3 * a, stride);
    nm2real(P, C, pt, PPt + 3 * a, stride);
  }
  free(rt);
  free(pt);
}

/* Init_Langevin_NM, Langevin.f90:68-98 (temp is the bead temperature P*T) */
void orc_langevin_init(int P, double delt2, int centroid, double tau_bead, double temp,
                       const double *om, double *c1, double *c2) {
  for (int i = 0; i < P; ++i) {
    double g = i == 0 ? (centroid ? 1 / tau_bead : 0.0) : 2 * om[i];
    c1[i] = exp(-delt2 * g);
    c2[i] = sqrt(1 - c1[i] * c1[i]);
    c2[i] = sqrt(KB_AMU * temp) * c2[i];
  }
}

/* Langevin_NM, Langevin.f90:30-64; noise[atom][mode][xyz] standard normals */
void orc_langevin(int P, int natoms, const double *C, const double *msf, const double *c1,
                  const double *c2, double massO, double massH, const double *noise,
                  double *PPt) {
  double *pt = malloc(sizeof(double) * 3 * P);
  size_t stride = (size_t)natoms * 3;
  for (int a = 0; a < natoms; ++a) {
    real2nm(P, C, PPt + 3 * a, stride, pt);
    for (int j = 0; j < P; ++j) {
      double sm = (a % 3 == 0) ? sqrt(massO * msf[j]) : sqrt(massH * msf[j]);
      for (int c = 0; c < 3; ++c)
        pt[3 * j + c] = c1[j] * pt[3 * j + c] + c2[j] * (sm * noise[((size_t)a * P + j) * 3 + c]);
    }
    nm2real(P, C, pt, PPt + 3 * a, stride);
  }
  free(pt);
}

/* Nose_Hoover, Nose_Hoover.f90:33-112 (chain length 2) */
void orc_nose_hoover(double *s, double *uk, double *vxi, double tau, double delt2, int L,
                     double temp) {
  double delt4 = delt2 / 2.0, delt8 = delt2 / 4.0;
  double kT = KB_AMU * temp;
  double M1 = L * kT * (tau * tau), M2 = M1;
  vxi[1] = vxi[1] + (M1 * (vxi[0] * vxi[0]) - kT) * delt4 / M2;
  vxi[0] = vxi[0] * exp(-vxi[1] * delt8);
  vxi[0] = vxi[0] + (2 * *uk - L * kT) * delt4 / M1;
  vxi[0] = vxi[0] * exp(-vxi[1] * delt8);
  *s = exp(-vxi[0] * delt2);
  *uk = *uk * *s * *s;
  vxi[0] = vxi[0] * exp(-vxi[1] * delt8);
  vxi[0] = vxi[0] + (2 * *uk - L * kT) * delt4 / M1;
  vxi[0] = vxi[0] * exp(-vxi[1] * delt8);
  vxi[1] = vxi[1] + (M1 * (vxi[0] * vxi[0]) - kT) * delt4 / M2;
}
