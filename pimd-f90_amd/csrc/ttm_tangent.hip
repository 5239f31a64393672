// Derivatives of the TTM energy and gradient with respect to the model's vdW,
// polarisation, Thole and charge-surface parameters, for fitting.  No
// counterpart in the reference, whose parameters are fixed in pot_mod.f90.
//
// Forward (tangent) mode: one workgroup evaluates pot_ttm (pot_ttm.f90:3-630)
// for one configuration with dual numbers along one parameter direction, so
// the value parts reproduce the energy and gradient of the force path and the
// derivative parts are their exact derivatives along that parameter, through
// the SOR dipole iterations as executed (the same iteration count, branch
// decisions on the values).  With ew and gw the loss weights of the energy
// and of the gradient of each configuration, the loss derivative along
// parameter k is sum_b ew_b dE_b/dtheta_k + gw_b . dG_b/dtheta_k.
//
// Parameters (theta, PIMD_TTM_NPARAM = 16): vdwA..vdwE, polarO, polarH,
// polarM, polfacO, polfacH, polfacM, aDD, aCCaCD, dms_param1..3
// (potential.f90:84-113, 152-160).  Geometry (sites, k-space phases) does not
// depend on them and stays plain FP64; charges, dipoles, fields, potentials
// and the screened interaction tensors carry the derivative.  The M-site
// fraction gammaM (pot_mod.f90:12,24) also moves the M sites (pot_ttm.f90:
// 74-79): its direction runs an instance whose site positions and phases are
// dual numbers too (GD = true), with dM/dgammaM = (r_OH1 + r_OH2) / 2.
//
// Work layout: grid (configuration, direction), one 256-lane workgroup each;
// lanes stride over sites, k-vectors or molecules per phase, each lane owns
// what it writes (every site-pair term is evaluated from both ends), so no
// atomics.  Per configuration a preparatory kernel writes the sites and the
// per-axis phase tables once for all directions.
#include <type_traits>

#include "dual.h"
#include "ttm_device.h"

namespace pimd {

template <>
struct RM<Dual> {
  static __device__ __forceinline__ Dual sqrt_(Dual x) { return dsqrt(x); }
  static __device__ __forceinline__ Dual rcp(Dual x) { return Dual(1.0) / x; }
  static __device__ __forceinline__ Dual div(Dual a, Dual b) { return a / b; }
  static __device__ __forceinline__ Dual exp_(Dual x) { return dexp(x); }
  static __device__ __forceinline__ Dual log_(Dual x) { return dlog(x); }
  static __device__ __forceinline__ Dual abs_(Dual x) { return dabs(x); }
};

// Dual numbers over duals: v is a value with its derivative along the
// adjoint's position direction, a the derivative of v with respect to one
// Thole factor, itself with its position derivative.  The adjoint runs the
// smearing functions on them for d ts / d a along the direction.
struct HD {
  Dual v, a;
  HD() = default;
  __device__ constexpr HD(double x) : v(x), a(0.0) {}
  __device__ constexpr HD(Dual x, Dual y) : v(x), a(y) {}
};
__device__ __forceinline__ HD operator+(HD x, HD y) { return HD(x.v + y.v, x.a + y.a); }
__device__ __forceinline__ HD operator-(HD x, HD y) { return HD(x.v - y.v, x.a - y.a); }
__device__ __forceinline__ HD operator-(HD x) { return HD(-x.v, -x.a); }
__device__ __forceinline__ HD operator*(HD x, HD y) { return HD(x.v * y.v, x.a * y.v + x.v * y.a); }
__device__ __forceinline__ HD operator/(HD x, HD y) {
  const Dual q = x.v / y.v;
  return HD(q, (x.a - q * y.a) / y.v);
}
__device__ __forceinline__ bool operator<(HD x, HD y) { return x.v.v < y.v.v; }
__device__ __forceinline__ bool operator<=(HD x, HD y) { return x.v.v <= y.v.v; }
template <>
struct RM<HD> {
  static __device__ __forceinline__ HD sqrt_(HD x) {
    const Dual r = dsqrt(x.v);
    return HD(r, 0.5 * x.a / r);
  }
  static __device__ __forceinline__ HD rcp(HD x) { return HD(1.0) / x; }
  static __device__ __forceinline__ HD div(HD x, HD y) { return x / y; }
  static __device__ __forceinline__ HD exp_(HD x) {
    const Dual e = dexp(x.v);
    return HD(e, e * x.a);
  }
  static __device__ __forceinline__ HD log_(HD x) { return HD(dlog(x.v), x.a / x.v); }
  static __device__ __forceinline__ HD abs_(HD x) { return x.v.v < 0.0 ? -x : x; }
};

constexpr int TAN_BLOCK = 256;
enum TanMode { TAN_PARAMS = 0, TAN_GAMMA_M = 1, TAN_ADJOINT = 2 };

// Per-configuration FP64 data: sites [4N][3] (O, H1, H2, M blocks, as
// pot_ttm.f90:56-80), then the phase tables exp(2 pi i m x / L) per axis,
// [m][site] as (cos, sin): m = 0..hmax, -kmax..kmax, -lmax..lmax; then the
// sites' derivatives along the geometric direction of the dual-geometry
// instances, [4N][3]: dr/dgammaM (nonzero for the M sites only) or the
// position direction of the adjoint's force-loss term.
static size_t tangent_tab_doubles(const SysConst &c) {
  const size_t ns = 4 * (size_t)c.N;
  return 2 * ns * (size_t)((c.hmax + 1) + (2 * c.kmax + 1) + (2 * c.lmax + 1));
}
size_t tangent_cfg_doubles(const SysConst &c) {
  const size_t ns = 4 * (size_t)c.N;
  return ns * 3 + tangent_tab_doubles(c) + 3 * ns;
}

// Per (configuration, direction) dual workspace: charge, phi [4N], Efq, dR
// [4N][3], dip, olddip, Efd [nd N][3], grdq [N][27], S(k) of the charges and
// of the dipoles [nk][2], the monomer energy [1].
size_t tangent_ws_duals(const SysConst &c) {
  const size_t ns = 4 * (size_t)c.N, nds = (size_t)c.nd * c.N;
  return 8 * ns + 9 * nds + 27 * (size_t)c.N + 4 * (size_t)c.nk + 1;
}

// The workspace arrays of one (configuration, direction).
struct TanWs {
  Dual *q, *phi, *Efq, *dR, *dip, *old, *Efd, *grdq, *Sq, *Sm, *umon;
  __device__ TanWs(Dual *base, const SysConst &c) {
    const int ns = 4 * c.N, nds = c.nd * c.N;
    q = base;
    phi = q + ns;
    Efq = phi + ns;
    dR = Efq + 3 * ns;
    dip = dR + 3 * ns;
    old = dip + 3 * nds;
    Efd = old + 3 * nds;
    grdq = Efd + 3 * nds;
    Sq = grdq + 27 * c.N;
    Sm = Sq + 2 * c.nk;
    umon = Sm + 2 * c.nk;
  }
};

// mode 1: dS = dr/dgammaM; mode 2: dS = the atoms' direction gw [P][3N][3]
// (the M site moving with its atoms, pot_ttm.f90:74-79).
__global__ __launch_bounds__(TAN_BLOCK) void k_tangent_prep(SysConst c, const double *__restrict__ RR,
                                                           double *__restrict__ cfg, size_t stride,
                                                           size_t ds_off, int mode,
                                                           const double *__restrict__ gw) {
  const int b = blockIdx.x, N = c.N, ns = 4 * N;
  double *Rs = cfg + (size_t)b * stride;
  double *dS = Rs + ds_off;
  const double *rr = RR + (size_t)b * 9 * N, *g = gw + (size_t)b * 9 * N;
  const double gM = c.gammaM;
  for (int w = threadIdx.x; w < N; w += TAN_BLOCK) {
    double roh1[3], roh2[3];
    for (int k = 0; k < 3; ++k) {
      Rs[(0 * N + w) * 3 + k] = rr[9 * w + k];
      Rs[(1 * N + w) * 3 + k] = rr[9 * w + 3 + k];
      Rs[(2 * N + w) * 3 + k] = rr[9 * w + 6 + k];
      roh1[k] = rr[9 * w + 3 + k] - rr[9 * w + k];
      roh2[k] = rr[9 * w + 6 + k] - rr[9 * w + k];
    }
    minimg(c, roh1[0], roh1[1], roh1[2]);
    minimg(c, roh2[0], roh2[1], roh2[2]);
    for (int k = 0; k < 3; ++k) {
      Rs[(3 * N + w) * 3 + k] = 0.5 * gM * (roh1[k] + roh2[k]) + rr[9 * w + k];
      if (mode == 1) {
        for (int t = 0; t < 3; ++t) dS[(t * N + w) * 3 + k] = 0.0;
        dS[(3 * N + w) * 3 + k] = 0.5 * (roh1[k] + roh2[k]);
      } else if (mode == 2) {
        for (int t = 0; t < 3; ++t) dS[(t * N + w) * 3 + k] = g[9 * w + 3 * t + k];
        dS[(3 * N + w) * 3 + k] =
            (1.0 - gM) * g[9 * w + k] + 0.5 * gM * (g[9 * w + 3 + k] + g[9 * w + 6 + k]);
      }
    }
  }
  __syncthreads();
  double2 *tab = (double2 *)(Rs + (size_t)ns * 3);
  const int lens[3] = {c.hmax + 1, 2 * c.kmax + 1, 2 * c.lmax + 1};
  const int offs[3] = {0, c.kmax, c.lmax};
  for (int ax = 0; ax < 3; ++ax) {
    const int rows = lens[ax];
    for (int e = threadIdx.x; e < rows * ns; e += TAN_BLOCK) {
      const int m = e / ns - offs[ax], s = e % ns;
      double sv, cv;
      sincos(TWOPI * c.boxi[ax] * Rs[s * 3 + ax] * (double)m, &sv, &cv);
      tab[e] = make_double2(cv, sv);
    }
    tab += (size_t)rows * ns;
  }
}

// Model parameters as duals along one direction.
struct TanPar {
  Dual vdw[5], polar[4], apCC[4][4], raCC[4][4], apDD[4][4], lrc;
  double pf[4], dpf[4];  // polfac^(-1/2) per site type and its derivative by polfac
};

struct TanArgs {
  const double *RR;      // [P][Natoms][3]
  const double *pes;     // monomer PES parameters
  const double4 *kv;     // [nk] kx, ky, kz, 4 pi Ak / volume
  const int4 *khkl;      // [nk] h, k, l
  const double *cfg;     // per-configuration sites and tables
  size_t cfg_stride;
  Dual *ws;              // per (configuration, direction) workspace
  size_t ws_stride;
  const double *ew;      // [P] energy weights
  const double *gw;      // [P][Natoms][3] gradient weights
  double *out;           // [P][dirs][4]: E, dE/dtheta, loss term, SOR iterations
  double *dG;            // NULL or [P][dirs][Natoms][3] dG/dtheta
  double theta[TTM_NPARAM];
  double lrc_coef[3];    // dUvdw_lrc / d(vdwA, vdwB, vdwC)
  int dms_coef;          // 1: the directions are the DMS coefficients coefD
  int gamma_m;           // 1: the one direction is gammaM
  size_t ds_off;         // offset of the sites' directions dS [4N][3] in a configuration's data
  double *adj;           // adjoint: [P][TTM_ADJ_NPARAM] (dE/dtheta, gw . d(dE/dr)/dtheta)
};

// The derivative of the DMS charges and their gradients with respect to
// dms_param1..3 (which = 0..2), the linear corrections of nasa.f90:245-291,
// in the layout of ps_monomer (gq[disp][charge][xyz], order H1, H2, O).
__device__ inline void dms_tangent(const SysConst &c, const double r[3][3], int which,
                                   double dq3[3], double dgq[3][3][3]) {
  double R1[3], R2[3];
  for (int k = 0; k < 3; ++k) {
    R1[k] = r[1][k] - r[0][k];
    R2[k] = r[2][k] - r[0][k];
  }
  minimg(c, R1[0], R1[1], R1[2]);
  minimg(c, R2[0], R2[1], R2[2]);
  const double d1 = sqrt(R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2]);
  const double d2 = sqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
  const double costh = (R1[0] * R2[0] + R1[1] * R2[1] + R1[2] * R2[2]) / (d1 * d2);
  const double ax = R1[1] * R2[2] - R1[2] * R2[1];
  const double ay = -R1[0] * R2[2] + R1[2] * R2[0];
  const double az = R1[0] * R2[1] - R1[1] * R2[0];
  const double sinth = sqrt(ax * ax + ay * ay + az * az) / (d1 * d2);
  const double ang = atan2(sinth, costh), ath0 = 1.82400520401572996557;
  double P1 = 0, P2 = 0, p1r1 = 0, p2r2 = 0, p1c = 0, p2c = 0;
  if (which == 0) {
    P1 = d1 - c.dms2;
    P2 = d2 - c.dms2;
    p1r1 = 1.0;
    p2r2 = 1.0;
  } else if (which == 1) {
    P1 = -c.dms1;
    P2 = -c.dms1;
  } else {
    P1 = ang - ath0;
    P2 = ang - ath0;
    p1c = -1.0 / sinth;
    p2c = -1.0 / sinth;
  }
  dq3[0] = -(P1 + P2);
  dq3[1] = P1;
  dq3[2] = P2;
  const double f1q1r13 = (p1r1 - (p1c * costh / d1)) / d1, f1q1r23 = p1c / (d1 * d2);
  const double f2q1r23 = -(p1c * costh / d2) / d2, f2q1r13 = p1c / (d2 * d1);
  const double f1q2r13 = -(p2c * costh / d1) / d1, f1q2r23 = p2c / (d1 * d2);
  const double f2q2r23 = (p2r2 - (p2c * costh / d2)) / d2, f2q2r13 = p2c / (d2 * d1);
  for (int k = 0; k < 3; ++k) {
    dgq[0][0][k] = f1q1r13 * R1[k] + f1q1r23 * R2[k];
    dgq[1][0][k] = f2q1r13 * R1[k] + f2q1r23 * R2[k];
    dgq[2][0][k] = -(dgq[0][0][k] + dgq[1][0][k]);
    dgq[0][1][k] = f1q2r13 * R1[k] + f1q2r23 * R2[k];
    dgq[1][1][k] = f2q2r13 * R1[k] + f2q2r23 * R2[k];
    dgq[2][1][k] = -(dgq[0][1][k] + dgq[1][1][k]);
    dgq[0][2][k] = -(dgq[0][0][k] + dgq[0][1][k]);
    dgq[1][2][k] = -(dgq[1][0][k] + dgq[1][1][k]);
    dgq[2][2][k] = -(dgq[2][0][k] + dgq[2][1][k]);
  }
}

// The derivative of the DMS charges and their gradients with respect to the
// DMS coefficient coefD(j+1) (nasa.f90:185-242): the charges are linear in
// the coefficients, so this is the surface of ps_monomer with that one
// coefficient set to 1 and every other term (the others, the constant part
// PC0 and the dms_param corrections) left out.  Layout as ps_monomer's.
__device__ inline void dms_coef_tangent(const SysConst &c, const double r[3][3], int j,
                                        double dq3[3], double dgq[3][3][3]) {
  for (int a = 0; a < 3; ++a)
    for (int q = 0; q < 3; ++q)
      for (int k = 0; k < 3; ++k) dgq[a][q][k] = 0.0;
  if (j == 0) {  // the constant term of both P1 and P2
    dq3[0] = -2.0;
    dq3[1] = 1.0;
    dq3[2] = 1.0;
    return;
  }
  double R1[3], R2[3];
  for (int k = 0; k < 3; ++k) {
    R1[k] = r[1][k] - r[0][k];
    R2[k] = r[2][k] - r[0][k];
  }
  minimg(c, R1[0], R1[1], R1[2]);
  minimg(c, R2[0], R2[1], R2[2]);
  const double d1 = sqrt(R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2]);
  const double d2 = sqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
  const double costh = (R1[0] * R2[0] + R1[1] * R2[1] + R1[2] * R2[2]) / (d1 * d2);
  const double costhe = -.24780227221366464506;
  const double x[3] = {(d1 - PS_REOH) / PS_REOH, (d2 - PS_REOH) / PS_REOH, costh - costhe};
  const int I = c_ps.idxd[j][0], J = c_ps.idxd[j][1], K = c_ps.idxd[j][2];
  // x^(n-1), with x^(-1) := 0
  auto pw = [](double v, int n) {
    if (n <= 0) return 0.0;
    double y = 1.0;
    for (int i = 1; i < n; ++i) y *= v;
    return y;
  };
  const double B2A = 0.529177249, bfac = 1.0 / (double)0.529177249f;
  const double efd = exp(-DMS_B1D * ((d1 - PS_REOH) * (d1 - PS_REOH) + (d2 - PS_REOH) * (d2 - PS_REOH)));
  const double defdr1 = -2.0 * DMS_B1D * (d1 - PS_REOH) * efd * B2A;
  const double defdr2 = -2.0 * DMS_B1D * (d2 - PS_REOH) * efd * B2A;
  const double t1 = pw(x[0], I) * pw(x[1], J) * pw(x[2], K);
  const double t2 = pw(x[0], J) * pw(x[1], I) * pw(x[2], K);
  double p1r1 = (double)(I - 1) * pw(x[0], I - 1) * pw(x[1], J) * pw(x[2], K) / PS_REOH * B2A;
  double p1r2 = (double)(J - 1) * pw(x[0], I) * pw(x[1], J - 1) * pw(x[2], K) / PS_REOH * B2A;
  double p1c = (double)(K - 1) * pw(x[0], I) * pw(x[1], J) * pw(x[2], K - 1);
  double p2r1 = (double)(J - 1) * pw(x[0], J - 1) * pw(x[1], I) * pw(x[2], K) / PS_REOH * B2A;
  double p2r2 = (double)(I - 1) * pw(x[0], J) * pw(x[1], I - 1) * pw(x[2], K) / PS_REOH * B2A;
  double p2c = (double)(K - 1) * pw(x[0], J) * pw(x[1], I) * pw(x[2], K - 1);
  p1r1 = (p1r1 * efd + t1 * defdr1) * bfac;
  p1r2 = (p1r2 * efd + t1 * defdr2) * bfac;
  p1c = p1c * efd;
  p2r1 = (p2r1 * efd + t2 * defdr1) * bfac;
  p2r2 = (p2r2 * efd + t2 * defdr2) * bfac;
  p2c = p2c * efd;
  const double P1 = t1 * efd, P2 = t2 * efd;
  dq3[0] = -(P1 + P2);
  dq3[1] = P1;
  dq3[2] = P2;
  const double f1q1r13 = (p1r1 - (p1c * costh / d1)) / d1, f1q1r23 = p1c / (d1 * d2);
  const double f2q1r23 = (p1r2 - (p1c * costh / d2)) / d2, f2q1r13 = p1c / (d2 * d1);
  const double f1q2r13 = (p2r1 - (p2c * costh / d1)) / d1, f1q2r23 = p2c / (d1 * d2);
  const double f2q2r23 = (p2r2 - (p2c * costh / d2)) / d2, f2q2r13 = p2c / (d2 * d1);
  for (int k = 0; k < 3; ++k) {
    dgq[0][0][k] = f1q1r13 * R1[k] + f1q1r23 * R2[k];
    dgq[1][0][k] = f2q1r13 * R1[k] + f2q1r23 * R2[k];
    dgq[2][0][k] = -(dgq[0][0][k] + dgq[1][0][k]);
    dgq[0][1][k] = f1q2r13 * R1[k] + f1q2r23 * R2[k];
    dgq[1][1][k] = f2q2r13 * R1[k] + f2q2r23 * R2[k];
    dgq[2][1][k] = -(dgq[0][1][k] + dgq[1][1][k]);
    dgq[0][2][k] = -(dgq[0][0][k] + dgq[0][1][k]);
    dgq[1][2][k] = -(dgq[1][0][k] + dgq[1][1][k]);
    dgq[2][2][k] = -(dgq[2][0][k] + dgq[2][1][k]);
  }
}

// exp(i k . r_s) from the per-axis tables (the products of ssincos,
// ewald.f90:403-437 and ewald.f90:20-33).
struct Phases {
  const double2 *tx, *ty, *tz;
  const double *dS;  // the sites' directions [4N][3]
  int ns, kmax, lmax, N;
  __device__ __forceinline__ void at(int s, int4 h, double &ck, double &sk) const {
    const double2 a = tx[(size_t)h.x * ns + s], b = ty[(size_t)(h.y + kmax) * ns + s],
                  z = tz[(size_t)(h.z + lmax) * ns + s];
    const double cxy = a.x * b.x - a.y * b.y, sxy = a.y * b.x + a.x * b.y;
    ck = cxy * z.x - sxy * z.y;
    sk = sxy * z.x + cxy * z.y;
  }
  // with moving sites: d exp(i k.r) = i (k . dr) exp(i k.r)
  __device__ __forceinline__ void at(int s, int4 h, const double4 &kv, Dual &ck, Dual &sk) const {
    double c, sn;
    at(s, h, c, sn);
    const double *d = dS + 3 * s;
    const double kd = kv.x * d[0] + kv.y * d[1] + kv.z * d[2];
    ck = Dual(c, -kd * sn);
    sk = Dual(sn, kd * c);
  }
  __device__ __forceinline__ void at(int s, int4 h, const double4 &, double &ck, double &sk) const {
    at(s, h, ck, sk);
  }
};

// Geometry of the GD = true instance: minimum image on the value part (the
// box shift is constant along gammaM).
__device__ __forceinline__ void minimg(const SysConst &c, Dual &x, Dual &y, Dual &z) {
  minimg(c, x.v, y.v, z.v);
}
__device__ __forceinline__ Dual as_dual(Dual x) { return x; }
__device__ __forceinline__ Dual as_dual(double x) { return Dual(x); }

// self1 / self2 (smear.f90:110-156) of a dual distance, the same steps as
// those of the plain ones.
__device__ __forceinline__ void self1(const SysConst &c, Dual drsq, Dual &ts0, Dual &ts1) {
  const Dual dd = dsqrt(drsq), drsqi = Dual(1.0) / drsq, dri = dsqrt(drsqi);
  const Dual expon = dexp(-(c.a2 * drsq));
  const Dual kr0 = erfc_as<Dual>(c.aewald * dd, expon) * dri;
  const Dual kr1 = (kr0 + c.cts1 * expon) * drsqi;
  const Dual ch0 = -dri, ch1 = ch0 * drsqi;
  ts0 = ch0 + kr0;
  ts1 = kr1 + ch1;
}
__device__ __forceinline__ void self2(const SysConst &c, Dual drsq, Dual &ts1, Dual &ts2) {
  const Dual dd = dsqrt(drsq), drsqi = Dual(1.0) / drsq, dri = dsqrt(drsqi);
  const Dual expon = dexp(-(c.a2 * drsq));
  const Dual kr0 = erfc_as<Dual>(c.aewald * dd, expon) * dri;
  const Dual kr1 = (kr0 + c.cts1 * expon) * drsqi;
  const Dual kr2 = (kr1 + c.cts2 * expon) * drsqi;
  const Dual ch0 = -dri, ch1 = ch0 * drsqi, ch2 = ch1 * drsqi;
  ts1 = kr1 + ch1;
  ts2 = kr2 + ch2;
}

__device__ __forceinline__ double block_sum(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = TAN_BLOCK / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Lanes per site of a loop over nsites sites whose per-site work is a sum
// over k-vectors: the largest power of two (at most a wave) such that the
// groups of all sites fit in the workgroup.  Each group splits the k-vectors
// of its site and combines its partial sums with group_sum.
__device__ __forceinline__ int site_lanes(int nsites) {
  int g = 1;
  while (g < 64 && 2 * g * nsites <= TAN_BLOCK) g *= 2;
  return g;
}

__device__ __forceinline__ Dual group_sum(Dual x, int G) {
  for (int o = G >> 1; o > 0; o >>= 1) {
    x.v += __shfl_xor(x.v, o);
    x.d += __shfl_xor(x.d, o);
  }
  return x;
}

// (3 t2 r r^T - t1) v, get_dd3 applied (pot_ttm.f90:633-647).
template <class G>
__device__ __forceinline__ void dd3d(const G r[3], Dual t1, Dual t2, const Dual v[3], Dual out[3]) {
  const Dual t = 3.0 * t2;
  const Dual rv = v[0] * r[0] + v[1] * r[1] + v[2] * r[2];
  for (int k = 0; k < 3; ++k) out[k] = t * r[k] * rv - t1 * v[k];
}

// The monomer phase of the tangent evaluation (pot_ttm.f90:84-151): DMS
// charges and their coordinate gradients with their derivatives along the
// direction, the PES gradient and the monomer energy of the configuration,
// into the (configuration, direction) workspace.  A kernel of its own: the
// PES and DMS polynomials hold more registers than any later phase.
template <int MODE>
__global__ __launch_bounds__(TAN_BLOCK) void k_tangent_monomer(SysConst c, TanArgs a) {
  using GT = std::conditional_t<MODE != TAN_PARAMS, Dual, double>;
  const int b = blockIdx.x, dir = blockIdx.y, ndir = gridDim.y, tid = threadIdx.x, N = c.N;
  const TanWs W(a.ws + ((size_t)b * ndir + dir) * a.ws_stride, c);
  Dual *q = W.q, *phi = W.phi, *Efq = W.Efq, *dR = W.dR, *grdq = W.grdq;
  __shared__ double red[TAN_BLOCK];
  // gammaM and the charge placement factors (pot_ttm.f90:134-137)
  GT gM = c.gammaM;
  if constexpr (MODE == TAN_GAMMA_M) gM.d = 1.0;
  const GT tmpq = 0.5 * gM / (1.0 - gM), qMf = GT(1.0) / (1.0 - gM);

  double umon = 0.0;
  for (int w = tid; w < N; w += TAN_BLOCK) {
    double r[3][3], e1, dr1[3][3], q3[3], gq[3][3][3], dq3[3] = {0, 0, 0}, dgq[3][3][3] = {};
    const double *rr = a.RR + ((size_t)b * 3 * N + 3 * w) * 3;
    for (int at = 0; at < 3; ++at)
      for (int k = 0; k < 3; ++k) r[at][k] = rr[3 * at + k];
    ps_monomer<true>(c, r, e1, dr1, q3, gq, a.pes);
    umon += e1;
    if constexpr (MODE == TAN_ADJOINT) {
      // the charges along the atoms' direction gw: dq = sum_atoms gq . gw, with
      // gq[disp][charge] in the order (H1, H2, O) and q3 in (O, H1, H2)
      const double *g = a.gw + ((size_t)b * 3 * N + 3 * w) * 3;
      const int atom[3] = {1, 2, 0};
      for (int i = 0; i < 3; ++i) {
        const int ch = i == 0 ? 2 : i - 1;
        double v = 0.0;
        for (int d = 0; d < 3; ++d)
          for (int k = 0; k < 3; ++k) v += gq[d][ch][k] * g[3 * atom[d] + k];
        dq3[i] = v;
      }
    } else if (MODE == TAN_PARAMS && a.dms_coef) {
      dms_coef_tangent(c, r, dir, dq3, dgq);
    } else if (MODE == TAN_PARAMS && dir >= TP_DMS1) {
      dms_tangent(c, r, dir - TP_DMS1, dq3, dgq);
    }
    Dual qd[3];
    for (int i = 0; i < 3; ++i) qd[i] = Dual(q3[i] * CHARGECON, dq3[i] * CHARGECON);
    q[w] = Dual(0.0);
    q[N + w] = qd[1] + tmpq * (qd[1] + qd[2]);
    q[2 * N + w] = qd[2] + tmpq * (qd[1] + qd[2]);
    q[3 * N + w] = qd[0] * qMf;
    Dual *g = grdq + 27 * w;
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 3; ++k) {
        const Dual g0 = Dual(gq[d][0][k] * CHARGECON, dgq[d][0][k] * CHARGECON);
        const Dual g1 = Dual(gq[d][1][k] * CHARGECON, dgq[d][1][k] * CHARGECON);
        const Dual g2 = Dual(gq[d][2][k] * CHARGECON, dgq[d][2][k] * CHARGECON);
        g[d * 9 + 0 * 3 + k] = g0 + tmpq * (g0 + g1);
        g[d * 9 + 1 * 3 + k] = g1 + tmpq * (g0 + g1);
        g[d * 9 + 2 * 3 + k] = g2 - 2.0 * tmpq * (g0 + g1);
      }
    for (int k = 0; k < 3; ++k) {
      dR[(0 * N + w) * 3 + k] = Dual(dr1[0][k]);
      dR[(1 * N + w) * 3 + k] = Dual(dr1[1][k]);
      dR[(2 * N + w) * 3 + k] = Dual(dr1[2][k]);
      dR[(3 * N + w) * 3 + k] = Dual(0.0);
    }
    for (int t = 0; t < 4; ++t) {
      phi[t * N + w] = Dual(0.0);
      for (int k = 0; k < 3; ++k) Efq[(t * N + w) * 3 + k] = Dual(0.0);
    }
  }
  const double u = block_sum(umon, red);
  if (tid == 0) *W.umon = Dual(u);
}

// The adjoint's position derivatives of the screening functions ts_n(r^2)
// follow the chain the gradient of pot_ttm uses, ts_(n+1) = -1/((2n+1) r)
// d ts_n/dr (the erfc-screened and Thole tensors of smear.f90), rather than
// differentiating the A&S erfc polynomial and the truncated Gamma(2/3) series
// themselves: the reference's gradient departs from the exact derivative of
// its energy by ~1e-5 relative there, and the adjoint differentiates the
// gradient as computed.  d ts_n = -(2n+1)/2 ts_(n+1) d(r^2).
__device__ __forceinline__ void chain_d(Dual &ts, int n, double ts_next, double dd2) {
  ts.d = -0.5 * (2 * n + 1) * ts_next * dd2;
}

// Sum over the workgroup of a dual.
__device__ __forceinline__ Dual block_sum(Dual v, double *red) {
  const double x = block_sum(v.v, red);
  return Dual(x, block_sum(v.d, red));
}

// The adjoint's outputs for configuration b: per parameter j of the 16 TTM
// parameters and the 84 DMS coefficients, adj[b][j] = (dE/dtheta_j,
// gw . d(dE/dr)/dtheta_j).  Beside the Thole and vdW partials gathered by the
// pair loops: the polarisabilities (-1/2 |mu|^2 / alpha^2 per dipole site,
// the alpha^-1 term of the stationary induction energy) and the charge
// parameters (sum of phi dq/dtheta over the charged sites, phi the potential
// of charges and dipoles, which is dE/dq).  Each is a dual along the atoms'
// direction gw, the second part the derivative of the first along it.
// (gammaM, the last of TTM_ADJ_NPARAM, moves the M sites: its slot is filled
// by the host from a tangent sweep.)
__device__ void adjoint_out(const SysConst &c, const TanArgs &a, const TanPar &P, const TanWs &W,
                            int b, const double vE[5], const double vF[5], Dual accCC, Dual accDD,
                            const Dual accPF[3], double *red) {
  const int N = c.N, tid = threadIdx.x, nds = c.nd * N, tdip = c.tdip;
  double *out = a.adj + (size_t)b * TTM_ADJ_NPARAM * 2;
  const Dual *phi = W.phi, *q = W.q, *dip = W.dip, *dR = W.dR;
  const double gM = c.gammaM, tq = 0.5 * gM / (1.0 - gM);
  // polarisabilities by site type (O, H, M)
  Dual pol[3] = {0.0, 0.0, 0.0};
  for (int i = tid; i < nds; i += TAN_BLOCK) {
    const int t = tdip + i / N, pt = t == 0 ? 0 : t == 3 ? 2 : 1;
    const double al = P.polar[t].v;
    const Dual *mu = dip + 3 * i;
    const Dual m2 = mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2];
    if (al > 0.0) pol[pt] -= 0.5 * m2 / (al * al);
  }
  const Dual sCC = block_sum(accCC, red), sDD = block_sum(accDD, red);
  Dual sPF[3], sPol[3];
  for (int i = 0; i < 3; ++i) {
    sPF[i] = block_sum(accPF[i], red);
    sPol[i] = block_sum(pol[i], red);
  }
  double sE[5], sF[5];
  for (int i = 0; i < 5; ++i) {
    sE[i] = block_sum(vE[i], red);
    sF[i] = block_sum(vF[i], red);
  }
  if (tid == 0) {
    auto put = [&](int j, Dual v) {
      out[2 * j] = v.v;
      out[2 * j + 1] = v.d;
    };
    for (int i = 0; i < 5; ++i) put(TP_VDWA + i, Dual(sE[i] + (i < 3 ? a.lrc_coef[i] : 0.0), sF[i]));
    for (int i = 0; i < 3; ++i) put(TP_POLARO + i, sPol[i]);
    for (int i = 0; i < 3; ++i) put(TP_POLFACO + i, sPF[i]);
    put(TP_ADD, sDD);
    put(TP_ACC, sCC);
    put(TTM_ADJ_NPARAM - 1, Dual(0.0));
  }
  // the charge parameters: dms_param1..3, then the 84 coefficients coefD;
  // one parameter per lane over the molecules
  for (int j = tid; j < 3 + DMS_NCOEF; j += TAN_BLOCK) {
    Dual acc = 0.0;
    for (int w = 0; w < N; ++w) {
      const double *rr = a.RR + ((size_t)b * 3 * N + 3 * w) * 3;
      const double *g = a.gw + ((size_t)b * 3 * N + 3 * w) * 3;
      double r[3][3], dq3[3], dgq[3][3][3];
      for (int at = 0; at < 3; ++at)
        for (int k = 0; k < 3; ++k) r[at][k] = rr[3 * at + k];
      if (j < 3) dms_tangent(c, r, j, dq3, dgq);
      else dms_coef_tangent(c, r, j - 3, dq3, dgq);
      // dq3 in (O, H1, H2), dgq[disp][charge] in (H1, H2, O); along gw
      const int atom[3] = {1, 2, 0};
      Dual dqd[3];
      for (int i = 0; i < 3; ++i) {
        const int ch = i == 0 ? 2 : i - 1;
        double v = 0.0;
        for (int d = 0; d < 3; ++d)
          for (int k = 0; k < 3; ++k) v += dgq[d][ch][k] * g[3 * atom[d] + k];
        dqd[i] = Dual(dq3[i] * CHARGECON, v * CHARGECON);
      }
      const Dual sH = dqd[1] + dqd[2];
      acc += phi[N + w] * (dqd[1] + tq * sH) + phi[2 * N + w] * (dqd[2] + tq * sH) +
             phi[3 * N + w] * dqd[0] * (1.0 / (1.0 - gM));
    }
    out[2 * (TP_DMS1 + j)] = acc.v;
    out[2 * (TP_DMS1 + j) + 1] = acc.d;
  }
}

// MODE TAN_PARAMS: one parameter direction per workgroup, plain FP64
// geometry (GT = double); TAN_GAMMA_M and TAN_ADJOINT: dual site positions
// and phases (GT = Dual) along dr/dgammaM or along the atoms' direction gw.
template <int MODE>
__global__ __launch_bounds__(TAN_BLOCK) void k_ttm_tangent(SysConst c, TanArgs a) {
  constexpr bool GD = MODE != TAN_PARAMS;
  using GT = std::conditional_t<GD, Dual, double>;
  const int b = blockIdx.x, dir = blockIdx.y, ndir = gridDim.y, tid = threadIdx.x;
  const int N = c.N, ns = 4 * N, nd = c.nd, tdip = c.tdip, nds = nd * N, nk = c.nk;
  const int tend = tdip + nd;  // dipole site types [tdip, tend): M (TTM3-F) or O, H1, H2 (TTM2.1-F)
  const double *Rs = a.cfg + (size_t)b * a.cfg_stride;
  Phases ph;
  ph.ns = ns;
  ph.kmax = c.kmax;
  ph.lmax = c.lmax;
  ph.N = N;
  ph.dS = Rs + a.ds_off;
  ph.tx = (const double2 *)(Rs + (size_t)ns * 3);
  ph.ty = ph.tx + (size_t)(c.hmax + 1) * ns;
  ph.tz = ph.ty + (size_t)(2 * c.kmax + 1) * ns;
  // site s's position, with its derivative along the geometric direction in
  // the dual-geometry instances
  auto site = [&](int s, GT r[3]) {
    for (int k = 0; k < 3; ++k) r[k] = Rs[3 * s + k];
    if constexpr (GD)
      for (int k = 0; k < 3; ++k) r[k].d = ph.dS[3 * s + k];
  };
  const TanWs W(a.ws + ((size_t)b * ndir + dir) * a.ws_stride, c);
  Dual *q = W.q, *phi = W.phi, *Efq = W.Efq, *dR = W.dR, *dip = W.dip, *old = W.old,
       *Efd = W.Efd, *grdq = W.grdq, *Sq = W.Sq, *Sm = W.Sm;

  __shared__ TanPar P;
  __shared__ double red[TAN_BLOCK];
  if (tid == 0) {
    const bool seed = MODE == TAN_PARAMS && !a.dms_coef;
    auto par = [&](int i) { return Dual(a.theta[i], (seed && i == dir) ? 1.0 : 0.0); };
    for (int i = 0; i < 5; ++i) P.vdw[i] = par(i);
    P.polar[0] = par(5);
    P.polar[1] = par(6);
    P.polar[2] = par(6);
    P.polar[3] = par(7);
    // polfac = (polfac^(1/6))^-3 (potential.f90:157-160)
    const Dual f6[4] = {dpow(par(8), 1.0 / 6.0), dpow(par(9), 1.0 / 6.0), dpow(par(9), 1.0 / 6.0),
                        dpow(par(10), 1.0 / 6.0)};
    Dual pf[4];
    for (int i = 0; i < 4; ++i) pf[i] = Dual(1.0) / (f6[i] * f6[i] * f6[i]);
    for (int i = 0; i < 4; ++i) {
      const double pfi = a.theta[TP_POLFACO + (i == 0 ? 0 : i == 3 ? 2 : 1)];
      P.pf[i] = pf[i].v;
      P.dpf[i] = -0.5 * pf[i].v / pfi;  // d polfac^(-1/2) / d polfac
    }
    const Dual aDD = par(11), aCC = par(12);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        P.apCC[i][j] = aCC * pf[i] * pf[j];
        P.raCC[i][j] = P.apCC[i][j].v > 0.0 ? dpow(P.apCC[i][j], 1.0 / 3.0) : Dual(0.0);
        P.apDD[i][j] = aDD * pf[i] * pf[j];
      }
    P.lrc = Dual(c.Uvdw_lrc, (seed && dir < 3) ? a.lrc_coef[dir] : 0.0);
  }
  __syncthreads();
  // gammaM, which redistributes the M-site gradient (pot_ttm.f90:597-602)
  GT gM = c.gammaM;
  if constexpr (MODE == TAN_GAMMA_M) gM.d = 1.0;
  // charges, their gradients, the monomer gradients and energy: k_tangent_monomer

  // ---- real space: vdW and charge-charge terms (pot_ttm.f90:159-228), per site
  Dual uvdw = 0.0;
  // adjoint: d Uvdw / d(vdwA..E) and gw . d G / d(vdwA..E) of the gradient as
  // computed (its vdW force, pot_ttm.f90:176, has no vdwA, vdwB terms)
  double vE[5] = {0, 0, 0, 0, 0}, vF[5] = {0, 0, 0, 0, 0};
  for (int s = tid; s < ns; s += TAN_BLOCK) {
    const int iw = s % N, ti = s / N;
    const double *ROi = Rs + 3 * iw;
    GT Ri[3];
    site(s, Ri);
    Dual pv = 0.0, ef[3] = {0.0, 0.0, 0.0}, gv[3] = {0.0, 0.0, 0.0};
    for (int jw = 0; jw < N; ++jw) {
      if (jw == iw) continue;
      const double *ROj = Rs + 3 * jw;
      double Rij[3] = {ROi[0] - ROj[0], ROi[1] - ROj[1], ROi[2] - ROj[2]};
      minimg(c, Rij[0], Rij[1], Rij[2]);
      const double d2 = norm2_exact(Rij[0], Rij[1], Rij[2]);
      if (!(d2 < c.rc2)) continue;
      if (ti == 0) {
        const double dRij = sqrt(d2), R2i = 1.0 / d2, R4i = R2i * R2i, R6i = R4i * R2i,
                     R8i = R6i * R2i, x = dRij - c.rc1;
        const Dual ex = P.vdw[3] * dexp(-(P.vdw[4] * dRij));
        const Dual uij = R6i * ((P.vdw[0] * R6i + P.vdw[1] * R4i) + P.vdw[2]) + ex;
        const Dual fsc = -(R8i * 6.0 * P.vdw[2] + ex * P.vdw[4] / dRij +
                           Dual(c.shiftA * (x * x) + c.shiftB * (x * x * x)));
        for (int k = 0; k < 3; ++k) gv[k] += fsc * Rij[k];
        if (iw < jw) uvdw += uij;
        if constexpr (MODE == TAN_ADJOINT) {
          const double Dv = P.vdw[3].v, Ev = P.vdw[4].v, e = exp(-(Ev * dRij));
          const double *gO = a.gw + ((size_t)b * 3 * N + 3 * iw) * 3;
          const double gr = gO[0] * Rij[0] + gO[1] * Rij[1] + gO[2] * Rij[2];
          if (iw < jw) {
            vE[0] += R6i * R6i;
            vE[1] += R6i * R4i;
            vE[2] += R6i;
            vE[3] += e;
            vE[4] -= Dv * dRij * e;
          }
          vF[2] -= 6.0 * R8i * gr;
          vF[3] -= e * Ev / dRij * gr;
          vF[4] -= Dv * e * (1.0 - Ev * dRij) / dRij * gr;
        }
      }
      GT Ra[3];
      for (int k = 0; k < 3; ++k) Ra[k] = Ri[k] + (Rij[k] - ROi[k] + ROj[k]);
      for (int tj = ti == 0 ? 1 : 0; tj < 4; ++tj) {
        const int sj = tj * N + jw;
        GT Rj[3];
        site(sj, Rj);
        GT r[3] = {Ra[0] - Rj[0], Ra[1] - Rj[1], Ra[2] - Rj[2]};
        minimg(c, r[0], r[1], r[2]);
        const GT d2s = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        Dual ts0, ts1;
        smear01<true, Dual>(c, as_dual(d2s), P.apCC[ti][tj], P.raCC[ti][tj], ts0, ts1);
        if constexpr (MODE == TAN_ADJOINT) {
          double u1, u2;
          smear2<true, double>(c, d2s.v, P.apCC[ti][tj].v, u1, u2);
          chain_d(ts0, 0, ts1.v, d2s.d);
          chain_d(ts1, 1, u2, d2s.d);
        }
        const Dual qj = q[sj];
        pv += ts0 * qj;
        const Dual f = ts1 * qj;
        for (int k = 0; k < 3; ++k) ef[k] += f * r[k];
      }
    }
    phi[s] += pv;
    for (int k = 0; k < 3; ++k) {
      Efq[3 * s + k] += ef[k];
      dR[3 * s + k] += gv[k];
    }
  }
  __syncthreads();

  // ---- Ewald self energy and intramolecular exclusions (pot_ttm.f90:232-261)
  for (int w = tid; w < N; w += TAN_BLOCK) {
    for (int t = 1; t < 4; ++t) phi[t * N + w] -= c.self_fac * q[t * N + w];
    for (int ti = 0; ti < 3; ++ti)
      for (int tj = ti + 1; tj < 4; ++tj) {
        const int si = ti * N + w, sj = tj * N + w;
        GT Ri[3], Rj[3];
        site(si, Ri);
        site(sj, Rj);
        GT r[3] = {Ri[0] - Rj[0], Ri[1] - Rj[1], Ri[2] - Rj[2]};
        minimg(c, r[0], r[1], r[2]);
        const GT d2x = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        GT ts0, ts1;
        self1(c, d2x, ts0, ts1);
        if constexpr (MODE == TAN_ADJOINT) {
          double u1, u2;
          self2(c, d2x.v, u1, u2);
          chain_d(ts0, 0, ts1.v, d2x.d);
          chain_d(ts1, 1, u2, d2x.d);
        }
        phi[si] += ts0 * q[sj];
        phi[sj] += ts0 * q[si];
        for (int k = 0; k < 3; ++k) {
          Efq[3 * si + k] += ts1 * q[sj] * r[k];
          Efq[3 * sj + k] -= ts1 * q[si] * r[k];
        }
      }
  }
  __syncthreads();

  // ---- reciprocal space, charges (ewald_std_qq, ewald.f90:1-52)
  for (int k = tid; k < nk; k += TAN_BLOCK) {
    const int4 h = a.khkl[k];
    const double4 kv = a.kv[k];
    Dual rQ = 0.0, iQ = 0.0;
    for (int s = N; s < ns; ++s) {
      GT ck, sk;
      ph.at(s, h, kv, ck, sk);
      rQ += q[s] * ck;
      iQ += q[s] * sk;
    }
    Sq[2 * k] = rQ;
    Sq[2 * k + 1] = iQ;
  }
  __syncthreads();
  const int Gs = site_lanes(ns), Gd = site_lanes(nds);
  for (int s = tid / Gs; s < ns; s += TAN_BLOCK / Gs) {
    Dual ef[3] = {0.0, 0.0, 0.0}, pv = 0.0;
    for (int k = tid % Gs; k < nk; k += Gs) {
      const int4 h = a.khkl[k];
      const double4 kv = a.kv[k];
      const double p4vAk = kv.w;
      GT ck, sk;
      ph.at(s, h, kv, ck, sk);
      const Dual rQ = Sq[2 * k], iQ = Sq[2 * k + 1];
      const Dual t = p4vAk * (ck * iQ - sk * rQ);
      ef[0] -= t * kv.x;
      ef[1] -= t * kv.y;
      ef[2] -= t * kv.z;
      pv += p4vAk * (ck * rQ + sk * iQ);
    }
    for (int k = 0; k < 3; ++k) ef[k] = group_sum(ef[k], Gs);
    pv = group_sum(pv, Gs);
    if (tid % Gs == 0) {
      for (int k = 0; k < 3; ++k) Efq[3 * s + k] += ef[k];
      if (s >= N) phi[s] += pv;
    }
  }
  __syncthreads();

  // ---- Uelec, the charge-field gradient (pot_ttm.f90:273-276), initial dipoles
  Dual uel = 0.0;
  for (int s = N + tid; s < ns; s += TAN_BLOCK) {
    uel += q[s] * phi[s];
    for (int k = 0; k < 3; ++k) dR[3 * s + k] -= q[s] * Efq[3 * s + k];
  }
  for (int i = tid; i < nds; i += TAN_BLOCK) {
    const int t = tdip + i / N, s = t * N + i % N;
    for (int k = 0; k < 3; ++k) {
      dip[3 * i + k] = P.polar[t] * Efq[3 * s + k];
      old[3 * i + k] = Dual(0.0);
    }
  }
  __syncthreads();

  // ---- SOR dipole iterations (pot_ttm.f90:306-392)
  int iters = c.maxiter;
  const double om = c.sor;
  for (int it = 1; it <= c.maxiter; ++it) {
    for (int i = tid; i < nds; i += TAN_BLOCK) {
      const int ti = tdip + i / N, iw = i % N, s = ti * N + iw;
      const double *ROi = Rs + 3 * iw;
      GT Ri[3];
      site(s, Ri);
      Dual e[3] = {0.0, 0.0, 0.0};
      for (int jw = 0; jw < N; ++jw) {
        if (jw == iw) continue;
        const double *ROj = Rs + 3 * jw;
        double Rij[3] = {ROi[0] - ROj[0], ROi[1] - ROj[1], ROi[2] - ROj[2]};
        minimg(c, Rij[0], Rij[1], Rij[2]);
        if (!(norm2_exact(Rij[0], Rij[1], Rij[2]) < c.rc2)) continue;
        double R0[3] = {Rij[0] - ROi[0] + ROj[0], Rij[1] - ROi[1] + ROj[1], Rij[2] - ROi[2] + ROj[2]};
        minimg(c, R0[0], R0[1], R0[2]);
        for (int tj = tdip; tj < tend; ++tj) {
          const int sj = tj * N + jw, j = (tj - tdip) * N + jw;
          GT Rj[3], r[3];
          site(sj, Rj);
          for (int k = 0; k < 3; ++k) r[k] = Ri[k] + R0[k] - Rj[k];
          minimg(c, r[0], r[1], r[2]);
          Dual t1, t2, o[3];
          const GT r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          smear2<true, Dual>(c, as_dual(r2), P.apDD[ti][tj], t1, t2);
          if constexpr (MODE == TAN_ADJOINT) {
            double u1, u2, u3;
            smear3<true, double>(c, r2.v, P.apDD[ti][tj].v, u1, u2, u3);
            chain_d(t1, 1, t2.v, r2.d);
            chain_d(t2, 2, u3, r2.d);
          }
          dd3d(r, t1, t2, dip + 3 * j, o);
          for (int k = 0; k < 3; ++k) e[k] += o[k];
        }
      }
      // dipole sites of the same molecule (TTM2.1-F, pot_ttm.f90:344-365)
      for (int tj = tdip; tj < tend; ++tj) {
        if (tj == ti) continue;
        const int sj = tj * N + iw, j = (tj - tdip) * N + iw;
        GT Rj[3];
        site(sj, Rj);
        GT r[3] = {Ri[0] - Rj[0], Ri[1] - Rj[1], Ri[2] - Rj[2]};
        minimg(c, r[0], r[1], r[2]);
        Dual t1, t2, o[3];
        const GT r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        smear2<true, Dual>(c, as_dual(r2), P.apDD[ti][tj], t1, t2);
        if constexpr (MODE == TAN_ADJOINT) {
          double u1, u2, u3;
          smear3<true, double>(c, r2.v, P.apDD[ti][tj].v, u1, u2, u3);
          chain_d(t1, 1, t2.v, r2.d);
          chain_d(t2, 2, u3, r2.d);
        }
        dd3d(r, t1, t2, dip + 3 * j, o);
        for (int k = 0; k < 3; ++k) e[k] += o[k];
      }
      for (int k = 0; k < 3; ++k) Efd[3 * i + k] = e[k];
    }
    // structure factor of the dipoles (ewald_std_dd, ewald.f90:56-119)
    for (int k = tid; k < nk; k += TAN_BLOCK) {
      const int4 h = a.khkl[k];
      const double4 kv = a.kv[k];
      Dual rQ = 0.0, iQ = 0.0;
      for (int i = 0; i < nds; ++i) {
        const int s = (tdip + i / N) * N + i % N;
        GT ck, sk;
        ph.at(s, h, kv, ck, sk);
        const Dual dK = dip[3 * i] * kv.x + dip[3 * i + 1] * kv.y + dip[3 * i + 2] * kv.z;
        rQ -= dK * sk;
        iQ += dK * ck;
      }
      Sm[2 * k] = rQ;
      Sm[2 * k + 1] = iQ;
    }
    __syncthreads();
    double ss = 0.0;
    for (int i = tid / Gd; i < nds; i += TAN_BLOCK / Gd) {
      const int ti = tdip + i / N, s = ti * N + i % N;
      Dual e[3] = {0.0, 0.0, 0.0};
      for (int k = tid % Gd; k < nk; k += Gd) {
        const int4 h = a.khkl[k];
        const double4 kv = a.kv[k];
        GT ck, sk;
        ph.at(s, h, kv, ck, sk);
        const Dual t = kv.w * (ck * Sm[2 * k + 1] - sk * Sm[2 * k]);
        e[0] -= kv.x * t;
        e[1] -= kv.y * t;
        e[2] -= kv.z * t;
      }
      for (int k = 0; k < 3; ++k) e[k] = Efd[3 * i + k] + group_sum(e[k], Gd);
      if (tid % Gd != 0) continue;
      for (int k = 0; k < 3; ++k) {
        const Dual mu = dip[3 * i + k];
        const Dual ek = e[k] + Efq[3 * s + k] + c.fac_dd * mu;
        const Dual nu = (1.0 - om) * mu + (om * P.polar[ti]) * ek;
        const double df = nu.v - old[3 * i + k].v;
        ss += df * df;
        dip[3 * i + k] = nu;
      }
    }
    const double sum = block_sum(ss, red);
    if (c.stath * sqrt(sum) < c.peps) {
      iters = it;
      break;
    }
    for (int i = tid; i < 3 * nds; i += TAN_BLOCK) old[i] = dip[i];
    __syncthreads();
  }
  __syncthreads();

  // ---- Uind (pot_ttm.f90:390) and the final real-space forces (pot_ttm.f90:399-542)
  Dual uind = 0.0;
  // adjoint: d U / d aCCaCD, aDD, polfacO/H/M at fixed dipoles, through the
  // Thole factors a = a0 polfac_i^(-1/2) polfac_j^(-1/2) of each site pair
  Dual accCC = 0.0, accDD = 0.0, accPF[3] = {0.0, 0.0, 0.0};
  auto ptype = [](int t) { return t == 0 ? 0 : t == 3 ? 2 : 1; };
  auto thole_acc = [&](Dual w, bool ddt, int ti, int tj) {
    const double pi = P.pf[ti], pj = P.pf[tj];
    if (ddt) accDD += w * (pi * pj);
    else accCC += w * (pi * pj);
    const Dual wa = w * a.theta[ddt ? TP_ADD : TP_ACC];
    accPF[ptype(ti)] += wa * (pj * P.dpf[ti]);
    accPF[ptype(tj)] += wa * (pi * P.dpf[tj]);
  };
  for (int i = tid; i < nds; i += TAN_BLOCK) {
    const int s = (tdip + i / N) * N + i % N;
    for (int k = 0; k < 3; ++k) uind += dip[3 * i + k] * Efq[3 * s + k];
  }
  for (int s = tid; s < ns; s += TAN_BLOCK) {
    const int iw = s % N, ti = s / N;
    const bool adip = ti >= tdip && ti < tend;
    const double *ROi = Rs + 3 * iw;
    GT Ri[3];
    site(s, Ri);
    const Dual qi = q[s];
    Dual di[3] = {0.0, 0.0, 0.0};
    if (adip)
      for (int k = 0; k < 3; ++k) di[k] = dip[3 * ((ti - tdip) * N + iw) + k];
    Dual g[3] = {0.0, 0.0, 0.0}, pv = 0.0;
    for (int jw = 0; jw < N; ++jw) {
      const bool same = jw == iw;
      double R0[3] = {0.0, 0.0, 0.0};
      if (!same) {
        const double *ROj = Rs + 3 * jw;
        double Rij[3] = {ROi[0] - ROj[0], ROi[1] - ROj[1], ROi[2] - ROj[2]};
        minimg(c, Rij[0], Rij[1], Rij[2]);
        if (!(norm2_exact(Rij[0], Rij[1], Rij[2]) < c.rc2)) continue;
        for (int k = 0; k < 3; ++k) R0[k] = Rij[k] - ROi[k] + ROj[k];
        minimg(c, R0[0], R0[1], R0[2]);
      }
      for (int tj = 0; tj < 4; ++tj) {
        if (same && tj == ti) continue;
        const int sj = tj * N + jw;
        const bool bdip = tj >= tdip && tj < tend;
        GT Rj[3], r[3];
        site(sj, Rj);
        for (int k = 0; k < 3; ++k) r[k] = same ? Ri[k] - Rj[k] : Ri[k] + R0[k] - Rj[k];
        minimg(c, r[0], r[1], r[2]);
        const GT d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        const Dual qj = q[sj];
        Dual dj[3] = {0.0, 0.0, 0.0};
        if (bdip)
          for (int k = 0; k < 3; ++k) dj[k] = dip[3 * ((tj - tdip) * N + jw) + k];
        if (adip && bdip) {
          // dipole-dipole: d_i . T3(r) . d_j with T3 the third-rank tensor of PBCsmear3
          Dual t1, t2, t3;
          smear3<true, Dual>(c, as_dual(d2), P.apDD[ti][tj], t1, t2, t3);
          const Dual dir_ = di[0] * r[0] + di[1] * r[1] + di[2] * r[2];
          const Dual djr = dj[0] * r[0] + dj[1] * r[1] + dj[2] * r[2];
          const Dual dd = di[0] * dj[0] + di[1] * dj[1] + di[2] * dj[2];
          const Dual a15 = 15.0 * t3 * dir_ * djr, a3 = 3.0 * t2;
          for (int k = 0; k < 3; ++k) g[k] += a15 * r[k] - a3 * (dir_ * dj[k] + djr * di[k] + dd * r[k]);
        }
        Dual qdqd[3];
        for (int k = 0; k < 3; ++k) qdqd[k] = qj * di[k] - qi * dj[k];
        Dual ts1 = 0.0, ts2 = 0.0;
        bool qd = true;
        if (same) {
          GT s1, s2;
          self2(c, d2, s1, s2);
          ts1 = s1;
          ts2 = s2;
        } else {
          const double qsq = qdqd[0].v * qdqd[0].v + qdqd[1].v * qdqd[1].v + qdqd[2].v * qdqd[2].v;
          qd = qsq > 1e-9;
          if (qd) smear2<true, Dual>(c, as_dual(d2), P.apCC[ti][tj], ts1, ts2);
        }
        // the potential of the dipoles (phi, dE/dq) along the direction; the
        // gradient terms (g) only feed dE/dr, which the adjoint does not use
        if constexpr (MODE == TAN_ADJOINT) chain_d(ts1, 1, ts2.v, as_dual(d2).d);
        if (qd) {
          Dual o[3];
          dd3d(r, ts1, ts2, qdqd, o);
          for (int k = 0; k < 3; ++k) g[k] += o[k];
        }
        pv += ts1 * (dj[0] * r[0] + dj[1] * r[1] + dj[2] * r[2]);
        if constexpr (MODE == TAN_ADJOINT) {
          // The energy is stationary in the converged dipoles, so its
          // derivative is the explicit one at fixed dipoles: per ordered site
          // pair 1/2 qi qj dts0/da (Uelec), -qj (di . r) dts1/da (-mu . Efq)
          // and -1/2 di . dT/da . dj (-1/2 mu T mu); only the Thole terms
          // (r < 5 A, smear.f90:23,57) depend on a.
          if (d2.v < 25.0) {
            const bool cc = !same && qi.v != 0.0 && qj.v != 0.0, cd = !same && adip && qj.v != 0.0;
            const double av = P.apCC[ti][tj].v, rav = P.raCC[ti][tj].v;
            const Dual d2d = as_dual(d2);
            if ((cc || cd) && av > 0.0) {
              HD t0h, t1h, u1h, u2h, u3h;
              smear01<true, HD>(c, HD(d2d, Dual(0.0)), HD(Dual(av), Dual(1.0)),
                                HD(Dual(rav), Dual(rav / (3.0 * av))), t0h, t1h);
              smear3<true, HD>(c, HD(d2d, Dual(0.0)), HD(Dual(av), Dual(1.0)), u1h, u2h, u3h);
              Dual da0 = t0h.a, da1 = u1h.a;
              chain_d(da0, 0, u1h.a.v, d2d.d);
              chain_d(da1, 1, u2h.a.v, d2d.d);
              Dual w = 0.0;
              if (cc) w += 0.5 * qi * qj * da0;
              if (cd) w -= qj * (di[0] * r[0] + di[1] * r[1] + di[2] * r[2]) * da1;
              thole_acc(w, false, ti, tj);
            }
            if (adip && bdip) {
              HD u1h, u2h, u3h;
              smear3<true, HD>(c, HD(d2d, Dual(0.0)), HD(Dual(P.apDD[ti][tj].v), Dual(1.0)), u1h,
                               u2h, u3h);
              Dual da1 = u1h.a, da2 = u2h.a;
              chain_d(da1, 1, u2h.a.v, d2d.d);
              chain_d(da2, 2, u3h.a.v, d2d.d);
              const Dual dir_ = di[0] * r[0] + di[1] * r[1] + di[2] * r[2];
              const Dual djr = dj[0] * r[0] + dj[1] * r[1] + dj[2] * r[2];
              const Dual dd = di[0] * dj[0] + di[1] * dj[1] + di[2] * dj[2];
              thole_acc(-0.5 * (3.0 * da2 * dir_ * djr - da1 * dd), true, ti, tj);
            }
          }
        }
      }
    }
    for (int k = 0; k < 3; ++k) dR[3 * s + k] += g[k];
    phi[s] += pv;
  }
  __syncthreads();

  // ---- reciprocal space, dipole forces (ewald_std_qd, ewald.f90:123-226); the
  // dipole structure factor is the one of the last SOR iteration's input
  for (int s = tid / Gs; s < ns; s += TAN_BLOCK / Gs) {
    const int iw = s % N, ti = s / N;
    const bool sdip = ti >= tdip && ti < tend, schg = s >= N;
    const Dual qs = q[s];
    Dual mu[3] = {0.0, 0.0, 0.0};
    if (sdip)
      for (int k = 0; k < 3; ++k) mu[k] = dip[3 * ((ti - tdip) * N + iw) + k];
    Dual g[3] = {0.0, 0.0, 0.0}, pv = 0.0;
    for (int k = tid % Gs; k < nk; k += Gs) {
      const int4 h = a.khkl[k];
      const double4 kv = a.kv[k];
      const double p4vAk = kv.w;
      GT ck, sk;
      ph.at(s, h, kv, ck, sk);
      const Dual rQm = Sm[2 * k], iQm = Sm[2 * k + 1];
      if (schg) {
        pv += p4vAk * (ck * rQm + sk * iQm);
        const Dual t = qs * p4vAk * (ck * iQm - sk * rQm);
        g[0] += t * kv.x;
        g[1] += t * kv.y;
        g[2] += t * kv.z;
      }
      if (sdip) {
        const Dual dK = mu[0] * kv.x + mu[1] * kv.y + mu[2] * kv.z;
        const Dual t = dK * p4vAk * (ck * (Sq[2 * k] + rQm) + sk * (Sq[2 * k + 1] + iQm));
        g[0] -= t * kv.x;
        g[1] -= t * kv.y;
        g[2] -= t * kv.z;
      }
    }
    for (int k = 0; k < 3; ++k) g[k] = group_sum(g[k], Gs);
    pv = group_sum(pv, Gs);
    if (tid % Gs == 0) {
      for (int k = 0; k < 3; ++k) dR[3 * s + k] += g[k];
      phi[s] += pv;
    }
  }
  __syncthreads();

  // ---- charge chain rule (pot_ttm.f90:556-584), M-site redistribution
  // (pot_ttm.f90:597-602), loss terms
  double lossg = 0.0;
  for (int w = tid; w < N; w += TAN_BLOCK) {
    const Dual *gq = grdq + 27 * w;
    const Dual p1 = phi[N + w], p2 = phi[2 * N + w], pm = phi[3 * N + w];
    Dual G[3][3];  // atoms O, H1, H2
    for (int k = 0; k < 3; ++k) {
      const Dual dg1 = gq[0 * 9 + 0 * 3 + k] * p1 + gq[0 * 9 + 1 * 3 + k] * p2 + gq[0 * 9 + 2 * 3 + k] * pm;
      const Dual dg2 = gq[1 * 9 + 0 * 3 + k] * p1 + gq[1 * 9 + 1 * 3 + k] * p2 + gq[1 * 9 + 2 * 3 + k] * pm;
      const Dual dg0 = gq[2 * 9 + 0 * 3 + k] * p1 + gq[2 * 9 + 1 * 3 + k] * p2 + gq[2 * 9 + 2 * 3 + k] * pm;
      const Dual dM = dR[(3 * N + w) * 3 + k];
      G[1][k] = (dR[(N + w) * 3 + k] + dg1) + 0.5 * gM * dM;
      G[2][k] = (dR[(2 * N + w) * 3 + k] + dg2) + 0.5 * gM * dM;
      G[0][k] = (dR[w * 3 + k] + dg0) + (1.0 - gM) * dM;
    }
    const size_t a0 = ((size_t)b * 3 * N + 3 * w) * 3;
    for (int at = 0; at < 3; ++at)
      for (int k = 0; k < 3; ++k) {
        lossg += a.gw[a0 + 3 * at + k] * G[at][k].d;
        if (a.dG) a.dG[((size_t)b * ndir + dir) * 9 * N + 9 * w + 3 * at + k] = G[at][k].d;
      }
  }
  const double Ev = W.umon->v + block_sum(uvdw.v, red) + 0.5 * block_sum(uel.v, red) -
                    0.5 * block_sum(uind.v, red) + P.lrc.v;
  const double Ed = block_sum(uvdw.d, red) + 0.5 * block_sum(uel.d, red) - 0.5 * block_sum(uind.d, red) +
                    P.lrc.d;
  const double lg = block_sum(lossg, red);
  if (tid == 0) {
    double *o = a.out + ((size_t)b * ndir + dir) * 4;
    o[0] = Ev;
    o[1] = Ed;
    o[2] = a.ew[b] * Ed + lg;
    o[3] = (double)iters;
  }
  if constexpr (MODE == TAN_ADJOINT) adjoint_out(c, a, P, W, b, vE, vF, accCC, accDD, accPF, red);
}

void launch_ttm_tangent(const SysConst &c, const TanLaunch &L, hipStream_t s) {
  const size_t ds_off = 3 * 4 * (size_t)c.N + tangent_tab_doubles(c);
  const int mode = L.adjoint ? 2 : L.gamma_m ? 1 : 0;
  hipLaunchKernelGGL(k_tangent_prep, dim3(c.P), dim3(TAN_BLOCK), 0, s, c, L.RR, L.cfg,
                     tangent_cfg_doubles(c), ds_off, mode, L.gw);
  TanArgs a;
  a.RR = L.RR;
  a.pes = L.pes;
  a.kv = L.kv;
  a.khkl = L.khkl;
  a.cfg = L.cfg;
  a.cfg_stride = tangent_cfg_doubles(c);
  a.ws = (Dual *)L.ws;
  a.ws_stride = tangent_ws_duals(c);
  a.ew = L.ew;
  a.gw = L.gw;
  a.out = L.out;
  a.dG = L.dG;
  for (int i = 0; i < TTM_NPARAM; ++i) a.theta[i] = L.theta[i];
  for (int i = 0; i < 3; ++i) a.lrc_coef[i] = L.lrc_coef[i];
  a.dms_coef = L.dms_coef;
  a.gamma_m = L.gamma_m;
  a.ds_off = ds_off;
  a.adj = L.adj;
  if (mode == TAN_ADJOINT) {
    hipLaunchKernelGGL(k_tangent_monomer<TAN_ADJOINT>, dim3(c.P, 1), dim3(TAN_BLOCK), 0, s, c, a);
    hipLaunchKernelGGL(k_ttm_tangent<TAN_ADJOINT>, dim3(c.P, 1), dim3(TAN_BLOCK), 0, s, c, a);
  } else if (mode == TAN_GAMMA_M) {
    hipLaunchKernelGGL(k_tangent_monomer<TAN_GAMMA_M>, dim3(c.P, 1), dim3(TAN_BLOCK), 0, s, c, a);
    hipLaunchKernelGGL(k_ttm_tangent<TAN_GAMMA_M>, dim3(c.P, 1), dim3(TAN_BLOCK), 0, s, c, a);
  } else {
    hipLaunchKernelGGL(k_tangent_monomer<TAN_PARAMS>, dim3(c.P, L.ndir), dim3(TAN_BLOCK), 0, s, c, a);
    hipLaunchKernelGGL(k_ttm_tangent<TAN_PARAMS>, dim3(c.P, L.ndir), dim3(TAN_BLOCK), 0, s, c, a);
  }
}

}  // namespace pimd
