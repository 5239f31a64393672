// Device-side scalar kernels of the TTM model: the A&S erfc, the truncated
// incomplete gamma function, the Ewald-screened Thole-smeared interaction
// tensors and the Partridge-Schwenke monomer PES / dipole surface.
#pragma once
#include "pimd_internal.h"
#include "ps_tables.h"

#include <type_traits>

namespace pimd {

// Bit 0 of *flags is set when a molecule's coordinates are not finite; the
// host reads it at the next synchronising call (the monomer kernels test
// their molecule as they load it).
__device__ __forceinline__ void flag_nonfinite(int *flags, const double (&r)[3][3]) {
  bool ok = true;
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) ok = ok && __builtin_isfinite(r[a][k]);
  if (!ok && flags) atomicOr(flags, 1);
}


// Exponent tables of the PES and DMS bases, built at compile time into the
// constant memory of each unit that uses them (the PES and DMS coefficients
// are per context, in its parameter array).
struct PsTables {
  int idx[PS_NTERM][3];
  int idxd[PS_NDMS][3];
};
constexpr PsTables make_ps_tables() {
  PsTables t{};
  for (int j = 0; j < PS_NTERM; ++j) {
    t.idx[j][0] = PS_IDX1[j];
    t.idx[j][1] = PS_IDX2[j];
    t.idx[j][2] = PS_IDX3[j];
  }
  for (int j = 0; j < PS_NDMS; ++j) {
    t.idxd[j][0] = PS_IDXD1[j];
    t.idxd[j][1] = PS_IDXD2[j];
    t.idxd[j][2] = PS_IDXD3[j];
  }
  return t;
}
static __constant__ PsTables c_ps = make_ps_tables();

// Nearest integer for the minimum image.  The reference's anint rounds halves
// away from zero; rint rounds them to even.  The two differ only when r/box is
// exactly n + 1/2, which coordinates never produce, and rint is one
// instruction (v_rndne_f64) where round() is a compare-and-select sequence.
__device__ __forceinline__ double anint(double x) { return __builtin_rint(x); }

// Minimum image r - box*anint(r/box) (pot_ttm.f90:64 and throughout).  box*n
// is exact for the integer n in {-1, 0, 1}, so contraction cannot change it.
__device__ __forceinline__ void minimg(const SysConst &c, double &x, double &y, double &z) {
  x = x - c.box[0] * anint(x * c.boxi[0]);
  y = y - c.box[1] * anint(y * c.boxi[1]);
  z = z - c.box[2] * anint(z * c.boxi[2]);
}

// Split sums over the MONO_W waves of a block (one part per wave, one
// molecule per lane): partials through LDS, added in part order by every wave
// alike.  Every thread of the block must call it.
template <int W>
struct WaveSplit {
  double *buf;  // [W][8][64] doubles of LDS
  __device__ void operator()(double *v, int n) const {
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    for (int i = 0; i < n; ++i) buf[(g * 8 + i) * 64 + lane] = v[i];
    __syncthreads();
    for (int i = 0; i < n; ++i) {
      double s = buf[i * 64 + lane];
      for (int h = 1; h < W; ++h) s += buf[(h * 8 + i) * 64 + lane];
      v[i] = s;
    }
    __syncthreads();  // the buffer is reused by the next call
  }
};

// Sum over the G lanes of a lane group (G a power of two, groups aligned).
template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// Visits list[first], list[first + G], ... below `end` in order, loading
// each next index while the current one is visited (the list is in global
// memory; its load latency would otherwise stall every step).
template <int G, class F>
__device__ __forceinline__ void walk_list(const int *__restrict__ list, int first, int end, F &&f) {
  if (first >= end) return;
  int j = list[first];
  for (int n = first; n < end; n += G) {
    const int jn = n + G < end ? list[n + G] : 0;
    f(j);
    j = jn;
  }
}

// The molecule pairs of a bead split evenly between their two molecules:
// molecule i owns the pair (i, j) when (j - i) mod N lies in [1, (N-1)/2], or
// equals N/2 (N even) and i < j.  Returns the last j of i's window unwrapped,
// i + E(i), so that i owns the partners j in (i, i + E(i)] (mod N).
__host__ __device__ __forceinline__ int pair_half_window(int i, int N) {
  return i + (N - 1) / 2 + ((N % 2 == 0 && i < N / 2) ? 1 : 0);
}

// walk_list over two ranges of the list, [s0, e0) then [s1, e1), as one
// index space: entry n of it (n = first, first + G, ...) is list[s0 + n]
// below e0 - s0 and list[s1 + n - (e0 - s0)] beyond.
template <int G, class F>
__device__ __forceinline__ void walk_ranges(const int *__restrict__ list, int first, int s0, int e0,
                                            int s1, int e1, F &&f) {
  const int n0 = e0 - s0, end = n0 + (e1 - s1);
  auto at = [&](int n) { return list[n < n0 ? s0 + n : s1 + (n - n0)]; };
  if (first >= end) return;
  int j = at(first);
  for (int n = first; n < end; n += G) {
    const int jn = n + G < end ? at(n + G) : 0;
    f(j);
    j = jn;
  }
}

// walk_list with D entries in flight, for visits too short to cover one
// global load: load(at) fetches entry `at` of the list (the partner index and
// what the visit reads with it) D entries ahead; f(entry, at) visits them in
// walk_list's order.
template <int G, int D, class L, class F>
__device__ __forceinline__ void walk_deep(int first, int end, L &&load, F &&f) {
  using E = decltype(load(0));
  E q[D];
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (first + d * G < end) q[d] = load(first + d * G);
  for (int n = first; n < end; n += D * G) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const E e = q[d];
      const int at = n + d * G, ahead = at + D * G;
      if (ahead < end) q[d] = load(ahead);
      if (at < end) f(e, at);
    }
  }
}

// Makes v an opaque register value, so that the arithmetic that follows is
// compiled the same whether v was just computed or loaded from memory.
template <typename R>
__device__ __forceinline__ void opaque(R &v) {
  asm("" : "+v"(v));
}

__device__ __forceinline__ void vadd9(double *t, const double *v) {
  for (int i = 0; i < 9; ++i) t[i] += v[i];
}

// x*x + y*y + z*z rounded exactly as the reference (no fused multiply-add), so
// cut-off tests (r^2 < Rc^2, r^2 < 25) decide identically.
__device__ __forceinline__ double norm2_exact(double x, double y, double z) {
  return __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

// Elementary functions of the compute type R.  FP64 is the reference's
// arithmetic; FP32 (the mixed-precision pair terms) uses the hardware
// square root, reciprocal and the ocml exp / log.
template <typename R>
struct RM;
template <>
struct RM<double> {
  static __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
  static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  static __device__ __forceinline__ double exp_(double x) { return exp(x); }
  static __device__ __forceinline__ double log_(double x) { return log(x); }
  static __device__ __forceinline__ double abs_(double x) { return fabs(x); }
};
template <>
struct RM<float> {
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  // e^x = 2^(x log2 e) on the hardware exp2 (one instruction plus the
  // scaling; about 1 ulp, with the relative error of the scaled argument, below
  // 1e-6 for the arguments here, -(alpha r)^2 and -E r up to ~15 in size)
  static __device__ __forceinline__ float exp_(float x) {
    return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  }
  // ln x = log2(x) ln 2 on the hardware log2 (x > 0 here: Q(2/3, x) for x > 0)
  static __device__ __forceinline__ float log_(float x) {
    return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
  }
  static __device__ __forceinline__ float abs_(float x) { return fabsf(x); }
};

// Abramowitz-Stegun 7.1.26 complementary error function (math.f90:287-314)
// with the exp(-x^2) factor supplied by the caller.
template <typename R>
__device__ __forceinline__ R erfc_as(R x, R expmx2) {
  const R A1 = (R)0.254829592, A2 = (R)-0.284496736, A3 = (R)1.421413741, A4 = (R)-1.453152027,
          A5 = (R)1.061405429, P = (R)0.3275911;
  R T = RM<R>::rcp((R)1.0 + P * x);
  return T * (A1 + T * (A2 + T * (A3 + T * (A4 + T * A5)))) * expmx2;
}

// Regularised upper incomplete gamma Q(2/3, x) with the reference's series /
// continued-fraction split, tolerance and iteration caps (math.f90:188-285).
template <typename R>
__device__ inline R gammq23(R x) {
  const R a = (R)(2.0 / 3.0), gln = (R)0.303150275147547, EPS = (R)3e-7, FPMIN = (R)1e-30;
  if (x < a + (R)1.0) {
    if (x <= (R)0.0) return (R)1.0;
    R ap = a, sum = RM<R>::rcp(a), del = sum;
    for (int n = 1; n <= 800; ++n) {
      ap = ap + (R)1.0;
      del = RM<R>::div(del * x, ap);
      sum = sum + del;
      if (RM<R>::abs_(del) < RM<R>::abs_(sum) * EPS) break;
    }
    return (R)1.0 - sum * RM<R>::exp_(-x + a * RM<R>::log_(x) - gln);
  }
  R b = x + (R)1.0 - a, cc = RM<R>::rcp(FPMIN), d = RM<R>::rcp(b), h = d;
  for (int i = 1; i <= 100; ++i) {
    R an = -((R)i * ((R)i - a));
    b = b + (R)2.0;
    d = an * d + b;
    if (RM<R>::abs_(d) < FPMIN) d = FPMIN;
    cc = b + RM<R>::div(an, cc);
    if (RM<R>::abs_(cc) < FPMIN) cc = FPMIN;
    d = RM<R>::rcp(d);
    R del = d * cc;
    h = h * del;
    if (RM<R>::abs_(del - (R)1.0) < EPS) break;
  }
  return RM<R>::exp_(-x + a * RM<R>::log_(x) - gln) * h;
}

// Screened Coulomb derivatives shared by PBCsmear01/2/3 (smear.f90:2-107):
// ts_n = erfc-screened 1/r^(2n+1) terms plus the Thole correction for r < 5.
// One exp(-alpha^2 r^2) serves both the explicit Gaussian and the erfc factor.
template <typename R>
struct Smear {
  R dd, dri, drsqi, expon, er;
};

template <typename R>
__device__ __forceinline__ Smear<R> smear_base(const SysConst &c, R drsq) {
  Smear<R> s;
  s.dd = RM<R>::sqrt_(drsq);
  s.dri = RM<R>::rcp(s.dd);
  s.drsqi = s.dri * s.dri;
  s.expon = RM<R>::exp_(-((R)c.a2 * drsq));
  s.er = erfc_as<R>((R)c.aewald * s.dd, s.expon);
  return s;
}

// PBCsmear01 (smear.f90:2-34): ts0, ts1 for the charge-charge terms.
// THOLE = false is for pairs known to be beyond the 5 A damping radius, where
// the r^2 < 25 branch cannot be taken.
template <bool THOLE = true, typename R>
__device__ __forceinline__ void smear01(const SysConst &c, R drsq, R apol, R ra, R &ts0, R &ts1) {
  Smear<R> s = smear_base<R>(c, drsq);
  R t0 = s.er * s.dri;
  R t1 = (t0 + (R)c.cts1 * s.expon) * s.drsqi;
  if (THOLE && drsq < (R)25.0) {
    R rA3 = s.dd * drsq * apol;
    R exp1 = RM<R>::exp_(-rA3);
    R ch1 = (R)-1.0 * s.dri * s.drsqi;
    R sr1 = ((R)1.0 - exp1) * s.dri * s.drsqi;
    t1 = t1 + ch1 + sr1;
    t0 = t0 - exp1 * s.dri + ra * (R)1.3541179394264 * gammq23<R>(rA3);
  }
  ts0 = t0;
  ts1 = t1;
}

// PBCsmear2 (smear.f90:36-69): ts1, ts2 for the charge-dipole and
// dipole-dipole field tensors.
template <bool THOLE = true, typename R>
__device__ __forceinline__ void smear2(const SysConst &c, R drsq, R apol, R &ts1, R &ts2) {
  Smear<R> s = smear_base<R>(c, drsq);
  R t1 = (s.er * s.dri + (R)c.cts1 * s.expon) * s.drsqi;
  R t2 = (t1 + (R)c.cts2 * s.expon) * s.drsqi;
  if (THOLE && drsq < (R)25.0) {
    R rA3 = (s.dd * drsq) * apol;
    R exp1 = RM<R>::exp_(-rA3);
    R ch1 = (R)-1.0 * s.dri * s.drsqi, ch2 = ch1 * s.drsqi;
    R sr1 = ((R)1.0 - exp1) * s.dri * s.drsqi;
    R sr2 = (sr1 - exp1 * apol) * s.drsqi;
    t1 = t1 + ch1 + sr1;
    t2 = t2 + ch2 + sr2;
  }
  ts1 = t1;
  ts2 = t2;
}

// PBCsmear3 (smear.f90:71-107): ts1, ts2, ts3 for the dipole-dipole forces.
template <bool THOLE = true, typename R>
__device__ __forceinline__ void smear3(const SysConst &c, R drsq, R apol, R &ts1, R &ts2, R &ts3) {
  Smear<R> s = smear_base<R>(c, drsq);
  R t1 = (s.er * s.dri + (R)c.cts1 * s.expon) * s.drsqi;
  R t2 = (t1 + (R)c.cts2 * s.expon) * s.drsqi;
  R t3 = (t2 + (R)c.cts3 * s.expon) * s.drsqi;
  if (THOLE && drsq < (R)25.0) {
    R rA3 = (s.dd * drsq) * apol;
    R exp1 = RM<R>::exp_(-rA3);
    R ch1 = (R)-1.0 * s.dri * s.drsqi, ch2 = ch1 * s.drsqi, ch3 = ch2 * s.drsqi;
    R sr1 = ((R)1.0 - exp1) * s.dri * s.drsqi;
    R sr2 = (sr1 - exp1 * apol) * s.drsqi;
    R sr3 = (sr2 - (R)0.6 * exp1 * s.dd * apol * apol) * s.drsqi;
    t1 = t1 + ch1 + sr1;
    t2 = t2 + ch2 + sr2;
    t3 = t3 + ch3 + sr3;
  }
  ts1 = t1;
  ts2 = t2;
  ts3 = t3;
}

// self1 / self2 (smear.f90:110-156): screened minus bare Coulomb for the
// intramolecular (excluded) pairs.
__device__ __forceinline__ void self1(const SysConst &c, double drsq, double &ts0, double &ts1) {
  double dd = sqrt(drsq), drsqi = 1.0 / drsq, dri = sqrt(drsqi);
  double expon = exp(-(c.a2 * drsq));
  double kr0 = erfc_as(c.aewald * dd, expon) * dri;
  double kr1 = (kr0 + c.cts1 * expon) * drsqi;
  double ch0 = -dri, ch1 = ch0 * drsqi;
  ts0 = ch0 + kr0;
  ts1 = kr1 + ch1;
}

__device__ __forceinline__ void self2(const SysConst &c, double drsq, double &ts1, double &ts2) {
  double dd = sqrt(drsq), drsqi = 1.0 / drsq, dri = sqrt(drsqi);
  double expon = exp(-(c.a2 * drsq));
  double kr0 = erfc_as(c.aewald * dd, expon) * dri;
  double kr1 = (kr0 + c.cts1 * expon) * drsqi;
  double kr2 = (kr1 + c.cts2 * expon) * drsqi;
  double ch0 = -dri, ch1 = ch0 * drsqi, ch2 = ch1 * drsqi;
  ts1 = kr1 + ch1;
  ts2 = kr2 + ch2;
}

// get_dd3 (pot_ttm.f90:633-647) applied to a vector: out = (3 ts2 r r^T - ts1) v.
template <typename R>
__device__ __forceinline__ void dd3_apply(const R r[3], R ts1, R ts2, const R v[3], R out[3]) {
  R t = (R)3.0 * ts2;
  R rv = r[0] * v[0] + r[1] * v[1] + r[2] * v[2];
  out[0] = t * r[0] * rv - ts1 * v[0];
  out[1] = t * r[1] * rv - ts1 * v[1];
  out[2] = t * r[2] * rv - ts1 * v[2];
}

// Partridge-Schwenke monomer: energy and gradient (nasa.f90:2-103) and the
// dipole-moment-surface charges with their gradients (nasa.f90:105-293).
// r[atom][xyz] for O, H1, H2.  gq[disp][charge][xyz] with disp and charge in
// the order (H1, H2, O), as the reference's gradq.
// pes: the fitted PES parameters, [0, 245) the combined coefficients c5z of
// nasa.f90:45, then deohA, alphaoh, roh, phh1A, phh2 (PES_* indices).
// DMS = false: the potential energy surface only (q3, gq untouched).
// Split sums: with a reducer, part `part` of L takes every L-th polynomial
// term and red(v, n) must replace v[0..n) by the sums over the L parts, the
// same on every part (NoSplit: one part, the terms in the reference's order).
struct NoSplit {
  __device__ void operator()(double *, int) const {}
};
template <bool DMS = true, int L = 1, class Red = NoSplit>
__device__ inline void ps_monomer(const SysConst &c, const double r[3][3], double &e1,
                                  double dr1[3][3], double q3[3], double gq[3][3][3],
                                  const double *__restrict__ pes, int part = 0,
                                  const Red &red = Red()) {
  const double deohA = pes[PES_DEOHA], alphaoh = pes[PES_ALPHAOH], roh = pes[PES_ROH],
               phh1A = pes[PES_PHH1A], phh2 = pes[PES_PHH2];
  double R1[3], R2[3], RH[3];
  for (int k = 0; k < 3; ++k) {
    R1[k] = r[1][k] - r[0][k];
    R2[k] = r[2][k] - r[0][k];
    RH[k] = r[1][k] - r[2][k];
  }
  minimg(c, R1[0], R1[1], R1[2]);
  minimg(c, R2[0], R2[1], R2[2]);
  minimg(c, RH[0], RH[1], RH[2]);
  double d1 = sqrt(R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2]);
  double d2 = sqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
  double dh = sqrt(RH[0] * RH[0] + RH[1] * RH[1] + RH[2] * RH[2]);
  double costh = (R1[0] * R2[0] + R1[1] * R2[1] + R1[2] * R2[2]) / (d1 * d2);
  const double costhe = -.24780227221366464506;
  double x1 = (d1 - PS_REOH) / PS_REOH, x2 = (d2 - PS_REOH) / PS_REOH, x3 = costh - costhe;
  // powers x^(n-1) for n = 0..15 with x^(-1) := 0 (fmat of nasa.f90:61-67)
  double f1[16], f2[16], f3[16];
  f1[0] = f2[0] = f3[0] = 0.0;
  f1[1] = f2[1] = f3[1] = 1.0;
#pragma unroll
  for (int j = 2; j < 16; ++j) {
    f1[j] = f1[j - 1] * x1;
    f2[j] = f2[j - 1] * x2;
    f3[j] = f3[j - 1] * x3;
  }
  // ---- potential energy surface
  double deoh = PS_F5Z * deohA;
  double phh1 = PS_F5Z * phh1A;
  phh1 = phh1 * exp(phh2);
  double ex1 = exp(-alphaoh * (d1 - roh));
  double ex2 = exp(-alphaoh * (d2 - roh));
  double Va = deoh * (ex1 * (ex1 - 2.0) + ex2 * (ex2 - 2.0));
  double Vb = phh1 * exp(-phh2 * dh);
  double dVa1 = 2.0 * alphaoh * deoh * ex1 * (1.0 - ex1) / d1;
  double dVa2 = 2.0 * alphaoh * deoh * ex2 * (1.0 - ex2) / d2;
  double dVb = -phh2 * Vb / dh;
  double efac = exp(-PS_B1 * ((d1 - PS_REOH) * (d1 - PS_REOH) + (d2 - PS_REOH) * (d2 - PS_REOH)));
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  // the loop is unrolled, so the table indices are constants; a part skips
  // the other parts' terms with a branch that is uniform across the wave
#pragma unroll
  for (int j = 1; j < PS_NTERM; ++j) {
    if (L > 1 && (j - 1) % L != part) continue;
    int I = c_ps.idx[j][0], J = c_ps.idx[j][1], K = c_ps.idx[j][2];
    double cz = pes[j];
    double fI1 = f1[I], fJ2 = f2[J], fJ1 = f1[J], fI2 = f2[I], fK3 = f3[K];
    s0 = s0 + cz * (fI1 * fJ2 + fJ1 * fI2) * fK3;
    s1 = s1 + cz * ((double)(I - 1) * f1[I - 1] * fJ2 + (double)(J - 1) * f1[J - 1] * fI2) * fK3;
    s2 = s2 + cz * ((double)(J - 1) * fI1 * f2[J - 1] + (double)(I - 1) * fJ1 * f2[I - 1]) * fK3;
    s3 = s3 + cz * (fI1 * fJ2 + fJ1 * fI2) * (double)(K - 1) * f3[K - 1];
  }
  if (L > 1) {
    double v[4] = {s0, s1, s2, s3};
    red(v, 4);
    s0 = v[0];
    s1 = v[1];
    s2 = v[2];
    s3 = v[3];
  }
  double Vc = 2.0 * pes[0] + efac * s0;
  double e = Va + Vb + Vc;
  e = e + 0.44739574026257;
  e1 = e * 0.00285914375100642899;
  double dVcdr1 = (-2.0 * PS_B1 * efac * (d1 - PS_REOH) * s0 + efac * s1 / PS_REOH) / d1;
  double dVcdr2 = (-2.0 * PS_B1 * efac * (d2 - PS_REOH) * s0 + efac * s2 / PS_REOH) / d2;
  double dVcdcth = efac * s3;
  for (int k = 0; k < 3; ++k) {
    double g1 = dVa1 * R1[k] + dVb * RH[k] + dVcdr1 * R1[k] +
                dVcdcth * (R2[k] / (d1 * d2) - costh * R1[k] / (d1 * d1));
    double g2 = dVa2 * R2[k] - dVb * RH[k] + dVcdr2 * R2[k] +
                dVcdcth * (R1[k] / (d1 * d2) - costh * R2[k] / (d2 * d2));
    dr1[1][k] = g1 * .00285914375100642899;
    dr1[2][k] = g2 * .00285914375100642899;
    dr1[0][k] = -(g1 + g2) * .00285914375100642899;
  }
  if (!DMS) return;
  // ---- dipole moment surface
  const double ath0 = 1.82400520401572996557;
  const double B2A = 0.529177249;
  const double bfac = 1.0 / (double)0.529177249f;  // real(4) literal, nasa.f90:218
  double efd = exp(-DMS_B1D * ((d1 - PS_REOH) * (d1 - PS_REOH) + (d2 - PS_REOH) * (d2 - PS_REOH)));
  double P1 = 0, P2 = 0, PL1 = costh, PL2 = 0.5 * (3.0 * PL1 * PL1 - 1.0);
  double p1r1 = 0, p1r2 = 0, p1c = 0, p2r1 = 0, p2r2 = 0, p2c = 0;
#pragma unroll
  for (int j = 1; j < PS_NDMS; ++j) {
    if (L > 1 && (j - 1) % L != part) continue;
    int I = c_ps.idxd[j][0], J = c_ps.idxd[j][1], K = c_ps.idxd[j][2];
    double cj = pes[PES_DMSC + j];
    P1 = P1 + cj * f1[I] * f2[J] * f3[K];
    P2 = P2 + cj * f1[J] * f2[I] * f3[K];
    p1r1 = p1r1 + cj * ((double)(I - 1) * f1[I - 1] * f2[J] * f3[K]);
    p1r2 = p1r2 + cj * ((double)(J - 1) * f1[I] * f2[J - 1] * f3[K]);
    p1c = p1c + cj * ((double)(K - 1) * f1[I] * f2[J] * f3[K - 1]);
    p2r1 = p2r1 + cj * ((double)(J - 1) * f1[J - 1] * f2[I] * f3[K]);
    p2r2 = p2r2 + cj * ((double)(I - 1) * f1[J] * f2[I - 1] * f3[K]);
    p2c = p2c + cj * ((double)(K - 1) * f1[J] * f2[I] * f3[K - 1]);
  }
  if (L > 1) {
    double v[8] = {P1, P2, p1r1, p1r2, p1c, p2r1, p2r2, p2c};
    red(v, 8);
    P1 = v[0];
    P2 = v[1];
    p1r1 = v[2];
    p1r2 = v[3];
    p1c = v[4];
    p2r1 = v[5];
    p2r2 = v[6];
    p2c = v[7];
  }
  p1r1 = p1r1 / PS_REOH * B2A;
  p1r2 = p1r2 / PS_REOH * B2A;
  p2r1 = p2r1 / PS_REOH * B2A;
  p2r2 = p2r2 / PS_REOH * B2A;
  double pw1 = pow(d1, DMS_B), pw2 = pow(d2, DMS_B);
  double poly = DMS_C0 + PL1 * DMS_C1 + PL2 * DMS_C2;
  double PC0 = DMS_A * (pw1 + pw2) * poly;
  double dpc0dr1 = DMS_A * (DMS_B * pow(d1, DMS_B - 1)) * poly * B2A * B2A;
  double dpc0dr2 = DMS_A * (DMS_B * pow(d2, DMS_B - 1)) * poly * B2A * B2A;
  double dpc0dc = DMS_A * (pw1 + pw2) * (DMS_C1 + 0.5 * (6.0 * PL1) * DMS_C2) * B2A;
  double defdr1 = -2.0 * DMS_B1D * (d1 - PS_REOH) * efd * B2A;
  double defdr2 = -2.0 * DMS_B1D * (d2 - PS_REOH) * efd * B2A;
  p1r1 = p1r1 * efd + P1 * defdr1 + dpc0dr1;
  p1r2 = p1r2 * efd + P1 * defdr2 + dpc0dr2;
  p1c = p1c * efd + dpc0dc;
  p2r1 = p2r1 * efd + P2 * defdr1 + dpc0dr1;
  p2r2 = p2r2 * efd + P2 * defdr2 + dpc0dr2;
  p2c = p2c * efd + dpc0dc;
  P1 = pes[PES_DMSC] + P1 * efd + PC0 * B2A;
  P2 = pes[PES_DMSC] + P2 * efd + PC0 * B2A;
  q3[0] = -(P1 + P2);
  q3[1] = P1;
  q3[2] = P2;
  p1r1 = p1r1 * bfac;
  p1r2 = p1r2 * bfac;
  p2r1 = p2r1 * bfac;
  p2r2 = p2r2 * bfac;
  double ax = R1[1] * R2[2] - R1[2] * R2[1];
  double ay = -R1[0] * R2[2] + R1[2] * R2[0];
  double az = R1[0] * R2[1] - R1[1] * R2[0];
  double sinth = sqrt(ax * ax + ay * ay + az * az) / (d1 * d2);
  double ang = atan2(sinth, costh);
  P1 = c.dms1 * (d1 - c.dms2) + c.dms3 * (ang - ath0);
  P2 = c.dms1 * (d2 - c.dms2) + c.dms3 * (ang - ath0);
  q3[0] = q3[0] - (P1 + P2);
  q3[1] = q3[1] + P1;
  q3[2] = q3[2] + P2;
  p1r1 = p1r1 + c.dms1;
  p2r2 = p2r2 + c.dms1;
  p1c = p1c - c.dms3 / sinth;
  p2c = p2c - c.dms3 / sinth;
  double f1q1r13 = (p1r1 - (p1c * costh / d1)) / d1;
  double f1q1r23 = p1c / (d1 * d2);
  double f2q1r23 = (p1r2 - (p1c * costh / d2)) / d2;
  double f2q1r13 = p1c / (d2 * d1);
  double f1q2r13 = (p2r1 - (p2c * costh / d1)) / d1;
  double f1q2r23 = p2c / (d1 * d2);
  double f2q2r23 = (p2r2 - (p2c * costh / d2)) / d2;
  double f2q2r13 = p2c / (d2 * d1);
  for (int k = 0; k < 3; ++k) {
    gq[0][0][k] = f1q1r13 * R1[k] + f1q1r23 * R2[k];
    gq[1][0][k] = f2q1r13 * R1[k] + f2q1r23 * R2[k];
    gq[2][0][k] = -(gq[0][0][k] + gq[1][0][k]);
    gq[0][1][k] = f1q2r13 * R1[k] + f1q2r23 * R2[k];
    gq[1][1][k] = f2q2r13 * R1[k] + f2q2r23 * R2[k];
    gq[2][1][k] = -(gq[0][1][k] + gq[1][1][k]);
    gq[0][2][k] = -(gq[0][0][k] + gq[0][1][k]);
    gq[1][2][k] = -(gq[1][0][k] + gq[1][1][k]);
    gq[2][2][k] = -(gq[2][0][k] + gq[2][1][k]);
  }
}

}  // namespace pimd
