// Ring-polymer propagation kernels: normal-mode free-ring step
// (NormalModes.f90:34-63), PILE Langevin thermostat in normal-mode space
// (Langevin.f90:30-64), velocity-Verlet kicks, centroid periodic wrap
// (main_stuff.f90:107-125), kinetic energy (main_stuff.f90:349-388) and the
// global Nose-Hoover chain (Nose_Hoover.f90:33-112).
//
// Layout: RRt/PPt [P][Natoms][3]; C [mode][bead]; A [4][P] = A11, A12, A21, A22.
// The mode transforms work on LDS slabs of P beads x CB (atom, component)
// columns with C resident in LDS.
#include <algorithm>
#include <cstdlib>

#include "pimd_internal.h"

namespace pimd {

// Columns (atom, component) per block so that one P x CB slab is 8 KB (the
// four slabs and C of P = 64 then fill 64 KB of LDS).
__host__ __device__ constexpr int nm_cols(int P) { return 1024 / P; }

// real2NM / NM2real (NormalModes.f90:308-341) around the exact harmonic
// propagator of each free-ring mode, for a slab of CB (atom, component)
// columns: X (P x CB) is staged in LDS, Y = C X, the 2x2 map is applied per
// mode, and X = C^T Y is written back.  Sums run over beads in the
// reference's order.
__global__ __launch_bounds__(256) void k_nm_evolve(int P, int natoms, const double *__restrict__ Cg,
                                                  const double *__restrict__ A,
                                                  const double *__restrict__ msf,
                                                  const double *__restrict__ masses, double *RR,
                                                  double *PP, const double *__restrict__ kick,
                                                  double delt2, const double *__restrict__ kick2,
                                                  double delt2b) {
  extern __shared__ double lds[];
  const int CB = nm_cols(P);
  double *Cs = lds, *xr = Cs + P * P, *xp = xr + P * CB, *yr = xp + P * CB, *yp = yr + P * CB;
  const int ncols = natoms * 3, col0 = blockIdx.x * CB;
  const size_t stride = (size_t)natoms * 3;
  for (int e = threadIdx.x; e < P * P; e += blockDim.x) Cs[e] = Cg[e];
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int k = e / CB, cc = e % CB, col = col0 + cc;
    xr[e] = col < ncols ? RR[k * stride + col] : 0.0;
    double p = col < ncols ? PP[k * stride + col] : 0.0;
    // with kick: the half kick PPt - MASSCON*dRRt*delt2 before the step
    // (MD.f90:49), as k_kick forms it
    if (kick && col < ncols) p = p - MASSCON * kick[k * stride + col] * delt2;
    if (kick2 && col < ncols) p = p - MASSCON * kick2[k * stride + col] * delt2b;
    xp[e] = p;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int j = e / CB, cc = e % CB, col = col0 + cc;
    double vr = 0.0, vp = 0.0;
    for (int k = 0; k < P; ++k) {
      vr = vr + Cs[j * P + k] * xr[k * CB + cc];
      vp = vp + Cs[j * P + k] * xp[k * CB + cc];
    }
    const double m = col < ncols ? masses[col / 3] : 1.0;
    const double t2 = A[j] * vr + A[P + j] * vp / (m * msf[j]);
    yp[e] = A[2 * P + j] * m * msf[j] * vr + A[3 * P + j] * vp;
    yr[e] = t2;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int j = e / CB, cc = e % CB, col = col0 + cc;
    if (col >= ncols) continue;
    double vr = 0.0, vp = 0.0;
    for (int k = 0; k < P; ++k) {
      vr = vr + Cs[k * P + j] * yr[k * CB + cc];
      vp = vp + Cs[k * P + j] * yp[k * CB + cc];
    }
    RR[j * stride + col] = vr;
    PP[j * stride + col] = vp;
  }
}

static bool nm_mfma();
static bool langevin_mfma();
static void launch_nm_evolve_mfma(int P, int natoms, DevBufs &b, double *RR, double *PP,
                                  hipStream_t s, const double *kick, double delt2,
                                  const double *kick2, double delt2b);

void launch_nm_evolve(int P, int natoms, DevBufs &b, double *RR, double *PP, hipStream_t s,
                      const double *kick, double delt2, const double *kick2, double delt2b) {
  if (nm_mfma() && P <= 64) {
    launch_nm_evolve_mfma(P, natoms, b, RR, PP, s, kick, delt2, kick2, delt2b);
    return;
  }
  const int CB = nm_cols(P), n = natoms * 3;
  const size_t lds = sizeof(double) * ((size_t)P * P + 4 * (size_t)P * CB);
  hipLaunchKernelGGL(k_nm_evolve, dim3((n + CB - 1) / CB), dim3(256), lds, s, P, natoms, b.C,
                     b.A, b.msf, b.masses, RR, PP, kick, delt2, kick2, delt2b);
}

// ---------------------------------------------------------------------------
// The same transforms on FP64 matrix cores.  v_mfma_f64_16x16x4_f64 adds
// A (16 x 4) B (4 x 16) to D (16 x 16) across one wave: lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16], and D[l / 16 + 4 i][l % 16] in
// element i (rows interleaved over the four lane groups).  A block is one wave and 16 (atom, component) columns; P is
// padded with zeros to PT (a multiple of 16) modes and beads, C sits in LDS
// with rows padded by one double (no bank conflicts between the 16 rows a
// wave reads at once), and the mode-space slab goes through LDS between the
// two transforms (the D layout of one product is not the B layout of the
// next).  Products are summed in the matrix cores' order, not the
// reference's; the difference is rounding.
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d d) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
}

constexpr int NM_TILE = 16;  // columns per block

// C (P x P) into LDS as PT x PT with row stride PT + 1, zero padded.
template <int PT>
__device__ __forceinline__ void stage_C(double *Cs, const double *__restrict__ Cg, int P) {
  for (int e = threadIdx.x; e < PT * PT; e += blockDim.x) {
    const int m = e / PT, k = e % PT;
    Cs[m * (PT + 1) + k] = (m < P && k < P) ? Cg[m * P + k] : 0.0;
  }
}

// Y[rt] = C[rt rows] X for the RT row tiles, X given as B operands x[kk]
// (bead kk*4 + l/16, column l%16).
template <int PT>
__device__ __forceinline__ void mfma_fwd(const double *Cs, const double *x, v4d *y) {
  constexpr int KS = PT / 4, RT = PT / 16;
  const int j = threadIdx.x & 15, q = (threadIdx.x >> 4) & 3;
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    v4d d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) d = mfma_f64(Cs[(rt * 16 + j) * (PT + 1) + kk * 4 + q], x[kk], d);
    y[rt] = d;
  }
}

// X[rt] = C^T[rt rows] Y with Y a PT x 16 slab in LDS (row stride 17).
template <int PT>
__device__ __forceinline__ void mfma_bwd(const double *Cs, const double *Ys, v4d *xo) {
  constexpr int KS = PT / 4, RT = PT / 16;
  const int j = threadIdx.x & 15, q = (threadIdx.x >> 4) & 3;
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    v4d d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < KS; ++kk)
      d = mfma_f64(Cs[(kk * 4 + q) * (PT + 1) + rt * 16 + j], Ys[(kk * 4 + q) * 17 + j], d);
    xo[rt] = d;
  }
}

template <int PT>
__global__ __launch_bounds__(64) void k_nm_evolve_mfma(int P, int natoms,
                                                       const double *__restrict__ Cg,
                                                       const double *__restrict__ A,
                                                       const double *__restrict__ msf,
                                                       const double *__restrict__ masses,
                                                       double *RR, double *PP,
                                                       const double *__restrict__ kick,
                                                       double delt2,
                                                       const double *__restrict__ kick2,
                                                       double delt2b) {
  constexpr int KS = PT / 4, RT = PT / 16;
  __shared__ double Cs[PT * (PT + 1)], Yr[PT * 17], Yp[PT * 17];
  const int ncols = natoms * 3;
  const size_t stride = (size_t)ncols;
  const int j = threadIdx.x & 15, q = threadIdx.x >> 4;
  const int col = blockIdx.x * NM_TILE + j;
  const bool vc = col < ncols;
  stage_C<PT>(Cs, Cg, P);
  double xr[KS], xp[KS];
#pragma unroll
  for (int kk = 0; kk < KS; ++kk) {
    const int bead = kk * 4 + q;
    const bool in = vc && bead < P;
    const size_t o = (size_t)bead * stride + col;
    xr[kk] = in ? RR[o] : 0.0;
    double p = in ? PP[o] : 0.0;
    // the half kicks before the step (MD.f90:49), rounded as k_kick does
    if (kick && in) p = p - MASSCON * kick[o] * delt2;
    if (kick2 && in) p = p - MASSCON * kick2[o] * delt2b;
    xp[kk] = p;
  }
  __syncthreads();
  v4d yr[RT], yp[RT];
  mfma_fwd<PT>(Cs, xr, yr);
  mfma_fwd<PT>(Cs, xp, yp);
  // exact harmonic propagator of each mode (NormalModes.f90:34-63)
  const double m = vc ? masses[col / 3] : 1.0;
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int mode = rt * 16 + q + 4 * i;
      double nr = 0.0, np = 0.0;
      if (mode < P) {
        const double vr = yr[rt][i], vp = yp[rt][i];
        nr = A[mode] * vr + A[P + mode] * vp / (m * msf[mode]);
        np = A[2 * P + mode] * m * msf[mode] * vr + A[3 * P + mode] * vp;
      }
      Yr[mode * 17 + j] = nr;
      Yp[mode * 17 + j] = np;
    }
  __syncthreads();
  v4d orr[RT], opp[RT];
  mfma_bwd<PT>(Cs, Yr, orr);
  mfma_bwd<PT>(Cs, Yp, opp);
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int bead = rt * 16 + q + 4 * i;
      if (vc && bead < P) {
        RR[(size_t)bead * stride + col] = orr[rt][i];
        PP[(size_t)bead * stride + col] = opp[rt][i];
      }
    }
}

// PT: P rounded up to a multiple of 16 (P <= 64).
static int nm_pt(int P) { return ((P + 15) / 16) * 16; }

// The free ring-polymer step runs on the FP64 matrix cores unless
// PIMD_NM_MFMA=0 (13 us against 21 us at 512 x 32).  The Langevin step keeps
// the FMA slabs unless PIMD_LANGEVIN_MFMA=1: its time goes to the normal
// deviates, which the slab layout spreads over twice as many lanes (18 us
// against 20 us).
static int env_flag(const char *name, int dflt) {
  const char *v = std::getenv(name);
  return (v && *v) ? std::atoi(v) : dflt;
}
static bool nm_mfma() { return env_flag("PIMD_NM_MFMA", 1) != 0; }
static bool langevin_mfma() { return env_flag("PIMD_LANGEVIN_MFMA", 0) != 0; }

static void launch_nm_evolve_mfma(int P, int natoms, DevBufs &b, double *RR, double *PP,
                                  hipStream_t s, const double *kick, double delt2,
                                  const double *kick2, double delt2b) {
  const int n = natoms * 3;
  const dim3 grid((n + NM_TILE - 1) / NM_TILE);
  switch (nm_pt(P)) {
    case 16:
      hipLaunchKernelGGL(k_nm_evolve_mfma<16>, grid, dim3(64), 0, s, P, natoms, b.C, b.A, b.msf,
                         b.masses, RR, PP, kick, delt2, kick2, delt2b);
      break;
    case 32:
      hipLaunchKernelGGL(k_nm_evolve_mfma<32>, grid, dim3(64), 0, s, P, natoms, b.C, b.A, b.msf,
                         b.masses, RR, PP, kick, delt2, kick2, delt2b);
      break;
    case 48:
      hipLaunchKernelGGL(k_nm_evolve_mfma<48>, grid, dim3(64), 0, s, P, natoms, b.C, b.A, b.msf,
                         b.masses, RR, PP, kick, delt2, kick2, delt2b);
      break;
    default:
      hipLaunchKernelGGL(k_nm_evolve_mfma<64>, grid, dim3(64), 0, s, P, natoms, b.C, b.A, b.msf,
                         b.masses, RR, PP, kick, delt2, kick2, delt2b);
      break;
  }
}

// Philox4x32-10 counter-based generator (J. K. Salmon et al., "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), written from the published round
// function, and a Box-Muller transform to standard normals.
__device__ __forceinline__ void philox_round(uint32_t ctr[4], const uint32_t key[2]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint64_t p0 = (uint64_t)M0 * ctr[0], p1 = (uint64_t)M1 * ctr[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ ctr[1] ^ key[0], n2 = hi0 ^ ctr[3] ^ key[1];
  ctr[0] = n0;
  ctr[1] = lo1;
  ctr[2] = n2;
  ctr[3] = lo0;
}

__device__ __forceinline__ void philox4x32(uint32_t ctr[4], uint32_t key[2]) {
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    philox_round(ctr, key);
    key[0] += W0;
    key[1] += W1;
  }
}

// Two standard normals for stream index idx under (seed, counter).
__device__ __forceinline__ void normal2(uint64_t seed, uint64_t counter, uint64_t idx, double &z0,
                                        double &z1) {
  uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)counter,
                     (uint32_t)(counter >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  philox4x32(ctr, key);
  const double two53 = 1.0 / 9007199254740992.0;
  const uint64_t a = ((uint64_t)ctr[0] << 21) ^ (ctr[1] >> 11);
  const uint64_t bb = ((uint64_t)ctr[2] << 21) ^ (ctr[3] >> 11);
  const double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * two53;  // (0, 1]
  const double u2 = (double)(bb & ((1ull << 53) - 1)) * two53;          // [0, 1)
  const double rr = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincos(TWOPI * u2, &sn, &cs);
  z0 = rr * cs;
  z1 = rr * sn;
}

// Langevin_NM: p~_j <- c1_j p~_j + c2_j sqrt(m s_j) xi (Langevin.f90:41-60),
// on the same LDS slabs as k_nm_evolve.  noise [Natoms][P][3] (the
// reference's consumption order) or drawn from (seed, counter).
__global__ __launch_bounds__(256) void k_langevin(int P, int natoms, const double *__restrict__ Cg,
                                                 const double *__restrict__ msf,
                                                 const double *__restrict__ c1,
                                                 const double *__restrict__ c2, double massO,
                                                 double massH, double *PP,
                                                 const double *__restrict__ noise, uint64_t seed,
                                                 uint64_t counter, const int *__restrict__ clk,
                                                 int half, const double *__restrict__ kick,
                                                 double delt2) {
  extern __shared__ double lds[];
  // in MD steps the counter follows the device step clock: step t, half 0 / 1
  if (clk) counter = (uint64_t)clk[0] * 2 + half;
  const int CB = nm_cols(P);
  double *Cs = lds, *xp = Cs + P * P, *yp = xp + P * CB;
  const int ncols = natoms * 3, col0 = blockIdx.x * CB;
  const size_t stride = (size_t)natoms * 3;
  for (int e = threadIdx.x; e < P * P; e += blockDim.x) Cs[e] = Cg[e];
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int k = e / CB, cc = e % CB, col = col0 + cc;
    double p = col < ncols ? PP[k * stride + col] : 0.0;
    // with kick: the half kick PPt - MASSCON*dRRt*delt2 first (MD.f90:94)
    if (kick && col < ncols) p = p - MASSCON * kick[k * stride + col] * delt2;
    xp[e] = p;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int j = e / CB, cc = e % CB, col = col0 + cc;
    double v = 0.0;
    for (int k = 0; k < P; ++k) v = v + Cs[j * P + k] * xp[k * CB + cc];
    if (col < ncols) {
      const int a = col / 3, comp = col % 3;
      const double m = (a % 3 == 0) ? massO : massH;
      const uint64_t n = ((uint64_t)a * P + j) * 3 + comp;
      double xi;
      if (noise) {
        xi = noise[n];
      } else {
        double z0, z1;
        normal2(seed, counter, n >> 1, z0, z1);
        xi = (n & 1) ? z1 : z0;
      }
      const double sm = sqrt(m * msf[j]);
      v = c1[j] * v + c2[j] * (sm * xi);
    }
    yp[e] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int j = e / CB, cc = e % CB, col = col0 + cc;
    if (col >= ncols) continue;
    double v = 0.0;
    for (int k = 0; k < P; ++k) v = v + Cs[k * P + j] * yp[k * CB + cc];
    PP[j * stride + col] = v;
  }
}

// k_langevin on FP64 matrix cores (layout as k_nm_evolve_mfma).
template <int PT>
__global__ __launch_bounds__(64) void k_langevin_mfma(int P, int natoms,
                                                      const double *__restrict__ Cg,
                                                      const double *__restrict__ msf,
                                                      const double *__restrict__ c1,
                                                      const double *__restrict__ c2, double massO,
                                                      double massH, double *PP,
                                                      const double *__restrict__ noise,
                                                      uint64_t seed, uint64_t counter,
                                                      const int *__restrict__ clk, int half,
                                                      const double *__restrict__ kick,
                                                      double delt2) {
  constexpr int KS = PT / 4, RT = PT / 16;
  __shared__ double Cs[PT * (PT + 1)], Yp[PT * 17];
  if (clk) counter = (uint64_t)clk[0] * 2 + half;
  const int ncols = natoms * 3;
  const size_t stride = (size_t)ncols;
  const int j = threadIdx.x & 15, q = threadIdx.x >> 4;
  const int col = blockIdx.x * NM_TILE + j;
  const bool vc = col < ncols;
  stage_C<PT>(Cs, Cg, P);
  double xp[KS];
#pragma unroll
  for (int kk = 0; kk < KS; ++kk) {
    const int bead = kk * 4 + q;
    const bool in = vc && bead < P;
    const size_t o = (size_t)bead * stride + col;
    double p = in ? PP[o] : 0.0;
    // with kick: the half kick PPt - MASSCON*dRRt*delt2 first (MD.f90:94)
    if (kick && in) p = p - MASSCON * kick[o] * delt2;
    xp[kk] = p;
  }
  __syncthreads();
  v4d yp[RT];
  mfma_fwd<PT>(Cs, xp, yp);
  const int a = col / 3, comp = col % 3;
  const double m = (a % 3 == 0) ? massO : massH;
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int mode = rt * 16 + q + 4 * i;
      double v = 0.0;
      if (mode < P) {
        v = yp[rt][i];
        if (vc) {
          const uint64_t n = ((uint64_t)a * P + mode) * 3 + comp;
          double xi;
          if (noise) {
            xi = noise[n];
          } else {
            double z0, z1;
            normal2(seed, counter, n >> 1, z0, z1);
            xi = (n & 1) ? z1 : z0;
          }
          const double sm = sqrt(m * msf[mode]);
          v = c1[mode] * v + c2[mode] * (sm * xi);
        }
      }
      Yp[mode * 17 + j] = v;
    }
  __syncthreads();
  v4d op[RT];
  mfma_bwd<PT>(Cs, Yp, op);
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int bead = rt * 16 + q + 4 * i;
      if (vc && bead < P) PP[(size_t)bead * stride + col] = op[rt][i];
    }
}

void launch_langevin(int P, int natoms, DevBufs &b, double *PP, const double *noise,
                     uint64_t seed, uint64_t counter, double massO, double massH,
                     hipStream_t s, const int *clk, int half, const double *kick, double delt2) {
  if (langevin_mfma() && P <= 64) {
    const dim3 grid((natoms * 3 + NM_TILE - 1) / NM_TILE);
#define PIMD_LANG_MFMA(PT)                                                                    \
  hipLaunchKernelGGL(k_langevin_mfma<PT>, grid, dim3(64), 0, s, P, natoms, b.C, b.msf, b.c1, \
                     b.c2, massO, massH, PP, noise, seed, counter, clk, half, kick, delt2)
    switch (nm_pt(P)) {
      case 16: PIMD_LANG_MFMA(16); break;
      case 32: PIMD_LANG_MFMA(32); break;
      case 48: PIMD_LANG_MFMA(48); break;
      default: PIMD_LANG_MFMA(64); break;
    }
#undef PIMD_LANG_MFMA
    return;
  }
  const int CB = nm_cols(P), n = natoms * 3;
  const size_t lds = sizeof(double) * ((size_t)P * P + 2 * (size_t)P * CB);
  hipLaunchKernelGGL(k_langevin, dim3((n + CB - 1) / CB), dim3(256), lds, s, P, natoms, b.C,
                     b.msf, b.c1, b.c2, massO, massH, PP, noise, seed, counter, clk, half, kick,
                     delt2);
}

// The device step clock {MD step t, record slot} read by the force and
// thermostat kernels of an MD step, so that one captured step replays for
// every step.
__global__ void k_set_clock(int *clk, int t, int slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    clk[0] = t;
    clk[1] = slot;
  }
}
// End of an MD step: its record takes the final kinetic energy (rec[3]) when
// rec0 is non-NULL, then the clock moves to the next step and record slot.
__global__ void k_advance_clock(int *clk, const double *__restrict__ uk, double *rec0) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (rec0) rec0[(size_t)clk[1] * MD_REC + 3] = uk[0];
    clk[0] += 1;
    clk[1] += 1;
  }
}
void launch_set_clock(int *clk, int t, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(64), 0, s, clk, t, slot);
}
void launch_advance_clock(int *clk, hipStream_t s, const double *uk, double *rec0) {
  hipLaunchKernelGGL(k_advance_clock, dim3(1), dim3(64), 0, s, clk, uk, rec0);
}

// PPt = PPt - MASSCON*dRRt*delt2 (MD.f90:49, 94).
__global__ void k_kick(int n, double *PP, const double *__restrict__ dRR, double delt2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) PP[i] = PP[i] - MASSCON * dRR[i] * delt2;
}

void launch_kick(int n, double *PP, const double *dRR, double delt2, hipStream_t s) {
  hipLaunchKernelGGL(k_kick, dim3((n + 255) / 256), dim3(256), 0, s, n, PP, dRR, delt2);
}

// PPt = PPt*s with s on the device (MD.f90:45, 105).
__global__ void k_scale(int n, double *PP, const double *__restrict__ sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) PP[i] = PP[i] * sc[0];
}

void launch_scale(int n, double *PP, const double *sc, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3((n + 255) / 256), dim3(256), 0, s, n, PP, sc);
}

// Free flight for P = 1 (MD.f90:60-68).
__global__ void k_free_flight(int n, double *RR, const double *__restrict__ PP,
                              const double *__restrict__ masses, double delt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) RR[i] = RR[i] + (1.0 / masses[i / 3]) * PP[i] * delt;
}

void launch_free_flight(int n, double *RR, const double *PP, const double *masses, double delt,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_free_flight, dim3((n + 255) / 256), dim3(256), 0, s, n, RR, PP, masses,
                     delt);
}

// RRc = sum(RRt,3)/Nbeads, then the centroid wrap of PBCs (MD.f90:70;
// main_stuff.f90:107-125): shifts use box(1) on every axis.
__global__ void k_centroid_pbc(int P, int natoms, double box0, double boxi0, double *RR,
                               double *RRc) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= natoms * 3) return;
  const size_t stride = (size_t)natoms * 3;
  // beads summed in order 1..P; loads issued 8 at a time
  double s = 0.0;
  int k = 0;
  for (; k + 8 <= P; k += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = RR[(k + u) * stride + id];
#pragma unroll
    for (int u = 0; u < 8; ++u) s = s + v[u];
  }
  for (; k < P; ++k) s = s + RR[k * stride + id];
  const double c = s / P;
  const double sh = box0 * round(c * boxi0);
  RRc[id] = c - sh;
  if (sh == 0.0) return;  // x - 0 = x
  for (k = 0; k < P; ++k) RR[k * stride + id] = RR[k * stride + id] - sh;
}

void launch_centroid_pbc(int P, int natoms, double box0, double boxi0, double *RR, double *RRc,
                         hipStream_t s) {
  const int n = natoms * 3;
  hipLaunchKernelGGL(k_centroid_pbc, dim3((n + 63) / 64), dim3(64), 0, s, P, natoms, box0,
                     boxi0, RR, RRc);
}

// uk = 1/2 sum_beads sum_atoms |p|^2 / m (calc_uk, RPMD branch): per-block
// partial sums, then one block adds them in a fixed order.
constexpr int KIN_BLOCKS = 256;
__global__ __launch_bounds__(256) void k_kinetic_part(int P, int natoms,
                                                     const double *__restrict__ masses,
                                                     const double *__restrict__ PP, double *part) {
  __shared__ double red[256];
  const size_t n = (size_t)P * natoms;
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const int a = (int)(i % natoms);
    const double *p = PP + i * 3;
    s += (1.0 / masses[a]) * (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_kinetic_final(const double *__restrict__ part, int nb,
                                                      double *uk) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) uk[0] = 0.5 * red[0];
}

void launch_kinetic(int P, int natoms, const double *masses, const double *PP, double *uk,
                    double *part, hipStream_t s) {
  const size_t n = (size_t)P * natoms;
  const int nb = (int)std::min<size_t>(KIN_BLOCKS, (n + 255) / 256);
  hipLaunchKernelGGL(k_kinetic_part, dim3(nb), dim3(256), 0, s, P, natoms, masses, PP, part);
  hipLaunchKernelGGL(k_kinetic_final, dim3(1), dim3(256), 0, s, part, nb, uk);
}

// calc_uk with centroid-MD mass scaling (main_stuff.f90:366-385): per atom
// the normal-mode momenta p~_j = sum_k C[j][k] p_k (real2NM,
// NormalModes.f90:308-322), uk = 1/2 sum_atoms sum_j |p~_j|^2 / (m msf_j).
// One thread per atom, per-block partials, then the fixed-order final sum.
__global__ __launch_bounds__(256) void k_kinetic_nm_part(int P, int natoms,
                                                        const double *__restrict__ Cg,
                                                        const double *__restrict__ msf,
                                                        const double *__restrict__ masses,
                                                        const double *__restrict__ PP, double *part) {
  __shared__ double red[256];
  const size_t stride = (size_t)natoms * 3;
  double s = 0.0;
  for (int a = blockIdx.x * 256 + threadIdx.x; a < natoms; a += gridDim.x * 256) {
    const double im = 1.0 / masses[a];
    for (int j = 0; j < P; ++j) {
      double q0 = 0.0, q1 = 0.0, q2 = 0.0;
      for (int k = 0; k < P; ++k) {
        const double cj = Cg[j * P + k];
        const double *p = PP + k * stride + 3 * (size_t)a;
        q0 = q0 + cj * p[0];
        q1 = q1 + cj * p[1];
        q2 = q2 + cj * p[2];
      }
      s += im * (q0 * q0 + q1 * q1 + q2 * q2) / msf[j];
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

void launch_kinetic_nm(int P, int natoms, const DevBufs &b, const double *PP, double *uk,
                       double *part, hipStream_t s) {
  const int nb = std::min(KIN_BLOCKS, (natoms + 255) / 256);
  hipLaunchKernelGGL(k_kinetic_nm_part, dim3(nb), dim3(256), 0, s, P, natoms, b.C, b.msf,
                     b.masses, PP, part);
  hipLaunchKernelGGL(k_kinetic_final, dim3(1), dim3(256), 0, s, part, nb, uk);
}

// Per-step observables of write_out (InputOutput.f90:556-621) from the step
// record rec[0..2] = (Upot, virial, virialc) of the force call and the
// kinetic energy after it:
//   obs[0] sys_temp = TEMPFACTOR uk / (Natoms P^2)
//   obs[1..2] quantum_virial_estimators energy, pressure (estimators.f90:30-49)
//   obs[3..4] simple_quantum_estimators energy, pressure (estimators.f90:55-94)
//   obs[5..7] total molecular dipole, obs[8] sum of molecular dipole
//   magnitudes, Debye, molecular dipoles averaged over beads
//   (InputOutput.f90:593-607); NaN when dipI is NULL.
// k_obs_part: per block, partial sums of the spring term
// K0 = sum_k sum_atoms m |R_k - R_(k-1)|^2 (bead 0 against P-1) and of the
// dipole sums over a slice of molecules; k_obs_final adds the partials in
// block order and forms the observables.
constexpr int OBS_BLOCKS = 64;
__global__ __launch_bounds__(256) void k_obs_part(ObsArgs a, const double *__restrict__ RR,
                                                 const double *__restrict__ dipI,
                                                 const double *__restrict__ masses,
                                                 double *__restrict__ part) {
  __shared__ double red[5][256];
  const int P = a.P, na = a.natoms, N = na / 3;
  const size_t stride = (size_t)na * 3;
  const size_t gsz = (size_t)gridDim.x * 256, g0 = (size_t)blockIdx.x * 256 + threadIdx.x;
  double k0 = 0.0;
  for (size_t e = g0; e < (size_t)P * na; e += gsz) {
    const int k = (int)(e / na), at = (int)(e % na);
    const int km = k == 0 ? P - 1 : k - 1;
    const double *r = RR + k * stride + 3 * (size_t)at, *rm = RR + km * stride + 3 * (size_t)at;
    const double m = masses[at];
    for (int c = 0; c < 3; ++c) {
      const double d = r[c] - rm[c];
      k0 += m * (d * d);
    }
  }
  double dx = 0.0, dy = 0.0, dz = 0.0, dm = 0.0;
  if (dipI) {
    const double f = DEBYE / CHARGECON;
    for (size_t w = g0; w < (size_t)N; w += gsz) {
      double v[3] = {0.0, 0.0, 0.0};
      for (int b = 0; b < P; ++b)
        for (int c = 0; c < 3; ++c) v[c] += dipI[((size_t)b * N + w) * 3 + c] * f;
      for (int c = 0; c < 3; ++c) v[c] = v[c] / P;
      dx += v[0];
      dy += v[1];
      dz += v[2];
      dm += sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
  }
  red[0][threadIdx.x] = k0;
  red[1][threadIdx.x] = dx;
  red[2][threadIdx.x] = dy;
  red[3][threadIdx.x] = dz;
  red[4][threadIdx.x] = dm;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int v = 0; v < 5; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 5) part[blockIdx.x * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// With Upot non-NULL it first forms the step record rec[0..2] from the
// per-bead outputs (force_calc.f90:94-97); otherwise rec[0..2] is already set.
__global__ __launch_bounds__(64) void k_obs_final(ObsArgs a, const double *__restrict__ part,
                                                 int nb, int has_dip,
                                                 const double *__restrict__ uk,
                                                 const double *__restrict__ Upot,
                                                 const double *__restrict__ virt,
                                                 const double *__restrict__ virialc, double *rec0,
                                                 const int *__restrict__ clk) {
  __shared__ double red[8][64];
  const int P = a.P, na = a.natoms, N = na / 3, tid = threadIdx.x;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = tid; i < nb; i += 64)
    for (int k = 0; k < 5; ++k) v[k] += part[i * 5 + k];
  if (Upot)
    for (int b = tid; b < P; b += 64) {
      v[5] += Upot[b];
      v[6] += virt[b * 9 + 0] + virt[b * 9 + 4] + virt[b * 9 + 8];
      v[7] += virialc[b];
    }
  for (int k = 0; k < 8; ++k) red[k][tid] = v[k];
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) {
    if (tid < off)
      for (int k = 0; k < 8; ++k) red[k][tid] += red[k][tid + off];
    __syncthreads();
  }
  if (tid != 0) return;
  double *rec = rec0 + (size_t)clk[1] * MD_REC;
  if (Upot) {
    rec[0] = red[5][0];
    rec[1] = red[6][0];
    rec[2] = red[7][0] / P;
  }
  const double Up = rec[0], virial = rec[1], virialc_ = rec[2];
  const double temp = TEMPFACTOR * uk[0] / (double)(na * P * P);
  const double p3v = PRESSCON2 * (1.0 / (3.0 * a.volume));
  const double ke = (double)(1.5f * (float)na) * KB_KCAL * temp;
  double *o = rec + 4;
  o[0] = temp;
  o[1] = (ke + 0.5 * (virial - virialc_) / P + Up / P) / N;
  o[2] = p3v * (2 * ke - virialc_ / P) / N;
  const double kes = (double)(1.5f * (float)na * (float)P) * KB_KCAL * temp * P;
  const double K0 = 0.5 * red[0][0] * (a.omegan * a.omegan) * MASSCONI;
  o[3] = (kes - K0 + Up) / (double)(N * P);
  o[4] = p3v * (2 * (kes - K0) - virial) / N;
  const double nan = __builtin_nan("");
  for (int c = 0; c < 4; ++c) o[5 + c] = has_dip ? red[1 + c][0] : nan;
}

void launch_observables(const ObsArgs &a, const double *RR, const double *dipI,
                        const double *masses, const double *uk, const double *Upot,
                        const double *virt, const double *virialc, double *part, double *rec0,
                        const int *clk, hipStream_t s) {
  const size_t n = (size_t)a.P * a.natoms;
  const int nb = (int)std::min<size_t>(OBS_BLOCKS, (n + 255) / 256);
  hipLaunchKernelGGL(k_obs_part, dim3(nb), dim3(256), 0, s, a, RR, dipI, masses, part);
  hipLaunchKernelGGL(k_obs_final, dim3(1), dim3(64), 0, s, a, part, nb, dipI ? 1 : 0, uk, Upot,
                     virt, virialc, rec0, clk);
}

// Nose_Hoover half step, chain length 2 (Nose_Hoover.f90:158-198), scalars on device.
__global__ void k_nose_hoover(double *uk, double *vxi, double *sc, double tau, double delt2, int L,
                              double temp) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double delt4 = delt2 / 2.0, delt8 = delt2 / 4.0;
  const double kT = KB_AMU * temp;
  const double M1 = L * kT * (tau * tau), M2 = M1;
  double v0 = vxi[0], v1 = vxi[1], u = uk[0];
  v1 = v1 + (M1 * (v0 * v0) - kT) * delt4 / M2;
  v0 = v0 * exp(-v1 * delt8);
  v0 = v0 + (2 * u - L * kT) * delt4 / M1;
  v0 = v0 * exp(-v1 * delt8);
  const double s = exp(-v0 * delt2);
  u = u * s * s;
  v0 = v0 * exp(-v1 * delt8);
  v0 = v0 + (2 * u - L * kT) * delt4 / M1;
  v0 = v0 * exp(-v1 * delt8);
  v1 = v1 + (M1 * (v0 * v0) - kT) * delt4 / M2;
  vxi[0] = v0;
  vxi[1] = v1;
  uk[0] = u;
  sc[0] = s;
}

void launch_nose_hoover(double *uk, double *vxi, double *sc, double tau, double delt2, int L,
                        double temp, hipStream_t s) {
  hipLaunchKernelGGL(k_nose_hoover, dim3(1), dim3(64), 0, s, uk, vxi, sc, tau, delt2, L, temp);
}

}  // namespace pimd
