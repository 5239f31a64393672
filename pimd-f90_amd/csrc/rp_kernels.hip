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
__device__ __forceinline__ double kicked(const Rep &R, int kick, size_t o, double p) {
  if (kick == KICK_DRR) return p - MASSCON * R.b.dRR[o] * R.m.delt2;
  if (kick == KICK_FAST) return p - MASSCON * R.b.dRRfast[o] * R.m.delt2f;
  return p;
}

__global__ __launch_bounds__(256) void k_nm_evolve(const Rep *__restrict__ reps, BlockMap bm,
                                                  int kick, int kick2) {
  extern __shared__ double lds[];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, natoms = 3 * R.c.N;
  const double *__restrict__ Cg = R.b.C, *__restrict__ A = R.b.A, *__restrict__ msf = R.b.msf;
  const double *__restrict__ masses = R.b.masses;
  double *RR = R.b.RR, *PP = R.b.PP;
  const int CB = nm_cols(P);
  double *Cs = lds, *xr = Cs + P * P, *xp = xr + P * CB, *yr = xp + P * CB, *yp = yr + P * CB;
  const int ncols = natoms * 3, col0 = bk.x * CB;
  const size_t stride = (size_t)natoms * 3;
  for (int e = threadIdx.x; e < P * P; e += blockDim.x) Cs[e] = Cg[e];
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int k = e / CB, cc = e % CB, col = col0 + cc;
    xr[e] = col < ncols ? RR[k * stride + col] : 0.0;
    double p = col < ncols ? PP[k * stride + col] : 0.0;
    // with kicks: the half kick PPt - MASSCON*dRRt*delt2 before the step
    // (MD.f90:49), as k_kick forms it
    if (col < ncols) p = kicked(R, kick2, k * stride + col, kicked(R, kick, k * stride + col, p));
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
__global__ __launch_bounds__(64) void k_nm_evolve_mfma(const Rep *__restrict__ reps, BlockMap bm,
                                                       int kick, int kick2) {
  constexpr int KS = PT / 4, RT = PT / 16;
  __shared__ double Cs[PT * (PT + 1)], Yr[PT * 17], Yp[PT * 17];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, natoms = 3 * R.c.N;
  const double *__restrict__ Cg = R.b.C, *__restrict__ A = R.b.A, *__restrict__ msf = R.b.msf;
  const double *__restrict__ masses = R.b.masses;
  double *RR = R.b.RR, *PP = R.b.PP;
  const int ncols = natoms * 3;
  const size_t stride = (size_t)ncols;
  const int j = threadIdx.x & 15, q = threadIdx.x >> 4;
  const int col = bk.x * NM_TILE + j;
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
    if (in) p = kicked(R, kick2, o, kicked(R, kick, o, p));
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

// The free ring-polymer step of the replicas of the group with more than one
// bead (all of them with with_p1; the 3 Natoms EvolveRing calls of
// MD.f90:52-58), with the half kicks kick, kick2 first.  On the matrix cores
// one launch per padded bead count PT.
void launch_nm_evolve(const Group &g, hipStream_t s, int kick, int kick2, bool with_p1) {
  auto on = [&](const SysConst &c) { return with_p1 || c.P > 1; };
  if (nm_mfma()) {
    launch_over(
        g,
        [&](const SysConst &c, const MDConst &) {
          return Plan{on(c) ? nm_pt(c.P) : -1, (3 * 3 * c.N + NM_TILE - 1) / NM_TILE, 1, 1, 0};
        },
        [&](int pt, dim3 grid, const BlockMap &m, size_t) {
          switch (pt) {
            case 16: hipLaunchKernelGGL(k_nm_evolve_mfma<16>, grid, dim3(64), 0, s, g.d, m, kick, kick2); break;
            case 32: hipLaunchKernelGGL(k_nm_evolve_mfma<32>, grid, dim3(64), 0, s, g.d, m, kick, kick2); break;
            case 48: hipLaunchKernelGGL(k_nm_evolve_mfma<48>, grid, dim3(64), 0, s, g.d, m, kick, kick2); break;
            default: hipLaunchKernelGGL(k_nm_evolve_mfma<64>, grid, dim3(64), 0, s, g.d, m, kick, kick2); break;
          }
        });
    return;
  }
  launch_over(
      g,
      [&](const SysConst &c, const MDConst &) {
        const int CB = nm_cols(c.P), n = 3 * 3 * c.N;
        return Plan{on(c) ? 0 : -1, (n + CB - 1) / CB, 1, 1,
                    sizeof(double) * ((size_t)c.P * c.P + 4 * (size_t)c.P * CB)};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t lds) {
        hipLaunchKernelGGL(k_nm_evolve, grid, dim3(256), lds, s, g.d, m, kick, kick2);
      });
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
// on the same LDS slabs as k_nm_evolve.  Normals: host_noise, the replica's
// noise buffer [Natoms][P][3] (the reference's consumption order); otherwise
// drawn from (seed, counter): with use_clk the replica's own seed and the
// counter 2 t + half of its device step clock (MD steps), else the given ones.
__device__ __forceinline__ double langevin_xi(const Rep &R, int host_noise, int use_clk,
                                              uint64_t seed, uint64_t counter, uint64_t n) {
  if (host_noise) return R.b.noise[n];
  double z0, z1;
  normal2(use_clk ? R.m.seed : seed, counter, n >> 1, z0, z1);
  return (n & 1) ? z1 : z0;
}

__global__ __launch_bounds__(256) void k_langevin(const Rep *__restrict__ reps, BlockMap bm,
                                                 int host_noise, int use_clk, uint64_t seed,
                                                 uint64_t counter, int half, int kick) {
  extern __shared__ double lds[];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, natoms = 3 * R.c.N;
  const double *__restrict__ Cg = R.b.C, *__restrict__ msf = R.b.msf;
  const double *__restrict__ c1 = R.b.c1, *__restrict__ c2 = R.b.c2;
  const double massO = R.m.massO, massH = R.m.massH;
  double *PP = R.b.PP;
  // in MD steps the counter follows the device step clock: step t, half 0 / 1
  if (use_clk) counter = (uint64_t)R.b.clk[0] * 2 + half;
  const int CB = nm_cols(P);
  double *Cs = lds, *xp = Cs + P * P, *yp = xp + P * CB;
  const int ncols = natoms * 3, col0 = bk.x * CB;
  const size_t stride = (size_t)natoms * 3;
  for (int e = threadIdx.x; e < P * P; e += blockDim.x) Cs[e] = Cg[e];
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int k = e / CB, cc = e % CB, col = col0 + cc;
    double p = col < ncols ? PP[k * stride + col] : 0.0;
    // with kick: the half kick PPt - MASSCON*dRRt*delt2 first (MD.f90:94)
    if (col < ncols) p = kicked(R, kick, k * stride + col, p);
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
      const double xi = langevin_xi(R, host_noise, use_clk, seed, counter, n);
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
__global__ __launch_bounds__(64) void k_langevin_mfma(const Rep *__restrict__ reps, BlockMap bm,
                                                      int host_noise, int use_clk, uint64_t seed,
                                                      uint64_t counter, int half, int kick) {
  constexpr int KS = PT / 4, RT = PT / 16;
  __shared__ double Cs[PT * (PT + 1)], Yp[PT * 17];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, natoms = 3 * R.c.N;
  const double *__restrict__ Cg = R.b.C, *__restrict__ msf = R.b.msf;
  const double *__restrict__ c1 = R.b.c1, *__restrict__ c2 = R.b.c2;
  const double massO = R.m.massO, massH = R.m.massH;
  double *PP = R.b.PP;
  if (use_clk) counter = (uint64_t)R.b.clk[0] * 2 + half;
  const int ncols = natoms * 3;
  const size_t stride = (size_t)ncols;
  const int j = threadIdx.x & 15, q = threadIdx.x >> 4;
  const int col = bk.x * NM_TILE + j;
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
    if (in) p = kicked(R, kick, o, p);
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
          const double xi = langevin_xi(R, host_noise, use_clk, seed, counter, n);
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

// Langevin_NM of the replicas whose bead thermostat is Langevin (every
// replica with all = true, as the pimd_langevin entry point asks).
void launch_langevin(const Group &g, hipStream_t s, bool all, int host_noise, int use_clk,
                     uint64_t seed, uint64_t counter, int half, int kick) {
  auto on = [&](const MDConst &m) { return all || m.bead_tt == 1; };
  if (langevin_mfma()) {
    launch_over(
        g,
        [&](const SysConst &c, const MDConst &m) {
          return Plan{on(m) ? nm_pt(c.P) : -1, (3 * 3 * c.N + NM_TILE - 1) / NM_TILE, 1, 1, 0};
        },
        [&](int pt, dim3 grid, const BlockMap &m, size_t) {
#define PIMD_LANG_MFMA(PT)                                                                 \
  hipLaunchKernelGGL(k_langevin_mfma<PT>, grid, dim3(64), 0, s, g.d, m, host_noise, use_clk, \
                     seed, counter, half, kick)
          switch (pt) {
            case 16: PIMD_LANG_MFMA(16); break;
            case 32: PIMD_LANG_MFMA(32); break;
            case 48: PIMD_LANG_MFMA(48); break;
            default: PIMD_LANG_MFMA(64); break;
          }
#undef PIMD_LANG_MFMA
        });
    return;
  }
  launch_over(
      g,
      [&](const SysConst &c, const MDConst &m) {
        const int CB = nm_cols(c.P), n = 3 * 3 * c.N;
        return Plan{on(m) ? 0 : -1, (n + CB - 1) / CB, 1, 1,
                    sizeof(double) * ((size_t)c.P * c.P + 2 * (size_t)c.P * CB)};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t lds) {
        hipLaunchKernelGGL(k_langevin, grid, dim3(256), lds, s, g.d, m, host_noise, use_clk, seed,
                           counter, half, kick);
      });
}

// The device step clock {MD step t, record slot} read by the force and
// thermostat kernels of an MD step, so that one captured step replays for
// every step.  One thread per replica.
__global__ void k_set_clock(const Rep *__restrict__ reps, BlockMap bm, int t, int slot) {
  const Blk bk = block_of(bm);
  if (threadIdx.x != 0) return;
  int *clk = reps[bk.r].b.clk;
  clk[0] = t;
  clk[1] = slot;
}
// End of an MD step: with with_uk its record takes the final kinetic energy
// (rec[3]), then the clock moves to the next step and record slot.
__global__ void k_advance_clock(const Rep *__restrict__ reps, BlockMap bm, int with_uk) {
  const Blk bk = block_of(bm);
  if (threadIdx.x != 0) return;
  const Rep &R = reps[bk.r];
  int *clk = R.b.clk;
  if (with_uk) R.m.rec[(size_t)clk[1] * MD_REC + 3] = R.b.uk[0];
  clk[0] += 1;
  clk[1] += 1;
}

void launch_set_clock(const Group &g, int t, int slot, hipStream_t s) {
  launch_over(
      g, [](const SysConst &, const MDConst &) { return Plan{0, 1, 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_set_clock, grid, dim3(64), 0, s, g.d, m, t, slot);
      });
}
void launch_advance_clock(const Group &g, bool with_uk, hipStream_t s) {
  launch_over(
      g, [](const SysConst &, const MDConst &) { return Plan{0, 1, 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_advance_clock, grid, dim3(64), 0, s, g.d, m, with_uk ? 1 : 0);
      });
}

// PPt = PPt - MASSCON*dRRt*delt2 (MD.f90:49, 94) with the replica's gradient
// of kind `kick` (KICK_DRR or KICK_FAST).
__global__ void k_kick(const Rep *__restrict__ reps, BlockMap bm, int kick) {
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const size_t n = (size_t)R.c.P * 3 * R.c.N * 3;
  const size_t i = (size_t)bk.x * blockDim.x + threadIdx.x;
  if (i < n) R.b.PP[i] = kicked(R, kick, i, R.b.PP[i]);
}

void launch_kick(const Group &g, int kick, int sel, hipStream_t s) {
  launch_over(
      g,
      [&](const SysConst &c, const MDConst &m) {
        const long n = (long)c.P * 3 * c.N * 3;
        const bool on = sel == SEL_ALL || (sel == SEL_P1 && c.P == 1) ||
                        (sel == SEL_NO_BEAD_TT && m.bead_tt == 0);
        return Plan{on ? 0 : -1, (int)((n + 255) / 256), 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_kick, grid, dim3(256), 0, s, g.d, m, kick);
      });
}

// PPt = PPt*s with the global chain's s on the device (MD.f90:45, 105).
__global__ void k_scale(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const size_t n = (size_t)R.c.P * 3 * R.c.N * 3;
  const size_t i = (size_t)bk.x * blockDim.x + threadIdx.x;
  if (i < n) R.b.PP[i] = R.b.PP[i] * R.b.nh_s[0];
}

// Free flight of one-bead replicas (MD.f90:60-68), over the step (fast = 0)
// or the inner step of contracted MD (fast = 1).
__global__ void k_free_flight(const Rep *__restrict__ reps, BlockMap bm, int fast) {
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const size_t n = (size_t)R.c.P * 3 * R.c.N * 3;
  const size_t i = (size_t)bk.x * blockDim.x + threadIdx.x;
  const double delt = fast ? R.m.deltf : R.m.delt;
  if (i < n) R.b.RR[i] = R.b.RR[i] + (1.0 / R.b.masses[i / 3]) * R.b.PP[i] * delt;
}

void launch_free_flight(const Group &g, bool fast, hipStream_t s) {
  launch_over(
      g,
      [&](const SysConst &c, const MDConst &) {
        const long n = (long)c.P * 3 * c.N * 3;
        return Plan{c.P == 1 ? 0 : -1, (int)((n + 255) / 256), 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_free_flight, grid, dim3(256), 0, s, g.d, m, fast ? 1 : 0);
      });
}

// RRc = sum(RRt,3)/Nbeads, then the centroid wrap of PBCs (MD.f90:70;
// main_stuff.f90:107-125): shifts use box(1) on every axis.
__global__ void k_centroid_pbc(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, natoms = 3 * R.c.N;
  const double box0 = R.m.box0, boxi0 = R.m.boxi0;
  double *RR = R.b.RR, *RRc = R.b.RRc;
  const int id = bk.x * blockDim.x + threadIdx.x;
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

void launch_centroid_pbc(const Group &g, hipStream_t s) {
  launch_over(
      g, [](const SysConst &c, const MDConst &) { return Plan{0, (9 * c.N + 63) / 64, 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_centroid_pbc, grid, dim3(64), 0, s, g.d, m);
      });
}

// uk = 1/2 sum_beads sum_atoms |p|^2 / m (calc_uk, RPMD branch): per-block
// partial sums, then one block adds them in a fixed order.
constexpr int KIN_BLOCKS = 256;
__host__ __device__ inline int kin_blocks(const SysConst &c, int cmd) {
  const long n = cmd ? 3L * c.N : (long)c.P * 3 * c.N;
  return (int)std::min<long>(KIN_BLOCKS, (n + 255) / 256);
}

__global__ __launch_bounds__(256) void k_kinetic_part(const Rep *__restrict__ reps, BlockMap bm) {
  __shared__ double red[256];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, natoms = 3 * R.c.N;
  const double *__restrict__ masses = R.b.masses, *__restrict__ PP = R.b.PP;
  const size_t n = (size_t)P * natoms;
  double s = 0.0;
  for (size_t i = (size_t)bk.x * 256 + threadIdx.x; i < n; i += (size_t)bk.nx * 256) {
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
  if (threadIdx.x == 0) R.b.ukpart[bk.x] = red[0];
}

// calc_uk with centroid-MD mass scaling (main_stuff.f90:366-385): per atom
// the normal-mode momenta p~_j = sum_k C[j][k] p_k (real2NM,
// NormalModes.f90:308-322), uk = 1/2 sum_atoms sum_j |p~_j|^2 / (m msf_j).
// One thread per atom, per-block partials, then the fixed-order final sum.
__global__ __launch_bounds__(256) void k_kinetic_nm_part(const Rep *__restrict__ reps, BlockMap bm) {
  __shared__ double red[256];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, natoms = 3 * R.c.N;
  const double *__restrict__ Cg = R.b.C, *__restrict__ msf = R.b.msf;
  const double *__restrict__ masses = R.b.masses, *__restrict__ PP = R.b.PP;
  const size_t stride = (size_t)natoms * 3;
  double s = 0.0;
  for (int a = bk.x * 256 + threadIdx.x; a < natoms; a += bk.nx * 256) {
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
  if (threadIdx.x == 0) R.b.ukpart[bk.x] = red[0];
}

__global__ __launch_bounds__(256) void k_kinetic_final(const Rep *__restrict__ reps, BlockMap bm) {
  __shared__ double red[256];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int nb = kin_blocks(R.c, R.m.cmd);
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += R.b.ukpart[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) R.b.uk[0] = 0.5 * red[0];
}

// calc_uk (main_stuff.f90:349-388) of every replica into its uk: mode-space
// for centroid MD (MDConst::cmd), bead-space otherwise.
void launch_kinetic(const Group &g, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &m) { return Plan{m.cmd ? 1 : 0, kin_blocks(c, m.cmd), 1, 1, 0}; },
      [&](int v, dim3 grid, const BlockMap &m, size_t) {
        if (v) hipLaunchKernelGGL(k_kinetic_nm_part, grid, dim3(256), 0, s, g.d, m);
        else hipLaunchKernelGGL(k_kinetic_part, grid, dim3(256), 0, s, g.d, m);
      });
  launch_over(
      g, [](const SysConst &, const MDConst &) { return Plan{0, 1, 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_kinetic_final, grid, dim3(256), 0, s, g.d, m);
      });
}

// Per-step observables of write_out (InputOutput.f90:556-621) from the step
// record rec[0..2] = (Upot, virial, virialc) of the force call and the
// kinetic energy after it:
//   obs[0] sys_temp = TEMPFACTOR uk / (Natoms P^2)
//   obs[1..2] quantum_virial_estimators energy, pressure (estimators.f90:30-49)
//   obs[3..4] simple_quantum_estimators energy, pressure (estimators.f90:55-94)
//   obs[5..7] total molecular dipole, obs[8] sum of molecular dipole
//   magnitudes, Debye, molecular dipoles averaged over beads
//   (InputOutput.f90:593-607); NaN without bead dipoles (has_dip = 0).
// k_obs_part: per block, partial sums of the spring term
// K0 = sum_k sum_atoms m |R_k - R_(k-1)|^2 (bead 0 against P-1) and of the
// dipole sums over a slice of molecules; k_obs_final adds the partials in
// block order and forms the observables.
constexpr int OBS_BLOCKS = 64;
__host__ __device__ inline int obs_blocks(const SysConst &c) {
  return (int)std::min<long>(OBS_BLOCKS, ((long)c.P * 3 * c.N + 255) / 256);
}

__global__ __launch_bounds__(256) void k_obs_part(const Rep *__restrict__ reps, BlockMap bm,
                                                 int has_dip) {
  __shared__ double red[5][256];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, na = 3 * R.c.N, N = R.c.N;
  const double *__restrict__ RR = R.b.RR, *__restrict__ masses = R.b.masses;
  const double *__restrict__ dipI = has_dip ? R.b.dipI : nullptr;
  const size_t stride = (size_t)na * 3;
  const size_t gsz = (size_t)bk.nx * 256, g0 = (size_t)bk.x * 256 + threadIdx.x;
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
  if (threadIdx.x < 5) R.b.obs_part[bk.x * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// With with_forces it first forms the step record rec[0..2] from the
// per-bead outputs of the force call (force_calc.f90:94-97); otherwise
// rec[0..2] is already set.
__global__ __launch_bounds__(64) void k_obs_final(const Rep *__restrict__ reps, BlockMap bm,
                                                 int has_dip, int with_forces) {
  __shared__ double red[8][64];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, na = 3 * R.c.N, N = R.c.N, tid = threadIdx.x;
  const int nb = obs_blocks(R.c);
  const double *__restrict__ part = R.b.obs_part;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = tid; i < nb; i += 64)
    for (int k = 0; k < 5; ++k) v[k] += part[i * 5 + k];
  if (with_forces)
    for (int b = tid; b < P; b += 64) {
      v[5] += R.b.Upot[b];
      v[6] += R.b.virt[b * 9 + 0] + R.b.virt[b * 9 + 4] + R.b.virt[b * 9 + 8];
      v[7] += R.b.virialc[b];
    }
  for (int k = 0; k < 8; ++k) red[k][tid] = v[k];
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) {
    if (tid < off)
      for (int k = 0; k < 8; ++k) red[k][tid] += red[k][tid + off];
    __syncthreads();
  }
  if (tid != 0) return;
  double *rec = R.m.rec + (size_t)R.b.clk[1] * MD_REC;
  if (with_forces) {
    rec[0] = red[5][0];
    rec[1] = red[6][0];
    rec[2] = red[7][0] / P;
  }
  const double Up = rec[0], virial = rec[1], virialc_ = rec[2];
  const double temp = TEMPFACTOR * R.b.uk[0] / (double)(na * P * P);
  const double p3v = PRESSCON2 * (1.0 / (3.0 * R.c.vol));
  const double ke = (double)(1.5f * (float)na) * KB_KCAL * temp;
  double *o = rec + 4;
  o[0] = temp;
  o[1] = (ke + 0.5 * (virial - virialc_) / P + Up / P) / N;
  o[2] = p3v * (2 * ke - virialc_ / P) / N;
  const double kes = (double)(1.5f * (float)na * (float)P) * KB_KCAL * temp * P;
  const double K0 = 0.5 * red[0][0] * (R.m.omegan * R.m.omegan) * MASSCONI;
  o[3] = (kes - K0 + Up) / (double)(N * P);
  o[4] = p3v * (2 * (kes - K0) - virial) / N;
  const double nan = __builtin_nan("");
  for (int c = 0; c < 4; ++c) o[5 + c] = has_dip ? red[1 + c][0] : nan;
}

void launch_observables(const Group &g, bool has_dip, bool with_forces, hipStream_t s) {
  launch_over(
      g, [](const SysConst &c, const MDConst &) { return Plan{0, obs_blocks(c), 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_obs_part, grid, dim3(256), 0, s, g.d, m, has_dip ? 1 : 0);
      });
  launch_over(
      g, [](const SysConst &, const MDConst &) { return Plan{0, 1, 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_obs_final, grid, dim3(64), 0, s, g.d, m, has_dip ? 1 : 0,
                           with_forces ? 1 : 0);
      });
}

// One Nose_Hoover call with a chain of length 2 (Nose_Hoover.f90:33-112):
// advances the chain velocities v0, v1 by half a step, scales u (2 x kinetic
// energy / 2) by s^2 and returns the momentum scale s.
__device__ __forceinline__ double nh_chain2(double &u, double &v0, double &v1, double tau,
                                            double delt2, double L, double temp) {
  const double delt4 = delt2 / 2.0, delt8 = delt2 / 4.0;
  const double kT = KB_AMU * temp;
  const double M1 = L * kT * (tau * tau), M2 = M1;
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
  return s;
}

// The global chain (MD.f90:43-46, 99-102) of each replica, scalars on
// device: L = 3 Natoms P degrees of freedom at P T.
__global__ void k_nose_hoover(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  if (threadIdx.x != 0) return;
  const Rep &R = reps[bk.r];
  double *vxi = R.b.nh_vxi;
  double v0 = vxi[0], v1 = vxi[1], u = R.b.uk[0];
  const int L = 3 * 3 * R.c.N * R.c.P;
  const double s = nh_chain2(u, v0, v1, R.m.tau, R.m.delt2, (double)L, R.m.tempP);
  vxi[0] = v0;
  vxi[1] = v1;
  R.b.uk[0] = u;
  R.b.nh_s[0] = s;
}

// The global chain, then PPt = PPt*s, for the replicas that have it.
void launch_global_nh(const Group &g, hipStream_t s) {
  launch_over(
      g, [](const SysConst &, const MDConst &m) { return Plan{m.global_nh ? 0 : -1, 1, 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_nose_hoover, grid, dim3(64), 0, s, g.d, m);
      });
  launch_over(
      g,
      [](const SysConst &c, const MDConst &m) {
        const long n = (long)c.P * 3 * c.N * 3;
        return Plan{m.global_nh ? 0 : -1, (int)((n + 255) / 256), 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_scale, grid, dim3(256), 0, s, g.d, m);
      });
}

// Nose-Hoover bead thermostat in normal-mode space (main_stuff.f90:145-171):
// p~ = C p per (atom, component) (real2NM); every mode momentum p~_j has a
// chain of its own (vxi [Natoms][P][3][2], Nose_Hoover with L = 1 at the bead
// temperature P T): the centroid (j = 0) only with the centroid thermostat,
// tau = tau_centroid and u = p~^2 / 2m; the other modes tau = 1 / omega_j and
// u = p~^2 / (2 m MassScaleFactor_j); p~_j is scaled by the chain's s and
// p = C^T p~ (NM2real).  LDS slabs and sum orders as k_langevin; with kick,
// the half kick PPt - MASSCON dRRt delt2 comes first (MD.f90:94).
__global__ __launch_bounds__(256) void k_bead_nh(const Rep *__restrict__ reps, BlockMap bm,
                                                int kick) {
  extern __shared__ double lds[];
  const Blk bk = block_of(bm);
  const Rep &R = reps[bk.r];
  const int P = R.c.P, natoms = 3 * R.c.N;
  const double *__restrict__ Cg = R.b.C, *__restrict__ msf = R.b.msf;
  const double *__restrict__ tau_mode = R.b.nh_tau;
  const int centroid = R.m.centroid;
  const double massO = R.m.massO, massH = R.m.massH, delt2 = R.m.delt2, tempP = R.m.tempP;
  double *vxi = R.b.vxi_beads, *PP = R.b.PP;
  const int CB = nm_cols(P);
  double *Cs = lds, *xp = Cs + P * P, *yp = xp + P * CB;
  const int ncols = natoms * 3, col0 = bk.x * CB;
  const size_t stride = (size_t)natoms * 3;
  for (int e = threadIdx.x; e < P * P; e += blockDim.x) Cs[e] = Cg[e];
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int k = e / CB, cc = e % CB, col = col0 + cc;
    double p = col < ncols ? PP[k * stride + col] : 0.0;
    if (col < ncols) p = kicked(R, kick, k * stride + col, p);
    xp[e] = p;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < P * CB; e += blockDim.x) {
    const int j = e / CB, cc = e % CB, col = col0 + cc;
    double v = 0.0;
    for (int k = 0; k < P; ++k) v = v + Cs[j * P + k] * xp[k * CB + cc];
    if (col < ncols && (j > 0 || centroid)) {
      const int a = col / 3, comp = col % 3;
      const double im = 1.0 / ((a % 3 == 0) ? massO : massH);
      double u = 0.5 * im * (v * v);
      if (j > 0) u = u / msf[j];
      double *ch = vxi + (((size_t)a * P + j) * 3 + comp) * 2;
      double v0 = ch[0], v1 = ch[1];
      const double sc = nh_chain2(u, v0, v1, tau_mode[j], delt2, 1.0, tempP);
      ch[0] = v0;
      ch[1] = v1;
      v = v * sc;
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

void launch_bead_nh(const Group &g, int kick, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &m) {
        const int CB = nm_cols(c.P), n = 3 * 3 * c.N;
        return Plan{m.bead_tt == 2 ? 0 : -1, (n + CB - 1) / CB, 1, 1,
                    sizeof(double) * ((size_t)c.P * c.P + 2 * (size_t)c.P * CB)};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t lds) {
        hipLaunchKernelGGL(k_bead_nh, grid, dim3(256), lds, s, g.d, m, kick);
      });
}

}  // namespace pimd
