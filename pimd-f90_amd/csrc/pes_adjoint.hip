// Adjoint of the Partridge-Schwenke monomer PES with respect to its fitted
// parameters, for PES fitting.  No counterpart in the reference, which fits
// the monomer offline; the PES itself is nasa.f90:2-103.
//
// For a loss L = sum_b (E_b - E*_b)^2 + sum_{b,atoms} |dE/dr - (dE/dr)*|^2 over
// configurations (beads) b, the PES parameters theta enter E only through the
// monomer energies, so
//   dL/dtheta = sum_{b,molecules} [ ew_b dEmon/dtheta + D_w (dEmon/dtheta) ]
// with ew_b = 2 (E_b - E*_b) and w = 2 (dE/dr - (dE/dr)*) the gradient residual
// of the molecule's atoms: D_w is the derivative along w of the function
// r -> dEmon/dtheta(r), which equals w . d(dE/dr)/dtheta.  It is evaluated
// exactly with dual numbers (value, derivative along w) through the closed
// forms of dEmon/dtheta.
#include "dual.h"
#include <mutex>

#include "ttm_device.h"

namespace pimd {

constexpr double KCAL = 0.00285914375100642899;  // cm-1 -> kcal/mol (nasa.f90:101)
constexpr int ADJ_BLOCK = 64;

// Exponent table of the PES basis (nasa_mod.f90 idx), this unit's own copy.
__constant__ int c_idx_adj[PS_NTERM][3];

// Once per process; false if the copy failed.
static bool set_adjoint_tables() {
  static std::once_flag once;
  static bool ok = false;
  std::call_once(once, [] {
    int idx[PS_NTERM][3];
    for (int j = 0; j < PS_NTERM; ++j) {
      idx[j][0] = PS_IDX1[j];
      idx[j][1] = PS_IDX2[j];
      idx[j][2] = PS_IDX3[j];
    }
    ok = hipMemcpyToSymbol(HIP_SYMBOL(c_idx_adj), idx, sizeof idx) == hipSuccess;
  });
  return ok;
}

// One thread per (bead, molecule); each block sums its molecules' terms for
// every parameter (wave reduction, fixed order) into partial[block][250].
__global__ __launch_bounds__(ADJ_BLOCK) void k_pes_adjoint(SysConst c, const double *__restrict__ RR,
                                                          const double *__restrict__ ew,
                                                          const double *__restrict__ gw,
                                                          const double *__restrict__ pes,
                                                          double *__restrict__ partial) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = gid < c.P * c.N;
  const int b = valid ? gid / c.N : 0, m = valid ? gid % c.N : 0;
  const size_t a0 = ((size_t)b * 3 * c.N + 3 * m) * 3;
  // atoms O, H1, H2 as duals along the residual direction w
  Dual R[3][3];
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k)
      R[a][k] = valid ? dmk(RR[a0 + 3 * a + k], gw[a0 + 3 * a + k]) : dmk(1.0 + a + k);
  const double ewb = valid ? ew[b] : 0.0;
  // minimum-image shifts from the values, as the forward pass (nasa.f90:31-38)
  Dual R1[3], R2[3], RH[3];
  for (int k = 0; k < 3; ++k) {
    R1[k] = R[1][k] - R[0][k];
    R2[k] = R[2][k] - R[0][k];
    RH[k] = R[1][k] - R[2][k];
    R1[k].v -= c.box[k] * anint(R1[k].v * c.boxi[k]);
    R2[k].v -= c.box[k] * anint(R2[k].v * c.boxi[k]);
    RH[k].v -= c.box[k] * anint(RH[k].v * c.boxi[k]);
  }
  const Dual d1 = dsqrt(R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2]);
  const Dual d2 = dsqrt(R2[0] * R2[0] + R2[1] * R2[1] + R2[2] * R2[2]);
  const Dual dh = dsqrt(RH[0] * RH[0] + RH[1] * RH[1] + RH[2] * RH[2]);
  const Dual costh = (R1[0] * R2[0] + R1[1] * R2[1] + R1[2] * R2[2]) / (d1 * d2);
  const double costhe = -.24780227221366464506;
  const Dual x1 = (1.0 / PS_REOH) * (d1 - dmk(PS_REOH)), x2 = (1.0 / PS_REOH) * (d2 - dmk(PS_REOH));
  const Dual x3 = costh - dmk(costhe);
  Dual f1[16], f2[16], f3[16];
  f1[0] = f2[0] = f3[0] = dmk(0.0);
  f1[1] = f2[1] = f3[1] = dmk(1.0);
  for (int j = 2; j < 16; ++j) {
    f1[j] = f1[j - 1] * x1;
    f2[j] = f2[j - 1] * x2;
    f3[j] = f3[j - 1] * x3;
  }
  const Dual u1 = d1 - dmk(PS_REOH), u2 = d2 - dmk(PS_REOH);
  const Dual efac = dexp(-PS_B1 * (u1 * u1 + u2 * u2));
  const double deohA = pes[PES_DEOHA], alphaoh = pes[PES_ALPHAOH], roh = pes[PES_ROH],
               phh1A = pes[PES_PHH1A], phh2 = pes[PES_PHH2];
  const double deoh = PS_F5Z * deohA;
  const double phh1 = PS_F5Z * phh1A * exp(phh2);
  const Dual r1 = d1 - dmk(roh), r2 = d2 - dmk(roh);
  const Dual e1 = dexp(-alphaoh * r1), e2 = dexp(-alphaoh * r2);
  const Dual ehh = dexp(-phh2 * dh);

  __shared__ double acc[PES_NPARAM];
  for (int j = threadIdx.x; j < PES_NPARAM; j += blockDim.x) acc[j] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  auto add = [&](int j, Dual g) {
    double v = valid ? KCAL * (ewb * g.v + g.d) : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) acc[j] += v;  // one wave per block: a fixed order
  };
  // dEmon/dc5z(1) = 2 (nasa.f90:90); dEmon/dc5z(j) = efac * basis_j (nasa.f90:72-78)
  add(0, dmk(2.0));
  for (int j = 1; j < PES_NC; ++j) {
    const int I = c_idx_adj[j][0], J = c_idx_adj[j][1], K = c_idx_adj[j][2];
    add(j, efac * ((f1[I] * f2[J] + f1[J] * f2[I]) * f3[K]));
  }
  // Va = deoh (e1 (e1 - 2) + e2 (e2 - 2)), e_i = exp(-alphaoh (d_i - roh))
  const Dual two = dmk(2.0);
  add(PES_DEOHA, PS_F5Z * (e1 * (e1 - two) + e2 * (e2 - two)));
  // dVa/de_i = deoh (2 e_i - 2); de_i/dalphaoh = -(d_i - roh) e_i; de_i/droh = alphaoh e_i
  const Dual g1 = deoh * (2.0 * e1 - two), g2 = deoh * (2.0 * e2 - two);
  add(PES_ALPHAOH, (g1 * (r1 * e1) + g2 * (r2 * e2)) * dmk(-1.0));
  add(PES_ROH, alphaoh * (g1 * e1 + g2 * e2));
  // Vb = f5z phh1A exp(phh2) exp(-phh2 dHH)
  add(PES_PHH1A, (PS_F5Z * exp(phh2)) * ehh);
  add(PES_PHH2, phh1 * (ehh * (dmk(1.0) - dh)));
  __syncthreads();
  for (int j = threadIdx.x; j < PES_NPARAM; j += blockDim.x)
    partial[(size_t)blockIdx.x * PES_NPARAM + j] = acc[j];
}

__global__ __launch_bounds__(256) void k_pes_adjoint_sum(const double *__restrict__ partial,
                                                        int nblocks, double *__restrict__ grad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= PES_NPARAM) return;
  double v = 0.0;
  for (int q = 0; q < nblocks; ++q) v += partial[(size_t)q * PES_NPARAM + j];
  grad[j] = v;
}

int pes_adjoint_blocks(const SysConst &c) { return (c.P * c.N + ADJ_BLOCK - 1) / ADJ_BLOCK; }

bool launch_pes_adjoint(const SysConst &c, DevBufs &b, const double *RR, const double *ew,
                        const double *gw, double *partial, int nblocks, double *grad,
                        hipStream_t s) {
  if (!set_adjoint_tables()) return false;
  hipLaunchKernelGGL(k_pes_adjoint, dim3(nblocks), dim3(ADJ_BLOCK), 0, s, c, RR, ew, gw, b.pes,
                     partial);
  hipLaunchKernelGGL(k_pes_adjoint_sum, dim3(1), dim3(256), 0, s, partial, nblocks, grad);
  return true;
}

}  // namespace pimd
