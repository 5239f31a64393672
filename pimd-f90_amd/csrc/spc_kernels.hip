// The fixed-charge water models of pot_spc (pot_ttm/pot_spc.f90:1-277):
// q-SPC/Fw (pot_model 4, harmonic bonds and angle) and SPC/F (pot_model 5,
// bond, bond and H-H terms), with vdW between oxygens and Ewald-summed point
// charges on O, H1, H2.
//
// They run on the TTM path's molecule records, neighbour lists, phase tables
// and reciprocal-space charge kernels.  A record keeps O in slot 0 with no
// charge (the O-O vdW term and the neighbour search read it there), the H
// atoms in slots 1 and 2, and O's charge in slot 3: an M site with
// gammaM = 0, which coincides with O.  The reciprocal-space kernels sum the
// charges of slots 1..3, which are then exactly the model's three charges.
#include "pimd_internal.h"
#include "ttm_device.h"

namespace pimd {

// r - box*anint(r/box): the division form pot_spc.f90 uses for vectors
// inside a molecule (lines 74-75, 113-115, 211, 261, 264).
__device__ __forceinline__ void minimg_div(const SysConst &c, double *r) {
  for (int k = 0; k < 3; ++k) r[k] = r[k] - c.box[k] * anint(r[k] / c.box[k]);
}

__device__ __forceinline__ double dot3d(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Intramolecular energy and gradient dr1[O, H1, H2] of one water and its
// virial contribution vt[col][row] (pot_spc.f90:69-137).
__device__ inline void spc_intra(const SysConst &c, const double r[3][3], double &e1,
                                 double dr1[3][3], double vt[9]) {
  if (c.model == PIMD_QSPCFW) {
    double o1[3], o2[3], axb[3], a1[3], a2[3], dh1[3], dh2[3];
    for (int k = 0; k < 3; ++k) {
      o1[k] = r[1][k] - r[0][k];
      o2[k] = r[2][k] - r[0][k];
    }
    minimg_div(c, o1);
    minimg_div(c, o2);
    const double d1 = sqrt(dot3d(o1, o1)), d2 = sqrt(dot3d(o2, o2));
    axb[0] = o1[1] * o2[2] - o1[2] * o2[1];
    axb[1] = -o1[0] * o2[2] + o1[2] * o2[0];
    axb[2] = o1[0] * o2[1] - o1[1] * o2[0];
    const double daxb = sqrt(dot3d(axb, axb));
    const double costh = dot3d(o1, o2) / (d1 * d2), sinth = daxb / (d1 * d2);
    const double ang = atan2(sinth, costh);
    const double b1 = d1 - c.spc_roh0, b2 = d2 - c.spc_roh0, da = ang - c.spc_th0;
    e1 = c.spc_kb * (b1 * b1 + b2 * b2) + c.spc_ka * (da * da);
    const double t1 = 2.0 * c.spc_kb * b1 / d1, t2 = 2.0 * c.spc_kb * b2 / d2;
    const double ta = 2.0 * c.spc_ka * da;
    for (int k = 0; k < 3; ++k) {
      a1[k] = -(o2[k] / (d1 * d2) - costh * o1[k] / (d1 * d1)) / sinth;
      a2[k] = -(o1[k] / (d1 * d2) - costh * o2[k] / (d2 * d2)) / sinth;
      dh1[k] = t1 * o1[k] + ta * a1[k];
      dh2[k] = t2 * o2[k] + ta * a2[k];
    }
    for (int k = 0; k < 3; ++k) {
      dr1[1][k] = dh1[k];
      dr1[2][k] = dh2[k];
      dr1[0][k] = -dh1[k] - dh2[k];
    }
    for (int iy = 0; iy < 3; ++iy)
      for (int ix = 0; ix < 3; ++ix) vt[iy * 3 + ix] = vt[iy * 3 + ix] + o1[ix] * dh1[iy] + o2[ix] * dh2[iy];
  } else {
    double o1[3], o2[3], hh[3];
    for (int k = 0; k < 3; ++k) {
      o1[k] = r[0][k] - r[1][k];
      o2[k] = r[0][k] - r[2][k];
      hh[k] = r[1][k] - r[2][k];
    }
    minimg_div(c, o1);
    minimg_div(c, o2);
    minimg_div(c, hh);
    const double d1 = sqrt(dot3d(o1, o1)), d2 = sqrt(dot3d(o2, o2)), d3 = sqrt(dot3d(hh, hh));
    const double x1 = d1 - c.spc_roh0, x2 = d2 - c.spc_roh0, x3 = d3 - c.spc_rhh0;
    e1 = c.spc_A * (x1 * x1 + x2 * x2) + c.spc_B * x3 * x3 + c.spc_C * (x1 + x2) * x3 +
         c.spc_D * x1 * x2;
    const double v1 = 2.0 * c.spc_A * x1 + c.spc_C * x3 + c.spc_D * x2;
    const double v2 = 2.0 * c.spc_A * x2 + c.spc_C * x3 + c.spc_D * x1;
    const double v3 = 2.0 * c.spc_B * x3 + c.spc_C * (x1 + x2);
    for (int k = 0; k < 3; ++k) {
      dr1[0][k] = v1 * o1[k] / d1 + v2 * o2[k] / d2;
      dr1[1][k] = -v1 * o1[k] / d1 + v3 * hh[k] / d3;
      dr1[2][k] = -v2 * o2[k] / d2 - v3 * hh[k] / d3;
    }
    for (int iy = 0; iy < 3; ++iy)
      for (int ix = 0; ix < 3; ++ix)
        vt[iy * 3 + ix] = vt[iy * 3 + ix] - o1[ix] * dr1[1][iy] - o2[ix] * dr1[2][iy];
  }
}

// ---------------------------------------------------------------------------
// Monomer stage: molecule records, charges, intramolecular terms, Ewald self
// energy and the screened intramolecular charge pairs (pot_spc.f90:55-137,
// 202-233).  One thread per (bead, molecule).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_spc_monomer(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const double *__restrict__ RR = Rp.b.RR;
  Site *site = Rp.b.site;
  double *phi = Rp.b.phi, *Efq = Rp.b.Efq, *dR = Rp.b.dR, *tally = Rp.b.tally;
  double *chg_out = Rp.b.chg, *radius = Rp.b.radius;
  int *flags = Rp.b.flags;
  const int gid = bk.x * blockDim.x + threadIdx.x;
  if (gid >= c.P * c.N) return;
  const int b = gid / c.N, w = gid % c.N, N = c.N;
  const double *rr = RR + ((size_t)b * 3 * N + 3 * w) * 3;
  double r[3][3];
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) r[a][k] = rr[3 * a + k];
  flag_nonfinite(flags, r);
  const double qO = c.spc_q[0], qH = c.spc_q[1];
  if (chg_out) {
    double *co = chg_out + (size_t)b * 3 * N + 3 * w;
    co[0] = qO;
    co[1] = qH;
    co[2] = qH;
  }
  Site *sm = site + ((size_t)b * N + w) * 4;
  sm[0] = Site{r[0][0], r[0][1], r[0][2], 0.0};
  sm[1] = Site{r[1][0], r[1][1], r[1][2], qH};
  sm[2] = Site{r[2][0], r[2][1], r[2][2], qH};
  sm[3] = Site{r[0][0], r[0][1], r[0][2], qO};
  radius[(size_t)b * N + w] = 0.0;
  double e1, dr1[3][3], vt[9] = {0};
  spc_intra(c, r, e1, dr1, vt);
  // atoms O, H1, H2 and their record slots
  const int slot[3] = {3, 1, 2};
  const double q[3] = {qO, qH, qH};
  double ph[3], ef[3][3];
  for (int a = 0; a < 3; ++a) {
    ph[a] = -c.self_fac * q[a];
    ef[a][0] = ef[a][1] = ef[a][2] = 0.0;
  }
  for (int i = 0; i < 2; ++i)
    for (int j = i + 1; j < 3; ++j) {
      double v[3] = {r[i][0] - r[j][0], r[i][1] - r[j][1], r[i][2] - r[j][2]};
      minimg_div(c, v);
      const double d2 = norm2_exact(v[0], v[1], v[2]);
      double ts0, ts1;
      self1(c, d2, ts0, ts1);
      ph[i] = ph[i] + ts0 * q[j];
      ph[j] = ph[j] + ts0 * q[i];
      double d[3];
      for (int k = 0; k < 3; ++k) {
        ef[i][k] = ef[i][k] + ts1 * q[j] * v[k];
        ef[j][k] = ef[j][k] - ts1 * q[i] * v[k];
        d[k] = -(ts1 * q[i] * q[j] * v[k]);
      }
      for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) vt[col * 3 + row] += d[col] * v[row];
    }
  const size_t base = ((size_t)b * N + w) * 4;
  phi[base] = 0.0;
  for (int k = 0; k < 3; ++k) {
    Efq[base * 3 + k] = 0.0;
    dR[base * 3 + k] = dr1[0][k];            // O: intramolecular (and later vdW) gradient
    dR[(base + 3) * 3 + k] = 0.0;            // O charge site
    dR[(base + 1) * 3 + k] = dr1[1][k];
    dR[(base + 2) * 3 + k] = dr1[2][k];
  }
  for (int a = 0; a < 3; ++a) {
    phi[base + slot[a]] = ph[a];
    for (int k = 0; k < 3; ++k) Efq[(base + slot[a]) * 3 + k] = ef[a][k];
  }
  double *tl = tally + ((size_t)b * N + w) * T_COUNT;
  tl[T_UMON] = e1;
  tl[T_UVDW] = 0.0;
  tl[T_UELEC] = 0.0;
  tl[T_UIND] = 0.0;
  tl[T_VIRC] = 0.0;
  for (int i = 0; i < 9; ++i) tl[T_VIRT + i] = vt[i];
}

void launch_spc_monomer(const Group &g, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        return Plan{c.model >= PIMD_QSPCFW ? 0 : -1, (c.P * c.N + 63) / 64, 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_spc_monomer, grid, dim3(64), 0, s, g.d, m);
      });
}

// ---------------------------------------------------------------------------
// Real-space pass over O-O neighbours (pot_spc.f90:142-191): the vdW term
// A/r^12 + C/r^6 between oxygens and the erfc-screened charge pairs of the
// nine site pairs (no Thole damping).  A block holds 256/G molecules of one
// bead, whose molecule records are staged in LDS; each molecule's G lanes
// walk its neighbour list, a pair being evaluated from both sides, and
// scalars are tallied from the lower molecule index.  R: pair arithmetic
// (double, or float in the mixed form with FP64 site vectors, cut-off tests
// and accumulation).
// ---------------------------------------------------------------------------
template <int G, typename R, bool LDS>
__global__ __launch_bounds__(256) void k_spc_pair(const Rep *__restrict__ reps, BlockMap bm) {
  extern __shared__ Site lds_site[];
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const int *__restrict__ nbr = Rp.b.nbr, *__restrict__ nbrc = Rp.b.nbrc, *__restrict__ nbrn = Rp.b.nbrn;
  double *phi = Rp.b.phi, *Efq = Rp.b.Efq, *dR = Rp.b.dR, *tally = Rp.b.tally;
  const int N = c.N, bb = bk.y;
  if (LDS) {
    for (int e = threadIdx.x; e < 4 * N; e += blockDim.x) lds_site[e] = site[(size_t)bb * 4 * N + e];
    __syncthreads();
  }
  // the bead's molecule records: staged in LDS, or read from global memory
  // when they do not fit
  const Site *recs = LDS ? lds_site : site + (size_t)bb * 4 * N;
  const int aa = bk.x * (256 / G) + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const bool valid = aa < N;
  const int a = valid ? aa : 0;
  const long grp = (long)bb * N + a;
  Site A[4];
  for (int t = 0; t < 4; ++t) A[t] = recs[a * 4 + t];
  double ph[4] = {0, 0, 0, 0}, ef[4][3] = {{0}}, dro[3] = {0, 0, 0};
  // the energy and virial terms of a pair are the same from either
  // molecule: every lane adds them for all its pairs, halved at the end
  double uvdw = 0.0, vt[9] = {0};
  const int nf = valid ? nbrc[grp] : 0, nn = valid ? nbrn[grp] : 0;
  const int *list = nbr + (size_t)grp * c.nbr_cap;
  auto visit = [&](const int j) {
    Site B[4];
    for (int t = 0; t < 4; ++t) B[t] = recs[j * 4 + t];
    double r[3] = {A[0].x - B[0].x, A[0].y - B[0].y, A[0].z - B[0].z};
    minimg(c, r[0], r[1], r[2]);
    const double r2 = norm2_exact(r[0], r[1], r[2]);
    const R r2r = (R)r2;
    const R R2i = RM<R>::rcp(r2r), R6i = R2i * R2i * R2i, R8i = R6i * R2i;
    const R uij = R6i * ((R)c.vdwA * R6i + (R)c.vdwC);
    const R fsc = -(R8i * ((R)12.0 * (R)c.vdwA * R6i + (R)6.0 * (R)c.vdwC));
    const double d[3] = {(double)(fsc * (R)r[0]), (double)(fsc * (R)r[1]),
                         (double)(fsc * (R)r[2])};
    for (int k = 0; k < 3; ++k) dro[k] += d[k];
    uvdw += (double)uij;
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row < 3; ++row) vt[col * 3 + row] += d[col] * r[row];
#pragma unroll
    for (int sa = 1; sa < 4; ++sa) {
#pragma unroll
      for (int sb = 1; sb < 4; ++sb) {
        double v[3] = {A[sa].x - B[sb].x, A[sa].y - B[sb].y, A[sa].z - B[sb].z};
        minimg(c, v[0], v[1], v[2]);
        const double d2 = norm2_exact(v[0], v[1], v[2]);
        const R vr[3] = {(R)v[0], (R)v[1], (R)v[2]};
        const Smear<R> sm = smear_base<R>(c, (R)d2);
        const R ts0 = sm.er * sm.dri;
        const R ts1 = (ts0 + (R)c.cts1 * sm.expon) * sm.drsqi;
        const R qb = (R)B[sb].q;
        ph[sa] += (double)(ts0 * qb);
        for (int k = 0; k < 3; ++k) ef[sa][k] += (double)(ts1 * qb * vr[k]);
        {
          const R qa = (R)A[sa].q;
          double dd[3];
          for (int k = 0; k < 3; ++k) dd[k] = (double)(-(ts1 * qa * qb * vr[k]));
          for (int col = 0; col < 3; ++col)
            for (int row = 0; row < 3; ++row) vt[col * 3 + row] += dd[col] * v[row];
        }
      }
    }
  };
  // near and far partners alike: there is no damping radius here
  walk_list<G>(list, lane, nn, visit);
  walk_list<G>(list, c.nbr_cap - nf + lane, c.nbr_cap, visit);
  for (int t = 1; t < 4; ++t) ph[t] = group_sum<G>(ph[t]);
  for (int t = 1; t < 4; ++t)
    for (int k = 0; k < 3; ++k) ef[t][k] = group_sum<G>(ef[t][k]);
  for (int k = 0; k < 3; ++k) dro[k] = group_sum<G>(dro[k]);
  uvdw = group_sum<G>(uvdw);
  for (int i = 0; i < 9; ++i) vt[i] = group_sum<G>(vt[i]);
  if (valid && lane == 0) {
    const size_t base = ((size_t)bb * N + a) * 4;
    for (int t = 1; t < 4; ++t) {
      phi[base + t] += ph[t];
      for (int k = 0; k < 3; ++k) Efq[(base + t) * 3 + k] += ef[t][k];
    }
    for (int k = 0; k < 3; ++k) dR[base * 3 + k] += dro[k];
    for (int i = 0; i < 9; ++i) vt[i] *= 0.5;
    double *tl = tally + ((size_t)bb * N + a) * T_COUNT;
    tl[T_UVDW] += 0.5 * uvdw;
    vadd9(tl + T_VIRT, vt);
  }
}

// Variant: bit 0 FP32 pair arithmetic, bit 1 records staged in LDS, bits 2..
// the lanes per molecule (4, 8, 16, 32).
template <typename R, bool LDS>
static void launch_spc_pair_v(int gi, dim3 grid, const BlockMap &m, size_t lds, const Group &g,
                              hipStream_t s) {
  switch (gi) {
    case 0: hipLaunchKernelGGL((k_spc_pair<4, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    case 2: hipLaunchKernelGGL((k_spc_pair<16, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    case 3: hipLaunchKernelGGL((k_spc_pair<32, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    default: hipLaunchKernelGGL((k_spc_pair<8, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
  }
}

void launch_spc_pair(const Group &g, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        if (c.model < PIMD_QSPCFW) return Plan{-1, 0, 0, 0, 0};
        const int G = (c.g_qq == 4 || c.g_qq == 16 || c.g_qq == 32) ? c.g_qq : 8;
        const int gi = G == 4 ? 0 : G == 8 ? 1 : G == 16 ? 2 : 3;
        const size_t lds = sizeof(Site) * 4 * c.N;
        const bool in_lds = lds <= 150 * 1024;
        return Plan{(c.mixed ? 1 : 0) | (in_lds ? 2 : 0) | (gi << 2),
                    (c.N + 256 / G - 1) / (256 / G), c.P, 1, in_lds ? lds : 0};
      },
      [&](int v, dim3 grid, const BlockMap &m, size_t lds) {
        const int gi = v >> 2;
        switch (v & 3) {
          case 0: launch_spc_pair_v<double, false>(gi, grid, m, lds, g, s); break;
          case 1: launch_spc_pair_v<float, false>(gi, grid, m, lds, g, s); break;
          case 2: launch_spc_pair_v<double, true>(gi, grid, m, lds, g, s); break;
          default: launch_spc_pair_v<float, true>(gi, grid, m, lds, g, s); break;
        }
      });
}

// ---------------------------------------------------------------------------
// Per-molecule finish (pot_spc.f90:242-275): Uelec partials, the charge
// gradient -q Efq, the gradient in OHH order and the molecular dipoles.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_spc_finalize(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ phi = Rp.b.phi, *__restrict__ Efq = Rp.b.Efq;
  const double *__restrict__ kphi = Rp.b.kphi, *__restrict__ kE = Rp.b.kE;
  const double *__restrict__ dR = Rp.b.dR;
  double *tally = Rp.b.tally, *dRR = Rp.b.dRR, *dipI = Rp.b.dipI, *dipE = Rp.b.dipE;
  const int gid = bk.x * blockDim.x + threadIdx.x;
  if (gid >= c.P * c.N) return;
  const int b = gid / c.N, w = gid % c.N, N = c.N;
  const size_t base = ((size_t)b * N + w) * 4;
  double ue = 0.0, g[4][3];
  for (int t = 1; t < 4; ++t) {
    // real-space sums plus the reciprocal-space part (ewald.f90:270-276)
    const double q = site[base + t].q;
    const double pt = phi[base + t] + kphi[base + t];
    ue += q * pt;
    for (int k = 0; k < 3; ++k) {
      const double e = Efq[(base + t) * 3 + k] + kE[(base + t) * 3 + k];
      g[t][k] = dR[(base + t) * 3 + k] - q * e;
    }
  }
  double *out = dRR + ((size_t)b * 3 * N + 3 * w) * 3;
  for (int k = 0; k < 3; ++k) {
    out[k] = dR[base * 3 + k] + g[3][k];
    out[3 + k] = g[1][k];
    out[6 + k] = g[2][k];
  }
  const Site O = site[base], H1 = site[base + 1], H2 = site[base + 2];
  double a1[3] = {H1.x - O.x, H1.y - O.y, H1.z - O.z}, a2[3] = {H2.x - O.x, H2.y - O.y, H2.z - O.z};
  minimg_div(c, a1);
  minimg_div(c, a2);
  for (int k = 0; k < 3; ++k) {
    if (dipI) dipI[((size_t)b * N + w) * 3 + k] = H1.q * a1[k] + H2.q * a2[k];
    if (dipE) dipE[((size_t)b * N + w) * 3 + k] = 0.0;  // no induced dipoles
  }
  tally[((size_t)b * N + w) * T_COUNT + T_UELEC] = ue;
}

void launch_spc_finalize(const Group &g, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        return Plan{c.model >= PIMD_QSPCFW ? 0 : -1, (c.P * c.N + 63) / 64, 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_spc_finalize, grid, dim3(64), 0, s, g.d, m);
      });
}

}  // namespace pimd
