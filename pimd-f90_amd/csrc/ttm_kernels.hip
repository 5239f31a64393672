// HIP kernels of the TTM2.1-F / TTM3-F force path for all P beads of a system.
//
// Layout in HBM (per context):
//   site  [P][N][4]       Site records (x, y, z, q) of O, H1, H2, M
//   phi, Efq, dR          per-site potential, charge field, gradient
//   dip, Efd [P][N][nd][3] induced dipoles and their field (nd = 1 for TTM3-F
//                           (M site), 3 for TTM2.1-F (O, H1, H2))
//   tab   [P][T][4N]      e^{i 2 pi h x/L} per site for h = 0..hmax (x),
//                         0..kmax (y), 0..lmax (z)
//   Sq, Sm, Sv [P][nk]    structure factors of charges, dipoles, dipole vectors
//   tally [P][N][14]      per-molecule partial energies / virials
//
// Real-space kernels give each molecule a group of G lanes that walks its full
// neighbour list and accumulates only that molecule's sites, so every result is
// written by exactly one owner and sums are deterministic without atomics.
// Pair scalars (energies, virials) are counted once, from the lower-index
// molecule, in the reference's orientation.
#include <algorithm>

#include "ttm_device.h"
#include <cstdlib>

namespace pimd {

// c5z = f5z*c5zA + fbasis*cbasis + fcore*ccore + frest*crest (nasa.f90:45).
// The reference's PES parameters in the order of the pes arrays, then its
// DMS coefficients coefD (nasa_mod.f90).
void pes_default_params(double *theta) {
  for (int j = 0; j < PS_NTERM; ++j)
    theta[j] = PS_F5Z * PS_C5ZA[j] + PS_FBASIS * PS_CBASIS[j] + PS_FCORE * PS_CCORE[j] +
               PS_FREST * PS_CREST[j];
  theta[PES_DEOHA] = PS_DEOHA;
  theta[PES_ALPHAOH] = PS_ALPHAOH;
  theta[PES_ROH] = PS_ROH;
  theta[PES_PHH1A] = PS_PHH1A;
  theta[PES_PHH2] = PS_PHH2;
  for (int j = 0; j < DMS_NCOEF; ++j) theta[PES_DMSC + j] = PS_COEFD[j];
}


// Largest distance of H1, H2, M from O in a molecule record.
__device__ __forceinline__ double site_radius(const SysConst &c, const Site *m) {
  double r2 = 0.0;
  for (int t = 1; t < 4; ++t) {
    double x = m[t].x - m[0].x, y = m[t].y - m[0].y, z = m[t].z - m[0].z;
    minimg(c, x, y, z);
    r2 = fmax(r2, x * x + y * y + z * z);
  }
  return sqrt(r2);
}

// ---------------------------------------------------------------------------
// Monomer stage (pot_ttm.f90:56-151 and 232-261): M sites, PS energy and
// gradient, DMS charges and their gradients, intramolecular virial, Ewald
// self energy and the screened intramolecular charge pairs.  One thread per
// (bead, molecule).
// ---------------------------------------------------------------------------
// A block holds 64 (bead, molecule) pairs, one per lane, in MONO_W waves:
// wave g sums the g-th share of the PS and DMS polynomial terms (so the term
// index, and the table reads it drives, stay uniform across a wave) and the
// shares are added through LDS; wave 0 writes the results.
constexpr int MONO_W = 4;
__global__ __launch_bounds__(64 * MONO_W) void k_monomer(const Rep *__restrict__ reps, BlockMap bm) {
  __shared__ double lds_part[MONO_W * 8 * 64];
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const double *__restrict__ RR = Rp.b.RR, *__restrict__ RRc = Rp.b.RRc;
  Site *site = Rp.b.site;
  double *Rcs = Rp.b.Rcs, *grdq = Rp.b.grdq, *phi = Rp.b.phi, *Efq = Rp.b.Efq, *dR = Rp.b.dR;
  double *tally = Rp.b.tally, *chg_out = Rp.b.chg, *radius = Rp.b.radius;
  const double *__restrict__ pes = Rp.b.pes;
  int *flags = Rp.b.flags;
  const int wv = threadIdx.x >> 6;
  const int gid0 = bk.x * 64 + (threadIdx.x & 63);
  const bool valid = gid0 < c.P * c.N;
  const int gid = valid ? gid0 : 0;  // spare lanes take part in the sums, write nothing
  const int b = gid / c.N, w = gid % c.N, N = c.N;
  const double *rr = RR + ((size_t)b * 3 * N + 3 * w) * 3;
  double r[3][3], rc[3][3];
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) r[a][k] = rr[3 * a + k];
  double e1, dr1[3][3], q3[3], gq[3][3][3];
  ps_monomer<true, MONO_W>(c, r, e1, dr1, q3, gq, pes, wv, WaveSplit<MONO_W>{lds_part});
  if (wv != 0 || !valid) return;
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) rc[a][k] = RRc[(3 * w + a) * 3 + k];
  flag_nonfinite(flags, r);
  if (b == 0) flag_nonfinite(flags, rc);
  const double gM = c.gammaM;
  double pos[4][3], cen[4][3];
  {
    double o1[3], o2[3], c1[3], c2[3];
    for (int k = 0; k < 3; ++k) {
      o1[k] = r[1][k] - r[0][k];
      o2[k] = r[2][k] - r[0][k];
      c1[k] = rc[1][k] - rc[0][k];
      c2[k] = rc[2][k] - rc[0][k];
    }
    minimg(c, o1[0], o1[1], o1[2]);
    minimg(c, o2[0], o2[1], o2[2]);
    minimg(c, c1[0], c1[1], c1[2]);
    minimg(c, c2[0], c2[1], c2[2]);
    for (int k = 0; k < 3; ++k) {
      for (int a = 0; a < 3; ++a) {
        pos[a][k] = r[a][k];
        cen[a][k] = rc[a][k];
      }
      pos[3][k] = 0.5 * gM * (o1[k] + o2[k]) + r[0][k];
      cen[3][k] = 0.5 * gM * (c1[k] + c2[k]) + rc[0][k];
    }
  }
  if (b == 0)
    for (int t = 0; t < 4; ++t)
      for (int k = 0; k < 3; ++k) Rcs[(w * 4 + t) * 3 + k] = cen[t][k];

  if (c.contraction) {  // pot_ttm.f90:92-99: no monomer energy, gradient or virial
    e1 = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < 3; ++k) dr1[a][k] = 0.0;
  }
  for (int a = 0; a < 3; ++a) q3[a] = q3[a] * CHARGECON;
  if (chg_out)
    for (int a = 0; a < 3; ++a) chg_out[(size_t)b * 3 * N + 3 * w + a] = q3[a];
  const double tq = 0.5 * gM / (1.0 - gM);
  double q[4];
  q[0] = 0.0;
  q[1] = q3[1] + tq * (q3[1] + q3[2]);
  q[2] = q3[2] + tq * (q3[1] + q3[2]);
  q[3] = q3[0] / (1.0 - gM);
  Site *sm = site + ((size_t)b * N + w) * 4;
  for (int t = 0; t < 4; ++t) sm[t] = Site{pos[t][0], pos[t][1], pos[t][2], q[t]};
  radius[(size_t)b * N + w] = site_radius(c, sm);
  double *g = grdq + ((size_t)b * N + w) * 27;
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) {
      double d0 = gq[a][0][k] * CHARGECON, d1 = gq[a][1][k] * CHARGECON,
             d2 = gq[a][2][k] * CHARGECON;
      g[a * 9 + 0 * 3 + k] = d0 + tq * (d0 + d1);
      g[a * 9 + 1 * 3 + k] = d1 + tq * (d0 + d1);
      g[a * 9 + 2 * 3 + k] = d2 - 2.0 * tq * (d0 + d1);
    }
  // intramolecular virial of the PS gradient (pot_ttm.f90:109-131)
  double vt[9] = {0}, vc = 0.0;
  {
    double o1[3], o2[3], c1[3], c2[3];
    for (int k = 0; k < 3; ++k) {
      o1[k] = r[1][k] - r[0][k];
      o2[k] = r[2][k] - r[0][k];
      c1[k] = rc[1][k] - rc[0][k];
      c2[k] = rc[2][k] - rc[0][k];
    }
    minimg(c, o1[0], o1[1], o1[2]);
    minimg(c, o2[0], o2[1], o2[2]);
    minimg(c, c1[0], c1[1], c1[2]);
    minimg(c, c2[0], c2[1], c2[2]);
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row < 3; ++row)
        vt[col * 3 + row] += o1[row] * dr1[1][col] + o2[row] * dr1[2][col];
    vc = vc + (c1[0] * dr1[1][0] + c1[1] * dr1[1][1] + c1[2] * dr1[1][2]);
    vc = vc + (c2[0] * dr1[2][0] + c2[1] * dr1[2][1] + c2[2] * dr1[2][2]);
  }
  // Ewald self energy and screened intramolecular charge pairs (pot_ttm.f90:232-261)
  double ph[4], ef[4][3];
  for (int t = 0; t < 4; ++t) {
    ph[t] = t == 0 ? 0.0 : -c.self_fac * q[t];
    ef[t][0] = ef[t][1] = ef[t][2] = 0.0;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 4; ++j) {
      double x = pos[i][0] - pos[j][0], y = pos[i][1] - pos[j][1], z = pos[i][2] - pos[j][2];
      minimg(c, x, y, z);
      double d2 = norm2_exact(x, y, z), ts0, ts1;
      self1(c, d2, ts0, ts1);
      ph[i] = ph[i] + ts0 * q[j];
      ph[j] = ph[j] + ts0 * q[i];
      double rv[3] = {x, y, z}, d[3];
      for (int k = 0; k < 3; ++k) {
        ef[i][k] = ef[i][k] + ts1 * q[j] * rv[k];
        ef[j][k] = ef[j][k] - ts1 * q[i] * rv[k];
        d[k] = -(ts1 * q[i] * q[j] * rv[k]);
      }
      double cx = cen[i][0] - cen[j][0], cy = cen[i][1] - cen[j][1], cz = cen[i][2] - cen[j][2];
      minimg(c, cx, cy, cz);
      vc = vc + (cx * d[0] + cy * d[1] + cz * d[2]);
      for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) vt[col * 3 + row] += d[col] * rv[row];
    }
  const size_t base = ((size_t)b * N + w) * 4;
  for (int t = 0; t < 4; ++t) {
    phi[base + t] = ph[t];
    for (int k = 0; k < 3; ++k) {
      Efq[(base + t) * 3 + k] = ef[t][k];
      dR[(base + t) * 3 + k] = t < 3 ? dr1[t][k] : 0.0;
    }
  }
  double *tl = tally + ((size_t)b * N + w) * T_COUNT;
  tl[T_UMON] = e1;
  tl[T_UVDW] = 0.0;
  tl[T_UELEC] = 0.0;
  tl[T_UIND] = 0.0;
  tl[T_VIRC] = vc;
  for (int i = 0; i < 9; ++i) tl[T_VIRT + i] = vt[i];
}


// ---------------------------------------------------------------------------
// Contracted MD (MD.f90:276-449).
// ---------------------------------------------------------------------------

// Monomer dipole moment of every molecule of every bead from its dipole-
// moment-surface charges (force_calc.f90:205-264): the H charges with the
// M-site share folded in, times the H positions relative to the M site
// (TTM3-F, q-SPC/Fw) or to the O (TTM2.1-F, SPC/F), minimum image.
// One thread per (bead, molecule); RR [P][3N][3], out [P][N][3].
__device__ inline void monomer_dipole(const SysConst &c, const double *__restrict__ pes,
                                      const double *__restrict__ rr, double *__restrict__ out) {
  double r[3][3];
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) r[a][k] = rr[3 * a + k];
  double e1, dr1[3][3], q3[3], gq[3][3][3];
  ps_monomer<true>(c, r, e1, dr1, q3, gq, pes);
  const double gM = c.gammaM, tq = 0.5 * gM / (1.0 - gM);
  const double qh1 = q3[1] * CHARGECON + tq * (q3[1] * CHARGECON + q3[2] * CHARGECON);
  const double qh2 = q3[2] * CHARGECON + tq * (q3[1] * CHARGECON + q3[2] * CHARGECON);
  double o1[3], o2[3];
  for (int k = 0; k < 3; ++k) {
    o1[k] = r[1][k] - r[0][k];
    o2[k] = r[2][k] - r[0][k];
  }
  minimg(c, o1[0], o1[1], o1[2]);
  minimg(c, o2[0], o2[1], o2[2]);
  const bool msite = c.model == PIMD_TTM3F || c.model == PIMD_QSPCFW;
  double h1[3], h2[3];
  for (int k = 0; k < 3; ++k) {
    h1[k] = o1[k];
    h2[k] = o2[k];
  }
  if (msite) {
    for (int k = 0; k < 3; ++k) {
      const double m = 0.5 * gM * (o1[k] + o2[k]) + r[0][k];
      h1[k] = r[1][k] - m;
      h2[k] = r[2][k] - m;
    }
    minimg(c, h1[0], h1[1], h1[2]);
    minimg(c, h2[0], h2[1], h2[2]);
  }
  for (int k = 0; k < 3; ++k) out[k] = qh1 * h1[k] + qh2 * h2[k];
}

__global__ __launch_bounds__(64) void k_monomer_dipoles(SysConst c, const double *__restrict__ pes,
                                                        const double *__restrict__ RR,
                                                        double *__restrict__ out) {
  const int g = blockIdx.x * 64 + threadIdx.x;
  if (g >= c.P * c.N) return;
  monomer_dipole(c, pes, RR + (size_t)g * 9, out + (size_t)g * 3);
}

// The same for every replica of a group, from its coordinates RR into its
// dipI (contracted MD's bead dipoles, which its force call does not make).
__global__ __launch_bounds__(64) void k_monomer_dipoles_rep(const Rep *__restrict__ reps,
                                                            BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const int g = bk.x * 64 + threadIdx.x;
  if (g >= c.P * c.N) return;
  monomer_dipole(c, Rp.b.pes, Rp.b.RR + (size_t)g * 9, Rp.b.dipI + (size_t)g * 3);
}

void launch_monomer_dipoles(const Group &g, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        const bool ttm = c.model == PIMD_TTM21F || c.model == PIMD_TTM3F;
        return Plan{ttm ? 0 : -1, (int)(((long)c.P * c.N + 63) / 64), 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_monomer_dipoles_rep, grid, dim3(64), 0, s, g.d, m);
      });
}

void launch_monomer_dipoles(const SysConst &c, const double *pes, const double *RR, double *out,
                            hipStream_t s) {
  const int n = c.P * c.N;
  hipLaunchKernelGGL(k_monomer_dipoles, dim3((n + 63) / 64), dim3(64), 0, s, c, pes, RR, out);
}


// Fast (intramolecular) forces: the PES gradient of every monomer of every
// bead into dRR (MD.f90:339-353).  With record, also per molecule the energy,
// the bead virial and the centroid virial (MD.f90:355-372) into mol[P][N][3].
__global__ __launch_bounds__(64 * MONO_W) void k_pes_fast(const Rep *__restrict__ reps, BlockMap bm,
                                                         int record) {
  __shared__ double lds_part[MONO_W * 8 * 64];
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const double *__restrict__ RR = Rp.b.RR, *__restrict__ RRc = Rp.b.RRc;
  const double *__restrict__ pes = Rp.b.pes;
  double *dRR = Rp.b.dRRfast, *mol = Rp.b.mol;
  const int wv = threadIdx.x >> 6;
  const int gid0 = bk.x * 64 + (threadIdx.x & 63);
  const bool valid = gid0 < c.P * c.N;
  const int gid = valid ? gid0 : 0;  // spare lanes take part in the sums, write nothing
  const int b = gid / c.N, w = gid % c.N, N = c.N;
  const double *rr = RR + ((size_t)b * 3 * N + 3 * w) * 3;
  double r[3][3], e1, dr1[3][3], q3[3], gq[3][3][3];
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) r[a][k] = rr[3 * a + k];
  ps_monomer<false, MONO_W>(c, r, e1, dr1, q3, gq, pes, wv, WaveSplit<MONO_W>{lds_part});
  if (wv != 0 || !valid) return;
  double *out = dRR + ((size_t)b * 3 * N + 3 * w) * 3;
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 3; ++k) out[3 * a + k] = dr1[a][k];
  if (!record) return;
  double vb = 0.0, vc = 0.0;
  for (int h = 1; h <= 2; ++h) {
    double rb[3], rcv[3];
    for (int k = 0; k < 3; ++k) {
      rcv[k] = RRc[(3 * w + h) * 3 + k] - RRc[(3 * w) * 3 + k];
      rb[k] = r[h][k] - r[0][k];
    }
    minimg(c, rcv[0], rcv[1], rcv[2]);
    minimg(c, rb[0], rb[1], rb[2]);
    vc += rcv[0] * dr1[h][0] + rcv[1] * dr1[h][1] + rcv[2] * dr1[h][2];
    vb += rb[0] * dr1[h][0] + rb[1] * dr1[h][1] + rb[2] * dr1[h][2];
  }
  double *m = mol + ((size_t)b * N + w) * 3;
  m[0] = e1;
  m[1] = vb;
  m[2] = vc;
}


// The centroid's gradient for every bead (MD.f90:400-403).
__global__ void k_broadcast_beads(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const int P = Rp.c.P, n = 9 * Rp.c.N;
  const size_t i = (size_t)bk.x * blockDim.x + threadIdx.x;
  if (i < (size_t)P * n) Rp.b.dRR[i] = Rp.b.cen_dRR[i % n];
}


// Step record of a contracted step (MD.f90:405-418): sum of the bead
// energies (monomers of every bead + P times the centroid energy), virial,
// centroid virial, kinetic energy; molecule terms summed in a fixed order.
__global__ __launch_bounds__(256) void k_contract_record(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const int P = Rp.c.P, N = Rp.c.N;
  const double *__restrict__ mol = Rp.b.mol, *__restrict__ uk = Rp.b.uk;
  const double *__restrict__ Ucen = Rp.b.cen_Upot, *__restrict__ vtcen = Rp.b.cen_virt;
  const double *__restrict__ vccen = Rp.b.cen_virialc;
  double *rec = Rp.m.rec + (size_t)Rp.b.clk[1] * MD_REC;
  __shared__ double red[3][256];
  double s[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < P * N; i += 256)
    for (int k = 0; k < 3; ++k) s[k] += mol[(size_t)i * 3 + k];
  for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    rec[0] = red[0][0] + (double)P * Ucen[0];
    rec[1] = red[1][0] + (vtcen[0] + vtcen[4] + vtcen[8]);
    rec[2] = red[2][0] / (double)P + vccen[0];
    rec[3] = uk[0];
  }
}


// ---------------------------------------------------------------------------
// Neighbour lists (find_neigh.f90:14-120): for every bead and molecule i, all
// molecules j != i whose minimum-image O-O distance satisfies r^2 < Rc^2, in
// ascending j.  One wavefront per (bead, i) at a time; ballot + mbcnt
// compaction.
// ---------------------------------------------------------------------------
// Each list holds near partners from the front (ascending j) and far
// partners from the back of its capacity: a far partner's O lies at least
// 5 A + rad_i + rad_j away, so none of its site pairs is inside the 5 A Thole
// radius (smear.f90:23,57,92) and the damping branch can be compiled out.
// nbrn = number of near, nbrc = number of far partners.
// A block takes NB_I molecules i of one bead, a wave at a time; with LDS the
// bead's O positions and radii are staged there once per block.
constexpr int NB_I = 16;
template <bool LDS>
__global__ __launch_bounds__(256) void k_neighbors(const Rep *__restrict__ reps, BlockMap bm) {
  extern __shared__ double lds_o[];  // [N][4]: x, y, z, radius
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ radius = Rp.b.radius;
  int *nbr = Rp.b.nbr, *nbrc = Rp.b.nbrc, *nbrn = Rp.b.nbrn;
  const int b = bk.y, N = c.N;
  if (LDS) {
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
      const Site o = site[((size_t)b * N + j) * 4];
      lds_o[4 * j + 0] = o.x;
      lds_o[4 * j + 1] = o.y;
      lds_o[4 * j + 2] = o.z;
      lds_o[4 * j + 3] = radius[(size_t)b * N + j];
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  int *nbrh = Rp.b.nbrh;
  for (int ii = wv; ii < NB_I; ii += blockDim.x >> 6) {
    const int i = bk.x * NB_I + ii;
    if (i >= N) break;
    // the last partner of i's half window, wrapped (pair_half_window)
    const int uw = pair_half_window(i, N) % N;
    double xi, yi, zi, radi;
    if (LDS) {
      xi = lds_o[4 * i]; yi = lds_o[4 * i + 1]; zi = lds_o[4 * i + 2]; radi = lds_o[4 * i + 3];
    } else {
      const Site o = site[((size_t)b * N + i) * 4];
      xi = o.x; yi = o.y; zi = o.z; radi = radius[(size_t)b * N + i];
    }
    const size_t row = (size_t)b * N + i;
    int *out = nbr + row * c.nbr_cap;
    int nnear = 0, nfar = 0, nlo[4] = {0, 0, 0, 0};
    for (int j0 = 0; j0 < N; j0 += 64) {
      const int j = j0 + lane;
      bool in = false, near = false;
      if (j < N && j != i) {
        double xj, yj, zj, radj;
        if (LDS) {
          xj = lds_o[4 * j]; yj = lds_o[4 * j + 1]; zj = lds_o[4 * j + 2]; radj = lds_o[4 * j + 3];
        } else {
          const Site o = site[((size_t)b * N + j) * 4];
          xj = o.x; yj = o.y; zj = o.z; radj = radius[(size_t)b * N + j];
        }
        double x = xi - xj, y = yi - yj, z = zi - zj;
        minimg(c, x, y, z);
        const double r2 = norm2_exact(x, y, z);
        in = r2 < c.rc2;
        const double lim = 5.0 + radi + radj + 1e-6;
        near = r2 < lim * lim;
      }
      const unsigned long long mn = __ballot(in && near), mf = __ballot(in && !near);
      if (in && near) out[nnear + __popcll(mn & below)] = j;
      if (in && !near) out[c.nbr_cap - 1 - (nfar + __popcll(mf & below))] = j;
      nnear += __popcll(mn);
      nfar += __popcll(mf);
      // entries below i and up to the window's end: the lists are in
      // ascending partner order
      const unsigned long long li = __ballot(j < i), lw = __ballot(j <= uw);
      nlo[0] += __popcll(mn & li);
      nlo[1] += __popcll(mn & lw);
      nlo[2] += __popcll(mf & li);
      nlo[3] += __popcll(mf & lw);
    }
    if (lane == 0) {
      nbrn[row] = nnear;
      nbrc[row] = nfar;
      if (nbrh)
        for (int k = 0; k < 4; ++k) nbrh[row * 4 + k] = nlo[k];
    }
  }
}

// Cell list of each bead (one block per bead): the cell of every oxygen
// (its position wrapped into the box), a counting sort into cell_mol with
// cell_start[c] the first entry of cell c, and each cell's molecules put in
// ascending order so that the lists do not depend on the order of atomics.
__global__ __launch_bounds__(256) void k_cells(const Rep *__restrict__ reps, BlockMap bm) {
  extern __shared__ int lds_cells[];  // [ncells] counts, then fill pointers
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  int *cell_start = Rp.b.cell_start, *cell_mol = Rp.b.cell_mol;
  const int b = bk.x, N = c.N;
  const int nx = c.ncell[0], ny = c.ncell[1], nz = c.ncell[2], nc = nx * ny * nz;
  int *cnt = lds_cells, *fill = lds_cells + nc;
  int *start = cell_start + (size_t)b * (nc + 1);
  int *mol = cell_mol + (size_t)b * N;
  auto cell_of = [&](int m) {
    const Site o = site[((size_t)b * N + m) * 4];
    const double r[3] = {o.x, o.y, o.z};
    int id[3];
    for (int k = 0; k < 3; ++k) {
      const double w = r[k] - c.box[k] * floor(r[k] * c.boxi[k]);  // in [0, L]
      id[k] = min((int)(w * c.boxi[k] * c.ncell[k]), c.ncell[k] - 1);
      id[k] = max(id[k], 0);
    }
    return (id[0] * ny + id[1]) * nz + id[2];
  };
  for (int e = threadIdx.x; e < nc; e += blockDim.x) cnt[e] = 0;
  __syncthreads();
  for (int m = threadIdx.x; m < N; m += blockDim.x) atomicAdd(&cnt[cell_of(m)], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int e = 0; e < nc; ++e) {
      start[e] = acc;
      fill[e] = acc;
      acc += cnt[e];
    }
    start[nc] = acc;
  }
  __syncthreads();
  for (int m = threadIdx.x; m < N; m += blockDim.x) mol[atomicAdd(&fill[cell_of(m)], 1)] = m;
  __syncthreads();
  // ascending molecule order inside each cell (a few molecules each)
  for (int e = threadIdx.x; e < nc; e += blockDim.x) {
    const int lo = start[e], hi = start[e + 1];
    for (int i = lo + 1; i < hi; ++i) {
      const int v = mol[i];
      int k = i - 1;
      while (k >= lo && mol[k] > v) {
        mol[k + 1] = mol[k];
        --k;
      }
      mol[k + 1] = v;
    }
  }
}

// The neighbour search of k_neighbors over the 27 cells around molecule i's
// cell (cells at least Rc wide, at least three per axis, so every partner
// inside Rc is among them once).  Partners come in cell order, ascending
// inside each cell.
__global__ __launch_bounds__(256) void k_neighbors_cells(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ radius = Rp.b.radius;
  const int *__restrict__ cell_start = Rp.b.cell_start, *__restrict__ cell_mol = Rp.b.cell_mol;
  int *nbr = Rp.b.nbr, *nbrc = Rp.b.nbrc, *nbrn = Rp.b.nbrn;
  const int wave = (bk.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= c.P * c.N) return;
  const int b = wave / c.N, i = wave % c.N, N = c.N;
  const int nx = c.ncell[0], ny = c.ncell[1], nz = c.ncell[2], nc = nx * ny * nz;
  const int *start = cell_start + (size_t)b * (nc + 1);
  const int *mol = cell_mol + (size_t)b * N;
  const Site Oi = site[((size_t)b * N + i) * 4];
  const double radi = radius[(size_t)b * N + i];
  int id[3];
  {
    const double r[3] = {Oi.x, Oi.y, Oi.z};
    for (int k = 0; k < 3; ++k) {
      const double w = r[k] - c.box[k] * floor(r[k] * c.boxi[k]);
      id[k] = max(min((int)(w * c.boxi[k] * c.ncell[k]), c.ncell[k] - 1), 0);
    }
  }
  int *out = nbr + (size_t)wave * c.nbr_cap;
  int nnear = 0, nfar = 0;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int cx = (id[0] + dx + nx) % nx, cy = (id[1] + dy + ny) % ny,
                  cz = (id[2] + dz + nz) % nz;
        const int cell = (cx * ny + cy) * nz + cz;
        const int lo = start[cell], hi = start[cell + 1];
        for (int e0 = lo; e0 < hi; e0 += 64) {
          const int e = e0 + lane;
          const int j = e < hi ? mol[e] : -1;
          bool in = false, near = false;
          if (j >= 0 && j != i) {
            const Site Oj = site[((size_t)b * N + j) * 4];
            double x = Oi.x - Oj.x, y = Oi.y - Oj.y, z = Oi.z - Oj.z;
            minimg(c, x, y, z);
            const double r2 = norm2_exact(x, y, z);
            in = r2 < c.rc2;
            const double lim = 5.0 + radi + radius[(size_t)b * N + j] + 1e-6;
            near = r2 < lim * lim;
          }
          const unsigned long long mn = __ballot(in && near), mf = __ballot(in && !near);
          if (in && near) out[nnear + __popcll(mn & below)] = j;
          if (in && !near) out[c.nbr_cap - 1 - (nfar + __popcll(mf & below))] = j;
          nnear += __popcll(mn);
          nfar += __popcll(mf);
        }
      }
  if (lane == 0) {
    nbrn[wave] = nnear;
    nbrc[wave] = nfar;
  }
}


// ---------------------------------------------------------------------------
// Real-space vdW + charge-charge pass (pot_ttm.f90:159-228).
// ---------------------------------------------------------------------------
// R is the type of the pair arithmetic: double for the FP64 form; float for
// the mixed form, where site-site vectors and cut-off tests stay FP64 and each
// pair's contributions are accumulated in FP64.
// LDS = true: a block holds 256/G molecules of one bead (grid (molecule
// blocks, beads)) and first copies the bead's molecule records into LDS, so
// neighbour sites are read from LDS instead of gathered from L2.
template <int MODEL, int G, typename R, bool LDS = false>
__global__ __launch_bounds__(256) void k_pair_qq(const Rep *__restrict__ reps, BlockMap bm) {
  constexpr int SA0 = MODEL == 2 ? 0 : 1;  // TTM3-F: O carries neither charge nor dipole
  // sites of j: unrolled in the FP32 form (more independent work per lane
  // at the same occupancy), one at a time in the FP64 form (fewer live
  // registers: two waves per SIMD instead of one)
  constexpr int SB_UNROLL = sizeof(R) == 4 ? 3 : 1;
  extern __shared__ Site lds_site[];
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ Rcs = Rp.b.Rcs;
  const int *__restrict__ nbr = Rp.b.nbr, *__restrict__ nbrc = Rp.b.nbrc, *__restrict__ nbrn = Rp.b.nbrn;
  double *phi = Rp.b.phi, *Efq = Rp.b.Efq, *dR = Rp.b.dR, *tally = Rp.b.tally;
  const int N = c.N;
  long grp;
  int lane;
  bool valid;
  if (LDS) {
    const int bb = bk.y;
    for (int e = threadIdx.x; e < 4 * N; e += blockDim.x) lds_site[e] = site[(size_t)bb * 4 * N + e];
    __syncthreads();
    const int aa = bk.x * (256 / G) + threadIdx.x / G;
    lane = threadIdx.x % G;
    valid = aa < N;
    grp = (long)bb * N + (valid ? aa : 0);
  } else {
    const long gtid = (long)bk.x * blockDim.x + threadIdx.x;
    grp = gtid / G;
    lane = (int)(gtid % G);
    valid = grp < (long)c.P * N;
  }
  const int b = valid ? (int)(grp / N) : 0, a = valid ? (int)(grp % N) : 0;
  // molecule records of this bead: LDS or global
  const Site *bead_sites = LDS ? lds_site : site + (size_t)b * N * 4;
  Site A[4];
  for (int t = 0; t < 4; ++t) A[t] = bead_sites[a * 4 + t];
  const double RcO[3] = {Rcs[a * 12 + 0], Rcs[a * 12 + 1], Rcs[a * 12 + 2]};
  double ph[4] = {0, 0, 0, 0}, ef[4][3] = {{0}}, dro[3] = {0, 0, 0};
  // The energy and virial terms of a pair are the same seen from either
  // molecule, so every lane adds them for all its pairs and the sums are
  // halved at the end (no branch on which molecule is the lower one).  The
  // centroid virial's Rc(a, site) . F part is formed once per site at the end
  // from the site's field: its pair forces are -q(a, site) times the field terms.
  double uvdw = 0.0, vc = 0.0, vt[9] = {0};
  const int nf = valid ? nbrc[grp] : 0, nn = valid ? nbrn[grp] : 0;
  const int *list = nbr + (size_t)grp * c.nbr_cap;
  auto visit = [&](const int j, auto thole_tag) {
    constexpr bool THOLE = decltype(thole_tag)::value;
    const bool lo = a < j;
    const Site B0 = bead_sites[j * 4];
    // vdW between oxygens
    double r[3] = {A[0].x - B0.x, A[0].y - B0.y, A[0].z - B0.z};
    minimg(c, r[0], r[1], r[2]);
    const double r2 = norm2_exact(r[0], r[1], r[2]);
    const R r2r = (R)r2;
    const R dRij = RM<R>::sqrt_(r2r);
    const R R2i = RM<R>::rcp(r2r), R4i = R2i * R2i, R6i = R2i * R2i * R2i, R8i = R6i * R2i;
    const R ex = (R)c.vdwD * RM<R>::exp_(-(R)c.vdwE * dRij);
    const R uij = R6i * (((R)c.vdwA * R6i + (R)c.vdwB * R4i) + (R)c.vdwC) + ex;
    const R xs = dRij - (R)c.rc1;
    const R fsc = -(R8i * (R)6.0 * (R)c.vdwC + RM<R>::div(ex * (R)c.vdwE, dRij) +
                    (R)c.shiftA * (xs * xs) + (R)c.shiftB * (xs * xs * xs));
    const double d[3] = {(double)(fsc * (R)r[0]), (double)(fsc * (R)r[1]),
                         (double)(fsc * (R)r[2])};
    for (int k = 0; k < 3; ++k) dro[k] += d[k];
    uvdw += (double)uij;
    {
      double cx = RcO[0] - Rcs[j * 12 + 0], cy = RcO[1] - Rcs[j * 12 + 1],
             cz = RcO[2] - Rcs[j * 12 + 2];
      minimg(c, cx, cy, cz);
      vc += cx * d[0] + cy * d[1] + cz * d[2];
    }
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row < 3; ++row) vt[col * 3 + row] += d[col] * r[row];
    // O-O image shift of the reference's charge loop, Rij - R(iO) + R(jO)
    // (pot_ttm.f90:193); it enters the centroid virial of the lower molecule.
    const double s0[3] = {r[0] - A[0].x + B0.x, r[1] - A[0].y + B0.y, r[2] - A[0].z + B0.z};
    // one site of j at a time (its record read in the loop), against every
    // site of a: each accumulator of a site of a still takes its terms in
    // the order (j, site of j)
#pragma unroll SB_UNROLL
    for (int sb = 1; sb < 4; ++sb) {
      const Site Bs = bead_sites[j * 4 + sb];
      const R qb = (R)Bs.q;
      // the pair-side part of the centroid virial's lever arm
      const double w[3] = {s0[0] - Rcs[(j * 4 + sb) * 3 + 0], s0[1] - Rcs[(j * 4 + sb) * 3 + 1],
                           s0[2] - Rcs[(j * 4 + sb) * 3 + 2]};
#pragma unroll
      for (int sa = SA0; sa < 4; ++sa) {
        // site-site minimum image, oriented from a to b (pot_ttm.f90:206-207)
        double v[3] = {A[sa].x - Bs.x, A[sa].y - Bs.y, A[sa].z - Bs.z};
        minimg(c, v[0], v[1], v[2]);
        const double d2 = norm2_exact(v[0], v[1], v[2]);
        const R vr[3] = {(R)v[0], (R)v[1], (R)v[2]};
        R ts0, ts1;
        smear01<THOLE>(c, (R)d2, (R)(lo ? c.apolCC[sa][sb] : c.apolCC[sb][sa]),
                       (R)(lo ? c.raCC[sa][sb] : c.raCC[sb][sa]), ts0, ts1);
        if (sa > 0) ph[sa] += (double)(ts0 * qb);
        for (int k = 0; k < 3; ++k) ef[sa][k] += (double)(ts1 * qb * vr[k]);
        if (sa > 0) {
          const R qa = (R)A[sa].q;
          double dd[3];
          for (int k = 0; k < 3; ++k) dd[k] = (double)(-(ts1 * qa * qb * vr[k]));
          vc += w[0] * dd[0] + w[1] * dd[1] + w[2] * dd[2];
          for (int col = 0; col < 3; ++col)
            for (int row = 0; row < 3; ++row) vt[col * 3 + row] += dd[col] * v[row];
        }
      }
    }
  };
  walk_list<G>(list, lane, nn, [&](int j) { visit(j, std::true_type{}); });
  walk_list<G>(list, c.nbr_cap - nf + lane, c.nbr_cap, [&](int j) { visit(j, std::false_type{}); });
  for (int t = 1; t < 4; ++t) ph[t] = group_sum<G>(ph[t]);
  for (int t = SA0; t < 4; ++t)
    for (int k = 0; k < 3; ++k) ef[t][k] = group_sum<G>(ef[t][k]);
  for (int k = 0; k < 3; ++k) dro[k] = group_sum<G>(dro[k]);
  uvdw = group_sum<G>(uvdw);
  vc = group_sum<G>(vc);
  for (int i = 0; i < 9; ++i) vt[i] = group_sum<G>(vt[i]);
  if (valid && lane == 0) {
    const size_t base = ((size_t)b * N + a) * 4;
    for (int t = 1; t < 4; ++t) phi[base + t] += ph[t];
    for (int t = SA0; t < 4; ++t)
      for (int k = 0; k < 3; ++k) Efq[(base + t) * 3 + k] += ef[t][k];
    for (int k = 0; k < 3; ++k) dR[base * 3 + k] += dro[k];
    // Rc(a, site) . (sum of the site's pair forces), the forces being
    // -q(a, site) times its field terms
    for (int t = 1; t < 4; ++t) {
      const double *rc = Rcs + (a * 4 + t) * 3;
      vc -= A[t].q * (rc[0] * ef[t][0] + rc[1] * ef[t][1] + rc[2] * ef[t][2]);
    }
    for (int i = 0; i < 9; ++i) vt[i] *= 0.5;
    double *tl = tally + ((size_t)b * N + a) * T_COUNT;
    tl[T_UVDW] += 0.5 * uvdw;
    tl[T_VIRC] += 0.5 * vc;
    vadd9(tl + T_VIRT, vt);
  }
}

// k_pair_qq with each molecule pair evaluated once, by the molecule that owns
// it (pair_half_window; lists in ascending partner order, nbrh giving the
// owned ranges).  The owner's side is summed as in k_pair_qq; the partner's
// potentials, fields and vdW force go into a per-block LDS table of
// two-word fixed-point sums (integer addition, so the result does not depend
// on the order the lanes add in), written out as the block's slab and added
// to the partner by k_pair_qq_partner in block order.  The block stages only
// the molecules it touches: its owners and the N/2 after them.  Site pairs: every charged
// site of one molecule with every charge-or-dipole site of the other, each
// feeding both sides (TTM2.1-F adds the pairs with the partner's oxygen,
// which k_pair_qq takes from the partner's side).  Energies and virials per
// pair are those k_pair_qq halves from its two sides, taken once.
// FP64 form: a value v goes in as h = rint(v 2^32) (range +-2^31) and the
// remainder v - h 2^-32, exact in FP64, as rint(. 2^64): each addition is
// exact to 2^-65, finer than FP64 rounding of the sums, so the result
// matches an FP64 sum and stays smooth in the model parameters.  FP32 form:
// one word, rint(v 2^40) (range +-2^23): its pair terms already round at
// 1e-7 relative, far above that quantum, and one LDS atomic per value in
// place of two.  Integer addition wraps modulo 2^64 and is associative: only
// totals must fit.
constexpr double QQ_HI = 4294967296.0;            // 2^32
constexpr double QQ_LO = 18446744073709551616.0;  // 2^64
constexpr double QQ_ONE = 1099511627776.0;        // 2^40
template <typename R>
struct QqWords {
  static constexpr int n = sizeof(R) == 8 ? 2 : 1;
};
// partner sites per step of the site loop: one at a time (unrolled, the
// per-site partner sums spill)
template <typename R>
struct QqUnroll {
  static constexpr int v = 1;
};
template <int NW>
__device__ __forceinline__ void qq_fix_add(unsigned long long *acc, double v) {
  if constexpr (NW == 2) {
    const double h = rint(fmin(fmax(v * QQ_HI, -9.2e18), 9.2e18));
    const double rem = fmin(fmax(v - h * (1.0 / QQ_HI), -1.0 / QQ_HI), 1.0 / QQ_HI);
    atomicAdd(acc, (unsigned long long)(long long)h);
    atomicAdd(acc + 1, (unsigned long long)__double2ll_rn(rem * QQ_LO));
  } else {
    atomicAdd(acc, (unsigned long long)__double2ll_rn(fmin(fmax(v * QQ_ONE, -9.2e18), 9.2e18)));
  }
}
template <int NW>
__device__ __forceinline__ double qq_fix_value(const unsigned long long *acc) {
  if constexpr (NW == 2)
    return (double)(long long)acc[0] * (1.0 / QQ_HI) + (double)(long long)acc[1] * (1.0 / QQ_LO);
  else
    return (double)(long long)acc[0] * (1.0 / QQ_ONE);
}

template <int MODEL, int G, typename R>
__global__ __launch_bounds__(QQ_HALF_BLOCK) void k_pair_qq_half(const Rep *__restrict__ reps,
                                                                BlockMap bm) {
  constexpr int SA0 = MODEL == 2 ? 0 : 1;  // first dipole-carrying (field-target) site type
  constexpr int OWN = QQ_HALF_BLOCK / G;  // owners per block
  constexpr int NW = QqWords<R>::n;       // words per partner value
  extern __shared__ Site lds_site[];  // [W][4] sites, then [W][QQ_NV][NW] partner sums
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ Rcs = Rp.b.Rcs;
  const int *__restrict__ nbr = Rp.b.nbr, *__restrict__ nbrc = Rp.b.nbrc,
            *__restrict__ nbrn = Rp.b.nbrn, *__restrict__ nbrh = Rp.b.nbrh;
  double *phi = Rp.b.phi, *Efq = Rp.b.Efq, *dR = Rp.b.dR, *tally = Rp.b.tally;
  const int N = c.N, b = bk.y, i0 = bk.x * OWN, W = qq_half_rows(N, G);
  unsigned long long *jacc = (unsigned long long *)(lds_site + 4 * W);
  for (int e = threadIdx.x; e < 4 * W; e += blockDim.x) {
    int m = i0 + e / 4;
    m -= m >= N ? N : 0;
    lds_site[e] = site[((size_t)b * N + m) * 4 + e % 4];
  }
  for (int e = threadIdx.x; e < W * QQ_NV * NW; e += blockDim.x) jacc[e] = 0ull;
  __syncthreads();
  const int a0 = i0 + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const bool valid = a0 < N;
  const int a = valid ? a0 : i0;
  const long grp = (long)b * N + a;
  // LDS row of a molecule: owners at a - i0, partners up to N/2 beyond
  auto row = [&](int j) { return j - i0 + (j < i0 ? N : 0); };
  Site A[4];
  for (int t = 0; t < 4; ++t) A[t] = lds_site[(a - i0) * 4 + t];
  const double RcO[3] = {Rcs[a * 12 + 0], Rcs[a * 12 + 1], Rcs[a * 12 + 2]};
  double ph[4] = {0, 0, 0, 0}, ef[4][3] = {{0}}, dro[3] = {0, 0, 0};
  double uvdw = 0.0, vc = 0.0, vt[9] = {0};
  const int *list = nbr + (size_t)grp * c.nbr_cap;
  // partner-side sums of molecule j: ph[1..3], then ef of sites SA0..3, then dro
  auto jslot = [&](int rj, int k) { return jacc + ((size_t)rj * QQ_NV + k) * NW; };
  auto visit = [&](const int j, auto thole_tag) {
    constexpr bool THOLE = decltype(thole_tag)::value;
    const bool lo = a < j;
    const int rj = row(j);
    const Site B0 = lds_site[rj * 4];
    double r[3] = {A[0].x - B0.x, A[0].y - B0.y, A[0].z - B0.z};
    minimg(c, r[0], r[1], r[2]);
    const double r2 = norm2_exact(r[0], r[1], r[2]);
    const R r2r = (R)r2;
    const R dRij = RM<R>::sqrt_(r2r);
    const R R2i = RM<R>::rcp(r2r), R4i = R2i * R2i, R6i = R2i * R2i * R2i, R8i = R6i * R2i;
    const R ex = (R)c.vdwD * RM<R>::exp_(-(R)c.vdwE * dRij);
    const R uij = R6i * (((R)c.vdwA * R6i + (R)c.vdwB * R4i) + (R)c.vdwC) + ex;
    const R xs = dRij - (R)c.rc1;
    const R fsc = -(R8i * (R)6.0 * (R)c.vdwC + RM<R>::div(ex * (R)c.vdwE, dRij) +
                    (R)c.shiftA * (xs * xs) + (R)c.shiftB * (xs * xs * xs));
    const double d[3] = {(double)(fsc * (R)r[0]), (double)(fsc * (R)r[1]),
                         (double)(fsc * (R)r[2])};
    for (int k = 0; k < 3; ++k) {
      dro[k] += d[k];
      qq_fix_add<NW>(jslot(rj, 3 + 3 * (4 - SA0) + k), -d[k]);
    }
    uvdw += (double)uij;
    {
      double cx = RcO[0] - Rcs[j * 12 + 0], cy = RcO[1] - Rcs[j * 12 + 1],
             cz = RcO[2] - Rcs[j * 12 + 2];
      minimg(c, cx, cy, cz);
      vc += cx * d[0] + cy * d[1] + cz * d[2];
    }
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row < 3; ++row) vt[col * 3 + row] += d[col] * r[row];
    const double s0[3] = {r[0] - A[0].x + B0.x, r[1] - A[0].y + B0.y, r[2] - A[0].z + B0.z};
#pragma unroll QqUnroll<R>::v
    for (int sb = SA0; sb < 4; ++sb) {
      const Site Bs = lds_site[rj * 4 + sb];
      const R qb = (R)Bs.q;
      const double *rcb = Rcs + (j * 4 + sb) * 3;
      double jp = 0.0, je[3] = {0.0, 0.0, 0.0};  // the partner's site sb
#pragma unroll
      for (int sa = SA0; sa < 4; ++sa) {
        if (sa == 0 && sb == 0) continue;  // the oxygens carry no charge
        double v[3] = {A[sa].x - Bs.x, A[sa].y - Bs.y, A[sa].z - Bs.z};
        minimg(c, v[0], v[1], v[2]);
        const double d2 = norm2_exact(v[0], v[1], v[2]);
        const R vr[3] = {(R)v[0], (R)v[1], (R)v[2]};
        R ts0, ts1;
        smear01<THOLE>(c, (R)d2, (R)(lo ? c.apolCC[sa][sb] : c.apolCC[sb][sa]),
                       (R)(lo ? c.raCC[sa][sb] : c.raCC[sb][sa]), ts0, ts1);
        const R qa = (R)A[sa].q;
        // this molecule's site sa from the partner's charge, and the
        // partner's site sb from this one's (the field along -v)
        if (sa > 0) ph[sa] += (double)(ts0 * qb);
        jp += (double)(ts0 * qa);
        for (int k = 0; k < 3; ++k) {
          ef[sa][k] += (double)(ts1 * qb * vr[k]);
          je[k] -= (double)(ts1 * qa * vr[k]);
        }
        if (sa > 0 && sb > 0) {
          double dd[3];
          for (int k = 0; k < 3; ++k) dd[k] = (double)(-(ts1 * qa * qb * vr[k]));
          // the mean of the two sides' lever arms (s0 - Rc(j, sb) from this
          // side, -s0 - Rc(a, sa) with the force reversed from the partner's)
          const double *rca = Rcs + (a * 4 + sa) * 3;
          double w[3];
          for (int k = 0; k < 3; ++k) w[k] = s0[k] + 0.5 * (rca[k] - rcb[k]);
          vc += w[0] * dd[0] + w[1] * dd[1] + w[2] * dd[2];
          for (int col = 0; col < 3; ++col)
            for (int row = 0; row < 3; ++row) vt[col * 3 + row] += dd[col] * v[row];
        }
      }
      if (sb > 0) qq_fix_add<NW>(jslot(rj, sb - 1), jp);
      for (int k = 0; k < 3; ++k) qq_fix_add<NW>(jslot(rj, 3 + 3 * (sb - SA0) + k), je[k]);
    }
  };
  if (valid) {
    // entries below a and up to the end of its window: near, then far
    const int4 h = make_int4(nbrh[grp * 4], nbrh[grp * 4 + 1], nbrh[grp * 4 + 2], nbrh[grp * 4 + 3]);
    const int nn = nbrn[grp], nf = nbrc[grp], cap = c.nbr_cap;
    const bool wrap = pair_half_window(a, N) >= N;
    // owned near entries: [h.x, h.y), or [h.x, nn) and [0, h.y) when the
    // window wraps; far entries are stored from the back (the k-th in
    // ascending partner order at cap - 1 - k): [cap - h.w, cap - h.z), or
    // [cap - nf, cap - h.z) and [cap - h.w, cap)
    walk_ranges<G>(list, lane, h.x, wrap ? nn : h.y, 0, wrap ? h.y : 0,
                   [&](int j) { visit(j, std::true_type{}); });
    walk_ranges<G>(list, lane, wrap ? cap - nf : cap - h.w, cap - h.z, cap - h.w, wrap ? cap : cap - h.w,
                   [&](int j) { visit(j, std::false_type{}); });
  }
  for (int t = 1; t < 4; ++t) ph[t] = group_sum<G>(ph[t]);
  for (int t = SA0; t < 4; ++t)
    for (int k = 0; k < 3; ++k) ef[t][k] = group_sum<G>(ef[t][k]);
  for (int k = 0; k < 3; ++k) dro[k] = group_sum<G>(dro[k]);
  uvdw = group_sum<G>(uvdw);
  vc = group_sum<G>(vc);
  for (int i = 0; i < 9; ++i) vt[i] = group_sum<G>(vt[i]);
  if (valid && lane == 0) {
    const size_t base = ((size_t)b * N + a) * 4;
    for (int t = 1; t < 4; ++t) phi[base + t] += ph[t];
    for (int t = SA0; t < 4; ++t)
      for (int k = 0; k < 3; ++k) Efq[(base + t) * 3 + k] += ef[t][k];
    for (int k = 0; k < 3; ++k) dR[base * 3 + k] += dro[k];
    // half of Rc(a, site) . (the site's pair forces) from this side's field;
    // the partner side's half follows in k_pair_qq_partner
    for (int t = 1; t < 4; ++t) {
      const double *rc = Rcs + (a * 4 + t) * 3;
      vc -= 0.5 * A[t].q * (rc[0] * ef[t][0] + rc[1] * ef[t][1] + rc[2] * ef[t][2]);
    }
    double *tl = tally + ((size_t)b * N + a) * T_COUNT;
    tl[T_UVDW] += uvdw;
    tl[T_VIRC] += vc;
    vadd9(tl + T_VIRT, vt);
  }
  __syncthreads();
  // the block's exact row sums, rounded once
  double *slab = Rp.b.qqslab + ((size_t)b * bk.nx + bk.x) * (size_t)W * QQ_NV;
  for (int e = threadIdx.x; e < W * QQ_NV; e += blockDim.x)
    slab[e] = qq_fix_value<NW>(jacc + (size_t)e * NW);
}

// The partner-side sums of k_pair_qq_half: 32 lanes per molecule, lane k
// adding value k over the blocks whose window holds the molecule, in block
// order, onto its potential, field or oxygen gradient; the lanes then sum the
// other half of Rc . F of the centroid virial.
template <int MODEL>
__global__ __launch_bounds__(256) void k_pair_qq_partner(const Rep *__restrict__ reps, BlockMap bm) {
  constexpr int SA0 = MODEL == 2 ? 0 : 1;
  constexpr int NVAL = 3 + 3 * (4 - SA0) + 3;  // values in use
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const int N = c.N;
  const long m = ((long)bk.x * blockDim.x + threadIdx.x) / 32;
  const int k = threadIdx.x % 32;
  const bool valid = m < (long)c.P * N;  // alike over the 32 lanes of a molecule
  const int b = valid ? (int)(m / N) : 0, a = valid ? (int)(m % N) : 0;
  const int G = half_lanes(c.g_qq), own = QQ_HALF_BLOCK / G;
  const int nblk = qq_half_blocks(N, G), W = qq_half_rows(N, G);
  double vc = 0.0;
  if (valid && k < NVAL) {
    const double *slab = Rp.b.qqslab + (size_t)b * nblk * W * QQ_NV + k;
    double v = 0.0;
    for (int q = 0; q < nblk; ++q) {
      int r = a - q * own;
      r += r < 0 ? N : 0;
      if (r < W) v += slab[((size_t)q * W + r) * QQ_NV];
    }
    const size_t base = ((size_t)b * N + a) * 4;
    if (k < 3) {
      Rp.b.phi[base + 1 + k] += v;
    } else if (k < 3 + 3 * (4 - SA0)) {
      const int t = SA0 + (k - 3) / 3, x = (k - 3) % 3;
      Rp.b.Efq[(base + t) * 3 + x] += v;
      if (t > 0) vc = -0.5 * Rp.b.site[base + t].q * Rp.b.Rcs[(a * 4 + t) * 3 + x] * v;
    } else {
      Rp.b.dR[base * 3 + (k - 3 - 3 * (4 - SA0))] += v;
    }
  }
  vc = group_sum<32>(vc);
  if (valid && k == 0) Rp.b.tally[((size_t)b * N + a) * T_COUNT + T_VIRC] += vc;
}

// Bytes of a bead's molecule records staged by the LDS pair kernels.
static size_t bead_sites_lds(const SysConst &c) { return sizeof(Site) * 4 * c.N; }
constexpr size_t PAIR_LDS_MAX = 150 * 1024;





// ---------------------------------------------------------------------------
// Ewald reciprocal space (ewald.f90:1-226, 403-437).
// ---------------------------------------------------------------------------

// Per-site phase tables, the reference's ssincos recurrence (ewald.f90:403-437).
__global__ __launch_bounds__(256) void k_tables(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  double2 *tab = Rp.b.tab;
  const int S = 4 * c.N;
  const int s = bk.x * blockDim.x + threadIdx.x;
  const int b = bk.y;
  if (s >= S) return;
  const Site p = site[(size_t)b * S + s];
  const double xyz[3] = {p.x, p.y, p.z};
  const int hm[3] = {c.hmax, c.kmax, c.lmax};
  int row = 0;
  double2 *tb = tab + (size_t)b * c.T * S + s;
  // the dipole sites' rows again, contiguous in dipole-site order (mol * nd +
  // j), for the passes over dipole sites only; FP32 too in the mixed form
  const int type = s & 3, MD = c.nd * c.N;
  const bool dsite = type >= c.tdip && type < c.tdip + c.nd;
  const size_t dm = (size_t)b * c.T * MD + (size_t)(s >> 2) * c.nd + (type - c.tdip);
  double2 *td = Rp.b.tabd + dm;
  float2 *tf = c.mixed ? Rp.b.tabdf + dm : nullptr;
  auto put = [&](int r, double2 v) {
    tb[(size_t)r * S] = v;
    if (dsite) {
      td[(size_t)r * MD] = v;
      if (tf) tf[(size_t)r * MD] = make_float2((float)v.x, (float)v.y);
    }
  };
  for (int ax = 0; ax < 3; ++ax) {
    const double arg = TWOPI * c.boxi[ax] * xyz[ax];
    const double c1 = cos(arg), s1 = sin(arg);
    double cc = 1.0, ss = 0.0;
    put(row++, make_double2(1.0, 0.0));
    for (int h = 1; h <= hm[ax]; ++h) {
      double nc, ns;
      if (h == 1) {
        nc = c1;
        ns = s1;
      } else {
        nc = cc * c1 - ss * s1;
        ns = cc * s1 + ss * c1;
      }
      cc = nc;
      ss = ns;
      put(row++, make_double2(cc, ss));
    }
  }
}


// Site index of the m-th member of a site set.  SET_CHG: H1, H2, M of every
// molecule; SET_DIP: the dipole sites; SET_ALL: all four.
// SET_CHG_T / SET_ALL_T: the same sites as SET_CHG / SET_ALL, type-major
// (all H1, then all H2, ...), so that a wave of the final reciprocal pass holds
// one site type and its dipole-site branch is wave-uniform.
enum SiteSet { SET_CHG = 0, SET_DIP = 1, SET_ALL = 2, SET_CHG_T = 3, SET_ALL_T = 4 };

__device__ __forceinline__ int set_size(const SysConst &c, int set) {
  return set == SET_CHG || set == SET_CHG_T ? 3 * c.N : set == SET_DIP ? c.nd * c.N : 4 * c.N;
}
__device__ __forceinline__ int set_site(const SysConst &c, int set, int m) {
  if (set == SET_CHG) return (m / 3) * 4 + 1 + m % 3;
  if (set == SET_DIP) return (m / c.nd) * 4 + c.tdip + m % c.nd;
  if (set == SET_CHG_T) return (m % c.N) * 4 + 1 + m / c.N;
  if (set == SET_ALL_T) return (m % c.N) * 4 + m / c.N;
  return m;
}

// Phase e^{i k.r} of one site from its table rows, in the reference's
// product order (ewald.f90:30-39).
template <class V, class R>
__device__ __forceinline__ void phase_from(V x, V y, V z, R &ck, R &sk) {
  const R cxy = x.x * y.x - x.y * y.y;
  const R sxy = x.y * y.x + x.x * y.y;
  ck = cxy * z.x - sxy * z.y;
  sk = sxy * z.x + cxy * z.y;
}

// Two- and four-component vectors of an arithmetic type.
template <class R> struct Vec;
template <> struct Vec<double> {
  using v2 = double2;
  using v4 = double4;
  static __device__ __forceinline__ double2 two(double a, double b) { return make_double2(a, b); }
  static __device__ __forceinline__ double4 four(double a, double b, double c, double d) {
    return make_double4(a, b, c, d);
  }
};
template <> struct Vec<float> {
  using v2 = float2;
  using v4 = float4;
  static __device__ __forceinline__ float2 two(double a, double b) {
    return make_float2((float)a, (float)b);
  }
  static __device__ __forceinline__ float4 four(double a, double b, double c, double d) {
    return make_float4((float)a, (float)b, (float)c, (float)d);
  }
};

// Phase-table rows of the m-th member of a site set: the dipole sites from
// their contiguous table (FP32 in the mixed form's SOR iterations), the
// other sets from the all-site table.
template <class R, int SET>
__device__ __forceinline__ const typename Vec<R>::v2 *set_rows(const Rep &Rp, int b, int &stride) {
  const SysConst &c = Rp.c;
  if constexpr (SET == SET_DIP) {
    stride = c.nd * c.N;
    const size_t off = (size_t)b * c.T * stride;
    if constexpr (sizeof(R) == 4) return Rp.b.tabdf + off;
    else return Rp.b.tabd + off;
  } else {
    static_assert(sizeof(R) == 8, "the all-site table is FP64");
    stride = 4 * c.N;
    return Rp.b.tab + (size_t)b * c.T * stride;
  }
}
template <int SET>
__device__ __forceinline__ int row_index(const SysConst &c, int m) {
  return SET == SET_DIP ? m : set_site(c, SET, m);
}

// Site chunk of the column structure-factor kernel and its padded LDS row
// (a multiple of 16 plus one double2, so rows start 4 banks apart).
__host__ __device__ inline void sfac_chunking(const SysConst &c, int mode, int M, int &nch,
                                              int &sc, int &scp) {
  // the dipole structure factor takes c.sf_sites per chunk; the charge and
  // virial ones at most 32, which keeps their blocks spread over the device
  const int sites = mode == SF_D ? c.sf_sites : (c.sf_sites < 32 ? c.sf_sites : 32);
  const int want = (M + sites - 1) / sites;
  nch = want < c.sf_chunks_max ? want : c.sf_chunks_max;
  if (nch < 1) nch = 1;
  sc = (M + nch - 1) / nch;
  if (sc < 1) sc = 1;  // an empty site set (no dipole sites): one empty chunk
  nch = (M + sc - 1) / sc;
  if (nch < 1) nch = 1;
  scp = ((sc + 15) / 16) * 16 + 1;
}

// Sites of a structure factor's site set.
__host__ __device__ inline int sfac_sites(const SysConst &c, int mode) {
  return mode == SF_Q ? 3 * c.N : c.nd * c.N;
}

// Structure factors with one kz-segment (fixed kx, ky, up to KZ consecutive kz)
// per lane: the phase of each site is formed once per segment and advanced
// along kz by e^{i 2 pi z/L}, accumulating every kz of the segment in
// registers.  A block takes a chunk of
// sites of one bead; their phase tables and weights are staged in LDS (rows
// padded so that lanes reading different rows of one site hit different
// banks).  Per-chunk partial sums are combined in a fixed order by
// k_sfac_reduce.  SF_Q and SF_D as above.
// SPLIT = 2: the block holds two lane sets for the same segments, each over
// half of the chunk's sites; the second hands its sums to the first through
// LDS, which adds them and writes (twice the waves for the same partials).
template <int MODE, int KZ, int SPLIT = 1, class R = double>
__global__ __launch_bounds__(512) void k_sfac_col(const Rep *__restrict__ reps, BlockMap bm) {
  using R2 = typename Vec<R>::v2;
  constexpr int SET = MODE == SF_Q ? SET_CHG : SET_DIP;
  constexpr int NW = MODE == SF_Q ? 1 : 3;
  constexpr int NV = MODE == SF_V ? 6 : 2;  // values per k
  // sites per unrolled step: the grid, not the registers, limits the waves
  // per SIMD, so more independent phase chains per wave hide FP64 latency
  constexpr int SITE_UNROLL = MODE == SF_V ? 2 : 4;
  extern __shared__ double2 lds_raw[];
  R2 *lds_sf = reinterpret_cast<R2 *>(lds_raw);  // [T][scp] tables, then [sc][NW] weights
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ dip = Rp.b.dip;
  const int4 *__restrict__ cols = Rp.b.segs;
  const int *__restrict__ colk = Rp.b.segk;
  const int *__restrict__ active = Rp.b.active;
  double *__restrict__ part = Rp.b.part;
  int nch, sc, scp;
  sfac_chunking(c, MODE, sfac_sites(c, MODE), nch, sc, scp);
  R *lw = reinterpret_cast<R *>(lds_sf + (size_t)c.T * scp);
  const int b = bk.z;
  if (MODE == SF_D && !active[b]) return;
  const int S = 4 * c.N, M = set_size(c, SET);
  const int chunk = bk.y;
  const int m0 = chunk * sc, nt = min(sc, M - m0);
  int RS;
  const R2 *__restrict__ tb = set_rows<R, SET>(Rp, b, RS);
  for (int e = threadIdx.x; e < c.T * sc; e += blockDim.x) {
    const int row = e / sc, l = e % sc;
    if (l < nt) lds_sf[row * scp + l] = tb[(size_t)row * RS + row_index<SET>(c, m0 + l)];
  }
  for (int l = threadIdx.x; l < nt; l += blockDim.x) {
    const int sidx = set_site(c, SET, m0 + l);
    if (MODE == SF_Q) {
      lw[l] = site[(size_t)b * S + sidx].q;
    } else {
      const double *mu = dip + ((size_t)b * c.N * c.nd + m0 + l) * 3;
      lw[l * 3 + 0] = (R)mu[0];
      lw[l * 3 + 1] = (R)mu[1];
      lw[l * 3 + 2] = (R)mu[2];
    }
  }
  __syncthreads();
  const int nl = blockDim.x / SPLIT, half_id = threadIdx.x / nl;
  const int col = bk.x * nl + threadIdx.x % nl;
  const bool live = col < c.nseg;
  const int hn = (nt + SPLIT - 1) / SPLIT;
  const int e0 = half_id * hn, e1 = min(nt, e0 + hn);
  const int4 cl = cols[live ? col : 0];
  const int n = cl.w;
  const int k0 = colk[live ? col : 0];
  const int rowx = cl.x, rowy = c.hmax + 1 + abs(cl.y), rowz = c.hmax + c.kmax + 2 + abs(cl.z);
  const int rowz1 = c.hmax + c.kmax + 3;
  const R sgy = cl.y < 0 ? -1.0 : 1.0, sgz = cl.z < 0 ? -1.0 : 1.0;
  const R kvx = TWOPI * c.boxi[0] * (double)cl.x, kvy = TWOPI * c.boxi[1] * (double)cl.y;
  const double tpz = TWOPI * c.boxi[2];
  // kz of each k of the segment
  R kzs[KZ];
#pragma unroll
  for (int i = 0; i < KZ; ++i) kzs[i] = (R)(tpz * (double)(cl.z + i));
  // SF_Q, SF_D: (re, im) per k; SF_V: (sum mu cos, sum mu sin) per component
  R acc[NV][KZ];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int i = 0; i < KZ; ++i) acc[v][i] = 0.0;
  #pragma unroll SITE_UNROLL
  for (int e = live ? e0 : e1; e < e1; ++e) {
    const R2 x = lds_sf[rowx * scp + e];
    R2 y = lds_sf[rowy * scp + e];
    R2 z = lds_sf[rowz * scp + e];
    const R2 z1 = c.lmax >= 1 ? lds_sf[rowz1 * scp + e] : Vec<R>::two(1.0, 0.0);
    y.y *= sgy;
    z.y *= sgz;
    R ck, sk;
    phase_from(x, y, z, ck, sk);
    R w = 0.0, wxy = 0.0, mz = 0.0;
    if (MODE == SF_Q) {
      w = lw[e];
    } else if (MODE == SF_D) {
      wxy = kvx * lw[e * 3 + 0] + kvy * lw[e * 3 + 1];
      mz = lw[e * 3 + 2];
    }
#pragma unroll
    for (int i = 0; i < KZ; ++i) {
      if (i < n) {
        if (i > 0) {
          const R nc = ck * z1.x - sk * z1.y;
          sk = sk * z1.x + ck * z1.y;
          ck = nc;
        }
        if (MODE == SF_Q) {
          acc[0][i] += w * ck;
          acc[1][i] += w * sk;
        } else if (MODE == SF_D) {
          const R dk = wxy + kzs[i] * mz;
          acc[0][i] -= dk * sk;
          acc[1][i] += dk * ck;
        } else {
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            acc[cc][i] += lw[e * 3 + cc] * ck;
            acc[3 + cc][i] += lw[e * 3 + cc] * sk;
          }
        }
      }
    }
  }
  if (SPLIT == 2) {
    // after the weights, 16-byte aligned: [NV][KZ][nl] sums of the second set
    constexpr int AL = 16 / sizeof(R);
    R *red = lw + ((sc * NW + AL - 1) & ~(AL - 1));
    if (half_id == 1 && live)
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int i = 0; i < KZ; ++i) red[(v * KZ + i) * nl + threadIdx.x % nl] = acc[v][i];
    __syncthreads();
    if (half_id == 0)
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int i = 0; i < KZ; ++i) acc[v][i] += red[(v * KZ + i) * nl + threadIdx.x];
  }
  if (!live || half_id != 0) return;
  double *o = part + (((size_t)b * nch + chunk) * c.nk + k0) * NV;
#pragma unroll
  for (int i = 0; i < KZ; ++i)
    if (i < n)
#pragma unroll
      for (int v = 0; v < NV; ++v) o[NV * i + v] = (double)acc[v][i];
}

// Structure factors with one k-vector per lane.  A block holds 256 k of one
// bead and a chunk of its sites; the sites' phase tables and weights are
// staged in LDS (rows padded so that lanes reading different rows of one site
// hit different banks; lanes on the same row read the same address).  Each
// lane forms e^{i k.r} = X(kx) Y(ky) Z(kz) of every site from the tables and
// accumulates its weighted sum; per-chunk partial sums are combined in a fixed
// order by k_sfac_reduce.
//   SF_Q: (sum q cos, sum q sin)                      (ewald.f90:41-44)
//   SF_D: (-sum (k.mu) sin, sum (k.mu) cos)           (ewald.f90:99-110)
//   SF_V: (sum mu cos, sum mu sin) per component      (ewald.f90:208-213)
template <int MODE>
__global__ __launch_bounds__(256) void k_sfac_k(const Rep *__restrict__ reps, BlockMap bm) {
  constexpr int SET = MODE == SF_Q ? SET_CHG : SET_DIP;
  constexpr int NW = MODE == SF_Q ? 1 : 3;
  constexpr int NV = MODE == SF_V ? 6 : 2;
  extern __shared__ double2 lds_sf[];  // [T][scp] tables, then [sc][NW] weights
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ dip = Rp.b.dip;
  const double4 *__restrict__ kv = Rp.b.kv;
  const int4 *__restrict__ khkl = Rp.b.khkl;
  const int *__restrict__ active = Rp.b.active;
  double *__restrict__ part = Rp.b.part;
  int nch, sc, scp;
  sfac_chunking(c, MODE, sfac_sites(c, MODE), nch, sc, scp);
  double *lw = reinterpret_cast<double *>(lds_sf + (size_t)c.T * scp);
  const int b = bk.z;
  if (MODE == SF_D && !active[b]) return;
  const int S = 4 * c.N, M = set_size(c, SET);
  const int chunk = bk.y;
  const int m0 = chunk * sc, nt = min(sc, M - m0);
  int RS;
  const double2 *__restrict__ tb = set_rows<double, SET>(Rp, b, RS);
  for (int e = threadIdx.x; e < c.T * sc; e += blockDim.x) {
    const int row = e / sc, l = e % sc;
    if (l < nt) lds_sf[row * scp + l] = tb[(size_t)row * RS + row_index<SET>(c, m0 + l)];
  }
  for (int l = threadIdx.x; l < nt; l += blockDim.x) {
    if (MODE == SF_Q) {
      lw[l] = site[(size_t)b * S + set_site(c, SET, m0 + l)].q;
    } else {
      const double *mu = dip + ((size_t)b * c.N * c.nd + m0 + l) * 3;
      lw[l * 3 + 0] = mu[0];
      lw[l * 3 + 1] = mu[1];
      lw[l * 3 + 2] = mu[2];
    }
  }
  __syncthreads();
  const int k = bk.x * blockDim.x + threadIdx.x;
  if (k >= c.nk) return;
  const int4 h = khkl[k];
  const double4 kvec = kv[k];
  const int rowx = h.x * scp, rowy = (c.hmax + 1 + abs(h.y)) * scp;
  const int rowz = (c.hmax + c.kmax + 2 + abs(h.z)) * scp;
  const double sy = h.y < 0 ? -1.0 : 1.0, sz = h.z < 0 ? -1.0 : 1.0;
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
#pragma unroll 2
  for (int e = 0; e < nt; ++e) {
    const double2 x = lds_sf[rowx + e];
    double2 y = lds_sf[rowy + e];
    double2 z = lds_sf[rowz + e];
    y.y *= sy;
    z.y *= sz;
    double ck, sk;
    phase_from(x, y, z, ck, sk);
    if (MODE == SF_Q) {
      const double q = lw[e];
      acc[0] += q * ck;
      acc[1] += q * sk;
    } else if (MODE == SF_D) {
      const double dk = (kvec.x * lw[e * 3] + kvec.y * lw[e * 3 + 1]) + kvec.z * lw[e * 3 + 2];
      acc[0] -= dk * sk;
      acc[1] += dk * ck;
    } else {
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        acc[cc] += lw[e * 3 + cc] * ck;
        acc[3 + cc] += lw[e * 3 + cc] * sk;
      }
    }
  }
  double *o = part + (((size_t)b * nch + chunk) * c.nk + k) * NV;
#pragma unroll
  for (int v = 0; v < NV; ++v) o[v] = acc[v];
}

// The per-chunk partial sums of structure factor `mode` added in chunk order.
__global__ __launch_bounds__(256) void k_sfac_reduce(const Rep *__restrict__ reps, BlockMap bm,
                                                    int mode) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const double *__restrict__ part = Rp.b.part;
  double *__restrict__ out = mode == SF_Q ? Rp.b.Sq : mode == SF_D ? Rp.b.Sm : Rp.b.Sv;
  const int nv = mode == SF_V ? 6 : 2;
  int nch, sc, scp;
  sfac_chunking(c, mode, sfac_sites(c, mode), nch, sc, scp);
  const int b = bk.y;
  if (mode == SF_D && !Rp.b.active[b]) return;
  const int e = bk.x * blockDim.x + threadIdx.x;  // (k, component)
  if (e >= nv * c.nk) return;
  double v = 0.0;
  const double *pp = part + (size_t)b * nch * c.nk * nv + e;
#pragma unroll 8
  for (int q = 0; q < nch; ++q) v += pp[(size_t)q * c.nk * nv];
  out[(size_t)b * c.nk * nv + e] = v;
}

static size_t sfac_lds(const SysConst &c, int mode, int sc, int scp, bool f32 = false) {
  const size_t r = f32 ? sizeof(float) : sizeof(double);
  return 2 * r * c.T * scp + r * sc * (mode == SF_Q ? 1 : 3);
}

// The most site chunks any structure factor of a force call uses (the
// partial sums hold P x this many chunks of nk k-vectors).
int sfac_chunks_used(const SysConst &c) {
  int most = 1;
  for (int mode : {SF_Q, SF_D, SF_V}) {
    int nch, sc, scp;
    sfac_chunking(c, mode, sfac_sites(c, mode), nch, sc, scp);
    most = std::max(most, nch);
  }
  return most;
}






// Reciprocal-space fields at each site.  A block holds 64 sites of one bead,
// one per lane, and CG waves; wave g walks kz-columns g, g + CG, ... (columns
// are sorted by length, so the waves get similar trip counts).  All lanes of a
// wave are on the same k at the same time, so the k data (k-vector, 4 pi A_k/V,
// structure factor) is wave-uniform and read with scalar loads; each lane's
// phase e^{i k.r} comes from its site's tables, staged in LDS, and is advanced
// along the column by e^{i 2 pi z/L}.  Along a column kx and ky are fixed, so
// the x and y field components are kx and ky times one column sum.  The CG
// partial fields of a site are combined in a fixed order through LDS.
//   FD_Q : Efq and phi from the charge structure factor (ewald.f90:45-49)
//   FD_D : Efd from the dipole structure factor (ewald.f90:114-117), then the
//          SOR update of the induced dipole (pot_ttm.f90:371-385)
//   FD_QD: phi and dR from the stale dipole and the total structure factors
//          (ewald.f90:184-193)
constexpr int FD_SB = 64;  // sites per block, one per lane
// iter: the SOR iteration (FD_D), or -1 for the mixed form's FP64
// refinement, the step after the replica's last (maxiter + 1).
template <int MODE, int MODEL, int CG, bool KL, class R = double>
__global__ __launch_bounds__(64 * CG) void k_field(const Rep *__restrict__ reps, BlockMap bm,
                                                   int iter) {
  using R2 = typename Vec<R>::v2;
  using R4 = typename Vec<R>::v4;
  constexpr int SET = MODE == FD_D    ? SET_DIP
                      : MODE == FD_QD ? (MODEL == 2 ? SET_ALL_T : SET_CHG_T)
                                      : (MODEL == 2 ? SET_ALL : SET_CHG);
  // [T][64] phase tables of the block's sites (later reused for the
  // reduction), then with KL the bead's k data: [nk] (w Re S, w Im S, kz, w)
  // with w = 4 pi A_k / V, and for FD_QD [nk] (w Re(Sq+Sm), w Im(Sq+Sm))
  extern __shared__ double2 lds_raw[];
  R2 *lds_tab = reinterpret_cast<R2 *>(lds_raw);
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double4 *__restrict__ kv = Rp.b.kv;
  const int4 *__restrict__ cols = Rp.b.cols;
  const int *__restrict__ colk = Rp.b.colk;
  const double *__restrict__ Sq = Rp.b.Sq, *__restrict__ Sm = Rp.b.Sm;
  // FD_Q and FD_QD write their reciprocal-space parts to kE / kphi, added by
  // their consumers (the real-space pair pass may be forming the rest meanwhile)
  double *__restrict__ phi = MODE == FD_D ? Rp.b.phi : Rp.b.kphi;
  double *__restrict__ Efq = MODE == FD_Q ? Rp.b.kE : Rp.b.Efq;
  double *__restrict__ dR = MODE == FD_QD ? Rp.b.kE : Rp.b.dR;
  double *__restrict__ dip = Rp.b.dip, *__restrict__ dip_prev = Rp.b.dip_prev;
  const double *__restrict__ Efd = Rp.b.Efd;
  double *__restrict__ diff2 = Rp.b.diff2;
  const int *__restrict__ active = Rp.b.active;
  if (iter < 0) iter = c.maxiter + 1;
  const int b = bk.y;
  if (MODE == FD_D && !active[b]) return;
  const int lane = threadIdx.x & 63;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int M = set_size(c, SET);
  const int S = 4 * c.N;
  const int mbase = bk.x * FD_SB;
  int RS;
  const R2 *__restrict__ tb = set_rows<R, SET>(Rp, b, RS);
  for (int e = threadIdx.x; e < c.T * FD_SB; e += 64 * CG) {
    const int row = e >> 6, l = e & 63;
    lds_tab[e] = tb[(size_t)row * RS + row_index<SET>(c, min(mbase + l, M - 1))];
  }
  R4 *kd = reinterpret_cast<R4 *>(lds_raw + c.fd_tab_d2);
  double2 *kq = reinterpret_cast<double2 *>(kd + c.nk);
  if (KL) {
    const double *sx = (MODE == FD_Q ? Sq : Sm) + (size_t)b * c.nk * 2;
    const double *sq2 = Sq + (size_t)b * c.nk * 2;
    for (int k = threadIdx.x; k < c.nk; k += 64 * CG) {
      const double4 kk = kv[k];
      kd[k] = Vec<R>::four(kk.w * sx[2 * k], kk.w * sx[2 * k + 1], kk.z, kk.w);
      if (MODE == FD_QD) kq[k] = make_double2(kk.w * (sq2[2 * k] + sx[2 * k]),
                                              kk.w * (sq2[2 * k + 1] + sx[2 * k + 1]));
    }
  }
  __syncthreads();
  const int m = mbase + lane;
  const bool valid = m < M;
  const int s = set_site(c, SET, valid ? m : M - 1);
  const int type = s & 3;
  const bool has_q = type != 0;
  const bool has_d = (type >= c.tdip && type < c.tdip + c.nd);
  const size_t dslot = ((size_t)b * c.N + (s >> 2)) * c.nd + (type - c.tdip);
  double mu[3] = {0, 0, 0};
  if ((MODE == FD_D || MODE == FD_QD) && has_d)
    for (int k = 0; k < 3; ++k) mu[k] = dip[dslot * 3 + k];
  const double q = (MODE == FD_QD) ? site[(size_t)b * S + s].q : 0.0;
  // e^{i 2 pi z/L}; with lmax = 0 every column holds one k and no step is taken
  const R2 zstep =
      c.lmax >= 1 ? lds_tab[(c.hmax + c.kmax + 3) * FD_SB + lane] : Vec<R>::two(1.0, 0.0);
  const double *sq = Sq + (size_t)b * c.nk * 2, *sm = Sm + (size_t)b * c.nk * 2;
  double e0 = 0, e1 = 0, e2 = 0, ph = 0;
  for (int col = g; col < c.ncol; col += CG) {
    const int4 cl = cols[col];
    const int k0 = colk[col];
    R2 x = lds_tab[cl.x * FD_SB + lane];
    R2 y = lds_tab[(c.hmax + 1 + abs(cl.y)) * FD_SB + lane];
    R2 z = lds_tab[(c.hmax + c.kmax + 2 + abs(cl.z)) * FD_SB + lane];
    if (cl.y < 0) y.y = -y.y;
    if (cl.z < 0) z.y = -z.y;
    R ck, sk;
    phase_from(x, y, z, ck, sk);
    const double kx = kv[k0].x, ky = kv[k0].y;
    // column sums of t and kz t (t carries the sign of each mode's field)
    R at = 0.0, az = 0.0;
    // FD_QD dipole sites: k.mu = (kx mux + ky muy) + kz muz
    const double dxy = (MODE == FD_QD) ? kx * mu[0] + ky * mu[1] : 0.0;
#pragma unroll 4
    for (int i = 0; i < cl.w; ++i) {
      const int k = k0 + i;
      if (KL) {
        const R4 q4 = kd[k];  // w Re S, w Im S, kz, w
        if (MODE == FD_Q) {
          const R t = ck * q4.y - sk * q4.x;
          at += t;
          az += q4.z * t;
          ph += ck * q4.x + sk * q4.y;
        } else if (MODE == FD_D) {
          const R t = ck * q4.y - sk * q4.x;
          at += t;
          az += q4.z * t;
        } else {
          R t = 0.0;
          if (has_q) {
            t = -(q * (ck * q4.y - sk * q4.x));
            ph += ck * q4.x + sk * q4.y;
          }
          if (has_d) {
            const double2 qq = kq[k];
            const double dk = dxy + q4.z * mu[2];
            t += dk * (ck * qq.x + sk * qq.y);
          }
          at += t;
          az += q4.z * t;
        }
        const R nc = ck * zstep.x - sk * zstep.y;
        sk = sk * zstep.x + ck * zstep.y;
        ck = nc;
        continue;
      }
      const double4 kk = kv[k];  // kx, ky, kz, 4 pi A_k / V
      const R kw = kk.w, kz = kk.z;
      if (MODE == FD_Q) {
        const double re = sq[2 * k], im = sq[2 * k + 1];
        const double t = kk.w * (ck * im - sk * re);
        at += t;
        az += kk.z * t;
        ph += kk.w * (ck * re + sk * im);
      } else if (MODE == FD_D) {
        const R re = sm[2 * k], im = sm[2 * k + 1];
        const R t = kw * (ck * im - sk * re);
        at += t;
        az += kz * t;
      } else {
        const double rqm = sm[2 * k], iqm = sm[2 * k + 1];
        double t = 0.0;
        if (has_q) {
          t = -(q * kk.w * (ck * iqm - sk * rqm));
          ph += kk.w * (ck * rqm + sk * iqm);
        }
        if (has_d) {
          const double rqqm = sq[2 * k] + rqm;
          const double iqqm = sq[2 * k + 1] + iqm;
          const double dk = dxy + kk.z * mu[2];
          t += dk * kk.w * (ck * rqqm + sk * iqqm);
        }
        at += t;
        az += kk.z * t;
      }
      // advance the phase to the next kz of the column
      const R nc = ck * zstep.x - sk * zstep.y;
      sk = sk * zstep.x + ck * zstep.y;
      ck = nc;
    }
    e0 -= kx * (double)at;
    e1 -= ky * (double)at;
    e2 -= (double)az;
  }
  // combine the CG column groups of each site in a fixed order
  __syncthreads();
  double *red = reinterpret_cast<double *>(lds_tab);  // [CG][4][64]
  red[(g * 4 + 0) * FD_SB + lane] = e0;
  red[(g * 4 + 1) * FD_SB + lane] = e1;
  red[(g * 4 + 2) * FD_SB + lane] = e2;
  red[(g * 4 + 3) * FD_SB + lane] = ph;
  __syncthreads();
  if (g != 0 || !valid) return;
  for (int w = 1; w < CG; ++w) {
    e0 += red[(w * 4 + 0) * FD_SB + lane];
    e1 += red[(w * 4 + 1) * FD_SB + lane];
    e2 += red[(w * 4 + 2) * FD_SB + lane];
    ph += red[(w * 4 + 3) * FD_SB + lane];
  }
  const size_t ss = (size_t)b * S + s;
  if (MODE == FD_Q || MODE == FD_QD) {
    // into their own buffers (kE, kphi): the consumer adds them to the
    // real-space sums, which the pair kernels may be forming meanwhile
    double *ko = MODE == FD_Q ? Efq : dR;
    ko[ss * 3 + 0] = e0;
    ko[ss * 3 + 1] = e1;
    ko[ss * 3 + 2] = e2;
    phi[ss] = has_q ? ph : 0.0;
  } else {
    const double e[3] = {e0, e1, e2};
    const double om = c.sor, pol = c.polar[type];
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double etot = (Efd[dslot * 3 + k] + e[k]) + Efq[ss * 3 + k] + c.fac_dd * mu[k];
      const double nm = (1.0 - om) * mu[k] + om * pol * etot;
      const double df = nm - (iter == 1 ? 0.0 : mu[k]);
      d2 += df * df;
      dip[dslot * 3 + k] = nm;
      if (c.mixed && iter <= c.maxiter) dip_prev[dslot * 3 + k] = mu[k];
    }
    diff2[(size_t)b * c.N * c.nd + m] = d2;
  }
}

// Shared memory of k_field: the phase tables (reused for the CG-way
// reduction; fd_tab_d2 double2 slots cover both), then the staged k data.
static size_t field_lds(const SysConst &c, int mode, bool kl, bool f32 = false) {
  return sizeof(double2) * c.fd_tab_d2 +
         (kl ? c.nk * ((f32 ? sizeof(float4) : sizeof(double4)) +
                       (mode == FD_QD ? sizeof(double2) : 0))
             : 0);
}




// ---------------------------------------------------------------------------
// After the charge pass: gradient of the charge energy, Uelec partials and the
// initial induced dipoles (pot_ttm.f90:273-297).  One thread per (bead, molecule).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_post_qq(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  double *phi = Rp.b.phi, *Efq = Rp.b.Efq, *dR = Rp.b.dR, *dip = Rp.b.dip, *tally = Rp.b.tally;
  const double *__restrict__ kE = Rp.b.kE, *__restrict__ kphi = Rp.b.kphi;
  const double *__restrict__ hist = Rp.b.hist;
  int *active = Rp.b.active, *iters = Rp.b.iters, *nactive = Rp.b.nactive;
  const int *__restrict__ clk = Rp.b.clk;
  const int gid = bk.x * blockDim.x + threadIdx.x;
  if (gid >= c.P * c.N) return;
  const int t = clk[0];  // MD step of this force call
  const int b = gid / c.N, w = gid % c.N, N = c.N;
  const size_t base = ((size_t)b * N + w) * 4;
  // real-space sums plus the reciprocal-space part (ewald.f90:45-49): the
  // field of every site of the charge pass (all four for TTM2.1-F, where O
  // holds a dipole), the potential of the charged sites
  for (int ty = c.model == 2 ? 0 : 1; ty < 4; ++ty) {
    for (int k = 0; k < 3; ++k) Efq[(base + ty) * 3 + k] = Efq[(base + ty) * 3 + k] + kE[(base + ty) * 3 + k];
    if (ty > 0) phi[base + ty] = phi[base + ty] + kphi[base + ty];
  }
  double ue = 0.0;
  for (int ty = 1; ty < 4; ++ty) {
    const double q = site[base + ty].q;
    ue += q * phi[base + ty];
    for (int k = 0; k < 3; ++k) dR[(base + ty) * 3 + k] -= q * Efq[(base + ty) * 3 + k];
  }
  tally[((size_t)b * N + w) * T_COUNT + T_UELEC] = ue;
  const int ndip = N * c.nd;
  double dp[4];
  const bool pred = t > 4 && c.guess;
  if (pred) {  // calc_dpcf (pot_ttm.f90:649-675) for md_step > 4
    const double tab4[4][4] = {{-1, 4, -6, 4}, {4, -1, 4, -6}, {-6, 4, -1, 4}, {4, -6, 4, -1}};
    for (int i = 0; i < 4; ++i) dp[i] = tab4[t % 4][i];
  }
  for (int u = 0; u < c.nd; ++u) {
    const int ty = c.tdip + u;
    const size_t dslot = ((size_t)b * N + w) * c.nd + u;
    for (int k = 0; k < 3; ++k) {
      double v;
      if (pred) {
        const size_t m = (size_t)w * c.nd + u;
        v = dp[0] * hist[(((size_t)b * 4 + 0) * ndip + m) * 3 + k];
        for (int i = 1; i < 4; ++i) v = v + dp[i] * hist[(((size_t)b * 4 + i) * ndip + m) * 3 + k];
      } else {
        v = c.polar[ty] * Efq[(base + ty) * 3 + k];
      }
      dip[dslot * 3 + k] = v;
    }
  }
  if (w == 0) {
    active[b] = 1;
    iters[b] = 0;
  }
  if (gid == 0)
    for (int i = 0; i <= c.maxiter; ++i) nactive[i] = 0;
}


// ---------------------------------------------------------------------------
// Real-space dipole field T.mu of one SOR iteration (pot_ttm.f90:311-365).
// ---------------------------------------------------------------------------
template <int MODEL, int G, typename R>
__global__ __launch_bounds__(256) void k_pair_dd(const Rep *__restrict__ reps, BlockMap bm) {
  constexpr int ND = MODEL == 2 ? 3 : 1;
  constexpr int T0 = MODEL == 2 ? 0 : 3;
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ dip = Rp.b.dip;
  const int *__restrict__ nbr = Rp.b.nbr, *__restrict__ nbrc = Rp.b.nbrc, *__restrict__ nbrn = Rp.b.nbrn;
  const int *__restrict__ active = Rp.b.active;
  double *Efd = Rp.b.Efd;
  const long gtid = (long)bk.x * blockDim.x + threadIdx.x;
  const long grp = gtid / G;
  const int lane = (int)(gtid % G);
  const int N = c.N;
  if (grp >= (long)c.P * N) return;
  const int b = (int)(grp / N), a = (int)(grp % N);
  if (!active[b]) return;
  Site A[4];
  for (int t = 0; t < 4; ++t) A[t] = site[((size_t)b * N + a) * 4 + t];
  double ef[ND][3];
  for (int u = 0; u < ND; ++u) ef[u][0] = ef[u][1] = ef[u][2] = 0.0;
  const int nf = nbrc[grp], nn = nbrn[grp];
  const int *list = nbr + (size_t)grp * c.nbr_cap;
  auto visit = [&](const int j, auto thole_tag) {
    constexpr bool THOLE = decltype(thole_tag)::value;
    const bool lo = a < j;
    const Site *Bs = site + ((size_t)b * N + j) * 4;
    double mub[ND][3];
    for (int u = 0; u < ND; ++u)
      for (int k = 0; k < 3; ++k) mub[u][k] = dip[(((size_t)b * N + j) * ND + u) * 3 + k];
#pragma unroll
    for (int ua = 0; ua < ND; ++ua) {
#pragma unroll
      for (int ub = 0; ub < ND; ++ub) {
        const int sa = T0 + ua, sb = T0 + ub;
        const Site B = Bs[sb];
        double v[3] = {A[sa].x - B.x, A[sa].y - B.y, A[sa].z - B.z};
        minimg(c, v[0], v[1], v[2]);
        const double d2 = norm2_exact(v[0], v[1], v[2]);
        R ts1, ts2;
        smear2<THOLE>(c, (R)d2, (R)(lo ? c.apolDD[sa][sb] : c.apolDD[sb][sa]), ts1, ts2);
        const R vr[3] = {(R)v[0], (R)v[1], (R)v[2]};
        const R mr[3] = {(R)mub[ub][0], (R)mub[ub][1], (R)mub[ub][2]};
        R o[3];
        dd3_apply(vr, ts1, ts2, mr, o);
        for (int k = 0; k < 3; ++k) ef[ua][k] += (double)o[k];
      }
    }
  };
  walk_list<G>(list, lane, nn, [&](int j) { visit(j, std::true_type{}); });
  walk_list<G>(list, c.nbr_cap - nf + lane, c.nbr_cap, [&](int j) { visit(j, std::false_type{}); });
  for (int u = 0; u < ND; ++u)
    for (int k = 0; k < 3; ++k) ef[u][k] = group_sum<G>(ef[u][k]);
  if (lane == 0) {
    if (MODEL == 2) {  // intramolecular dipole pairs of TTM2.1-F (pot_ttm.f90:344-365)
      double mua[ND][3];
      for (int u = 0; u < ND; ++u)
        for (int k = 0; k < 3; ++k) mua[u][k] = dip[(((size_t)b * N + a) * ND + u) * 3 + k];
      for (int u = 0; u < ND - 1; ++u)
        for (int v2 = u + 1; v2 < ND; ++v2) {
          const int sa = T0 + u, sb = T0 + v2;
          double r[3] = {A[sa].x - A[sb].x, A[sa].y - A[sb].y, A[sa].z - A[sb].z};
          minimg(c, r[0], r[1], r[2]);
          const double d2 = norm2_exact(r[0], r[1], r[2]);
          double ts1, ts2, o[3];
          smear2(c, d2, c.apolDD[sa][sb], ts1, ts2);
          dd3_apply(r, ts1, ts2, mua[v2], o);
          for (int k = 0; k < 3; ++k) ef[u][k] += o[k];
          dd3_apply(r, ts1, ts2, mua[u], o);
          for (int k = 0; k < 3; ++k) ef[v2][k] += o[k];
        }
    }
    for (int u = 0; u < ND; ++u)
      for (int k = 0; k < 3; ++k) Efd[(((size_t)b * N + a) * ND + u) * 3 + k] = ef[u][k];
  }
}

// Real-space dipole field with the bead's dipole sites staged in LDS: a block
// of up to 1024 threads takes 1024/G molecules of one bead and first copies
// every dipole site of that bead (position and dipole, 48 bytes) into LDS, so
// the neighbour loop reads LDS instead of gathering scattered records from
// L2.  Arithmetic and summation order are those of k_pair_dd.
constexpr int DD_BLOCK = 1024;
// TTM2.1-F (three dipole sites, nine products per neighbour) in blocks of 256:
// at 1024 lanes a block's registers are capped at 128 per lane and it spills
template <int MODEL>
constexpr int dd_block() {
  return MODEL == 2 ? 256 : DD_BLOCK;
}
// doubles per staged dipole site: 6 used, one of padding so that lanes
// gathering different sites spread over the LDS banks
constexpr int DD_REC = 7;
// neighbour indices requested ahead by the TTM3-F dipole pass
constexpr int DD_AHEAD = 4;
// The SOR convergence test of iteration `iter` for bead b (pot_ttm.f90:306,
// 380-386), as k_converge forms it: the first 256 threads of the block add
// the bead's squared dipole changes in the same order.  All blocks of the
// bead reach the same outcome; block `writer` records it.  Returns whether
// the bead is still active; every thread of the block must call it.
__device__ bool sor_test(const SysConst &c, const double *__restrict__ diff2, int *active,
                         int *iters, int *nactive, int b, int iter, bool writer, double *red,
                         int &act) {
  const int tid = threadIdx.x;
  if (tid == 0) act = active[b];  // one read per block: the writer block may change it
  __syncthreads();
  if (!act) return false;
  const int M = c.N * c.nd;
  if (tid < 256) {
    double s = 0.0;
    for (int m = tid; m < M; m += 256) s += diff2[(size_t)b * M + m];
    red[tid] = s;
  }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) red[tid] += red[tid + off];
    __syncthreads();
  }
  const double delta = c.stath * sqrt(red[0]);
  const bool conv = delta < c.peps, last = iter >= c.maxiter;
  if (writer && tid == 0) {
    if (conv) {
      active[b] = 0;
      iters[b] = iter;
    } else if (last) {
      active[b] = 0;
      iters[b] = -iter;
    } else {
      iters[b] = iter;
      atomicAdd(&nactive[iter], 1);
    }
  }
  return !(conv || last);
}

// (ts1, ts2) of one dipole-site pair in the SOR arithmetic type, and the
// context's cache of them
template <typename R>
struct TsPair;
template <>
struct TsPair<double> {
  using T = double2;
  static __device__ dptr<double2> cache(const DevBufs &b) { return b.ddts; }
};
template <>
struct TsPair<float> {
  using T = float2;
  static __device__ dptr<float2> cache(const DevBufs &b) { return b.ddtsf; }
};

// CACHED: the pair screening factors come from the cache the first SOR
// iteration wrote (store != 0 there); otherwise they are computed.  Either
// way the field is the same, bit for bit.
template <int MODEL, int G, typename R, bool CACHED>
__global__ __launch_bounds__(dd_block<MODEL>()) void k_pair_dd_lds(const Rep *__restrict__ reps,
                                                                   BlockMap bm, int test_iter,
                                                                   int store) {
  constexpr int ND = MODEL == 2 ? 3 : 1;
  constexpr int NP = ND * ND;
  constexpr int T0 = MODEL == 2 ? 0 : 3;
  using R2 = typename TsPair<R>::T;
  extern __shared__ double lds_dd[];  // [N][ND][DD_REC]: x, y, z, mu_x, mu_y, mu_z, pad
  __shared__ double red[256];
  __shared__ int act;
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ dip = Rp.b.dip;
  const int *__restrict__ nbr = Rp.b.nbr, *__restrict__ nbrc = Rp.b.nbrc, *__restrict__ nbrn = Rp.b.nbrn;
  int *active = Rp.b.active, *iters = Rp.b.iters, *nactive = Rp.b.nactive;
  double *__restrict__ Efd = Rp.b.Efd;
  const double *__restrict__ diff2 = Rp.b.diff2;
  const int b = bk.y;
  if (test_iter > 0) {
    // the convergence test of the previous iteration, in place of a
    // k_converge launch
    if (!sor_test(c, diff2, active, iters, nactive, b, test_iter, bk.x == 0, red, act))
      return;
  } else if (!active[b]) {
    return;
  }
  const int N = c.N;
  for (int e = threadIdx.x; e < N * ND; e += blockDim.x) {
    const int j = e / ND, u = e % ND;
    const Site sj = site[((size_t)b * N + j) * 4 + T0 + u];
    const double *mu = dip + (((size_t)b * N + j) * ND + u) * 3;
    double *o = lds_dd + (size_t)e * DD_REC;
    o[0] = sj.x;
    o[1] = sj.y;
    o[2] = sj.z;
    o[3] = mu[0];
    o[4] = mu[1];
    o[5] = mu[2];
  }
  __syncthreads();
  const int a = bk.x * (blockDim.x / G) + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  if (a >= N) return;
  const long grp = (long)b * N + a;
  double A[ND][3];
  for (int u = 0; u < ND; ++u)
    for (int k = 0; k < 3; ++k) A[u][k] = lds_dd[((size_t)a * ND + u) * DD_REC + k];
  double ef[ND][3];
  for (int u = 0; u < ND; ++u) ef[u][0] = ef[u][1] = ef[u][2] = 0.0;
  const int nf = nbrc[grp], nn = nbrn[grp];
  const int *list = nbr + (size_t)grp * c.nbr_cap;
  // this molecule's row of the cache: [cap][NP] (ts1, ts2)
  R2 *tsrow = (R2 *)TsPair<R>::cache(Rp.b) + (size_t)grp * c.nbr_cap * NP;
  // one dipole-site pair: the staged partner site B, the minimum-image
  // vector v to it and the field of its dipole
  auto apply = [&](const double *B, const double v[3], int ua, R ts1, R ts2) {
    const R vr[3] = {(R)v[0], (R)v[1], (R)v[2]};
    const R mr[3] = {(R)B[3], (R)B[4], (R)B[5]};
    R o[3];
    dd3_apply(vr, ts1, ts2, mr, o);
    for (int k = 0; k < 3; ++k) ef[ua][k] += (double)o[k];
  };
  // a TTM3-F visit is one short dipole-tensor product: keep DD_AHEAD
  // neighbour indices in flight (TTM2.1-F's nine products cover one load)
  constexpr int AHEAD = MODEL == 2 ? 1 : DD_AHEAD;
  if constexpr (CACHED) {
    struct E {
      int j;
      R2 t[NP];
    };
    auto load = [&](int at) {
      E e;
      e.j = list[at];
#pragma unroll
      for (int p = 0; p < NP; ++p) e.t[p] = tsrow[(size_t)at * NP + p];
      return e;
    };
    auto visit = [&](const E &e, int) {
#pragma unroll
      for (int ua = 0; ua < ND; ++ua)
#pragma unroll
        for (int ub = 0; ub < ND; ++ub) {
          const double *B = lds_dd + ((size_t)e.j * ND + ub) * DD_REC;
          double v[3] = {A[ua][0] - B[0], A[ua][1] - B[1], A[ua][2] - B[2]};
          minimg(c, v[0], v[1], v[2]);
          R2 t = e.t[ua * ND + ub];
          opaque(t.x);
          opaque(t.y);
          apply(B, v, ua, t.x, t.y);
        }
    };
    walk_deep<G, AHEAD>(lane, nn, load, visit);
    walk_deep<G, AHEAD>(c.nbr_cap - nf + lane, c.nbr_cap, load, visit);
  } else {
    auto visit = [&](const int j, const int at, auto thole_tag) {
      constexpr bool THOLE = decltype(thole_tag)::value;
      const bool lo = a < j;
#pragma unroll
      for (int ua = 0; ua < ND; ++ua) {
#pragma unroll
        for (int ub = 0; ub < ND; ++ub) {
          const int sa = T0 + ua, sb = T0 + ub;
          const double *B = lds_dd + ((size_t)j * ND + ub) * DD_REC;
          double v[3] = {A[ua][0] - B[0], A[ua][1] - B[1], A[ua][2] - B[2]};
          minimg(c, v[0], v[1], v[2]);
          const double d2 = norm2_exact(v[0], v[1], v[2]);
          R ts1, ts2;
          smear2<THOLE>(c, (R)d2, (R)(lo ? c.apolDD[sa][sb] : c.apolDD[sb][sa]), ts1, ts2);
          opaque(ts1);
          opaque(ts2);
          if (store) tsrow[(size_t)at * NP + ua * ND + ub] = R2{ts1, ts2};
          apply(B, v, ua, ts1, ts2);
        }
      }
    };
    auto idx = [&](int at) { return list[at]; };
    walk_deep<G, AHEAD>(lane, nn, idx, [&](int j, int at) { visit(j, at, std::true_type{}); });
    walk_deep<G, AHEAD>(c.nbr_cap - nf + lane, c.nbr_cap, idx,
                        [&](int j, int at) { visit(j, at, std::false_type{}); });
  }
  for (int u = 0; u < ND; ++u)
    for (int k = 0; k < 3; ++k) ef[u][k] = group_sum<G>(ef[u][k]);
  if (lane == 0) {
    if (MODEL == 2) {  // intramolecular dipole pairs of TTM2.1-F (pot_ttm.f90:344-365)
      for (int u = 0; u < ND - 1; ++u)
        for (int v2 = u + 1; v2 < ND; ++v2) {
          const int sa = T0 + u, sb = T0 + v2;
          double r[3] = {A[u][0] - A[v2][0], A[u][1] - A[v2][1], A[u][2] - A[v2][2]};
          minimg(c, r[0], r[1], r[2]);
          const double d2 = norm2_exact(r[0], r[1], r[2]);
          double ts1, ts2, o[3];
          smear2(c, d2, c.apolDD[sa][sb], ts1, ts2);
          const double *mu_v = lds_dd + ((size_t)a * ND + v2) * DD_REC + 3;
          const double *mu_u = lds_dd + ((size_t)a * ND + u) * DD_REC + 3;
          dd3_apply(r, ts1, ts2, mu_v, o);
          for (int k = 0; k < 3; ++k) ef[u][k] += o[k];
          dd3_apply(r, ts1, ts2, mu_u, o);
          for (int k = 0; k < 3; ++k) ef[v2][k] += o[k];
        }
    }
    for (int u = 0; u < ND; ++u)
      for (int k = 0; k < 3; ++k) Efd[(grp * ND + u) * 3 + k] = ef[u][k];
  }
}

static size_t dd_lds(const SysConst &c) { return sizeof(double) * DD_REC * c.N * c.nd; }



// The LDS-staged dipole pass holds a whole bead's dipole sites: up to
// PIMD_DD_LDS_KB (default 64) KiB of LDS per workgroup, the gather kernel
// beyond.
static size_t dd_lds_max() {
  static const size_t v = [] {
    const char *e = std::getenv("PIMD_DD_LDS_KB");
    long kb = e ? std::strtol(e, nullptr, 10) : 64;
    if (kb < 16 || kb > 150) kb = 64;
    return (size_t)kb * 1024;
  }();
  return v;
}
bool pair_dd_tests(const SysConst &c) { return c.g_dd > 0 && dd_lds(c) <= dd_lds_max(); }



// fp64 forces the FP64 form (the refinement pass of the mixed solve).

// Start of the FP64 refinement of the mixed solve: every bead's dipoles go
// back to the input of its last SOR step and every bead is active again.
__global__ __launch_bounds__(256) void k_refine_prep(const Rep *__restrict__ reps, BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const double *__restrict__ dip_prev = Rp.b.dip_prev;
  double *__restrict__ dip = Rp.b.dip;
  int *active = Rp.b.active;
  const size_t n = (size_t)c.P * c.N * c.nd * 3;
  const size_t i = (size_t)bk.x * blockDim.x + threadIdx.x;
  if (i < n) dip[i] = dip_prev[i];
  if (i < (size_t)c.P) active[i] = 1;
}



// Per-bead convergence test (pot_ttm.f90:380-385): deltadip =
// DEBYE/CHARGECON/sqrt(Natoms) * ||mu_new - mu_old||, summed in a fixed order.
__global__ __launch_bounds__(256) void k_converge(const Rep *__restrict__ reps, BlockMap bm,
                                                 int iter) {
  __shared__ double red[256];
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const double *__restrict__ diff2 = Rp.b.diff2;
  int *active = Rp.b.active, *iters = Rp.b.iters, *nactive = Rp.b.nactive;
  const int b = bk.x;
  if (!active[b]) return;
  const int M = c.N * c.nd;
  double s = 0.0;
  for (int m = threadIdx.x; m < M; m += 256) s += diff2[(size_t)b * M + m];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double delta = c.stath * sqrt(red[0]);
    if (delta < c.peps) {
      active[b] = 0;
      iters[b] = iter;
    } else if (iter >= c.maxiter) {
      active[b] = 0;
      iters[b] = -iter;
    } else {
      iters[b] = iter;
      atomicAdd(&nactive[iter], 1);
    }
  }
}


// ---------------------------------------------------------------------------
// Real-space forces of the induced dipoles (pot_ttm.f90:399-478): dipole-dipole
// (third-rank tensor) and charge-dipole terms, the dipole potential phi.
// ---------------------------------------------------------------------------
template <int MODEL>
__device__ __forceinline__ bool is_dip(int t) {
  return MODEL == 2 ? t < 3 : t == 3;
}

// LDS = true: as k_pair_qq, with the bead's dipoles staged beside the sites.
template <int MODEL, int G, typename R, bool LDS = false>
__global__ __launch_bounds__(256) void k_pair_final(const Rep *__restrict__ reps, BlockMap bm) {
  constexpr int ND = MODEL == 2 ? 3 : 1;
  constexpr int T0 = MODEL == 2 ? 0 : 3;
  constexpr int FIN_SB_UNROLL = MODEL == 2 && sizeof(R) == 8 ? 1 : 4;
  extern __shared__ Site lds_site[];  // [4N] sites, then [N][ND][3] dipoles
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ Rcs = Rp.b.Rcs, *__restrict__ dip = Rp.b.dip;
  const int *__restrict__ nbr = Rp.b.nbr, *__restrict__ nbrc = Rp.b.nbrc, *__restrict__ nbrn = Rp.b.nbrn;
  double *phi = Rp.b.phi, *dR = Rp.b.dR, *tally = Rp.b.tally;
  double *lds_dip = reinterpret_cast<double *>(lds_site + 4 * c.N);
  const int N = c.N;
  long grp;
  int lane;
  bool valid;
  if (LDS) {
    const int bb = bk.y;
    for (int e = threadIdx.x; e < 4 * N; e += blockDim.x) lds_site[e] = site[(size_t)bb * 4 * N + e];
    for (int e = threadIdx.x; e < N * ND * 3; e += blockDim.x)
      lds_dip[e] = dip[(size_t)bb * N * ND * 3 + e];
    __syncthreads();
    const int aa = bk.x * (256 / G) + threadIdx.x / G;
    lane = threadIdx.x % G;
    valid = aa < N;
    grp = (long)bb * N + (valid ? aa : 0);
  } else {
    const long gtid = (long)bk.x * blockDim.x + threadIdx.x;
    grp = gtid / G;
    lane = (int)(gtid % G);
    valid = grp < (long)c.P * N;
  }
  const int b = valid ? (int)(grp / N) : 0, a = valid ? (int)(grp % N) : 0;
  // molecule records and dipoles of this bead: LDS or global
  const Site *bead_sites = LDS ? lds_site : site + (size_t)b * N * 4;
  const double *bead_dip = LDS ? lds_dip : dip + (size_t)b * N * ND * 3;
  Site A[4];
  double mua[4][3];
  for (int t = 0; t < 4; ++t) {
    A[t] = bead_sites[a * 4 + t];
    for (int k = 0; k < 3; ++k)
      mua[t][k] = is_dip<MODEL>(t) ? bead_dip[((size_t)a * ND + (t - T0)) * 3 + k] : 0.0;
  }
  double ph[4] = {0, 0, 0, 0}, fr[4][3] = {{0}}, vc = 0.0, vt[9] = {0};
  const int nf = valid ? nbrc[grp] : 0, nn = valid ? nbrn[grp] : 0;
  const int *list = nbr + (size_t)grp * c.nbr_cap;
  auto visit = [&](const int j, auto thole_tag) {
    constexpr bool THOLE = decltype(thole_tag)::value;
    const bool lo = a < j;
    // one site of j at a time (read in the loop) against every site of a:
    // each accumulator of a site of a still takes its terms in the order
    // (j, site of j).  Unrolled except in the FP64 form of TTM2.1-F, where one
    // site at a time keeps it at two waves per SIMD (TTM2.1-F 128 x 32: 2 %
    // faster force calls; TTM3-F FP64 is faster unrolled).
#pragma unroll FIN_SB_UNROLL
    for (int sb = MODEL == 3 ? 1 : 0; sb < 4; ++sb) {
      const bool db = is_dip<MODEL>(sb);
      const Site Bs = bead_sites[j * 4 + sb];
      double mubs[3];
      for (int k = 0; k < 3; ++k)
        mubs[k] = db ? bead_dip[((size_t)j * ND + max(sb - T0, 0)) * 3 + k] : 0.0;
#pragma unroll
      for (int sa = 0; sa < 4; ++sa) {
        const bool da = is_dip<MODEL>(sa);
        // Pairs whose force and live potentials vanish identically: neither
        // site carries a dipole, or the only coupling is a zero charge.
        if (!da && !db) continue;
        if (MODEL == 3 && sa == 0) continue;
        // site-site minimum image oriented from a to b (pot_ttm.f90:419-420);
        // v is the reference's orientation (lower molecule first)
        double rd[3] = {A[sa].x - Bs.x, A[sa].y - Bs.y, A[sa].z - Bs.z};
        minimg(c, rd[0], rd[1], rd[2]);
        const double *v = rd;
        const double d2 = norm2_exact(rd[0], rd[1], rd[2]);
        const R r[3] = {(R)rd[0], (R)rd[1], (R)rd[2]};
        const R qa = (R)A[sa].q, qb = (R)Bs.q;
        R di[3], dj[3];
        for (int k = 0; k < 3; ++k) {
          di[k] = (R)mua[sa][k];
          dj[k] = (R)mubs[k];
        }
        R qd[3];
        for (int k = 0; k < 3; ++k) qd[k] = qb * di[k] - qa * dj[k];
        const R qsq = qd[0] * qd[0] + qd[1] * qd[1] + qd[2] * qd[2];
        R dijr[3] = {0, 0, 0}, ts1 = 0, ts2 = 0;
        bool have_ts1 = false;
        if (da && db) {
          R t1d, t2d, t3d;
          smear3<THOLE>(c, (R)d2, (R)(lo ? c.apol3DD[sa][sb] : c.apol3DD[sb][sa]), t1d, t2d, t3d);
          // di . T3(r) . dj with T3_abc = 15 r_a r_b r_c ts3 - 3 ts2 (r_a d_bc + r_b d_ac + r_c d_ab)
          const R rdi = r[0] * di[0] + r[1] * di[1] + r[2] * di[2];
          const R rdj = r[0] * dj[0] + r[1] * dj[1] + r[2] * dj[2];
          const R didj = di[0] * dj[0] + di[1] * dj[1] + di[2] * dj[2];
          for (int k = 0; k < 3; ++k)
            dijr[k] = (R)15.0 * t3d * rdi * rdj * r[k] -
                      (R)3.0 * t2d * (rdi * dj[k] + rdj * di[k] + didj * r[k]);
          if (qsq > (R)1e-9) {
            if (c.same_a) {
              ts1 = t1d;
              ts2 = t2d;
            } else {
              smear2<THOLE>(c, (R)d2, (R)(lo ? c.apolCC[sa][sb] : c.apolCC[sb][sa]), ts1, ts2);
            }
            have_ts1 = true;
            R o[3];
            dd3_apply(r, ts1, ts2, qd, o);
            for (int k = 0; k < 3; ++k) dijr[k] += o[k];
          }
        } else if (qsq > (R)1e-9) {
          smear2<THOLE>(c, (R)d2, (R)(lo ? c.apolCC[sa][sb] : c.apolCC[sb][sa]), ts1, ts2);
          have_ts1 = true;
          R o[3];
          dd3_apply(r, ts1, ts2, qd, o);
          for (int k = 0; k < 3; ++k) dijr[k] += o[k];
        }
        if (have_ts1 && db) ph[sa] += (double)(ts1 * (r[0] * dj[0] + r[1] * dj[1] + r[2] * dj[2]));
        const double dij[3] = {(double)dijr[0], (double)dijr[1], (double)dijr[2]};
        for (int k = 0; k < 3; ++k) fr[sa][k] += dij[k];
        if (lo) {
          double rcv[3];
          // Rij0 of the dipole loop is wrapped to ~0 (pot_ttm.f90:406-407,468)
          for (int k = 0; k < 3; ++k) rcv[k] = Rcs[(a * 4 + sa) * 3 + k] - Rcs[(j * 4 + sb) * 3 + k];
          vc += rcv[0] * dij[0] + rcv[1] * dij[1] + rcv[2] * dij[2];
          for (int col = 0; col < 3; ++col)
            for (int row = 0; row < 3; ++row) vt[col * 3 + row] += dij[row] * v[col];
        }
      }
    }
  };
  walk_list<G>(list, lane, nn, [&](int j) { visit(j, std::true_type{}); });
  walk_list<G>(list, c.nbr_cap - nf + lane, c.nbr_cap, [&](int j) { visit(j, std::false_type{}); });
  for (int t = 0; t < 4; ++t) {
    ph[t] = group_sum<G>(ph[t]);
    for (int k = 0; k < 3; ++k) fr[t][k] = group_sum<G>(fr[t][k]);
  }
  vc = group_sum<G>(vc);
  for (int i = 0; i < 9; ++i) vt[i] = group_sum<G>(vt[i]);
  if (valid && lane == 0) {
    const size_t base = ((size_t)b * N + a) * 4;
    for (int t = 0; t < 4; ++t) {
      phi[base + t] += ph[t];
      for (int k = 0; k < 3; ++k) dR[(base + t) * 3 + k] += fr[t][k];
    }
    double *tl = tally + ((size_t)b * N + a) * T_COUNT;
    tl[T_VIRC] += vc;
    vadd9(tl + T_VIRT, vt);
  }
}





// k_pair_final with each molecule pair evaluated once, by the molecule that
// owns it (as k_pair_qq_half): the partner's site potentials (its dipoles'
// field terms with r reversed) and forces (the pair force reversed) go into
// the block's two-word fixed-point LDS rows, added to the partner by
// k_pair_final_partner.  The virial terms are even under the exchange of the
// two molecules (site-centroid difference and force both reverse), so the
// owner adds them whichever molecule is the lower one.  Against the
// two-sided pass, with one-word partner sums in FP32: 512 H2O x 32 TTM3-F
// 0.104 -> 0.092 ms, 128 x 32 TTM2.1-F 0.060 -> 0.049 ms (two words: slower
// than two-sided).
constexpr int FIN_NV = 16;  // partner values: potentials of sites 0-3, then their forces
template <int MODEL, int G, typename R>
__global__ __launch_bounds__(QQ_HALF_BLOCK) void k_pair_final_half(const Rep *__restrict__ reps,
                                                                   BlockMap bm) {
  constexpr int ND = MODEL == 2 ? 3 : 1;
  constexpr int T0 = MODEL == 2 ? 0 : 3;
  constexpr int SB0 = MODEL == 3 ? 1 : 0;
  constexpr int OWN = QQ_HALF_BLOCK / G;
  constexpr int NW = QqWords<R>::n;  // words per partner value
  // [W][4] sites, [W][ND][3] dipoles, [W][FIN_NV][NW] partner sums
  extern __shared__ Site lds_site[];
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ Rcs = Rp.b.Rcs, *__restrict__ dip = Rp.b.dip;
  const int *__restrict__ nbr = Rp.b.nbr, *__restrict__ nbrc = Rp.b.nbrc,
            *__restrict__ nbrn = Rp.b.nbrn, *__restrict__ nbrh = Rp.b.nbrh;
  double *phi = Rp.b.phi, *dR = Rp.b.dR, *tally = Rp.b.tally;
  const int N = c.N, b = bk.y, i0 = bk.x * OWN, W = qq_half_rows(N, G);
  double *lds_dip = reinterpret_cast<double *>(lds_site + 4 * W);
  unsigned long long *jacc = reinterpret_cast<unsigned long long *>(lds_dip + (size_t)W * ND * 3);
  for (int e = threadIdx.x; e < 4 * W; e += blockDim.x) {
    int m = i0 + e / 4;
    m -= m >= N ? N : 0;
    lds_site[e] = site[((size_t)b * N + m) * 4 + e % 4];
  }
  for (int e = threadIdx.x; e < W * ND * 3; e += blockDim.x) {
    int m = i0 + e / (ND * 3);
    m -= m >= N ? N : 0;
    lds_dip[e] = dip[((size_t)b * N + m) * ND * 3 + e % (ND * 3)];
  }
  for (int e = threadIdx.x; e < W * FIN_NV * NW; e += blockDim.x) jacc[e] = 0ull;
  __syncthreads();
  const int a0 = i0 + threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const bool valid = a0 < N;
  const int a = valid ? a0 : i0;
  const long grp = (long)b * N + a;
  auto row = [&](int j) { return j - i0 + (j < i0 ? N : 0); };
  Site A[4];
  double mua[4][3];
  for (int t = 0; t < 4; ++t) {
    A[t] = lds_site[(a - i0) * 4 + t];
    for (int k = 0; k < 3; ++k)
      mua[t][k] = is_dip<MODEL>(t) ? lds_dip[((a - i0) * ND + (t - T0)) * 3 + k] : 0.0;
  }
  double ph[4] = {0, 0, 0, 0}, fr[4][3] = {{0}}, vc = 0.0, vt[9] = {0};
  const int *list = nbr + (size_t)grp * c.nbr_cap;
  auto jslot = [&](int rj, int k) { return jacc + ((size_t)rj * FIN_NV + k) * NW; };
  auto visit = [&](const int j, auto thole_tag) {
    constexpr bool THOLE = decltype(thole_tag)::value;
    const bool lo = a < j;
    const int rj = row(j);
#pragma unroll 1
    for (int sb = SB0; sb < 4; ++sb) {
      const bool db = is_dip<MODEL>(sb);
      const Site Bs = lds_site[rj * 4 + sb];
      double mubs[3];
      for (int k = 0; k < 3; ++k)
        mubs[k] = db ? lds_dip[(rj * ND + max(sb - T0, 0)) * 3 + k] : 0.0;
      double jp = 0.0, jf[3] = {0.0, 0.0, 0.0};  // the partner's site sb
#pragma unroll
      for (int sa = 0; sa < 4; ++sa) {
        const bool da = is_dip<MODEL>(sa);
        if (!da && !db) continue;
        if (MODEL == 3 && sa == 0) continue;
        double rd[3] = {A[sa].x - Bs.x, A[sa].y - Bs.y, A[sa].z - Bs.z};
        minimg(c, rd[0], rd[1], rd[2]);
        const double d2 = norm2_exact(rd[0], rd[1], rd[2]);
        const R r[3] = {(R)rd[0], (R)rd[1], (R)rd[2]};
        const R qa = (R)A[sa].q, qb = (R)Bs.q;
        R di[3], dj[3];
        for (int k = 0; k < 3; ++k) {
          di[k] = (R)mua[sa][k];
          dj[k] = (R)mubs[k];
        }
        R qd[3];
        for (int k = 0; k < 3; ++k) qd[k] = qb * di[k] - qa * dj[k];
        const R qsq = qd[0] * qd[0] + qd[1] * qd[1] + qd[2] * qd[2];
        R dijr[3] = {0, 0, 0}, ts1 = 0, ts2 = 0;
        bool have_ts1 = false;
        if (da && db) {
          R t1d, t2d, t3d;
          smear3<THOLE>(c, (R)d2, (R)(lo ? c.apol3DD[sa][sb] : c.apol3DD[sb][sa]), t1d, t2d, t3d);
          const R rdi = r[0] * di[0] + r[1] * di[1] + r[2] * di[2];
          const R rdj = r[0] * dj[0] + r[1] * dj[1] + r[2] * dj[2];
          const R didj = di[0] * dj[0] + di[1] * dj[1] + di[2] * dj[2];
          for (int k = 0; k < 3; ++k)
            dijr[k] = (R)15.0 * t3d * rdi * rdj * r[k] -
                      (R)3.0 * t2d * (rdi * dj[k] + rdj * di[k] + didj * r[k]);
          if (qsq > (R)1e-9) {
            if (c.same_a) {
              ts1 = t1d;
              ts2 = t2d;
            } else {
              smear2<THOLE>(c, (R)d2, (R)(lo ? c.apolCC[sa][sb] : c.apolCC[sb][sa]), ts1, ts2);
            }
            have_ts1 = true;
            R o[3];
            dd3_apply(r, ts1, ts2, qd, o);
            for (int k = 0; k < 3; ++k) dijr[k] += o[k];
          }
        } else if (qsq > (R)1e-9) {
          smear2<THOLE>(c, (R)d2, (R)(lo ? c.apolCC[sa][sb] : c.apolCC[sb][sa]), ts1, ts2);
          have_ts1 = true;
          R o[3];
          dd3_apply(r, ts1, ts2, qd, o);
          for (int k = 0; k < 3; ++k) dijr[k] += o[k];
        }
        // this site's potential from the partner's dipole, and the partner
        // site's from this one's (r reversed)
        if (have_ts1 && db) ph[sa] += (double)(ts1 * (r[0] * dj[0] + r[1] * dj[1] + r[2] * dj[2]));
        if (have_ts1 && da) jp -= (double)(ts1 * (r[0] * di[0] + r[1] * di[1] + r[2] * di[2]));
        const double dij[3] = {(double)dijr[0], (double)dijr[1], (double)dijr[2]};
        for (int k = 0; k < 3; ++k) {
          fr[sa][k] += dij[k];
          jf[k] -= dij[k];
        }
        // Rij0 of the dipole loop is wrapped to ~0 (pot_ttm.f90:406-407,468)
        double rcv[3];
        for (int k = 0; k < 3; ++k) rcv[k] = Rcs[(a * 4 + sa) * 3 + k] - Rcs[(j * 4 + sb) * 3 + k];
        vc += rcv[0] * dij[0] + rcv[1] * dij[1] + rcv[2] * dij[2];
        for (int col = 0; col < 3; ++col)
          for (int rw = 0; rw < 3; ++rw) vt[col * 3 + rw] += dij[rw] * rd[col];
      }
      qq_fix_add<NW>(jslot(rj, sb), jp);
      for (int k = 0; k < 3; ++k) qq_fix_add<NW>(jslot(rj, 4 + 3 * sb + k), jf[k]);
    }
  };
  if (valid) {
    const int4 h = make_int4(nbrh[grp * 4], nbrh[grp * 4 + 1], nbrh[grp * 4 + 2], nbrh[grp * 4 + 3]);
    const int nn = nbrn[grp], nf = nbrc[grp], cap = c.nbr_cap;
    const bool wrap = pair_half_window(a, N) >= N;
    // the owned entries, as in k_pair_qq_half
    walk_ranges<G>(list, lane, h.x, wrap ? nn : h.y, 0, wrap ? h.y : 0,
                   [&](int j) { visit(j, std::true_type{}); });
    walk_ranges<G>(list, lane, wrap ? cap - nf : cap - h.w, cap - h.z, cap - h.w, wrap ? cap : cap - h.w,
                   [&](int j) { visit(j, std::false_type{}); });
  }
  for (int t = 0; t < 4; ++t) {
    ph[t] = group_sum<G>(ph[t]);
    for (int k = 0; k < 3; ++k) fr[t][k] = group_sum<G>(fr[t][k]);
  }
  vc = group_sum<G>(vc);
  for (int i = 0; i < 9; ++i) vt[i] = group_sum<G>(vt[i]);
  if (valid && lane == 0) {
    const size_t base = ((size_t)b * N + a) * 4;
    for (int t = 0; t < 4; ++t) {
      phi[base + t] += ph[t];
      for (int k = 0; k < 3; ++k) dR[(base + t) * 3 + k] += fr[t][k];
    }
    double *tl = tally + ((size_t)b * N + a) * T_COUNT;
    tl[T_VIRC] += vc;
    vadd9(tl + T_VIRT, vt);
  }
  __syncthreads();
  // the block's exact row sums, rounded once (rows of QQ_NV, as the charge pass)
  double *slab = Rp.b.qqslab + ((size_t)b * bk.nx + bk.x) * (size_t)W * QQ_NV;
  for (int e = threadIdx.x; e < W * FIN_NV; e += blockDim.x) {
    const int rw = e / FIN_NV, k = e % FIN_NV;
    slab[(size_t)rw * QQ_NV + k] = qq_fix_value<NW>(jacc + NW * (size_t)e);
  }
}

// The partner-side sums of k_pair_final_half (as k_pair_qq_partner): lane k
// of a molecule's 32 adds value k over the blocks, in block order.
template <int MODEL>
__global__ __launch_bounds__(256) void k_pair_final_partner(const Rep *__restrict__ reps,
                                                            BlockMap bm) {
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const int N = c.N;
  const long m = ((long)bk.x * blockDim.x + threadIdx.x) / 32;
  const int k = threadIdx.x % 32;
  // TTM3-F: no values for the oxygen (slots 0 and 4-6)
  const bool used = k < FIN_NV && !(MODEL == 3 && (k == 0 || (k >= 4 && k < 7)));
  if (m >= (long)c.P * N || !used) return;
  const int b = (int)(m / N), a = (int)(m % N);
  const int G = half_lanes(c.g_fin), own = QQ_HALF_BLOCK / G;
  const int nblk = qq_half_blocks(N, G), W = qq_half_rows(N, G);
  const double *slab = Rp.b.qqslab + (size_t)b * nblk * W * QQ_NV + k;
  double v = 0.0;
  for (int q = 0; q < nblk; ++q) {
    int r = a - q * own;
    r += r < 0 ? N : 0;
    if (r < W) v += slab[((size_t)q * W + r) * QQ_NV];
  }
  const size_t base = ((size_t)b * N + a) * 4;
  if (k < 4) Rp.b.phi[base + k] += v;
  else Rp.b.dR[(base + (k - 4) / 3) * 3 + (k - 4) % 3] += v;
}

// Reciprocal-space virial (ewald.f90:194-224), partial sums per block of k:
// virk[b][blk][0..8] = sum Ureck (delta - coeff2 k k), [9..17] = dipole part.
__global__ __launch_bounds__(256) void k_virk(const Rep *__restrict__ reps, BlockMap bm) {
  __shared__ double red[18][256];
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const double4 *__restrict__ kv = Rp.b.kv;
  const double *__restrict__ Ak = Rp.b.red;
  const double *__restrict__ Sq = Rp.b.Sq, *__restrict__ Sm = Rp.b.Sm, *__restrict__ Sv = Rp.b.Sv;
  double *virk = Rp.b.virk;
  const int b = bk.y;
  const int k = bk.x * 256 + threadIdx.x;
  double v[18];
  for (int i = 0; i < 18; ++i) v[i] = 0.0;
  if (k < c.nk) {
    const double4 kk = kv[k];
    const double kx[3] = {kk.x, kk.y, kk.z};
    const double A = Ak[k];
    const size_t o = (size_t)b * c.nk + k;
    // models without induced dipoles (nd = 0) have the charge part only
    // (ewald.f90:267-279)
    const bool dip = c.nd > 0;
    const double rqqm = dip ? Sq[o * 2] + Sm[o * 2] : Sq[o * 2];
    const double iqqm = dip ? Sq[o * 2 + 1] + Sm[o * 2 + 1] : Sq[o * 2 + 1];
    const double ksq = kx[0] * kx[0] + kx[1] * kx[1] + kx[2] * kx[2];
    const double coeff2 = 2.0 / ksq + 0.5 / (c.aewald * c.aewald);
    const double U = A * (rqqm * rqqm + iqqm * iqqm);
    for (int col = 0; col < 3; ++col)
      for (int row = 0; row < 3; ++row)
        v[col * 3 + row] = (row == col ? U : 0.0) - U * coeff2 * kx[row] * kx[col];
    const double *sv = Sv + o * 6;  // sum mu cos (3), sum mu sin (3)
    for (int ix = 0; ix < 3 && dip; ++ix) {
      const double t = 2.0 * A * kx[ix];
      for (int iy = 0; iy < 3; ++iy)
        v[9 + iy * 3 + ix] = t * (rqqm * (-sv[3 + iy]) + iqqm * sv[iy]);
    }
  }
  for (int i = 0; i < 18; ++i) red[i][threadIdx.x] = v[i];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int i = 0; i < 18; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 18) virk[((size_t)b * bk.nx + bk.x) * 18 + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// Per-molecule finish (pot_ttm.f90:380-392, 482-542, 556-627): intramolecular
// dipole terms, Uind, predictor history, charge-derivative chain rule, M-site
// redistribution and molecular dipoles.  One thread per (bead, molecule).
// ---------------------------------------------------------------------------
template <int MODEL>
__global__ __launch_bounds__(64) void k_finalize(const Rep *__restrict__ reps, BlockMap bm) {
  constexpr int ND = MODEL == 2 ? 3 : 1;
  constexpr int T0 = MODEL == 2 ? 0 : 3;
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const Site *__restrict__ site = Rp.b.site;
  const double *__restrict__ Rcs = Rp.b.Rcs, *__restrict__ grdq = Rp.b.grdq;
  const double *__restrict__ dip = Rp.b.dip, *__restrict__ Efq = Rp.b.Efq;
  const double *__restrict__ kE = Rp.b.kE, *__restrict__ kphi = Rp.b.kphi;
  double *phi = Rp.b.phi, *dR = Rp.b.dR, *hist = Rp.b.hist, *tally = Rp.b.tally;
  double *dRR = Rp.b.dRR, *dipI = Rp.b.dipI, *dipE = Rp.b.dipE;
  const int *__restrict__ clk = Rp.b.clk;
  const int gid = bk.x * blockDim.x + threadIdx.x;
  if (gid >= c.P * c.N) return;
  const int t = clk[0];  // MD step of this force call
  const int b = gid / c.N, w = gid % c.N, N = c.N;
  const size_t base = ((size_t)b * N + w) * 4;
  Site A[4];
  double mu[4][3], cen[4][3], ph[4], fr[4][3];
  for (int ty = 0; ty < 4; ++ty) {
    A[ty] = site[base + ty];
    ph[ty] = phi[base + ty];
    for (int k = 0; k < 3; ++k) {
      mu[ty][k] = is_dip<MODEL>(ty) ? dip[(((size_t)b * N + w) * ND + (ty - T0)) * 3 + k] : 0.0;
      cen[ty][k] = Rcs[(w * 4 + ty) * 3 + k];
      fr[ty][k] = dR[(base + ty) * 3 + k];
    }
    // plus the reciprocal-space part of the last field pass (ewald.f90:184-193):
    // every site of the pass (all four for TTM2.1-F), the potential of the
    // charged sites
    if (MODEL == 2 || ty > 0) {
      for (int k = 0; k < 3; ++k) fr[ty][k] = fr[ty][k] + kE[(base + ty) * 3 + k];
      if (ty > 0) ph[ty] = ph[ty] + kphi[base + ty];
    }
  }
  double vc = 0.0, vt[9] = {0};
  // intramolecular dipole terms (pot_ttm.f90:482-542)
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 4; ++j) {
      double r[3] = {A[i].x - A[j].x, A[i].y - A[j].y, A[i].z - A[j].z};
      minimg(c, r[0], r[1], r[2]);
      double rc[3] = {cen[i][0] - cen[j][0], cen[i][1] - cen[j][1], cen[i][2] - cen[j][2]};
      minimg(c, rc[0], rc[1], rc[2]);
      const double d2 = norm2_exact(r[0], r[1], r[2]);
      const double *di = mu[i], *dj = mu[j];
      if (is_dip<MODEL>(i) && is_dip<MODEL>(j)) {
        double t1, t2, t3;
        smear3(c, d2, c.apol3DD[i][j], t1, t2, t3);
        const double rdi = r[0] * di[0] + r[1] * di[1] + r[2] * di[2];
        const double rdj = r[0] * dj[0] + r[1] * dj[1] + r[2] * dj[2];
        const double didj = di[0] * dj[0] + di[1] * dj[1] + di[2] * dj[2];
        double dij[3];
        for (int k = 0; k < 3; ++k)
          dij[k] = 15.0 * t3 * rdi * rdj * r[k] - 3.0 * t2 * (rdi * dj[k] + rdj * di[k] + didj * r[k]);
        for (int k = 0; k < 3; ++k) {
          fr[i][k] += dij[k];
          fr[j][k] -= dij[k];
        }
        vc += rc[0] * dij[0] + rc[1] * dij[1] + rc[2] * dij[2];
        for (int col = 0; col < 3; ++col)
          for (int row = 0; row < 3; ++row) vt[col * 3 + row] += dij[row] * r[col];
      }
      double qd[3];
      for (int k = 0; k < 3; ++k) qd[k] = A[j].q * di[k] - A[i].q * dj[k];
      double ts1, ts2, dij[3];
      self2(c, d2, ts1, ts2);
      dd3_apply(r, ts1, ts2, qd, dij);
      ph[i] += ts1 * (r[0] * dj[0] + r[1] * dj[1] + r[2] * dj[2]);
      ph[j] -= ts1 * (r[0] * di[0] + r[1] * di[1] + r[2] * di[2]);
      for (int k = 0; k < 3; ++k) {
        fr[i][k] += dij[k];
        fr[j][k] -= dij[k];
      }
      vc += rc[0] * dij[0] + rc[1] * dij[1] + rc[2] * dij[2];
      for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) vt[col * 3 + row] += dij[row] * r[col];
    }
  // Uind and predictor history (pot_ttm.f90:390-392)
  double ui = 0.0;
  for (int u = 0; u < ND; ++u)
    for (int k = 0; k < 3; ++k) ui += mu[T0 + u][k] * Efq[(base + T0 + u) * 3 + k];
  if (c.guess) {
    const int slot = t % 4, ndip = N * ND;
    for (int u = 0; u < ND; ++u)
      for (int k = 0; k < 3; ++k)
        hist[(((size_t)b * 4 + slot) * ndip + (size_t)w * ND + u) * 3 + k] = mu[T0 + u][k];
  }
  // chain rule through the DMS charges (pot_ttm.f90:556-584)
  const double *g = grdq + ((size_t)b * N + w) * 27;
  double dg[3][3];  // [O, H1, H2][xyz]
  for (int k = 0; k < 3; ++k) {
    dg[1][k] = g[0 * 9 + 0 * 3 + k] * ph[1] + g[0 * 9 + 1 * 3 + k] * ph[2] + g[0 * 9 + 2 * 3 + k] * ph[3];
    dg[2][k] = g[1 * 9 + 0 * 3 + k] * ph[1] + g[1 * 9 + 1 * 3 + k] * ph[2] + g[1 * 9 + 2 * 3 + k] * ph[3];
    dg[0][k] = g[2 * 9 + 0 * 3 + k] * ph[1] + g[2 * 9 + 1 * 3 + k] * ph[2] + g[2 * 9 + 2 * 3 + k] * ph[3];
  }
  for (int k = 0; k < 3; ++k) {
    fr[1][k] += dg[1][k];
    fr[2][k] += dg[2][k];
    fr[0][k] += dg[0][k];
  }
  {
    double o1[3] = {A[1].x - A[0].x, A[1].y - A[0].y, A[1].z - A[0].z};
    double o2[3] = {A[2].x - A[0].x, A[2].y - A[0].y, A[2].z - A[0].z};
    minimg(c, o1[0], o1[1], o1[2]);
    minimg(c, o2[0], o2[1], o2[2]);
    for (int ix = 0; ix < 3; ++ix)
      for (int iy = 0; iy < 3; ++iy) vt[iy * 3 + ix] += o1[ix] * dg[1][iy] + o2[ix] * dg[2][iy];
    double c1[3] = {cen[1][0] - cen[0][0], cen[1][1] - cen[0][1], cen[1][2] - cen[0][2]};
    double c2[3] = {cen[2][0] - cen[0][0], cen[2][1] - cen[0][1], cen[2][2] - cen[0][2]};
    minimg(c, c1[0], c1[1], c1[2]);
    minimg(c, c2[0], c2[1], c2[2]);
    for (int k = 0; k < 3; ++k) vc += c1[k] * dg[1][k] + c2[k] * dg[2][k];
  }
  // M-site redistribution (pot_ttm.f90:597-602)
  const double gM = c.gammaM;
  double *out = dRR + ((size_t)b * 3 * N + 3 * w) * 3;
  for (int k = 0; k < 3; ++k) {
    out[3 + k] = fr[1][k] + 0.5 * gM * fr[3][k];
    out[6 + k] = fr[2][k] + 0.5 * gM * fr[3][k];
    out[k] = fr[0][k] + (1.0 - gM) * fr[3][k];
  }
  // molecular dipoles (pot_ttm.f90:605-627)
  {
    double h1[3] = {A[1].x - A[3].x, A[1].y - A[3].y, A[1].z - A[3].z};
    double h2[3] = {A[2].x - A[3].x, A[2].y - A[3].y, A[2].z - A[3].z};
    minimg(c, h1[0], h1[1], h1[2]);
    minimg(c, h2[0], h2[1], h2[2]);
    double dm[3], de[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) dm[k] = A[1].q * h1[k] + A[2].q * h2[k];
    for (int u = 0; u < ND; ++u)
      for (int k = 0; k < 3; ++k) {
        dm[k] = mu[T0 + u][k] + dm[k];
        de[k] = mu[T0 + u][k];
      }
    if (dipI)
      for (int k = 0; k < 3; ++k) dipI[((size_t)b * N + w) * 3 + k] = dm[k];
    if (dipE)
      for (int k = 0; k < 3; ++k) dipE[((size_t)b * N + w) * 3 + k] = de[k];
  }
  double *tl = tally + ((size_t)b * N + w) * T_COUNT;
  tl[T_UIND] = ui;
  tl[T_VIRC] += vc;
  vadd9(tl + T_VIRT, vt);
}



// Per-bead totals (pot_ttm.f90:552-553, 588-591, 628; force_calc.f90:48).
__global__ __launch_bounds__(256) void k_reduce_bead(const Rep *__restrict__ reps, BlockMap bm) {
  __shared__ double red[T_COUNT][256];
  const Blk bk = block_of(bm);
  const Rep &Rp = reps[bk.r];
  const SysConst &c = Rp.c;
  const double *__restrict__ tally = Rp.b.tally, *__restrict__ virk = Rp.b.virk;
  const int nkb = (c.nk + 255) / 256;
  const int *__restrict__ iters_in = Rp.b.iters;
  double *Upot = Rp.b.Upot, *virt = Rp.b.virt, *virialc = Rp.b.virialc, *terms = Rp.b.terms;
  int *iters = Rp.b.iters_out;
  const int b = bk.x;
  double v[T_COUNT];
  for (int i = 0; i < T_COUNT; ++i) v[i] = 0.0;
  for (int w = threadIdx.x; w < c.N; w += 256)
    for (int i = 0; i < T_COUNT; ++i) v[i] += tally[((size_t)b * c.N + w) * T_COUNT + i];
  for (int i = 0; i < T_COUNT; ++i) red[i][threadIdx.x] = v[i];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int i = 0; i < T_COUNT; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double vr[18];
    for (int i = 0; i < 18; ++i) vr[i] = 0.0;
    for (int blk = 0; blk < nkb; ++blk)
      for (int i = 0; i < 18; ++i) vr[i] += virk[((size_t)b * nkb + blk) * 18 + i];
    double vrec[9];
    for (int i = 0; i < 9; ++i)
      vrec[i] = -0.5 * vr[i] * FOURPI / c.vol + -0.5 * vr[9 + i] * FOURPI / c.vol;
    const double umon = red[T_UMON][0], uvdw = red[T_UVDW][0];
    const double uelec = 0.5 * red[T_UELEC][0], uind = -0.5 * red[T_UIND][0];
    double vt[9];
    for (int i = 0; i < 9; ++i) vt[i] = red[T_VIRT + i][0] + vrec[i];
    vt[0] -= c.Uvdw_lrc;
    vt[4] -= c.Uvdw_lrc;
    vt[8] -= c.Uvdw_lrc;
    // the fixed-charge models leave virialc unset (pot_spc.f90:1-13): zero
    const double vcc =
        c.model >= PIMD_QSPCFW ? 0.0 : red[T_VIRC][0] + vrec[0] + vrec[4] + vrec[8] - 3 * c.Uvdw_lrc;
    if (Upot) Upot[b] = umon + uvdw + uelec + uind + c.Uvdw_lrc;
    if (virt)
      for (int i = 0; i < 9; ++i) virt[b * 9 + i] = vt[i];
    if (virialc) virialc[b] = vcc;
    if (terms) {
      terms[b * 5 + 0] = umon;
      terms[b * 5 + 1] = uvdw;
      terms[b * 5 + 2] = uelec;
      terms[b * 5 + 3] = uind;
      terms[b * 5 + 4] = c.Uvdw_lrc;
    }
    // the fixed-charge models have no SOR solve: zero iterations
    iters[b] = c.model >= PIMD_QSPCFW ? 0 : iters_in[b];
  }
}



// ---------------------------------------------------------------------------
// Launchers.  Each takes a group of replicas and launches one grid per kernel
// variant present (template instantiation: model, arithmetic, lanes per
// molecule, LDS staging), each replica with the grid and variant it would
// have alone.
// ---------------------------------------------------------------------------
static bool is_ttm(const SysConst &c) { return c.model == PIMD_TTM21F || c.model == PIMD_TTM3F; }
static int idiv(long a, long b) { return (int)((a + b - 1) / b); }

void launch_monomer(const Group &g, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        const bool ttm = c.model == PIMD_TTM21F || c.model == PIMD_TTM3F;
        return Plan{ttm ? 0 : -1, (int)(((long)c.P * c.N + 63) / 64), 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_monomer, grid, dim3(64 * MONO_W), 0, s, g.d, m);
      });
}

void launch_pes_fast(const Group &g, bool record, hipStream_t s) {
  launch_over(
      g, [](const SysConst &c, const MDConst &) { return Plan{0, idiv((long)c.P * c.N, 64), 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_pes_fast, grid, dim3(64 * MONO_W), 0, s, g.d, m, record ? 1 : 0);
      });
}

void launch_broadcast_beads(const Group &g, hipStream_t s) {
  launch_over(
      g, [](const SysConst &c, const MDConst &) { return Plan{0, idiv((long)c.P * 9 * c.N, 256), 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_broadcast_beads, grid, dim3(256), 0, s, g.d, m);
      });
}

void launch_contract_record(const Group &g, hipStream_t s) {
  launch_over(
      g, [](const SysConst &, const MDConst &) { return Plan{0, 1, 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_contract_record, grid, dim3(256), 0, s, g.d, m);
      });
}

void launch_neighbors(const Group &g, hipStream_t s) {
  // variant 0: all molecules through LDS, 1: the same from global memory,
  // 2: through the cell list
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        if (c.cells) return Plan{-1, 0, 0, 0, 0};
        const size_t lds = sizeof(double) * 4 * c.N;
        return Plan{lds <= 64 * 1024 ? 0 : 1, idiv(c.N, NB_I), c.P, 1, lds <= 64 * 1024 ? lds : 0};
      },
      [&](int v, dim3 grid, const BlockMap &m, size_t lds) {
        if (v == 0) hipLaunchKernelGGL(k_neighbors<true>, grid, dim3(256), lds, s, g.d, m);
        else hipLaunchKernelGGL(k_neighbors<false>, grid, dim3(256), 0, s, g.d, m);
      });
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        const int nc = c.ncell[0] * c.ncell[1] * c.ncell[2];
        return Plan{c.cells ? 0 : -1, c.P, 1, 1, sizeof(int) * 2 * (size_t)nc};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t lds) {
        hipLaunchKernelGGL(k_cells, grid, dim3(256), lds, s, g.d, m);
      });
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        return Plan{c.cells ? 0 : -1, idiv((long)c.P * c.N * 64, 256), 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_neighbors_cells, grid, dim3(256), 0, s, g.d, m);
      });
}

// Lanes per molecule as a variant index and back.
static int gidx(int G) { return G == 4 ? 0 : G == 8 ? 1 : G == 16 ? 2 : G == 32 ? 3 : 4; }

// Variant of the charge / final pair passes: bit 0 TTM3-F, bit 1 FP32 pair
// arithmetic, bit 2 LDS-staged records, bits 3.. the lanes-per-molecule index.
static Plan pair_plan(const SysConst &c, int g_lds, size_t lds) {
  if (!is_ttm(c)) return Plan{-1, 0, 0, 0, 0};
  const int v0 = (c.model == 3 ? 1 : 0) | (c.mixed ? 2 : 0);
  if (g_lds > 0 && lds <= PAIR_LDS_MAX) {
    const int G = (g_lds == 4 || g_lds == 8 || g_lds == 32) ? g_lds : 16;
    return Plan{v0 | 4 | (gidx(G) << 3), idiv(c.N, 256 / G), c.P, 1, lds};
  }
  const int G = c.g_nolds == 64 ? 64 : c.g_nolds == 32 ? 32 : 16;
  return Plan{v0 | (gidx(G) << 3), idiv((long)c.P * c.N * G, 256), 1, 1, 0};
}

template <int ML, typename R, bool LDS>
static void launch_qq_v(int gi, dim3 grid, const BlockMap &m, size_t lds, const Group &g, hipStream_t s) {
  switch (gi) {
    case 0: if (LDS) hipLaunchKernelGGL((k_pair_qq<ML, 4, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    case 1: if (LDS) hipLaunchKernelGGL((k_pair_qq<ML, 8, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    case 2: hipLaunchKernelGGL((k_pair_qq<ML, 16, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    case 3: hipLaunchKernelGGL((k_pair_qq<ML, 32, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    default: if (!LDS) hipLaunchKernelGGL((k_pair_qq<ML, 64, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
  }
}

// LDS of k_pair_qq_half: the bead's sites and the partner-side sums
static size_t qq_half_lds(const SysConst &c) {
  const size_t W = qq_half_rows(c.N, qq_half_lanes(c));
  return W * (sizeof(Site) * 4 + sizeof(long long) * QQ_NV * (c.mixed ? 1 : 2));
}
// The once-per-pair pass (qq_half 1) where its blocks fill at least half the
// device: one bead of 512 H2O, the centroid call of contracted MD, gets 8-16
// blocks and runs faster two-sided (1.12 against 1.04 ms per contracted
// step).  qq_half 2 takes it wherever it applies.
constexpr int QQ_HALF_MIN_BLOCKS = 128;
static bool qq_half_on(const SysConst &c) {
  return c.qq_half && is_ttm(c) && !c.cells && c.g_qq > 0 && qq_half_lds(c) <= PAIR_LDS_MAX &&
         (c.qq_half == 2 ||
          (long)c.P * qq_half_blocks(c.N, qq_half_lanes(c)) >= QQ_HALF_MIN_BLOCKS);
}

template <int ML, typename R>
static void launch_qq_half_v(int G, dim3 grid, const BlockMap &m, size_t lds, const Group &g,
                             hipStream_t s) {
  const dim3 B(QQ_HALF_BLOCK);
  if (G == 4) hipLaunchKernelGGL((k_pair_qq_half<ML, 4, R>), grid, B, lds, s, g.d, m);
  else if (G == 8) hipLaunchKernelGGL((k_pair_qq_half<ML, 8, R>), grid, B, lds, s, g.d, m);
  else hipLaunchKernelGGL((k_pair_qq_half<ML, 16, R>), grid, B, lds, s, g.d, m);
}

void launch_pair_qq(const Group &g, hipStream_t s) {
  // variant bit 6: the pass that evaluates each pair once (bits 3-5 its lanes)
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        if (qq_half_on(c)) {
          const int G = qq_half_lanes(c);
          const int v0 = (c.model == 3 ? 1 : 0) | (c.mixed ? 2 : 0);
          const int gi = G == 4 ? 1 : G == 8 ? 2 : 3;
          return Plan{v0 | 64 | (gi << 3), idiv(c.N, QQ_HALF_BLOCK / G), c.P, 1, qq_half_lds(c)};
        }
        return pair_plan(c, c.g_qq, bead_sites_lds(c));
      },
      [&](int v, dim3 grid, const BlockMap &m, size_t lds) {
        if (v & 64) {
          const int gi = (v >> 3) & 7, G = gi == 1 ? 4 : gi == 2 ? 8 : 16;
          switch (v & 3) {
            case 0: launch_qq_half_v<2, double>(G, grid, m, lds, g, s); break;
            case 1: launch_qq_half_v<3, double>(G, grid, m, lds, g, s); break;
            case 2: launch_qq_half_v<2, float>(G, grid, m, lds, g, s); break;
            default: launch_qq_half_v<3, float>(G, grid, m, lds, g, s); break;
          }
          return;
        }
        const int gi = v >> 3;
        switch (v & 7) {
          case 0: launch_qq_v<2, double, false>(gi, grid, m, lds, g, s); break;
          case 1: launch_qq_v<3, double, false>(gi, grid, m, lds, g, s); break;
          case 2: launch_qq_v<2, float, false>(gi, grid, m, lds, g, s); break;
          case 3: launch_qq_v<3, float, false>(gi, grid, m, lds, g, s); break;
          case 4: launch_qq_v<2, double, true>(gi, grid, m, lds, g, s); break;
          case 5: launch_qq_v<3, double, true>(gi, grid, m, lds, g, s); break;
          case 6: launch_qq_v<2, float, true>(gi, grid, m, lds, g, s); break;
          default: launch_qq_v<3, float, true>(gi, grid, m, lds, g, s); break;
        }
      });
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        if (!qq_half_on(c)) return Plan{-1, 0, 0, 0, 0};
        return Plan{c.model == 3 ? 1 : 0, idiv((long)c.P * c.N * 32, 256), 1, 1, 0};
      },
      [&](int v, dim3 grid, const BlockMap &m, size_t) {
        if (v == 1) hipLaunchKernelGGL(k_pair_qq_partner<3>, grid, dim3(256), 0, s, g.d, m);
        else hipLaunchKernelGGL(k_pair_qq_partner<2>, grid, dim3(256), 0, s, g.d, m);
      });
}

void launch_tables(const Group &g, hipStream_t s) {
  launch_over(
      g, [](const SysConst &c, const MDConst &) { return Plan{0, idiv(4 * c.N, 256), c.P, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_tables, grid, dim3(256), 0, s, g.d, m);
      });
}

// Structure factor `mode`.  Variant: bit 0 the column kernel (else one k per
// lane), bits 1-3 its segment-length bucket, bit 4 two lane sets per block,
// bits 5.. the lane count of a set in units of 64.
static const int KZ_BUCKET[6] = {4, 8, 12, 16, 24, 32};

template <int MODE, int KZ, class R>
static void launch_sfac_col_v(bool split, int threads, dim3 grid, const BlockMap &m, size_t lds,
                              const Group &g, hipStream_t s) {
  if (split) hipLaunchKernelGGL((k_sfac_col<MODE, KZ, 2, R>), grid, dim3(2 * threads), lds, s, g.d, m);
  else hipLaunchKernelGGL((k_sfac_col<MODE, KZ, 1, R>), grid, dim3(threads), lds, s, g.d, m);
}

template <int MODE, class R>
static void launch_sfac_kz(int v, dim3 grid, const BlockMap &m, size_t lds, const Group &g,
                           hipStream_t s) {
  const bool split = (v & 16) != 0;
  const int threads = ((v >> 5) & 7) * 64;
  switch ((v >> 1) & 7) {
    case 0: launch_sfac_col_v<MODE, 4, R>(split, threads, grid, m, lds, g, s); break;
    case 1: launch_sfac_col_v<MODE, 8, R>(split, threads, grid, m, lds, g, s); break;
    case 2: launch_sfac_col_v<MODE, 12, R>(split, threads, grid, m, lds, g, s); break;
    case 3: launch_sfac_col_v<MODE, 16, R>(split, threads, grid, m, lds, g, s); break;
    case 4: launch_sfac_col_v<MODE, 24, R>(split, threads, grid, m, lds, g, s); break;
    default: launch_sfac_col_v<MODE, 32, R>(split, threads, grid, m, lds, g, s); break;
  }
}

// refine: the FP64 dipole structure factor of the mixed form's refinement
// step (mixed replicas only); otherwise the mixed form's SOR iterations take
// it in FP32 (variant bit 8).
// part: 1 the chunk partial sums, 2 their fixed-order reduction, 3 both.
template <int MODE>
static void launch_sfac_mode(const Group &g, bool refine, int part, hipStream_t s) {
  constexpr int NV = MODE == SF_V ? 6 : 2;
  auto plan = [&](const SysConst &c) {
        if (MODE != SF_Q && !is_ttm(c)) return Plan{-1, 0, 0, 0, 0};
        if (refine && !c.mixed) return Plan{-1, 0, 0, 0, 0};
        const bool f32 = MODE == SF_D && c.mixed && !refine;
        int nch, sc, scp;
        sfac_chunking(c, MODE, sfac_sites(c, MODE), nch, sc, scp);
        const size_t lds = sfac_lds(c, MODE, sc, scp, f32);
        int kzb = -1;
        for (int i = 0; i < 6; ++i)
          if (kzb < 0 && c.segmax <= KZ_BUCKET[i]) kzb = i;
        const bool col = c.sf_kind == 1 && (MODE != SF_V || c.sfv_col) && lds <= 64 * 1024 && kzb >= 0;
        if (!col) return Plan{0, idiv(c.nk, 256), nch, c.P, sfac_lds(c, MODE, sc, scp)};
        const int threads = std::min(256, idiv(c.nseg, 64) * 64);
        // two lane sets per block where the second set's sums fit beside the tables
        const size_t red = (f32 ? sizeof(float) : sizeof(double)) *
                           ((size_t)NV * KZ_BUCKET[kzb] * threads + 4);
        const bool split = MODE != SF_V && c.sf_split && lds + red <= 96 * 1024;
        return Plan{1 | (kzb << 1) | (split ? 16 : 0) | ((threads / 64) << 5) | (f32 ? 256 : 0),
                    idiv(c.nseg, threads), nch, c.P, split ? lds + red : lds};
  };
  if (part & 1) launch_over(
      g, [&](const SysConst &c, const MDConst &) { return plan(c); },
      [&](int v, dim3 grid, const BlockMap &m, size_t lds) {
        if (!(v & 1)) {
          hipLaunchKernelGGL((k_sfac_k<MODE>), grid, dim3(256), lds, s, g.d, m);
          return;
        }
        if constexpr (MODE == SF_D) {
          if (v & 256) {
            launch_sfac_kz<MODE, float>(v, grid, m, lds, g, s);
            return;
          }
        }
        launch_sfac_kz<MODE, double>(v, grid, m, lds, g, s);
      });
  if (part & 2) launch_over(
      g,
      [&](const SysConst &c, const MDConst &) {
        if (plan(c).variant < 0) return Plan{-1, 0, 0, 0, 0};
        return Plan{0, idiv((long)NV * c.nk, 256), c.P, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_sfac_reduce, grid, dim3(256), 0, s, g.d, m, MODE);
      });
}

void launch_sfac(const Group &g, int mode, hipStream_t s, bool refine, int part) {
  if (mode == SF_Q) launch_sfac_mode<SF_Q>(g, false, part, s);
  else if (mode == SF_D) launch_sfac_mode<SF_D>(g, refine, part, s);
  else launch_sfac_mode<SF_V>(g, false, part, s);
}

// The FP32 dipole field of the mixed form's SOR iterations (variant bits as
// launch_field_mode's).
static void launch_field_f32(int v, bool kl, dim3 grid, const BlockMap &m, size_t lds, int iter,
                             const Group &g, hipStream_t s) {
#define PIMD_FIELD32(ML, CG)                                                                      \
  if (kl) hipLaunchKernelGGL((k_field<FD_D, ML, CG, true, float>), grid, dim3(64 * CG), lds, s, g.d, m, iter); \
  else hipLaunchKernelGGL((k_field<FD_D, ML, CG, false, float>), grid, dim3(64 * CG), lds, s, g.d, m, iter)
  switch (v & 7) {
    case 0: PIMD_FIELD32(2, 4); break;
    case 1: PIMD_FIELD32(3, 4); break;
    case 2: PIMD_FIELD32(2, 8); break;
    case 3: PIMD_FIELD32(3, 8); break;
    case 4: PIMD_FIELD32(2, 16); break;
    default: PIMD_FIELD32(3, 16); break;
  }
#undef PIMD_FIELD32
}

// Reciprocal-space field `mode`.  Variant: bit 0 TTM3-F site sets (the
// fixed-charge models share them), bits 1-2 the column groups (4, 8, 16),
// bit 3 k data staged in LDS.  sel_mixed: only the mixed-form replicas.
template <int MO>
static void launch_field_mode(const Group &g, int iter, bool sel_mixed, hipStream_t s) {
  launch_over(
      g,
      [&](const SysConst &c, const MDConst &) {
        if ((MO != FD_Q && !is_ttm(c)) || (sel_mixed && !c.mixed)) return Plan{-1, 0, 0, 0, 0};
        const int ML = c.model == 2 ? 2 : 3;
        const int M = MO == FD_D ? c.nd * c.N : (ML == 2 ? 4 * c.N : 3 * c.N);
        // as many column groups as the columns keep busy: more resident waves hide
        // the table and structure-factor reads (512 H2O x 32 beads: the charge
        // fields at 16 groups instead of 8 take 1.68 instead of 1.73 ms per force
        // call)
        int cg = c.fd_cg;
        if (cg == 0) cg = c.ncol < 8 ? 4 : c.ncol < 16 ? 8 : 16;
        const int cgi = cg == 4 ? 0 : cg == 8 ? 1 : 2;
        // the mixed form's SOR iterations in FP32 (its refinement, iter < 0, in FP64)
        const bool f32 = MO == FD_D && c.mixed && iter > 0;
        const bool kl = c.fd_klds && field_lds(c, MO, true, f32) <= 160 * 1024;
        return Plan{(ML == 3 ? 1 : 0) | (cgi << 1) | (kl ? 8 : 0) | (f32 ? 16 : 0), idiv(M, FD_SB),
                    c.P, 1, field_lds(c, MO, kl, f32)};
      },
      [&](int v, dim3 grid, const BlockMap &m, size_t lds) {
        const bool kl = (v & 8) != 0;
        if constexpr (MO == FD_D) {
          if (v & 16) {
            launch_field_f32(v, kl, grid, m, lds, iter, g, s);
            return;
          }
        }
#define PIMD_FIELD(ML, CG)                                                                     \
  if (kl) hipLaunchKernelGGL((k_field<MO, ML, CG, true>), grid, dim3(64 * CG), lds, s, g.d, m, iter); \
  else hipLaunchKernelGGL((k_field<MO, ML, CG, false>), grid, dim3(64 * CG), lds, s, g.d, m, iter)
        switch (v & 7) {
          case 0: PIMD_FIELD(2, 4); break;
          case 1: PIMD_FIELD(3, 4); break;
          case 2: PIMD_FIELD(2, 8); break;
          case 3: PIMD_FIELD(3, 8); break;
          case 4: PIMD_FIELD(2, 16); break;
          default: PIMD_FIELD(3, 16); break;
        }
#undef PIMD_FIELD
      });
}

void launch_field(const Group &g, int mode, int iter, hipStream_t s, bool sel_mixed) {
  if (mode == FD_Q) launch_field_mode<FD_Q>(g, iter, sel_mixed, s);
  else if (mode == FD_D) launch_field_mode<FD_D>(g, iter, sel_mixed, s);
  else launch_field_mode<FD_QD>(g, iter, sel_mixed, s);
}

void launch_post_qq(const Group &g, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        const bool ttm = c.model == PIMD_TTM21F || c.model == PIMD_TTM3F;
        return Plan{ttm ? 0 : -1, (int)(((long)c.P * c.N + 63) / 64), 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_post_qq, grid, dim3(64), 0, s, g.d, m);
      });
}

// Variant of the dipole pass: bit 0 TTM3-F, bit 1 FP32, bit 2 LDS-staged
// (with the folded convergence test), bits 3-5 the lanes-per-molecule index,
// bits 6-7 the screening-factor cache (0 none, 1 write, 2 read).
template <int ML, typename R, bool CACHED>
static void launch_dd_lds(int gi, dim3 grid, const BlockMap &m, size_t lds, int test_iter,
                          int store, const Group &g, hipStream_t s) {
  constexpr int B = dd_block<ML>();
  switch (gi) {
    case 1: hipLaunchKernelGGL((k_pair_dd_lds<ML, 8, R, CACHED>), grid, dim3(B), lds, s, g.d, m, test_iter, store); break;
    case 3: hipLaunchKernelGGL((k_pair_dd_lds<ML, 32, R, CACHED>), grid, dim3(B), lds, s, g.d, m, test_iter, store); break;
    case 4: hipLaunchKernelGGL((k_pair_dd_lds<ML, 64, R, CACHED>), grid, dim3(B), lds, s, g.d, m, test_iter, store); break;
    default: hipLaunchKernelGGL((k_pair_dd_lds<ML, 16, R, CACHED>), grid, dim3(B), lds, s, g.d, m, test_iter, store); break;
  }
}

template <int ML, typename R>
static void launch_dd_v(bool lds_kernel, int gi, int cache, dim3 grid, const BlockMap &m,
                        size_t lds, int test_iter, const Group &g, hipStream_t s) {
  if (lds_kernel) {
    if (cache == 2) launch_dd_lds<ML, R, true>(gi, grid, m, lds, test_iter, 0, g, s);
    else launch_dd_lds<ML, R, false>(gi, grid, m, lds, test_iter, cache == 1, g, s);
    return;
  }
  switch (gi) {
    case 3: hipLaunchKernelGGL((k_pair_dd<ML, 32, R>), grid, dim3(256), 0, s, g.d, m); break;
    case 4: hipLaunchKernelGGL((k_pair_dd<ML, 64, R>), grid, dim3(256), 0, s, g.d, m); break;
    default: hipLaunchKernelGGL((k_pair_dd<ML, 16, R>), grid, dim3(256), 0, s, g.d, m); break;
  }
}

// fp64 forces the FP64 form (the refinement pass of the mixed solve);
// sel_mixed: only the mixed-form replicas.
void launch_pair_dd(const Group &g, bool fp64, hipStream_t s, int test_iter, bool sel_mixed,
                    int sor_iter) {
  launch_over(
      g,
      [&](const SysConst &c, const MDConst &) {
        if (!is_ttm(c) || (sel_mixed && !c.mixed)) return Plan{-1, 0, 0, 0, 0};
        const int v0 = (c.model == 3 ? 1 : 0) | (c.mixed && !fp64 ? 2 : 0);
        if (pair_dd_tests(c)) {
          const int G = c.g_dd == 64 ? 64 : c.g_dd == 32 ? 32 : c.g_dd == 8 ? 8 : 16;
          const int B = c.model == 2 ? dd_block<2>() : dd_block<3>();
          // the cache holds the factors of the SOR iterations' arithmetic
          // type; the FP64 refinement of the mixed form computes its own
          const int cache = !c.dd_cache || sor_iter <= 0 || (c.mixed && fp64) ? 0
                            : sor_iter == 1                                 ? 1
                                                                            : 2;
          return Plan{v0 | 4 | (gidx(G) << 3) | (cache << 6), idiv(c.N, B / G), c.P, 1, dd_lds(c)};
        }
        const int G = c.g_nolds == 64 ? 64 : c.g_nolds == 32 ? 32 : 16;
        return Plan{v0 | (gidx(G) << 3), idiv((long)c.P * c.N * G, 256), 1, 1, 0};
      },
      [&](int v, dim3 grid, const BlockMap &m, size_t lds) {
        const bool lk = (v & 4) != 0;
        const int gi = (v >> 3) & 7, cache = v >> 6;
        switch (v & 3) {
          case 0: launch_dd_v<2, double>(lk, gi, cache, grid, m, lds, test_iter, g, s); break;
          case 1: launch_dd_v<3, double>(lk, gi, cache, grid, m, lds, test_iter, g, s); break;
          case 2: launch_dd_v<2, float>(lk, gi, cache, grid, m, lds, test_iter, g, s); break;
          default: launch_dd_v<3, float>(lk, gi, cache, grid, m, lds, test_iter, g, s); break;
        }
      });
}

void launch_refine_prep(const Group &g, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        const long n = std::max((long)c.P * c.N * c.nd * 3, (long)c.P);
        return Plan{is_ttm(c) && c.mixed ? 0 : -1, idiv(n, 256), 1, 1, 0};
      },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_refine_prep, grid, dim3(256), 0, s, g.d, m);
      });
}

void launch_converge(const Group &g, int iter, hipStream_t s) {
  launch_over(
      g, [](const SysConst &c, const MDConst &) { return Plan{is_ttm(c) ? 0 : -1, c.P, 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_converge, grid, dim3(256), 0, s, g.d, m, iter);
      });
}

template <int ML, typename R, bool LDS>
static void launch_fin_v(int gi, dim3 grid, const BlockMap &m, size_t lds, const Group &g, hipStream_t s) {
  switch (gi) {
    case 0: if (LDS) hipLaunchKernelGGL((k_pair_final<ML, 4, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    case 1: if (LDS) hipLaunchKernelGGL((k_pair_final<ML, 8, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    case 2: hipLaunchKernelGGL((k_pair_final<ML, 16, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    case 3: hipLaunchKernelGGL((k_pair_final<ML, 32, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
    default: if (!LDS) hipLaunchKernelGGL((k_pair_final<ML, 64, R, LDS>), grid, dim3(256), lds, s, g.d, m); break;
  }
}

// LDS of k_pair_final_half: sites, dipoles and partner sums of its rows
static size_t fin_half_lds(const SysConst &c) {
  const size_t W = qq_half_rows(c.N, half_lanes(c.g_fin));
  return W * (sizeof(Site) * 4 + sizeof(double) * c.nd * 3 +
              sizeof(long long) * FIN_NV * (c.mixed ? 1 : 2));
}
static bool fin_half_on(const SysConst &c) {
  return c.qq_half && c.fin_half && is_ttm(c) && !c.cells &&
         c.g_fin > 0 &&
         fin_half_lds(c) <= PAIR_LDS_MAX &&
         (c.qq_half == 2 ||
          (long)c.P * qq_half_blocks(c.N, half_lanes(c.g_fin)) >= QQ_HALF_MIN_BLOCKS);
}

template <int ML, typename R>
static void launch_fin_half_v(int G, dim3 grid, const BlockMap &m, size_t lds, const Group &g,
                              hipStream_t s) {
  const dim3 B(QQ_HALF_BLOCK);
  if (G == 4) hipLaunchKernelGGL((k_pair_final_half<ML, 4, R>), grid, B, lds, s, g.d, m);
  else if (G == 8) hipLaunchKernelGGL((k_pair_final_half<ML, 8, R>), grid, B, lds, s, g.d, m);
  else hipLaunchKernelGGL((k_pair_final_half<ML, 16, R>), grid, B, lds, s, g.d, m);
}

void launch_pair_final(const Group &g, hipStream_t s) {
  // variant bit 6: the pass that evaluates each pair once (bits 3-5 its lanes)
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        if (fin_half_on(c)) {
          const int G = half_lanes(c.g_fin);
          const int gi = G == 4 ? 1 : G == 8 ? 2 : 3;
          return Plan{(c.model == 3 ? 1 : 0) | (c.mixed ? 2 : 0) | 64 | (gi << 3),
                      idiv(c.N, QQ_HALF_BLOCK / G), c.P, 1, fin_half_lds(c)};
        }
        return pair_plan(c, c.g_fin, bead_sites_lds(c) + sizeof(double) * c.N * c.nd * 3);
      },
      [&](int v, dim3 grid, const BlockMap &m, size_t lds) {
        if (v & 64) {
          const int gi = (v >> 3) & 7, G = gi == 1 ? 4 : gi == 2 ? 8 : 16;
          switch (v & 3) {
            case 0: launch_fin_half_v<2, double>(G, grid, m, lds, g, s); break;
            case 1: launch_fin_half_v<3, double>(G, grid, m, lds, g, s); break;
            case 2: launch_fin_half_v<2, float>(G, grid, m, lds, g, s); break;
            default: launch_fin_half_v<3, float>(G, grid, m, lds, g, s); break;
          }
          return;
        }
        const int gi = v >> 3;
        switch (v & 7) {
          case 0: launch_fin_v<2, double, false>(gi, grid, m, lds, g, s); break;
          case 1: launch_fin_v<3, double, false>(gi, grid, m, lds, g, s); break;
          case 2: launch_fin_v<2, float, false>(gi, grid, m, lds, g, s); break;
          case 3: launch_fin_v<3, float, false>(gi, grid, m, lds, g, s); break;
          case 4: launch_fin_v<2, double, true>(gi, grid, m, lds, g, s); break;
          case 5: launch_fin_v<3, double, true>(gi, grid, m, lds, g, s); break;
          case 6: launch_fin_v<2, float, true>(gi, grid, m, lds, g, s); break;
          default: launch_fin_v<3, float, true>(gi, grid, m, lds, g, s); break;
        }
      });
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        if (!fin_half_on(c)) return Plan{-1, 0, 0, 0, 0};
        return Plan{c.model == 3 ? 1 : 0, idiv((long)c.P * c.N * 32, 256), 1, 1, 0};
      },
      [&](int v, dim3 grid, const BlockMap &m, size_t) {
        if (v == 1) hipLaunchKernelGGL(k_pair_final_partner<3>, grid, dim3(256), 0, s, g.d, m);
        else hipLaunchKernelGGL(k_pair_final_partner<2>, grid, dim3(256), 0, s, g.d, m);
      });
}

void launch_virk(const Group &g, hipStream_t s) {
  launch_over(
      g, [](const SysConst &c, const MDConst &) { return Plan{0, idiv(c.nk, 256), c.P, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_virk, grid, dim3(256), 0, s, g.d, m);
      });
}

void launch_finalize(const Group &g, hipStream_t s) {
  launch_over(
      g,
      [](const SysConst &c, const MDConst &) {
        return Plan{is_ttm(c) ? c.model : -1, idiv((long)c.P * c.N, 64), 1, 1, 0};
      },
      [&](int v, dim3 grid, const BlockMap &m, size_t) {
        if (v == 2) hipLaunchKernelGGL(k_finalize<2>, grid, dim3(64), 0, s, g.d, m);
        else hipLaunchKernelGGL(k_finalize<3>, grid, dim3(64), 0, s, g.d, m);
      });
}

void launch_reduce_bead(const Group &g, hipStream_t s) {
  launch_over(
      g, [](const SysConst &c, const MDConst &) { return Plan{0, c.P, 1, 1, 0}; },
      [&](int, dim3 grid, const BlockMap &m, size_t) {
        hipLaunchKernelGGL(k_reduce_bead, grid, dim3(256), 0, s, g.d, m);
      });
}

}  // namespace pimd
