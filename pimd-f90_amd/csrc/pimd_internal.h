// Internal declarations shared by the HIP kernels and the host-side C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pimd_ttm.h"

namespace pimd {

// consts.f90:22-42
constexpr double CHARGECON = 18.22234397655801030455;
constexpr double DEBYE = 4.8033324;
constexpr double MASSCON = 418.4;
constexpr double KB_AMU = 0.831447148;   // KB_amuA2ps2perK
constexpr double HBAR = 6.35077993041;   // amu*A^2/ps
constexpr double PI = 3.14159265358979323844;
constexpr double TWOPI = 6.28318530717958647688;
constexpr double FOURPI = 12.56637061435917295376;
constexpr double TSP = 1.12837916709551257390;   // 2/sqrt(pi)
constexpr double SQPI = 1.77245385090551602729;
constexpr double TEMPFACTOR = 0.80181484527;   // amu A^2/ps^2 -> K (consts.f90:31)
constexpr double PRESSCON2 = 69477.0;          // kcal/mol/A^3 -> bar (consts.f90:26)
constexpr double KB_KCAL = 0.0019872041;       // kcal/mol/K (consts.f90:32)
constexpr double MASSCONI = 1.0 / 418.4;       // consts.f90:25

// Fitted Partridge-Schwenke PES parameters of a context: the 245 combined
// coefficients c5z (nasa.f90:45), then these five (nasa_mod.f90 values).
constexpr int PES_NC = 245;
constexpr int PES_DEOHA = 245, PES_ALPHAOH = 246, PES_ROH = 247, PES_PHH1A = 248, PES_PHH2 = 249;
constexpr int PES_NPARAM = 250;
// The 84 coefficients of the dipole moment surface (coefD, nasa_mod.f90)
// follow the PES parameters in a context's parameter array.
constexpr int DMS_NCOEF = 84;
constexpr int PES_DMSC = PES_NPARAM;
constexpr int PES_NALL = PES_NPARAM + DMS_NCOEF;

// One site of a molecule record: position and charge (charge in
// CHARGECON-scaled units, as the reference's charge array).
struct __align__(32) Site {
  double x, y, z, q;
};

// Everything the kernels need about the model and the box, passed by value.
struct SysConst {
  int N, P, model;        // waters, beads, pot_model (2, 3 TTM; 4, 5 fixed-charge SPC)
  int nd, tdip;           // dipole sites per molecule and the first dipole type
  int nk, ncol, T;        // k-vectors, kz-columns, phase-table rows per site
  int hmax, kmax, lmax;
  int maxiter, guess;
  int mixed;               // PIMD_MIXED: FP32 real-space pair arithmetic
  int contraction;         // no monomer PES in the force call (pot_ttm.f90:92-132)
  int nbr_cap;
  int cells;               // 1: neighbour search through a cell list
  int ncell[3];            // cells per axis (>= 3 each when cells = 1), width >= Rc
  int kzmax;               // longest kz-column
  int nseg, segmax;        // kz-columns cut into segments of at most segmax k
  int sf_chunks_max;       // site chunks of the column structure-factor kernel
  int sf_sites;            // fewest sites per structure-factor chunk
  int fd_cg;               // column groups per field block (0: by occupancy)
  int fd_klds;             // 1: field kernels stage the bead's k data in LDS
  int fd_tab_d2;           // double2 slots of the field kernel's table / reduction area
  int sf_kind;             // structure-factor kernel: 0 one k per lane, 1 kz-columns
  int sfv_col;             // 1: the 6-component SF_V by kz-columns too
  int sf_split;            // 1: column structure factors with two lane sets per block
  int g_dd;                // lanes per molecule of the LDS-staged dipole kernel (0: gather kernel)
  int g_qq, g_fin;         // the same for the charge and final pair kernels
  int g_nolds;             // lanes per molecule of the pair kernels that do not stage in LDS
  int poll;                // 1: read the active-bead count back after every SOR iteration
  int overlap;             // 1: real-space dipole field beside the structure factor (2 streams)
  int dd_cache;            // 1: the SOR dipole pass keeps each pair's screening factors
  int qq_half;             // once-per-pair charge pass: 0 off, 1 where it fills the device, 2 always
  int fin_half;            // the same for the final dipole pass (with qq_half): 0 off, 1 on
  double box[3], boxi[3], vol;
  double rc2, rc1;
  double vdwA, vdwB, vdwC, vdwD, vdwE, shiftA, shiftB;
  double aewald, a2, cts1, cts2, cts3, self_fac, fac_dd;
  double gammaM;
  double polar[4];
  double apolCC[4][4], raCC[4][4], apolDD[4][4], apol3DD[4][4];
  int same_a;             // |aCCaCD - aDD| < 1e-12 (pot_ttm.f90:443)
  double dms1, dms2, dms3;
  double sor, stath, peps;
  double Uvdw_lrc;
  double p4v;             // FOURPI / volume
  // q-SPC/Fw and SPC/F (pot_mod.f90:31-44): charges qO, qH (times CHARGECON),
  // equilibrium geometry and force constants
  double spc_q[2], spc_roh0, spc_th0, spc_rhh0, spc_kb, spc_ka, spc_A, spc_B, spc_C, spc_D;
};

// Tally slots per (bead, molecule): scalar partial sums reduced per bead.
enum Tally { T_UMON = 0, T_UVDW, T_UELEC, T_UIND, T_VIRC, T_VIRT /* 9 slots, [col][row] */,
             T_COUNT = T_VIRT + 9 };

// Buffer addresses as the kernels see them: global-memory pointers, so that
// their loads are global ones (scalar ones where the address is the same on
// every lane) rather than flat.  The host sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using dptr = __attribute__((address_space(1))) T *;
#else
template <class T>
using dptr = T *;
#endif

// Device buffers of one context.
struct DevBufs {
  dptr<double> RR = nullptr, RRc = nullptr;  // staging of host inputs
  dptr<Site> site = nullptr;  // [P][N][4]
  dptr<double> Rcs = nullptr;  // [N][4][3] centroid sites
  dptr<double> grdq = nullptr;  // [P][N][27]
  dptr<double> phi = nullptr;  // [P][N][4]
  dptr<double> Efq = nullptr;  // [P][N][4][3]
  dptr<double> dR = nullptr;  // [P][N][4][3]
  dptr<double> dip = nullptr;  // [P][N][nd][3]
  dptr<double> dip_prev = nullptr;  // [P][N][nd][3] input of the last SOR step
  dptr<double> Efd = nullptr;  // [P][N][nd][3]
  dptr<double> diff2 = nullptr;  // [P][N*nd]
  dptr<double> tally = nullptr;  // [P][N][T_COUNT]
  dptr<double> hist = nullptr;  // [P][4][N*nd][3]
  dptr<int> nbr = nullptr, nbrc = nullptr;  // [P][N][cap], [P][N]
  dptr<int> nbrn = nullptr;  // [P][N] near partners (listed first)
  // [P][N][4]: per part of the list (near, far) the entries with a partner
  // below the molecule and those up to the end of its half window (the
  // partners whose pair it owns, see pair_half_window)
  dptr<int> nbrh = nullptr;
  // the charge pair pass's partner-side sums, one slab per block of a bead:
  // [P][blocks][rows][QQ_NV]: a block's partner sums, rounded once to FP64
  dptr<double> qqslab = nullptr;
  dptr<int> cell_start = nullptr, cell_mol = nullptr;  // [P][ncells + 1], [P][N] cell list
  dptr<double2> tab = nullptr;  // [P][T][4N]
  dptr<double2> tabd = nullptr;  // [P][T][nd N] the dipole sites' rows
  dptr<float2> tabdf = nullptr;  // the same in FP32 (mixed form)
  // screening factors (ts1, ts2) of every dipole-site pair of the neighbour
  // lists, [P][N][cap][nd*nd], written by the first SOR iteration and read by
  // the others: FP64 (ddts) or, in the mixed form, FP32 (ddtsf, same memory)
  dptr<double2> ddts = nullptr;
  dptr<float2> ddtsf = nullptr;
  dptr<double> Sq = nullptr, Sm = nullptr;  // [P][nk][2]
  dptr<double> Sv = nullptr;  // [P][nk][6]
  dptr<double> virk = nullptr;  // [P][nkb][9]
  dptr<double> part = nullptr;  // [P][sf_chunks_max][nk][2]
  dptr<int> active = nullptr, iters = nullptr;  // [P]
  dptr<int> nactive = nullptr;  // [1]
  dptr<double4> kv = nullptr;  // [nk] kx, ky, kz, Ak
  dptr<int4> khkl = nullptr;  // [nk] h, k, l, 0
  dptr<int4> cols = nullptr;  // [ncol] h, k, l_first, count
  dptr<int> colk = nullptr;  // [ncol] first k index
  dptr<int4> segs = nullptr;  // [nseg] h, k, l_first, count
  dptr<int> segk = nullptr;  // [nseg] first k index
  // outputs of the last force call
  dptr<double> Upot = nullptr, virt = nullptr, virialc = nullptr, terms = nullptr;  // [P],[P][9],[P],[P][5]
  dptr<double> dRR = nullptr, dipI = nullptr, dipE = nullptr, chg = nullptr;
  dptr<int> iters_out = nullptr;
  // ring polymer
  dptr<double> C = nullptr, A = nullptr;  // [P][P], [4][P] (A11,A12,A21,A22)
  dptr<double> msf = nullptr, c1 = nullptr, c2 = nullptr;
  dptr<double> masses = nullptr;  // [Natoms]
  dptr<double> PP = nullptr, noise = nullptr;  // [P][Natoms][3], [2][Natoms][P][3]
  dptr<double> uk = nullptr;  // [1] kinetic energy
  dptr<double> ukpart = nullptr;  // [256] kinetic partial sums
  dptr<double> radius = nullptr;  // [P][N] site radius about O
  dptr<double> pes = nullptr;  // [PES_NALL] PES parameters, then coefD
  dptr<double> adj = nullptr;  // adjoint inputs and partial sums
  dptr<double> dRRfast = nullptr, mol = nullptr;  // contracted MD: [P][Natoms][3], [P][N][3]
  dptr<int> clk = nullptr;  // device step clock {t, record slot}
  dptr<double> obs_part = nullptr;  // [64][5] partial sums of the observables
  dptr<int> flags = nullptr;  // [1] bit 0: non-finite coordinates seen
  dptr<double> vxi_beads = nullptr;  // [Natoms][P][3][2] bead Nose-Hoover chains
  dptr<double> nh_vxi = nullptr, nh_s = nullptr;  // global chain velocities [2], its scale [1]
  // contracted MD: the outputs of the centroid's force call (one bead)
  dptr<const double> cen_Upot = nullptr, cen_virt = nullptr, cen_virialc = nullptr;
  dptr<const double> cen_dRR = nullptr;
  dptr<double> nh_tau = nullptr;  // [P] their time constants per mode
  dptr<int> clk_own = nullptr;  // this context's clock (clk may be its owner's)
  dptr<double> tan = nullptr;  // TTM parameter tangents: inputs, tables, workspace
  size_t tan_bytes = 0;
  dptr<double> red = nullptr;  // scratch for reductions
  // reciprocal-space field, potential (FD_Q) and gradient, potential (FD_QD)
  // of the last field pass, [P][N][4][3] and [P][N][4]; their consumers add
  // them to Efq / dR and phi, so the real-space pair pass may run beside
  dptr<double> kE = nullptr, kphi = nullptr;
};

// ---------------------------------------------------------------------------
// Replica dimension.  Every kernel of an MD step takes a table of replicas
// (independent systems of any size, model and bead count) and covers several
// of them in one grid: the blocks of replica i of a launch are off[i] ..
// off[i+1]-1, numbered x-fastest over that replica's own nx x ny x nz grid
// (the grid it would get alone), so each replica's arithmetic and summation
// order do not depend on its companions.  One system alone is a table of one.
// ---------------------------------------------------------------------------
constexpr int MAXREP = 64;   // replicas per launch; more are cut into several launches

// Ring-polymer and thermostat constants of a replica's MD step.
struct MDConst {
  double delt, delt2;         // step and half step (ps)
  double deltf, delt2f;       // contracted MD: inner step and its half
  double tempP;               // bead temperature P T (K)
  double massO, massH;
  double omegan, tau;         // ring frequency kB T P / hbar; global chain tau (ps)
  double box0, boxi0;         // box(1), 1 / box(1) of the centroid wrap (main_stuff.f90:115)
  uint64_t seed;              // device normals of the Langevin thermostat
  int bead_tt, centroid, global_nh, cmd;
  double *rec;                // per-step records [steps][MD_REC]
};

struct Rep {
  SysConst c;
  DevBufs b;
  MDConst m;
};

struct BlockMap {
  int n;
  int xcd;  // 1: XCD-contiguous block numbering (xcd_block)
  int off[MAXREP + 1];
  int nx[MAXREP], ny[MAXREP];
  int rep[MAXREP];            // entry of the Rep table
};

struct Blk {
  int r, x, y, z;             // Rep table entry; block coordinates in its own grid
  int nx, ny;                 // its own grid's x and y extents
};

// The hardware hands workgroup w of a launch to XCD w % 8.  Within each
// replica's range of the grid the block index is renumbered so that
// consecutive blocks of the replica (the blocks of one bead, which stage the
// same bead records, tables and k data) run on one XCD and share its L2: of
// the replica's n blocks, those on the x-th XCD class take the x-th
// contiguous eighth.  A replica's blocks still spread over all eight XCDs,
// so replicas of different cost do not pile onto one of them.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_block(int p, int n) {
  const int q = n / NXCD, r = n % NXCD, x = p % NXCD;
  return x * q + min(x, r) + p / NXCD;
}

__device__ __forceinline__ Blk block_of(const BlockMap &m) {
  const int g = blockIdx.x;
  int lo = 0, hi = m.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const int n = m.off[lo + 1] - m.off[lo];
  const int l = m.xcd ? xcd_block(g - m.off[lo], n) : g - m.off[lo];
  const int nx = m.nx[lo], ny = m.ny[lo];
  return Blk{m.rep[lo], l % nx, (l / nx) % ny, l / (nx * ny), nx, ny};
}

// Host side: the replicas of one launch sequence.  hc[i] mirrors the SysConst
// of Rep table entry i (launch decisions and grids are made from it).
struct Group {
  std::vector<SysConst> hc;
  std::vector<MDConst> hm;
  Rep *d = nullptr;
  int n() const { return (int)hc.size(); }
};

// A replica's part of a launch: the kernel variant it takes (< 0: none), its
// own grid and its dynamic LDS.
struct Plan {
  int variant;
  int nx, ny, nz;
  size_t lds;
};

// Launches one grid per variant (and per MAXREP replicas of it), replicas in
// table order.  plan(const SysConst &, const MDConst &) -> Plan;
// launch(int variant, dim3 grid, const BlockMap &, size_t lds).
// Whether launches number their blocks XCD-contiguously (PIMD_XCD_MAP, default 1).
int xcd_numbering();

template <class PlanFn, class LaunchFn>
void launch_over(const Group &g, PlanFn plan, LaunchFn launch) {
  const int n = g.n();
  std::vector<Plan> pl(n);
  for (int i = 0; i < n; ++i) pl[i] = plan(g.hc[i], g.hm[i]);
  std::vector<char> done(n, 0);
  for (int i = 0; i < n; ++i) {
    if (done[i]) continue;
    if (pl[i].variant < 0 || (long)pl[i].nx * pl[i].ny * pl[i].nz == 0) {
      done[i] = 1;
      continue;
    }
    BlockMap m{};
    m.xcd = xcd_numbering();
    size_t lds = 0;
    long total = 0;
    for (int j = i; j < n && m.n < MAXREP; ++j) {
      if (done[j] || pl[j].variant != pl[i].variant) continue;
      const long nb = (long)pl[j].nx * pl[j].ny * pl[j].nz;
      if (nb == 0) {
        done[j] = 1;
        continue;
      }
      m.off[m.n] = (int)total;
      m.nx[m.n] = pl[j].nx;
      m.ny[m.n] = pl[j].ny;
      m.rep[m.n] = j;
      ++m.n;
      total += nb;
      lds = std::max(lds, pl[j].lds);
      done[j] = 1;
    }
    m.off[m.n] = (int)total;
    launch(pl[i].variant, dim3((unsigned)total), m, lds);
  }
}

// Host-side launchers of the force path (ttm_kernels.hip, spc_kernels.hip).
// Each covers every replica of the group it applies to; a force call reads
// each replica's coordinates RR, RRc and writes its outputs (Upot, dRR,
// virt, virialc, dipI, dipE, chg, iters_out) in its own buffers.
void launch_monomer(const Group &g, hipStream_t s);
void launch_neighbors(const Group &g, hipStream_t s);
void launch_pair_qq(const Group &g, hipStream_t s);
void launch_tables(const Group &g, hipStream_t s);
// refine: SF_D in FP64 for the mixed replicas' refinement step.  part: 1 the
// per-chunk partial sums, 2 their reduction into the structure factor, 3 both.
void launch_sfac(const Group &g, int mode, hipStream_t s, bool refine = false, int part = 3);
int sfac_chunks_used(const SysConst &c);
void launch_converge(const Group &g, int iter, hipStream_t s);
// iter: the SOR iteration of FD_D, -1 for the FP64 refinement (maxiter + 1);
// sel_mixed: only the replicas of the mixed form
void launch_field(const Group &g, int mode, int iter, hipStream_t s, bool sel_mixed = false);
void launch_post_qq(const Group &g, hipStream_t s);
// test_iter > 0: the kernel first runs the convergence test of SOR iteration
// test_iter (in place of launch_converge); only where pair_dd_tests(c) holds
// for every replica.  sor_iter: the SOR iteration (1 writes the pair
// screening factors of replicas with dd_cache, later ones read them), 0 for
// the FP64 refinement, which computes its own.
void launch_pair_dd(const Group &g, bool fp64, hipStream_t s, int test_iter = 0,
                    bool sel_mixed = false, int sor_iter = 0);
bool pair_dd_tests(const SysConst &c);
// (ts1, ts2) pairs of the SOR dipole pass's screening-factor cache
inline size_t dd_cache_elems(const SysConst &c) {
  return (size_t)c.P * c.N * (size_t)c.nbr_cap * c.nd * c.nd;
}
// The charge pair pass that evaluates each molecule pair once: blocks of
// QQ_HALF_BLOCK lanes, at most 16 lanes per molecule, and per molecule QQ_NV
// partner-side values (potentials of the three charged sites, fields of up to
// four dipole-carrying sites, the oxygen's vdW force)
constexpr int QQ_HALF_BLOCK = 512;
constexpr int QQ_NV = 18;
// lanes per molecule of the once-per-pair passes for a lane setting g
__host__ __device__ inline int half_lanes(int g) { return g <= 4 ? 4 : g <= 8 ? 8 : 16; }
__host__ __device__ inline int qq_half_lanes(const SysConst &c) { return half_lanes(c.g_qq); }
// Molecules one block touches with G lanes per molecule: its owners
// [i0, i0 + QQ_HALF_BLOCK / G) and their partners, which lie at most N/2
// after their owner; rows are counted from i0 modulo N.
__host__ __device__ inline int qq_half_rows(int N, int G) {
  const int w = QQ_HALF_BLOCK / G + N / 2;
  return w < N ? w : N;
}
__host__ __device__ inline int qq_half_blocks(int N, int G) {
  return (N + QQ_HALF_BLOCK / G - 1) / (QQ_HALF_BLOCK / G);
}
// [P][blocks][rows][QQ_NV]: the largest over the lane counts
inline size_t qq_slab_elems(const SysConst &c) {
  size_t m = 0;
  for (int G = 4; G <= 16; G *= 2) {
    const size_t e = (size_t)qq_half_blocks(c.N, G) * qq_half_rows(c.N, G);
    m = e > m ? e : m;
  }
  return (size_t)c.P * m * QQ_NV;
}
void launch_refine_prep(const Group &g, hipStream_t s);
void launch_pair_final(const Group &g, hipStream_t s);
void launch_virk(const Group &g, hipStream_t s);
void launch_finalize(const Group &g, hipStream_t s);
void launch_reduce_bead(const Group &g, hipStream_t s);
void launch_spc_monomer(const Group &g, hipStream_t s);
void launch_spc_pair(const Group &g, hipStream_t s);
void launch_spc_finalize(const Group &g, hipStream_t s);
// contracted MD (MD.f90:276-449): monomer forces of every bead into the fast
// gradient (with record, also the energies and virials of the step record),
// the centroid's gradient to every bead, and the step record
void launch_pes_fast(const Group &g, bool record, hipStream_t s);
void launch_monomer_dipoles(const SysConst &c, const double *pes, const double *RR, double *out,
                            hipStream_t s);
// every replica's monomer dipoles from its RR into its dipI
void launch_monomer_dipoles(const Group &g, hipStream_t s);
void launch_broadcast_beads(const Group &g, hipStream_t s);
void launch_contract_record(const Group &g, hipStream_t s);

// theta[PES_NALL]: the reference's PES parameters and DMS coefficients
void pes_default_params(double *theta);
bool launch_pes_adjoint(const SysConst &c, DevBufs &b, const double *RR, const double *ew,
                        const double *gw, double *partial, int nblocks, double *grad,
                        hipStream_t s);
int pes_adjoint_blocks(const SysConst &c);

enum SfacMode { SF_Q = 0, SF_D = 1, SF_V = 2 };
enum FieldMode { FD_Q = 0, FD_D = 1, FD_QD = 2 };

// TTM model parameters for fitting, in the order of pimd_ttm_params_get.
constexpr int TTM_NPARAM = 16;
// parameters of the adjoint: the TTM parameters, the DMS coefficients, gammaM
constexpr int TTM_ADJ_NPARAM = TTM_NPARAM + DMS_NCOEF + 1;
enum TtmParam { TP_VDWA = 0, TP_VDWB, TP_VDWC, TP_VDWD, TP_VDWE, TP_POLARO, TP_POLARH, TP_POLARM,
                TP_POLFACO, TP_POLFACH, TP_POLFACM, TP_ADD, TP_ACC, TP_DMS1, TP_DMS2, TP_DMS3 };

// Buffers and parameters of one tangent launch (ttm_tangent.hip).
struct TanLaunch {
  const double *RR, *pes;
  const double4 *kv;
  const int4 *khkl;
  double *cfg;
  void *ws;
  const double *ew, *gw;
  double *out, *dG;
  double theta[TTM_NPARAM];
  double lrc_coef[3];
  int dms_coef;  // 0: the TTM parameters (ndir = TTM_NPARAM), 1: the DMS coefficients
  int gamma_m;   // 1: the one direction is the M-site fraction gammaM (moves the M sites)
  int adjoint;   // 1: the adjoint of all TTM_ADJ_NPARAM parameters (one sweep along gw)
  double *adj;   // adjoint output [P][TTM_ADJ_NPARAM][2]
  int ndir;
};
size_t tangent_cfg_doubles(const SysConst &c);
size_t tangent_ws_duals(const SysConst &c);
void launch_ttm_tangent(const SysConst &c, const TanLaunch &L, hipStream_t s);

// Ring-polymer launchers (rp_kernels.hip), over every replica of the group.
constexpr int MD_REC = 13;   // per-step record: Upot, virial, virialc, uk, then 9 observables
// Kicks folded into a ring-polymer kernel: PP - MASSCON*src*dt2 of the
// replica's gradient (KICK_DRR: dRR and the half step; KICK_FAST: the fast
// gradient of contracted MD and the inner half step).
enum KickSrc { KICK_NONE = 0, KICK_DRR = 1, KICK_FAST = 2 };
// with_p1: one-bead replicas too (in MD steps their free flight is a kick and
// k_free_flight, MD.f90:60-68)
void launch_nm_evolve(const Group &g, hipStream_t s, int kick = KICK_NONE, int kick2 = KICK_NONE,
                      bool with_p1 = false);
// Langevin_NM of the replicas with the Langevin bead thermostat (all of them
// with all = true).  host_noise: the replica's noise buffer holds the normals;
// else drawn from (seed, counter), with use_clk the replica's seed and its
// step clock.
void launch_langevin(const Group &g, hipStream_t s, bool all, int host_noise, int use_clk,
                     uint64_t seed, uint64_t counter, int half, int kick);
void launch_bead_nh(const Group &g, int kick, hipStream_t s);
void launch_global_nh(const Group &g, hipStream_t s);
// Replica selections of launch_kick: all, the one-bead ones (whose free flight
// replaces the ring-polymer step), those without a bead thermostat (whose
// second half kick no thermostat kernel carries).
enum KickSel { SEL_ALL = 0, SEL_P1 = 1, SEL_NO_BEAD_TT = 2 };
void launch_kick(const Group &g, int kick, int sel, hipStream_t s);
void launch_free_flight(const Group &g, bool fast, hipStream_t s);
void launch_centroid_pbc(const Group &g, hipStream_t s);
void launch_kinetic(const Group &g, hipStream_t s);
// with_forces: form the step record's Upot, virial, virialc from the force
// call's per-bead outputs first
void launch_observables(const Group &g, bool has_dip, bool with_forces, hipStream_t s);
void launch_set_clock(const Group &g, int t, int slot, hipStream_t s);
void launch_advance_clock(const Group &g, bool with_uk, hipStream_t s);

}  // namespace pimd
