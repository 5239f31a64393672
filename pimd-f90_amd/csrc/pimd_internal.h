// Internal declarations shared by the HIP kernels and the host-side C ABI.
#pragma once
#include <hip/hip_runtime.h>

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
  int poll;                // 1: read the active-bead count back after every SOR iteration
  int overlap;             // 1: real-space dipole field beside the structure factor (2 streams)
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

// Device buffers of one context.
struct DevBufs {
  double *RR = nullptr, *RRc = nullptr;         // staging of host inputs
  Site *site = nullptr;                         // [P][N][4]
  double *Rcs = nullptr;                        // [N][4][3] centroid sites
  double *grdq = nullptr;                       // [P][N][27]
  double *phi = nullptr;                        // [P][N][4]
  double *Efq = nullptr;                        // [P][N][4][3]
  double *dR = nullptr;                         // [P][N][4][3]
  double *dip = nullptr;                        // [P][N][nd][3]
  double *dip_prev = nullptr;                   // [P][N][nd][3] input of the last SOR step
  double *Efd = nullptr;                        // [P][N][nd][3]
  double *diff2 = nullptr;                      // [P][N*nd]
  double *tally = nullptr;                      // [P][N][T_COUNT]
  double *hist = nullptr;                       // [P][4][N*nd][3]
  int *nbr = nullptr, *nbrc = nullptr;          // [P][N][cap], [P][N]
  int *nbrn = nullptr;                          // [P][N] near partners (listed first)
  int *cell_start = nullptr, *cell_mol = nullptr;  // [P][ncells + 1], [P][N] cell list
  double2 *tab = nullptr;                       // [P][T][4N]
  double *Sq = nullptr, *Sm = nullptr;          // [P][nk][2]
  double *Sv = nullptr;                         // [P][nk][6]
  double *virk = nullptr;                       // [P][nkb][9]
  double *part = nullptr;                       // [P][sf_chunks_max][nk][2]
  int *active = nullptr, *iters = nullptr;      // [P]
  int *nactive = nullptr;                       // [1]
  double4 *kv = nullptr;                        // [nk] kx, ky, kz, Ak
  int4 *khkl = nullptr;                         // [nk] h, k, l, 0
  int4 *cols = nullptr;                         // [ncol] h, k, l_first, count
  int *colk = nullptr;                          // [ncol] first k index
  int4 *segs = nullptr;                         // [nseg] h, k, l_first, count
  int *segk = nullptr;                          // [nseg] first k index
  // outputs of the last force call
  double *Upot = nullptr, *virt = nullptr, *virialc = nullptr, *terms = nullptr;  // [P],[P][9],[P],[P][5]
  double *dRR = nullptr, *dipI = nullptr, *dipE = nullptr, *chg = nullptr;
  int *iters_out = nullptr;
  // ring polymer
  double *C = nullptr, *A = nullptr;            // [P][P], [4][P] (A11,A12,A21,A22)
  double *msf = nullptr, *c1 = nullptr, *c2 = nullptr;
  double *masses = nullptr;                     // [Natoms]
  double *PP = nullptr, *noise = nullptr;       // [P][Natoms][3], [2][Natoms][P][3]
  double *uk = nullptr;                         // [1] kinetic energy
  double *ukpart = nullptr;                     // [256] kinetic partial sums
  double *radius = nullptr;                     // [P][N] site radius about O
  double *pes = nullptr;                        // [PES_NPARAM] PES parameters
  double *adj = nullptr;                        // adjoint inputs and partial sums
  double *dRRfast = nullptr, *mol = nullptr;    // contracted MD: [P][Natoms][3], [P][N][3]
  int *clk = nullptr;                           // device step clock {t, record slot}
  double *obs_part = nullptr;                   // [64][5] partial sums of the observables
  int *clk_own = nullptr;                       // this context's clock (clk may be its owner's)
  double *tan = nullptr;                        // TTM parameter tangents: inputs, tables, workspace
  size_t tan_bytes = 0;
  double *red = nullptr;                        // scratch for reductions
  // reciprocal-space field, potential (FD_Q) and gradient, potential (FD_QD)
  // of the last field pass, [P][N][4][3] and [P][N][4]; their consumers add
  // them to Efq / dR and phi, so the real-space pair pass may run beside
  double *kE = nullptr, *kphi = nullptr;
};

// Host-side launchers (ttm_kernels.hip / rp_kernels.hip).
void launch_monomer(const SysConst &c, DevBufs &b, const double *RR, const double *RRc,
                    double *chg_out, hipStream_t s);
void launch_neighbors(const SysConst &c, DevBufs &b, hipStream_t s);
void launch_pair_qq(const SysConst &c, DevBufs &b, int group, hipStream_t s);
void launch_tables(const SysConst &c, DevBufs &b, hipStream_t s);
void launch_sfac(const SysConst &c, DevBufs &b, int mode, hipStream_t s);
void launch_converge(const SysConst &c, DevBufs &b, int iter, hipStream_t s);
void launch_field(const SysConst &c, DevBufs &b, int mode, int iter, hipStream_t s);
void launch_post_qq(const SysConst &c, DevBufs &b, hipStream_t s);
// test_iter > 0: the kernel first runs the convergence test of SOR iteration
// test_iter (in place of launch_converge); only where pair_dd_tests(c).
void launch_pair_dd(const SysConst &c, DevBufs &b, int group, bool fp64, hipStream_t s,
                    int test_iter = 0);
bool pair_dd_tests(const SysConst &c);
void launch_refine_prep(const SysConst &c, DevBufs &b, hipStream_t s);
void launch_pair_final(const SysConst &c, DevBufs &b, int group, hipStream_t s);
void launch_virk(const SysConst &c, DevBufs &b, hipStream_t s);
void launch_finalize(const SysConst &c, DevBufs &b, double *dRR, double *dipI,
                     double *dipE, hipStream_t s);
void launch_reduce_bead(const SysConst &c, DevBufs &b, double *Upot, double *virt,
                        double *virialc, int *iters, hipStream_t s);
void launch_spc_monomer(const SysConst &c, DevBufs &b, const double *RR, double *chg_out,
                        hipStream_t s);
void launch_spc_pair(const SysConst &c, DevBufs &b, hipStream_t s);
void launch_spc_finalize(const SysConst &c, DevBufs &b, double *dRR, double *dipI, double *dipE,
                         hipStream_t s);
void pes_default_params(double *theta);
void launch_pes_adjoint(const SysConst &c, DevBufs &b, const double *RR, const double *ew,
                        const double *gw, double *partial, int nblocks, double *grad,
                        hipStream_t s);
int pes_adjoint_blocks(const SysConst &c);

enum SfacMode { SF_Q = 0, SF_D = 1, SF_V = 2 };
enum FieldMode { FD_Q = 0, FD_D = 1, FD_QD = 2 };

// Ring polymer launchers (rp_kernels.hip).
// TTM model parameters for fitting, in the order of pimd_ttm_params_get.
constexpr int TTM_NPARAM = 16;
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
};
size_t tangent_cfg_doubles(const SysConst &c);
size_t tangent_ws_duals(const SysConst &c);
void launch_ttm_tangent(const SysConst &c, const TanLaunch &L, hipStream_t s);

// Per-step observables (write_out): sizes and constants of the system.
struct ObsArgs {
  int P, natoms;
  double volume, omegan;
};
constexpr int MD_REC = 13;   // per-step record: Upot, virial, virialc, uk, then 9 observables
void launch_kinetic_nm(int P, int natoms, const DevBufs &b, const double *PP, double *uk,
                       double *part, hipStream_t s);
// Upot, virt, virialc: the per-bead outputs of the step's force call, or NULL
// when the step record rec[0..2] has been written already.
void launch_observables(const ObsArgs &a, const double *RR, const double *dipI,
                        const double *masses, const double *uk, const double *Upot,
                        const double *virt, const double *virialc, double *part, double *rec0,
                        const int *clk, hipStream_t s);
void launch_set_clock(int *clk, int t, int slot, hipStream_t s);
void launch_advance_clock(int *clk, hipStream_t s, const double *uk = nullptr,
                          double *rec0 = nullptr);

void launch_pes_fast(const SysConst &c, DevBufs &b, const double *RR, const double *RRc,
                     double *dRR, double *mol, bool record, hipStream_t s);
void launch_broadcast_beads(int P, int n, const double *g, double *dRR, hipStream_t s);
void launch_contract_record(int P, int N, const double *mol, const double *Ucen,
                            const double *vtcen, const double *vccen, const double *uk,
                            double *rec0, const int *clk, hipStream_t s);
// kick, kick2 (optional): PP first takes the half kicks PP - MASSCON*kick*delt2,
// then - MASSCON*kick2*delt2b, each rounded as a k_kick launch would.
void launch_nm_evolve(int P, int natoms, DevBufs &b, double *RR, double *PP, hipStream_t s,
                      const double *kick = nullptr, double delt2 = 0.0,
                      const double *kick2 = nullptr, double delt2b = 0.0);
void launch_langevin(int P, int natoms, DevBufs &b, double *PP, const double *noise,
                     uint64_t seed, uint64_t counter, double massO, double massH,
                     hipStream_t s, const int *clk = nullptr, int half = 0,
                     const double *kick = nullptr, double delt2 = 0.0);
void launch_kick(int n, double *PP, const double *dRR, double f, hipStream_t s);
void launch_scale(int n, double *PP, const double *sc, hipStream_t s);
void launch_centroid_pbc(int P, int natoms, double box0, double boxi0, double *RR,
                         double *RRc, hipStream_t s);
void launch_kinetic(int P, int natoms, const double *masses, const double *PP, double *uk,
                    double *part, hipStream_t s);
void launch_nose_hoover(double *uk, double *vxi, double *sc, double tau, double delt2, int L,
                        double temp, hipStream_t s);
void launch_free_flight(int n, double *RR, const double *PP, const double *masses,
                        double delt, hipStream_t s);

}  // namespace pimd
