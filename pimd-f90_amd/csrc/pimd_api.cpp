// C ABI (include/pimd_ttm.h): context set-up, the all-beads force pipeline,
// ring-polymer entry points, the device-resident PIMD_VV loop and timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "pimd_internal.h"

using namespace pimd;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(PIMD_EDEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),      \
                  __FILE__, __LINE__);                                               \
  } while (0)

struct Pending {
  int id;
  hipEvent_t a, b;
};

}  // namespace

struct pimd_ctx {
  pimd_params p{};
  SysConst c{};
  DevBufs d{};
  hipStream_t stream = nullptr;
  bool stream_borrowed = false;  // the stream belongs to the context that owns this one
  // contracted MD: the centroid's intermolecular forces (one bead, no
  // monomer PES), computed on this context's stream
  pimd_ctx *cen = nullptr;
  // second stream (made when first needed): real-space pair kernels run on
  // it beside the reciprocal-space kernels they do not depend on (SOR
  // overlap, graph branches); fork/join through the two events
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int group = 64;
  int natoms = 0;
  std::vector<double> kvec, Ak, p4vAk;
  double polfac[4] = {0, 0, 0, 0};
  double ttm[TTM_NPARAM] = {};  // model parameters (vdW, polarisabilities, Thole, DMS)
  std::vector<int> hkl;
  // ring polymer
  bool nm_ready = false, lang_ready = false, md_ready = false, state_ready = false;
  std::vector<double> C, A11, A12, A21, A22, om, msf, c1, c2;
  pimd_md_params md{};
  bool cmd = false;       // centroid MD (setNMfreq != 0, P > 1): kinetic energy in mode space
  // one captured MD step, replayed for every step of a run
  hipGraphExec_t graph = nullptr;
  bool graph_ok = false;
  bool graph_off = false;  // capture failed once: eager launches from then on
  uint64_t graph_seed = 0;
  // the force call of the all-beads entry points, captured for its buffer
  // addresses fkey and replayed while they and the model stay the same
  hipGraphExec_t fgraph = nullptr;
  bool fgraph_ok = false, fgraph_off = false;
  const void *fkey[10] = {};
  double omegan = 0.0;    // ring frequency kB T P / hbar of the MD temperature
  double *d_vxi = nullptr, *d_s = nullptr, *d_step = nullptr;
  double *d_rec = nullptr;  // per-step records of md runs
  int rec_cap = 0;
  int rec_steps = 0;        // steps recorded by the last run
  int *h_nactive = nullptr;
  // pinned staging of the host-array force call: coordinates in, all outputs
  // out (made on the first call)
  double *h_stage = nullptr;
  // profiling
  bool prof = false;
  // inside a graph capture: independent real- and reciprocal-space kernels
  // of a force call go on two streams (graph branches; in eager launches the
  // event fork and join cost more than the overlap gains)
  bool branches = false;
  std::vector<std::string> names;
  std::map<std::string, int> name_id;
  std::vector<long> launches;
  std::vector<double> total_ms;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;

  int kid(const char *n) {
    auto it = name_id.find(n);
    if (it != name_id.end()) return it->second;
    int id = (int)names.size();
    names.push_back(n);
    name_id[n] = id;
    launches.push_back(0);
    total_ms.push_back(0.0);
    return id;
  }
  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
  void resolve() {
    if (pending.empty()) return;
    (void)hipStreamSynchronize(stream);
    for (auto &q : pending) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, q.a, q.b);
      launches[q.id] += 1;
      total_ms[q.id] += ms;
      pool.push_back(q.a);
      pool.push_back(q.b);
    }
    pending.clear();
  }
  template <class F>
  void timed(const char *n, F &&f) {
    if (!prof) {
      f();
      return;
    }
    if (pending.size() > 4096) resolve();
    Pending q{kid(n), ev(), ev()};
    (void)hipEventRecord(q.a, stream);
    f();
    (void)hipEventRecord(q.b, stream);
    pending.push_back(q);
  }
};

namespace {

// Captured launches hold launch shapes, model constants and buffer addresses:
// after a change to any of them both graphs are captured again.
void invalidate_graphs(pimd_ctx *x) {
  x->graph_ok = false;
  x->fgraph_ok = false;
}

template <class T>
int dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) return fail(PIMD_ENOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T),
                                   hipGetErrorString(e));
  return PIMD_OK;
}

void free_all(pimd_ctx *x) {
  DevBufs &d = x->d;
  if (x->cen) {
    free_all(x->cen);
    delete x->cen;
    x->cen = nullptr;
  }
  void *ptrs[] = {d.RR,   d.RRc,   d.site,   d.Rcs,    d.grdq,   d.phi,   d.Efq,   d.dR,
                  d.dip,  d.dip_prev, d.Efd,   d.diff2,  d.tally,  d.hist,   d.nbr,   d.nbrc,  d.nbrn,  d.tab,
                  d.Sq,   d.Sm,    d.Sv,     d.virk,   d.part,   d.active, d.iters, d.nactive, d.kv,
                  d.khkl, d.cols,  d.colk,   d.segs,   d.segk,   d.Upot,   d.virt,   d.virialc, d.terms, d.dRR,
                  d.dipI, d.dipE,  d.chg,    d.iters_out, d.C,  d.A,     d.msf,   d.c1,
                  d.c2,   d.masses, d.PP,    d.noise,  d.uk,  d.ukpart, d.radius,     d.red, d.pes, d.adj, d.dRRfast, d.mol, d.tan, d.clk_own, d.obs_part, d.kE, d.kphi, d.cell_start, d.cell_mol, x->d_vxi, x->d_s,
                  x->d_step};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (x->h_nactive) (void)hipHostFree(x->h_nactive);
  if (x->h_stage) (void)hipHostFree(x->h_stage);
  for (auto &q : x->pending) {
    (void)hipEventDestroy(q.a);
    (void)hipEventDestroy(q.b);
  }
  for (auto e : x->pool) (void)hipEventDestroy(e);
  if (x->graph) (void)hipGraphExecDestroy(x->graph);
  if (x->fgraph) (void)hipGraphExecDestroy(x->fgraph);
  if (x->stream && !x->stream_borrowed) (void)hipStreamDestroy(x->stream);
  if (x->stream2) (void)hipStreamDestroy(x->stream2);
  if (x->d_rec) (void)hipFree(x->d_rec);
  if (x->ev_fork) (void)hipEventDestroy(x->ev_fork);
  if (x->ev_join) (void)hipEventDestroy(x->ev_join);
}

float f32(double v) { return (float)v; }

// The parameter-dependent constants of init_pot (potential.f90:84-113,
// 152-160) from the context's theta: vdW terms and their long-range
// correction, polarisabilities per site type (O, H, H, M), the Thole products
// aCCaCD or aDD * polfac_i * polfac_j with polfac = (polfac^(1/6))^-3, and the
// DMS corrections.
void apply_ttm_params(pimd_ctx *x) {
  SysConst &c = x->c;
  const double *th = x->ttm, Rc = x->p.rc;
  c.vdwA = th[TP_VDWA]; c.vdwB = th[TP_VDWB]; c.vdwC = th[TP_VDWC];
  c.vdwD = th[TP_VDWD]; c.vdwE = th[TP_VDWE];
  c.polar[0] = th[TP_POLARO]; c.polar[1] = th[TP_POLARH]; c.polar[2] = th[TP_POLARH];
  c.polar[3] = th[TP_POLARM];
  c.dms1 = th[TP_DMS1]; c.dms2 = th[TP_DMS2]; c.dms3 = th[TP_DMS3];
  c.Uvdw_lrc = TWOPI * (double)(c.N * c.N) / c.vol *
               (c.vdwA / (9.0 * std::pow(Rc, 9)) + c.vdwB / (7.0 * std::pow(Rc, 7)) +
                c.vdwC / (3.0 * (Rc * Rc * Rc)));
  const double pf6[4] = {std::pow(th[TP_POLFACO], 1.0 / 6.0), std::pow(th[TP_POLFACH], 1.0 / 6.0),
                         std::pow(th[TP_POLFACH], 1.0 / 6.0), std::pow(th[TP_POLFACM], 1.0 / 6.0)};
  const double aDD = th[TP_ADD], aCC = th[TP_ACC];
  double pf[4];
  for (int i = 0; i < 4; ++i) pf[i] = 1.0 / (pf6[i] * pf6[i] * pf6[i]);
  for (int i = 0; i < 4; ++i) x->polfac[i] = pf[i];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      c.apolCC[i][j] = aCC * pf[i] * pf[j];
      c.raCC[i][j] = std::pow(c.apolCC[i][j], 1.0 / 3.0);
      c.apolDD[i][j] = aDD * pf[i] * pf[j];
      c.apol3DD[i][j] = c.apolDD[i][j];
    }
  c.same_a = std::fabs(aCC - aDD) < 1e-12;
}

int env_int(const char *name, int dflt, int lo, int hi);

// init_pot + ewald_set (potential.f90:43-199, ewald.f90:293-378) on the host.
int setup_system(pimd_ctx *x) {
  const pimd_params &p = x->p;
  SysConst &c = x->c;
  c.N = p.nwaters;
  c.P = p.nbeads;
  c.model = p.pot_model;
  for (int k = 0; k < 3; ++k) {
    c.box[k] = p.box[k];
    c.boxi[k] = 1.0 / p.box[k];
  }
  c.vol = p.box[0] * p.box[1] * p.box[2];
  const double Rc = p.rc, rc1 = p.rc1;
  c.rc2 = Rc * Rc;
  c.rc1 = rc1;
  double *th = x->ttm;
  if (p.pot_model == PIMD_TTM21F) {  // pot_mod.f90:7-14
    const double v[TTM_NPARAM] = {-1.329565985e+6, 3.632560798e+5, -2.147141323e+3, 1e13, 13.2,
                                  0.837, 0.496, 0.837, 0.837, 0.496, 0.837, 0.3, 0.2,
                                  0.0, 0.0, 0.0};
    std::copy(v, v + TTM_NPARAM, th);
    c.gammaM = 0.426706882;
    c.nd = 3; c.tdip = 0;
    c.shiftA = c.shiftB = 0.0;
  } else if (p.pot_model >= PIMD_QSPCFW) {
    // q-SPC/Fw (pot_mod.f90:31-35) and SPC/F (pot_mod.f90:40-44): vdwA and
    // vdwC only; polfac set to 1 (no polarisation in these models)
    const bool q = p.pot_model == PIMD_QSPCFW;
    const double v[TTM_NPARAM] = {q ? 629326.57249877361701443513 : 628213.72817108343399678768,
                                  0.0, q ? -625.50206521141523463400 : -624.97858519194071637260,
                                  0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    std::copy(v, v + TTM_NPARAM, th);
    c.gammaM = 0.0;
    c.nd = 0; c.tdip = 3;
    c.shiftA = c.shiftB = 0.0;
    c.spc_q[0] = (q ? -0.84 : -0.82) * CHARGECON;  // potential.f90:186-192
    c.spc_q[1] = (q ? 0.42 : 0.41) * CHARGECON;
    c.spc_roh0 = q ? 1.012 : 1.0;
    c.spc_th0 = q ? 1.95476876223364912614 : 1.91061193215819258784;
    c.spc_rhh0 = 2.0 * std::sin(0.5 * c.spc_th0) * c.spc_roh0;  // pot_spc.f90:108
    c.spc_kb = q ? 529.581 : 0.0;
    c.spc_ka = q ? 37.95 : 0.0;
    c.spc_A = q ? 0.0 : 671.288262912;
    c.spc_B = q ? 0.0 : 164.376;
    c.spc_C = q ? 0.0 : -211.536;
    c.spc_D = q ? 0.0 : 111.744;
  } else {  // pot_mod.f90:19-26
    const double v[TTM_NPARAM] = {0.0, 0.0, -0.72298855e+03, 0.10211829e+06, 0.37170376e+01,
                                  0.0, 0.0, 1.444, 0.837, 0.496, 0.837, 0.175, 0.175,
                                  0.5, 0.9578, 0.012};
    std::copy(v, v + TTM_NPARAM, th);
    c.gammaM = 0.46;
    c.nd = 1; c.tdip = 3;
    if (Rc == rc1) {
      c.shiftA = c.shiftB = 0.0;
    } else {  // potential.f90:107-113
      const double Rc8 = std::pow(Rc, 8);
      c.shiftA = -(10 * Rc - 7 * rc1) / (Rc8 * ((Rc - rc1) * (Rc - rc1)));
      c.shiftB = (9 * Rc - 7 * rc1) / (Rc8 * ((Rc - rc1) * (Rc - rc1) * (Rc - rc1)));
    }
  }
  apply_ttm_params(x);
  // ewald_set(.true.)
  const double PP = -std::log(p.eps_ewald);
  c.aewald = std::sqrt(PP) / Rc;
  const double kcut = 2.0 * c.aewald * std::sqrt(PP), kcut2 = kcut * kcut;
  c.hmax = (int)std::floor(kcut / TWOPI * p.box[0]);
  c.kmax = (int)std::floor(kcut / TWOPI * p.box[1]);
  c.lmax = (int)std::floor(kcut / TWOPI * p.box[2]);
  const double fac = -0.25 / (c.aewald * c.aewald);
  x->kvec.clear();
  x->Ak.clear();
  x->hkl.clear();
  for (int kx = 0; kx <= c.hmax; ++kx) {
    const double kv0 = TWOPI * c.boxi[0] * (double)kx;
    const int ky1 = kx == 0 ? 0 : -c.kmax;
    for (int ky = ky1; ky <= c.kmax; ++ky) {
      const double kv1 = TWOPI * c.boxi[1] * (double)ky;
      const int kz1 = (kx == 0 && ky == 0) ? 1 : -c.lmax;
      const double kvxy = kv0 * kv0 + kv1 * kv1;
      for (int kz = kz1; kz <= c.lmax; ++kz) {
        const double kv2 = TWOPI * c.boxi[2] * (double)kz;
        const double kvsq = kvxy + kv2 * kv2;
        if (kvsq < kcut2 && kvsq > 1e-16) {
          x->hkl.insert(x->hkl.end(), {kx, ky, kz});
          x->kvec.insert(x->kvec.end(), {kv0, kv1, kv2});
          x->Ak.push_back(2.0 * std::exp(fac * kvsq) / kvsq);
        }
      }
    }
  }
  c.nk = (int)x->Ak.size();
  if (c.nk == 0) return fail(PIMD_EINVAL, "no k-vectors inside the Ewald cut-off");
  c.T = (c.hmax + 1) + (c.kmax + 1) + (c.lmax + 1);
  // the reciprocal-space kernels stage T table rows of up to 64 sites in LDS
  // (T KiB of it, within the 160 KiB of a workgroup)
  if (c.T > 128)
    return fail(PIMD_EINVAL, "%d Ewald table rows (hmax+kmax+lmax+3) exceed the limit of 128; "
                "raise eps_ewald or the cut-off", c.T);
  c.a2 = c.aewald * c.aewald;
  c.cts1 = 2.0 * c.aewald / SQPI;
  c.cts2 = 4.0 * c.aewald * c.aewald * c.aewald / (3.0 * SQPI);
  c.cts3 = 8.0 * std::pow(c.aewald, 5) / (15.0 * SQPI);
  c.self_fac = c.aewald * TSP;
  c.fac_dd = 4.0 * (c.aewald * c.aewald * c.aewald) / (3.0 * std::sqrt(PI));
  c.maxiter = p.polar_maxiter;
  c.guess = p.guess_initdip;
  c.mixed = p.precision == PIMD_MIXED ? 1 : 0;
  c.contraction = p.contraction;
  c.sor = p.polar_sor;
  c.peps = p.polar_eps;
  c.stath = DEBYE / CHARGECON / std::sqrt((double)(3 * c.N));
  c.p4v = FOURPI / c.vol;
  c.nbr_cap = std::max(c.N - 1, 1);
  // a cell list for the neighbour search needs every axis to hold at least
  // three cells of width >= Rc (else the 27 neighbouring cells repeat); it
  // pays once those 27 cells are at most a quarter of the box (timed at
  // N = 2048, 4^3 cells: 0.12 ms against 0.08 ms for the full search;
  // N = 4096, 5^3 cells: 0.20 ms against 0.28 ms). PIMD_CELLS=0 turns it
  // off, 2 uses it wherever the geometry allows.
  for (int k = 0; k < 3; ++k) c.ncell[k] = (int)std::floor(p.box[k] / Rc);
  {
    const long nc = (long)c.ncell[0] * c.ncell[1] * c.ncell[2];
    const int mode = env_int("PIMD_CELLS", 1, 0, 2);
    const bool fits = c.ncell[0] >= 3 && c.ncell[1] >= 3 && c.ncell[2] >= 3 && nc <= 4096;
    c.cells = fits && (mode == 2 || (mode == 1 && nc >= 4 * 27));
  }
  x->p4vAk.resize(c.nk);
  for (int k = 0; k < c.nk; ++k) x->p4vAk[k] = FOURPI * x->Ak[k] / c.vol;
  return PIMD_OK;
}

int env_int(const char *name, int dflt, int lo, int hi) {
  const char *v = std::getenv(name);
  if (!v || !*v) return dflt;
  const int n = std::atoi(v);
  return n < lo ? lo : n > hi ? hi : n;
}

int alloc_system(pimd_ctx *x) {
  SysConst &c = x->c;
  DevBufs &d = x->d;
  const size_t P = c.P, N = c.N, S = 4 * N, na = 3 * N, nk = c.nk, nd = c.nd;
  int rc;
#define ALLOC(ptr, n) \
  if ((rc = dalloc(&(ptr), (n))) != PIMD_OK) return rc
  ALLOC(d.RR, P * na * 3);
  ALLOC(d.RRc, na * 3);
  ALLOC(d.site, P * S);
  ALLOC(d.Rcs, S * 3);
  ALLOC(d.grdq, P * N * 27);
  ALLOC(d.phi, P * S);
  ALLOC(d.Efq, P * S * 3);
  ALLOC(d.kE, P * S * 3);
  ALLOC(d.kphi, P * S);
  ALLOC(d.dR, P * S * 3);
  ALLOC(d.dip, P * N * nd * 3);
  ALLOC(d.dip_prev, P * N * nd * 3);
  ALLOC(d.Efd, P * N * nd * 3);
  ALLOC(d.diff2, P * N * nd);
  ALLOC(d.tally, P * N * T_COUNT);
  ALLOC(d.hist, P * 4 * N * nd * 3);
  ALLOC(d.nbr, P * N * (size_t)c.nbr_cap);
  ALLOC(d.nbrc, P * N);
  ALLOC(d.nbrn, P * N);
  ALLOC(d.tab, P * (size_t)c.T * S);
  ALLOC(d.Sq, P * nk * 2);
  ALLOC(d.Sm, P * nk * 2);
  ALLOC(d.Sv, P * nk * 6);
  ALLOC(d.virk, P * ((nk + 255) / 256) * 18);
  ALLOC(d.active, P);
  ALLOC(d.iters, P);
  ALLOC(d.nactive, (size_t)c.maxiter + 1);
  ALLOC(d.kv, nk);
  ALLOC(d.khkl, nk);
  ALLOC(d.red, nk);  // A_k
  ALLOC(d.Upot, P);
  ALLOC(d.virt, P * 9);
  ALLOC(d.virialc, P);
  ALLOC(d.terms, P * 5);
  ALLOC(d.dRR, P * na * 3);
  ALLOC(d.dipI, P * N * 3);
  ALLOC(d.dipE, P * N * 3);
  ALLOC(d.chg, P * na);
  ALLOC(d.iters_out, P);
  ALLOC(d.C, P * P);
  ALLOC(d.A, 4 * P);
  ALLOC(d.msf, P);
  ALLOC(d.c1, P);
  ALLOC(d.c2, P);
  ALLOC(d.masses, na);
  ALLOC(d.PP, P * na * 3);
  ALLOC(d.noise, 2 * na * P * 3);
  ALLOC(d.uk, 1);
  ALLOC(d.ukpart, 256);
  ALLOC(d.radius, P * N);
  if (c.cells) {
    ALLOC(d.cell_start, P * ((size_t)c.ncell[0] * c.ncell[1] * c.ncell[2] + 1));
    ALLOC(d.cell_mol, P * N);
  }
  ALLOC(d.pes, PES_NPARAM);
  ALLOC(d.adj, 2 * P * na * 3 + P + (size_t)pes_adjoint_blocks(c) * PES_NPARAM + PES_NPARAM);
  ALLOC(x->d_vxi, 2);
  ALLOC(x->d_s, 1);
  ALLOC(x->d_step, 4);
#undef ALLOC
  HIPCHK(hipMemset(d.hist, 0, sizeof(double) * P * 4 * N * nd * 3));
  {
    double theta[PES_NPARAM];
    pes_default_params(theta);
    HIPCHK(hipMemcpy(d.pes, theta, sizeof theta, hipMemcpyHostToDevice));
  }
  // k-vectors and kz-columns
  std::vector<double4> kv(nk);
  std::vector<int4> kh(nk);
  std::vector<int4> cols;
  std::vector<int> colk;
  for (size_t k = 0; k < nk; ++k) {
    kv[k] = make_double4(x->kvec[3 * k], x->kvec[3 * k + 1], x->kvec[3 * k + 2], x->p4vAk[k]);
    kh[k] = make_int4(x->hkl[3 * k], x->hkl[3 * k + 1], x->hkl[3 * k + 2], 0);
    const bool cont = k > 0 && x->hkl[3 * k] == x->hkl[3 * k - 3] &&
                      x->hkl[3 * k + 1] == x->hkl[3 * k - 2] &&
                      x->hkl[3 * k + 2] == x->hkl[3 * k - 1] + 1;
    if (cont) {
      cols.back().w += 1;
    } else {
      cols.push_back(make_int4(x->hkl[3 * k], x->hkl[3 * k + 1], x->hkl[3 * k + 2], 1));
      colk.push_back((int)k);
    }
  }
  // longest columns first: lanes and column groups that run together get
  // similar trip counts
  {
    std::vector<int> order(cols.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b2) { return cols[a].w > cols[b2].w; });
    std::vector<int4> cs(cols.size());
    std::vector<int> ck(colk.size());
    for (size_t i = 0; i < order.size(); ++i) {
      cs[i] = cols[order[i]];
      ck[i] = colk[order[i]];
    }
    cols.swap(cs);
    colk.swap(ck);
  }
  c.ncol = (int)cols.size();
  c.kzmax = 0;
  for (auto &cl : cols) c.kzmax = std::max(c.kzmax, cl.w);
  // segments for the structure-factor kernel: the columns cut into pieces of
  // at most segmax consecutive kz, longest first
  std::vector<int4> segs;
  std::vector<int> segk;
  {
    const int seg = env_int("PIMD_SFAC_SEG", 8, 1, 32);
    for (size_t i = 0; i < cols.size(); ++i)
      for (int o = 0; o < cols[i].w; o += seg) {
        segs.push_back(make_int4(cols[i].x, cols[i].y, cols[i].z + o, std::min(seg, cols[i].w - o)));
        segk.push_back(colk[i] + o);
      }
    std::vector<int> order(segs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b2) { return segs[a].w > segs[b2].w; });
    std::vector<int4> ss(segs.size());
    std::vector<int> sk(segk.size());
    for (size_t i = 0; i < order.size(); ++i) {
      ss[i] = segs[order[i]];
      sk[i] = segk[order[i]];
    }
    segs.swap(ss);
    segk.swap(sk);
  }
  c.nseg = (int)segs.size();
  c.segmax = 0;
  for (auto &sg : segs) c.segmax = std::max(c.segmax, sg.w);
  // site chunks of at most 64 sites keep the staged tables within LDS
  c.sf_chunks_max = std::max(32, (3 * c.N + 63) / 64);
  // launch-shape tuning (results do not depend on them)
  // default launch shapes by size: a system of few molecule-beads (a
  // contracted run's one-bead centroid, small replicas) gives each molecule
  // more lanes and the structure factor more site chunks, so that its
  // kernels still spread over the device
  const long mb = (long)c.P * c.N;
  auto pow2_clamp = [](long v, int lo, int hi) {
    int g = lo;
    while (g < hi && g < v) g <<= 1;
    return g;
  };
  const int dflt_qq = pow2_clamp(131072 / std::max(mb, 1L), 8, 32);
  // the FP32 dipole pass runs lighter lanes: twice the lanes per molecule
  // (512 H2O x 32 beads: 32 instead of 16, 1.685 -> 1.657 ms per force call);
  // TTM2.1-F's three dipole sites at most 32 (128 x 32: 1.50 -> 1.37 ms FP64)
  const int dflt_dd = pow2_clamp((c.mixed ? 524288 : 262144) / std::max(mb, 1L), 16,
                                 c.model == 2 ? 32 : 64);
  // structure-factor site chunks: about 8 per bead from 4096 molecule-beads
  // on (with two lane sets per block; 512 H2O x 32 beads: 1.66 -> 1.61 ms per
  // mixed force call, TTM2.1-F 128 x 32: 1.75 -> 1.66 ms FP64), the staged
  // tables of a chunk within 60 KB; 8 sites below, to spread small systems
  int dflt_sites = 8;
  if (mb >= 4096) {
    const int lds_cap = std::max(8, 60 * 1024 / (16 * c.T) - 17);
    dflt_sites = std::max(8, std::min({(c.nd * c.N + 7) / 8, 128, lds_cap}));
  }
  c.sf_sites = env_int("PIMD_SFAC_SITES", dflt_sites, 1, 128);
  c.fd_cg = env_int("PIMD_FIELD_CG", 0, 0, 16);
  c.sf_kind = env_int("PIMD_SFAC_KIND", 1, 0, 1);
  c.sfv_col = env_int("PIMD_SFV_COL", 1, 0, 1);
  c.sf_split = env_int("PIMD_SFAC_SPLIT", 1, 0, 1);
  c.fd_klds = env_int("PIMD_FIELD_KLDS", 1, 0, 1);
  // phase tables of 64 sites, or the 16-group reduction if larger (in double2)
  c.fd_tab_d2 = std::max(c.T * 64, 16 * 4 * 64 / 2);
  c.g_dd = env_int("PIMD_G_DD", dflt_dd, 0, 64);
  c.g_qq = env_int("PIMD_G_QQ", dflt_qq, 0, 32);
  c.g_fin = env_int("PIMD_G_FIN", dflt_qq, 0, 32);
  c.poll = env_int("PIMD_SOR_POLL", 0, 0, 1);
  c.overlap = env_int("PIMD_SOR_OVERLAP", 0, 0, 1);
  if (c.fd_cg != 0 && c.fd_cg != 4 && c.fd_cg != 8 && c.fd_cg != 16) c.fd_cg = 0;
  int rc2;
  if ((rc2 = dalloc(&d.part, P * (size_t)c.sf_chunks_max * nk * 6)) != PIMD_OK) return rc2;
  if ((rc2 = dalloc(&d.cols, cols.size())) != PIMD_OK) return rc2;
  if ((rc2 = dalloc(&d.colk, colk.size())) != PIMD_OK) return rc2;
  HIPCHK(hipMemcpy(d.kv, kv.data(), sizeof(double4) * nk, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d.khkl, kh.data(), sizeof(int4) * nk, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d.red, x->Ak.data(), sizeof(double) * nk, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d.cols, cols.data(), sizeof(int4) * cols.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d.colk, colk.data(), sizeof(int) * colk.size(), hipMemcpyHostToDevice));
  if ((rc2 = dalloc(&d.segs, segs.size())) != PIMD_OK) return rc2;
  if ((rc2 = dalloc(&d.segk, segk.size())) != PIMD_OK) return rc2;
  HIPCHK(hipMemcpy(d.segs, segs.data(), sizeof(int4) * segs.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d.segk, segk.data(), sizeof(int) * segk.size(), hipMemcpyHostToDevice));
  // the device step clock {t, record slot}
  if ((rc2 = dalloc(&d.obs_part, 64 * 5)) != PIMD_OK) return rc2;
  if ((rc2 = dalloc(&d.clk_own, 2)) != PIMD_OK) return rc2;
  d.clk = d.clk_own;
  HIPCHK(hipMemset(d.clk, 0, sizeof(int) * 2));
  std::vector<double> masses(na);
  for (size_t a = 0; a < na; ++a) masses[a] = (a % 3 == 0) ? x->p.mass_o : x->p.mass_h;
  HIPCHK(hipMemcpy(d.masses, masses.data(), sizeof(double) * na, hipMemcpyHostToDevice));
  HIPCHK(hipHostMalloc((void **)&x->h_nactive, sizeof(int) * (c.maxiter + 1), hipHostMallocDefault));
  return PIMD_OK;
}

// The pair kernels give each molecule G lanes; fewer lanes for short lists.
int pick_group(const SysConst &c) {
  const double frac = (4.0 / 3.0) * PI * std::pow(std::sqrt(c.rc2), 3) / c.vol;
  const double nnb = std::min((double)(c.N - 1), frac * c.N);
  if (nnb >= 96) return 64;
  if (nnb >= 40) return 32;
  return 16;
}

// The second stream, made on first use (outside any capture: by the caller
// before a capture begins).
int ensure_stream2(pimd_ctx *x) {
  if (x->stream2) return PIMD_OK;
  if (hipStreamCreateWithFlags(&x->stream2, hipStreamNonBlocking) != hipSuccess)
    return fail(PIMD_EDEVICE, "hipStreamCreate failed");
  return PIMD_OK;
}

// a(stream2) beside b(s); s waits for both.
template <class A, class B>
int beside(pimd_ctx *x, hipStream_t s, A &&a, B &&b) {
  if (!x->stream2) return fail(PIMD_ESTATE, "second stream missing");
  HIPCHK(hipEventRecord(x->ev_fork, s));
  HIPCHK(hipStreamWaitEvent(x->stream2, x->ev_fork, 0));
  a(x->stream2);
  HIPCHK(hipEventRecord(x->ev_join, x->stream2));
  b(s);
  HIPCHK(hipStreamWaitEvent(s, x->ev_join, 0));
  return PIMD_OK;
}

// The real-space dipole field on the second stream beside the dipole
// structure factor on the first; the first stream waits for both.
int dd_beside_sfac(pimd_ctx *x, hipStream_t s) {
  return beside(x, s, [&](hipStream_t s2) { launch_pair_dd(x->c, x->d, x->group, false, s2); },
                [&](hipStream_t s1) { launch_sfac(x->c, x->d, SF_D, s1); });
}

__global__ void k_zero_int(int *v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0;
}

// One pot_ttm evaluation per bead on device arrays (pot_ttm.f90:3-630).
// The MD step t of the call (dipole predictor, history slot) is read from the
// device clock d.clk, set by the caller.
int run_forces(pimd_ctx *x, const double *RR, const double *RRc, double *Upot,
               double *dRR, double *virt, double *virialc, double *dipI, double *dipE,
               double *chg, int *iters) {
  const SysConst &c = x->c;
  DevBufs &d = x->d;
  hipStream_t s = x->stream;
  const bool two = x->branches && !x->prof;
  if (c.overlap && !x->stream2) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return fail(PIMD_ESTATE, "second stream missing");
    const int err = ensure_stream2(x);
    if (err != PIMD_OK) return err;
  }
  // the reciprocal-space charge pass, beside the real-space pair pass when
  // two streams are in use (it writes kE / kphi, read after the join)
  auto kspace_q = [&](hipStream_t st) {
    x->timed("ewald_tables", [&] { launch_tables(c, d, st); });
    x->timed("sfac_q", [&] { launch_sfac(c, d, SF_Q, st); });
    x->timed("field_q", [&] { launch_field(c, d, FD_Q, 0, st); });
  };
  if (c.model >= PIMD_QSPCFW) {  // pot_spc (pot_spc.f90:1-277)
    x->timed("monomer", [&] { launch_spc_monomer(c, d, RR, chg, s); });
    x->timed("neighbors", [&] { launch_neighbors(c, d, s); });
    auto pair = [&](hipStream_t st) { x->timed("pair_qq", [&] { launch_spc_pair(c, d, st); }); };
    if (two) {
      const int err = beside(x, s, pair, kspace_q);
      if (err != PIMD_OK) return err;
    } else {
      pair(s);
      kspace_q(s);
    }
    x->timed("virial_k", [&] { launch_virk(c, d, s); });
    x->timed("finalize", [&] { launch_spc_finalize(c, d, dRR, dipI, dipE, s); });
    x->timed("reduce_bead", [&] { launch_reduce_bead(c, d, Upot, virt, virialc, nullptr, s); });
    // no SOR loop: zero iterations
    hipLaunchKernelGGL(k_zero_int, dim3((c.P + 63) / 64), dim3(64), 0, s, iters, c.P);
    HIPCHK(hipGetLastError());
    return PIMD_OK;
  }
  x->timed("monomer", [&] { launch_monomer(c, d, RR, RRc, chg, s); });
  x->timed("neighbors", [&] { launch_neighbors(c, d, s); });
  {
    auto pair = [&](hipStream_t st) { x->timed("pair_qq", [&] { launch_pair_qq(c, d, x->group, st); }); };
    if (two) {
      const int err = beside(x, s, pair, kspace_q);
      if (err != PIMD_OK) return err;
    } else {
      pair(s);
      kspace_q(s);
    }
  }
  x->timed("post_qq", [&] { launch_post_qq(c, d, s); });
  // the convergence test of iteration it - 1 runs at the start of the
  // dipole pair kernel of iteration it, where that kernel can hold it (and
  // nothing runs beside it); the last iteration's test stays a launch
  const bool fold = pair_dd_tests(c) && !c.overlap && !two && !c.poll && !x->prof;
  for (int it = 1; it <= c.maxiter; ++it) {
    if (c.overlap || two) {
      // pair_dd on the second stream, concurrent with the structure factor
      int err = PIMD_OK;
      x->timed("dd_sfac_d", [&] { err = dd_beside_sfac(x, s); });
      if (err != PIMD_OK) return err;
    } else {
      x->timed("pair_dd", [&] { launch_pair_dd(c, d, x->group, false, s, fold && it > 1 ? it - 1 : 0); });
      x->timed("sfac_d", [&] { launch_sfac(c, d, SF_D, s); });
    }
    x->timed("field_d_sor", [&] { launch_field(c, d, FD_D, it, s); });
    if (!fold || it == c.maxiter) x->timed("converge", [&] { launch_converge(c, d, it, s); });
    // No host round trip: converged beads are switched off on the device
    // and every kernel of the loop skips them, so the remaining iterations
    // of a converged system are near-empty launches.
    // (with per-kernel timing on, poll so only live iterations are launched
    // and timed)
    if ((c.poll || x->prof) && it < c.maxiter) {
      HIPCHK(hipMemcpyAsync(x->h_nactive + it, d.nactive + it, sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (x->h_nactive[it] == 0) break;
    }
  }
  if (c.mixed) {
    // Iterative refinement: redo each bead's last SOR step in FP64 from the
    // dipoles it started from.  The reciprocal-space dipole structure factor
    // of that step is still in place (beads stop updating it once converged).
    x->timed("refine_dd", [&] {
      launch_refine_prep(c, d, s);
      launch_pair_dd(c, d, x->group, true, s);
      launch_field(c, d, FD_D, c.maxiter + 1, s);
    });
  }
  {
    // the final real-space pass beside the reciprocal-space forces and virial
    auto pair = [&](hipStream_t st) { x->timed("pair_final", [&] { launch_pair_final(c, d, x->group, st); }); };
    auto kspace = [&](hipStream_t st) {
      x->timed("sfac_v", [&] { launch_sfac(c, d, SF_V, st); });
      x->timed("field_qd", [&] { launch_field(c, d, FD_QD, 0, st); });
      x->timed("virial_k", [&] { launch_virk(c, d, st); });
    };
    if (two) {
      const int err = beside(x, s, pair, kspace);
      if (err != PIMD_OK) return err;
    } else {
      pair(s);
      kspace(s);
    }
  }
  x->timed("finalize", [&] { launch_finalize(c, d, dRR, dipI, dipE, s); });
  x->timed("reduce_bead", [&] { launch_reduce_bead(c, d, Upot, virt, virialc, iters, s); });
  HIPCHK(hipGetLastError());
  return PIMD_OK;
}


// run_forces as one replay of a captured graph, captured again when the
// buffer addresses change (or after invalidate_graphs).  Per-kernel timing,
// polling and the two-stream overlap keep eager launches, as does a device
// where capture fails.  PIMD_FORCE_GRAPH=0 turns replay off.
int force_call(pimd_ctx *x, const double *RR, const double *RRc, double *Upot, double *dRR,
               double *virt, double *virialc, double *dipI, double *dipE, double *chg,
               int *iters) {
  const SysConst &c = x->c;
  const bool on = env_int("PIMD_FORCE_GRAPH", 1, 0, 1) != 0;
  if (!on || x->fgraph_off || x->prof || c.poll || c.overlap)
    return run_forces(x, RR, RRc, Upot, dRR, virt, virialc, dipI, dipE, chg, iters);
  const void *key[10] = {RR, RRc, Upot, dRR, virt, virialc, dipI, dipE, chg, iters};
  if (!x->fgraph_ok || !std::equal(key, key + 10, x->fkey)) {
    if (x->fgraph) (void)hipGraphExecDestroy(x->fgraph);
    x->fgraph = nullptr;
    x->fgraph_ok = false;
    hipStream_t s = x->stream;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = run_forces(x, RR, RRc, Upot, dRR, virt, virialc, dipI, dipE, chg, iters);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != PIMD_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    hipError_t e2 = e;
    if (e == hipSuccess) e2 = hipGraphInstantiate(&x->fgraph, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      // capture is not available here: launch eagerly from now on
      x->fgraph = nullptr;
      x->fgraph_off = true;
      (void)hipGetLastError();
      return run_forces(x, RR, RRc, Upot, dRR, virt, virialc, dipI, dipE, chg, iters);
    }
    std::copy(key, key + 10, x->fkey);
    x->fgraph_ok = true;
  }
  HIPCHK(hipGraphLaunch(x->fgraph, x->stream));
  return PIMD_OK;
}
}  // namespace

extern "C" {

int pimd_last_error(char *buf, size_t len) {
  if (!buf || len == 0) return PIMD_EINVAL;
  snprintf(buf, len, "%s", g_err.c_str());
  return PIMD_OK;
}

int pimd_ctx_create(const pimd_params *params, pimd_ctx **out) {
  if (!params || !out) return fail(PIMD_EINVAL, "null argument");
  *out = nullptr;
  pimd_params p = *params;
  if (p.pot_model < PIMD_TTM21F || p.pot_model > PIMD_SPCF)
    return fail(PIMD_EINVAL, "pot_model must be 2 (TTM2.1-F), 3 (TTM3-F), 4 (q-SPC/Fw) or "
                "5 (SPC/F), got %d", p.pot_model);
  if (p.nwaters < 2) return fail(PIMD_EINVAL, "need at least 2 waters, got %d", p.nwaters);
  if (p.nbeads < 1 || p.nbeads > 64) return fail(PIMD_EINVAL, "nbeads must be 1..64, got %d", p.nbeads);
  for (int k = 0; k < 3; ++k)
    if (!(p.box[k] > 0.0)) return fail(PIMD_EINVAL, "box must be positive");
  if (p.rc <= 0.0) p.rc = std::min(p.box[0], std::min(p.box[1], p.box[2])) / 2.0;
  if (p.rc1 < 0.0) p.rc1 = (double)f32(0.8) * p.rc;
  const double lmin = std::min(p.box[0], std::min(p.box[1], p.box[2]));
  if (p.rc > lmin / 2.0 * (1.0 + 1e-12))  // InputOutput.f90:277-281
    return fail(PIMD_EINVAL, "cut-off %.6f exceeds half the box (%.6f)", p.rc, lmin / 2.0);
  if (!(p.eps_ewald > 0.0 && p.eps_ewald < 1.0)) return fail(PIMD_EINVAL, "eps_ewald must be in (0,1)");
  if (p.polar_maxiter < 1) return fail(PIMD_EINVAL, "polar_maxiter must be >= 1");
  if (p.contraction != 0 && p.contraction != 1)
    return fail(PIMD_EINVAL, "contraction must be 0 or 1, got %d", p.contraction);
  // contracted MD takes its fast forces from the PS monomer surface
  // (MD.f90:318-395), which only the TTM models share
  if (p.contraction && p.pot_model >= PIMD_QSPCFW)
    return fail(PIMD_EINVAL, "contraction needs a TTM model (pot_model 2 or 3)");
  if (p.precision != PIMD_FP64 && p.precision != PIMD_MIXED)
    return fail(PIMD_EINVAL, "precision must be PIMD_FP64 (0) or PIMD_MIXED (1), got %d", p.precision);
  if (p.mass_o <= 0.0) p.mass_o = (double)f32(15.994);
  if (p.mass_h <= 0.0) p.mass_h = (double)f32(1.008);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PIMD_EDEVICE, "no HIP device available");
  if (p.device < 0 || p.device >= ndev) return fail(PIMD_EINVAL, "device %d out of range", p.device);
  if (hipSetDevice(p.device) != hipSuccess) return fail(PIMD_EDEVICE, "hipSetDevice failed");
  pimd_ctx *x = new pimd_ctx();
  x->p = p;
  x->natoms = 3 * p.nwaters;
  int rc = setup_system(x);
  // one stream per context (the second is made when first needed): each
  // context of a batch then gets a hardware queue of its own
  if (rc == PIMD_OK && hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(PIMD_EDEVICE, "hipStreamCreate failed");
  if (rc == PIMD_OK && (hipEventCreateWithFlags(&x->ev_fork, hipEventDisableTiming) != hipSuccess ||
                        hipEventCreateWithFlags(&x->ev_join, hipEventDisableTiming) != hipSuccess))
    rc = fail(PIMD_EDEVICE, "hipEventCreate failed");
  if (rc == PIMD_OK) rc = alloc_system(x);
  if (rc == PIMD_OK) {
    x->group = pick_group(x->c);
  }
  if (rc != PIMD_OK) {
    free_all(x);
    delete x;
    return rc;
  }
  *out = x;
  return PIMD_OK;
}

int pimd_ctx_destroy(pimd_ctx *x) {
  if (!x) return PIMD_OK;
  (void)hipSetDevice(x->p.device);
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  free_all(x);
  delete x;
  return PIMD_OK;
}

int pimd_ctx_info(const pimd_ctx *x, double *out) {
  if (!x || !out) return fail(PIMD_EINVAL, "null argument");
  const SysConst &c = x->c;
  out[0] = c.aewald;
  out[1] = c.Uvdw_lrc;
  out[2] = c.shiftA;
  out[3] = c.shiftB;
  for (int i = 0; i < 4; ++i) out[4 + i] = x->polfac[i];
  out[8] = c.nk;
  out[9] = c.hmax;
  out[10] = c.kmax;
  out[11] = c.lmax;
  out[12] = c.cells;
  return PIMD_OK;
}

int pimd_forces_all_beads_dev(pimd_ctx *x, const double *RR, const double *RRc, int t,
                              double *Upot, double *dRR, double *virt, double *virialc,
                              double *dipI, double *dipE, double *chg, int *iters) {
  if (!x || !RR || !RRc) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipSetDevice(x->p.device));
  DevBufs &d = x->d;
  launch_set_clock(d.clk, t, 0, x->stream);
  return force_call(x, RR, RRc, Upot ? Upot : d.Upot, dRR ? dRR : d.dRR,
                    virt ? virt : d.virt, virialc ? virialc : d.virialc, dipI ? dipI : d.dipI,
                    dipE ? dipE : d.dipE, chg ? chg : d.chg, iters ? iters : d.iters_out);
}

int pimd_forces_all_beads(pimd_ctx *x, const double *RRt, const double *RRc, int t,
                          double *Upott, double *dRRt, double *virt, double *virialct,
                          double *dipI, double *dipE, double *chg, int *iters) {
  if (!x || !RRt || !RRc) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipSetDevice(x->p.device));
  const size_t P = x->c.P, na = x->natoms, N = x->c.N;
  DevBufs &d = x->d;
  hipStream_t s = x->stream;
  // the caller's arrays go through a pinned staging block: DMA copies at
  // full link speed instead of the runtime's pageable path
  const size_t nR = P * na * 3, nc = na * 3;
  const size_t o_U = nR + nc, o_dR = o_U + P, o_vt = o_dR + nR, o_vc = o_vt + P * 9,
               o_dI = o_vc + P, o_dE = o_dI + P * N * 3, o_ch = o_dE + P * N * 3,
               o_it = o_ch + P * na, n_stage = o_it + P;
  if (!x->h_stage)
    HIPCHK(hipHostMalloc((void **)&x->h_stage, sizeof(double) * n_stage, hipHostMallocDefault));
  double *h = x->h_stage;
  std::memcpy(h, RRt, sizeof(double) * nR);
  std::memcpy(h + nR, RRc, sizeof(double) * nc);
  HIPCHK(hipMemcpyAsync(d.RR, h, sizeof(double) * nR, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.RRc, h + nR, sizeof(double) * nc, hipMemcpyHostToDevice, s));
  launch_set_clock(d.clk, t, 0, s);
  int rc = force_call(x, d.RR, d.RRc, d.Upot, d.dRR, d.virt, d.virialc, d.dipI, d.dipE, d.chg,
                      d.iters_out);
  if (rc != PIMD_OK) return rc;
  struct Out {
    void *dst;
    const void *src;
    size_t off, bytes;
  } outs[] = {{Upott, d.Upot, o_U, sizeof(double) * P},
              {dRRt, d.dRR, o_dR, sizeof(double) * nR},
              {virt, d.virt, o_vt, sizeof(double) * P * 9},
              {virialct, d.virialc, o_vc, sizeof(double) * P},
              {dipI, d.dipI, o_dI, sizeof(double) * P * N * 3},
              {dipE, d.dipE, o_dE, sizeof(double) * P * N * 3},
              {chg, d.chg, o_ch, sizeof(double) * P * na},
              {iters, d.iters_out, o_it, sizeof(int) * P}};
  for (const Out &o : outs)
    if (o.dst) HIPCHK(hipMemcpyAsync(h + o.off, o.src, o.bytes, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (const Out &o : outs)
    if (o.dst) std::memcpy(o.dst, h + o.off, o.bytes);
  return PIMD_OK;
}

int pimd_energy_terms(pimd_ctx *x, double *terms) {
  if (!x || !terms) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipMemcpyAsync(terms, x->d.terms, sizeof(double) * x->c.P * 5, hipMemcpyDeviceToHost,
                        x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return PIMD_OK;
}

// InitNormalModes (NormalModes.f90:68-218), on the host; uploads C and A.
int pimd_nm_init(pimd_ctx *x, double omegan, double delta, double setnm) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  const int P = x->c.P;
  const int l = (P % 2 == 0) ? (P - 2) / 2 : (P - 1) / 2;
  std::vector<double> C(P * P), A11(P), A12(P), A21(P), A22(P), om(P), msf(P, 1.0);
  for (int i = 1; i <= P; ++i) {
    C[0 * P + (i - 1)] = std::sqrt(1.0 / P);
    for (int k = 1; k <= l; ++k) {
      C[k * P + (i - 1)] = std::sqrt(2.0 / P) * std::cos((double)(k * i) * 2.0 * PI / P);
      C[(k + l) * P + (i - 1)] = std::sqrt(2.0 / P) * std::sin((double)(k * i) * 2.0 * PI / P);
    }
    if (P % 2 == 0) C[(2 * l + 1) * P + (i - 1)] = std::sqrt(1.0 / P) * ((i % 2) ? -1.0 : 1.0);
  }
  auto setmode = [&](int j, double w) {
    om[j] = w;
    A11[j] = std::cos(w * delta);
    A12[j] = (1 / w) * std::sin(w * delta);
    A21[j] = -w * std::sin(w * delta);
    A22[j] = std::cos(w * delta);
  };
  om[0] = 0.0; A11[0] = 1.0; A12[0] = delta; A21[0] = 0.0; A22[0] = 1.0;
  for (int k = 1; k <= l; ++k) {
    const double w = 2 * omegan * std::sin(PI * (double)(f32(k) / f32(P)));  // REAL(4) ratio
    setmode(k, w);
    setmode(k + l, w);
  }
  if (P % 2 == 0) setmode(2 * l + 1, 2 * omegan * std::sin(PI * (double)(f32(l + 1) / f32(P))));
  if (setnm != 0.0) {  // NormalModes.f90:168-206
    const double w = (2 * PI) * setnm / 33.33333;
    for (int i = 0; i < P; ++i) msf[i] = om[i] == 0 ? 1.0 : (om[i] / w) * (om[i] / w);
    const int last = (P % 2 == 0) ? 2 * l + 1 : 2 * l;
    for (int j = 1; j <= last; ++j) setmode(j, w);
  }
  x->C = C; x->A11 = A11; x->A12 = A12; x->A21 = A21; x->A22 = A22; x->om = om; x->msf = msf;
  std::vector<double> A(4 * P);
  for (int j = 0; j < P; ++j) {
    A[j] = A11[j];
    A[P + j] = A12[j];
    A[2 * P + j] = A21[j];
    A[3 * P + j] = A22[j];
  }
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipMemcpy(x->d.C, C.data(), sizeof(double) * P * P, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(x->d.A, A.data(), sizeof(double) * 4 * P, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(x->d.msf, msf.data(), sizeof(double) * P, hipMemcpyHostToDevice));
  x->nm_ready = true;
  return PIMD_OK;
}

int pimd_nm_get(pimd_ctx *x, double *C, double *A11, double *A12, double *A21, double *A22,
                double *om, double *msf) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  if (!x->nm_ready) return fail(PIMD_ESTATE, "pimd_nm_init has not been called");
  const size_t P = x->c.P;
  if (C) std::memcpy(C, x->C.data(), sizeof(double) * P * P);
  if (A11) std::memcpy(A11, x->A11.data(), sizeof(double) * P);
  if (A12) std::memcpy(A12, x->A12.data(), sizeof(double) * P);
  if (A21) std::memcpy(A21, x->A21.data(), sizeof(double) * P);
  if (A22) std::memcpy(A22, x->A22.data(), sizeof(double) * P);
  if (om) std::memcpy(om, x->om.data(), sizeof(double) * P);
  if (msf) std::memcpy(msf, x->msf.data(), sizeof(double) * P);
  return PIMD_OK;
}

int pimd_nm_evolve(pimd_ctx *x, double *RRt, double *PPt, const double *masses) {
  if (!x || !RRt || !PPt) return fail(PIMD_EINVAL, "null argument");
  if (!x->nm_ready) return fail(PIMD_ESTATE, "pimd_nm_init has not been called");
  HIPCHK(hipSetDevice(x->p.device));
  const size_t P = x->c.P, na = x->natoms, n = P * na * 3;
  hipStream_t s = x->stream;
  DevBufs &d = x->d;
  if (masses) HIPCHK(hipMemcpyAsync(d.masses, masses, sizeof(double) * na, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.RR, RRt, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.PP, PPt, sizeof(double) * n, hipMemcpyHostToDevice, s));
  x->timed("nm_evolve", [&] { launch_nm_evolve((int)P, (int)na, d, d.RR, d.PP, s); });
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(RRt, d.RR, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(PPt, d.PP, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return PIMD_OK;
}

// Init_Langevin_NM (Langevin.f90:68-98).
int pimd_langevin_init(pimd_ctx *x, double delt2, int centroid, double tau_bead, double temp) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  if (!x->nm_ready) return fail(PIMD_ESTATE, "pimd_nm_init has not been called");
  const int P = x->c.P;
  x->c1.assign(P, 0.0);
  x->c2.assign(P, 0.0);
  for (int i = 0; i < P; ++i) {
    const double g = i == 0 ? (centroid ? 1 / tau_bead : 0.0) : 2 * x->om[i];
    x->c1[i] = std::exp(-delt2 * g);
    x->c2[i] = std::sqrt(1 - x->c1[i] * x->c1[i]);
    x->c2[i] = std::sqrt(KB_AMU * temp) * x->c2[i];
  }
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipMemcpy(x->d.c1, x->c1.data(), sizeof(double) * P, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(x->d.c2, x->c2.data(), sizeof(double) * P, hipMemcpyHostToDevice));
  x->lang_ready = true;
  return PIMD_OK;
}

int pimd_langevin(pimd_ctx *x, double *PPt, const double *noise, uint64_t seed, uint64_t counter) {
  if (!x || !PPt) return fail(PIMD_EINVAL, "null argument");
  if (!x->lang_ready) return fail(PIMD_ESTATE, "pimd_langevin_init has not been called");
  HIPCHK(hipSetDevice(x->p.device));
  const size_t P = x->c.P, na = x->natoms, n = P * na * 3;
  hipStream_t s = x->stream;
  DevBufs &d = x->d;
  HIPCHK(hipMemcpyAsync(d.PP, PPt, sizeof(double) * n, hipMemcpyHostToDevice, s));
  if (noise) HIPCHK(hipMemcpyAsync(d.noise, noise, sizeof(double) * n, hipMemcpyHostToDevice, s));
  x->timed("langevin", [&] {
    launch_langevin((int)P, (int)na, d, d.PP, noise ? d.noise : nullptr, seed, counter,
                    x->p.mass_o, x->p.mass_h, s);
  });
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(PPt, d.PP, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return PIMD_OK;
}

int pimd_nose_hoover(double *s, double *uk, double *vxi, double tau, double delt2, int L,
                     double temp) {
  if (!s || !uk || !vxi) return fail(PIMD_EINVAL, "null argument");
  const double delt4 = delt2 / 2.0, delt8 = delt2 / 4.0;
  const double kT = KB_AMU * temp;
  const double M1 = L * kT * (tau * tau), M2 = M1;
  vxi[1] = vxi[1] + (M1 * (vxi[0] * vxi[0]) - kT) * delt4 / M2;
  vxi[0] = vxi[0] * std::exp(-vxi[1] * delt8);
  vxi[0] = vxi[0] + (2 * *uk - L * kT) * delt4 / M1;
  vxi[0] = vxi[0] * std::exp(-vxi[1] * delt8);
  *s = std::exp(-vxi[0] * delt2);
  *uk = *uk * *s * *s;
  vxi[0] = vxi[0] * std::exp(-vxi[1] * delt8);
  vxi[0] = vxi[0] + (2 * *uk - L * kT) * delt4 / M1;
  vxi[0] = vxi[0] * std::exp(-vxi[1] * delt8);
  vxi[1] = vxi[1] + (M1 * (vxi[0] * vxi[0]) - kT) * delt4 / M2;
  return PIMD_OK;
}

int pimd_md_init(pimd_ctx *x, const pimd_md_params *md) {
  if (!x || !md) return fail(PIMD_EINVAL, "null argument");
  if (!(md->delt_fs > 0.0) || !(md->temp > 0.0)) return fail(PIMD_EINVAL, "delt_fs and temp must be > 0");
  if (md->bead_thermostat != 0 && md->bead_thermostat != 1)
    return fail(PIMD_EINVAL, "bead_thermostat must be 0 (none) or 1 (Langevin)");
  if (md->intra_timesteps < 0 || md->intra_timesteps > 1000)
    return fail(PIMD_EINVAL, "intra_timesteps must be 0 (plain PIMD) or 1..1000, got %d",
                md->intra_timesteps);
  if (md->intra_timesteps > 0 && x->p.contraction)
    return fail(PIMD_EINVAL, "contracted MD needs a context created with contraction = 0");
  HIPCHK(hipSetDevice(x->p.device));
  invalidate_graphs(x);
  x->md = *md;
  const int P = x->c.P;
  const double delt = md->delt_fs / 1000.0, delt2 = delt / 2.0;
  const double omegan = KB_AMU * md->temp * P / HBAR;  // InputOutput.f90:140
  x->omegan = omegan;
  x->cmd = md->setNMfreq != 0.0 && P > 1;  // calc_uk's branch (main_stuff.f90:353)
  // contracted MD propagates the ring polymer with the fast step
  // delt / intra_timesteps (InputOutput.f90:379-385)
  const double delt_nm = md->intra_timesteps > 0 ? delt / md->intra_timesteps : delt;
  int rc = pimd_nm_init(x, omegan, delt_nm, md->setNMfreq);
  if (rc != PIMD_OK) return rc;
  if (x->cen) {
    (void)hipStreamSynchronize(x->stream);
    free_all(x->cen);
    delete x->cen;
    x->cen = nullptr;
  }
  if (md->intra_timesteps > 0) {
    // the centroid's intermolecular forces: one bead, monomer PES off
    // (pot_ttm.f90:92-132), same model, cut-offs, Ewald and SOR settings
    pimd_params pc = x->p;
    pc.nbeads = 1;
    pc.contraction = 1;
    pimd_ctx *cen = nullptr;
    if ((rc = pimd_ctx_create(&pc, &cen)) != PIMD_OK) return rc;
    (void)hipStreamDestroy(cen->stream);
    cen->stream = x->stream;
    cen->stream_borrowed = true;
    x->cen = cen;
    cen->d.clk = x->d.clk;  // the centroid's force call reads the step of this context's clock
    std::copy(x->ttm, x->ttm + TTM_NPARAM, cen->ttm);  // parameters set on this context
    apply_ttm_params(cen);
    const size_t n = (size_t)P * x->natoms * 3;
    if (!x->d.dRRfast) {
      if ((rc = dalloc(&x->d.dRRfast, n)) != PIMD_OK) return rc;
      if ((rc = dalloc(&x->d.mol, (size_t)P * x->c.N * 3)) != PIMD_OK) return rc;
    }
    // the fast forces start from zero (MD.f90:293)
    HIPCHK(hipMemset(x->d.dRRfast, 0, sizeof(double) * n));
  }
  if (md->bead_thermostat == 1) {
    rc = pimd_langevin_init(x, delt2, md->centroid_thermostat, md->tau_centroid, P * md->temp);
    if (rc != PIMD_OK) return rc;
  }
  const double vxi0[2] = {1.0, 1.0}, one = 1.0;  // vxi_global = 1 (InputOutput.f90:399)
  HIPCHK(hipMemcpy(x->d_vxi, vxi0, sizeof vxi0, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(x->d_s, &one, sizeof one, hipMemcpyHostToDevice));
  x->md_ready = true;
  return PIMD_OK;
}

int pimd_state_set(pimd_ctx *x, const double *RRt, const double *PPt, const double *dRRt) {
  if (!x || !RRt || !PPt) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipSetDevice(x->p.device));
  const size_t n = (size_t)x->c.P * x->natoms * 3;
  DevBufs &d = x->d;
  HIPCHK(hipMemcpy(d.RR, RRt, sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d.PP, PPt, sizeof(double) * n, hipMemcpyHostToDevice));
  if (dRRt)
    HIPCHK(hipMemcpy(d.dRR, dRRt, sizeof(double) * n, hipMemcpyHostToDevice));
  else
    HIPCHK(hipMemset(d.dRR, 0, sizeof(double) * n));
  // centroid for the first step's bookkeeping
  std::vector<double> rc(x->natoms * 3, 0.0);
  for (int k = 0; k < x->c.P; ++k)
    for (size_t i = 0; i < rc.size(); ++i) rc[i] += RRt[k * rc.size() + i];
  for (auto &v : rc) v /= x->c.P;
  HIPCHK(hipMemcpy(d.RRc, rc.data(), sizeof(double) * rc.size(), hipMemcpyHostToDevice));
  x->state_ready = true;
  return PIMD_OK;
}

int pimd_state_get(pimd_ctx *x, double *RRt, double *PPt, double *dRRt, double *RRc) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  const size_t n = (size_t)x->c.P * x->natoms * 3;
  DevBufs &d = x->d;
  if (RRt) HIPCHK(hipMemcpy(RRt, d.RR, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (PPt) HIPCHK(hipMemcpy(PPt, d.PP, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (dRRt) HIPCHK(hipMemcpy(dRRt, d.dRR, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (RRc) HIPCHK(hipMemcpy(RRc, d.RRc, sizeof(double) * x->natoms * 3, hipMemcpyDeviceToHost));
  return PIMD_OK;
}

}  // extern "C"

// Checks that a context is ready for device-resident MD steps.
static int md_check(pimd_ctx *x) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  if (!x->md_ready) return fail(PIMD_ESTATE, "pimd_md_init has not been called");
  if (!x->state_ready) return fail(PIMD_ESTATE, "pimd_state_set has not been called");
  return PIMD_OK;
}

// Device buffer for the per-step records of nsteps steps (grown, never shrunk).
static int md_records(pimd_ctx *x, int nsteps) {
  if (nsteps <= x->rec_cap) return PIMD_OK;
  if (x->d_rec) (void)hipFree(x->d_rec);
  x->d_rec = nullptr;
  x->rec_cap = 0;
  invalidate_graphs(x);  // captured steps hold the record buffer's address
  int rc = dalloc(&x->d_rec, (size_t)nsteps * MD_REC);
  if (rc == PIMD_OK) x->rec_cap = nsteps;
  return rc;
}

// calc_uk (main_stuff.f90:349-388) into d.uk.
static void md_kinetic(pimd_ctx *x) {
  DevBufs &d = x->d;
  x->timed("kinetic", [&] {
    if (x->cmd)
      launch_kinetic_nm(x->c.P, x->natoms, d, d.PP, d.uk, d.ukpart, x->stream);
    else
      launch_kinetic(x->c.P, x->natoms, d.masses, d.PP, d.uk, d.ukpart, x->stream);
  });
}

// Start of a run at MD step t0: the clock, and the kinetic energy of the
// loaded state (needed by the first thermostat call).
static void md_begin(pimd_ctx *x, int t0) {
  launch_set_clock(x->d.clk, t0, 0, x->stream);
  md_kinetic(x);
}

// write_out's observables of the step from its force record and the kinetic
// energy after the force call.  dipI is NULL when the step has no bead
// dipoles; Upot (with virt, virialc) is NULL when the step has written its
// record's first three entries already, else the kernel forms them.
static void md_observables(pimd_ctx *x, const double *dipI, const double *Upot,
                           const double *virt, const double *virialc) {
  DevBufs &d = x->d;
  const ObsArgs a{x->c.P, x->natoms, x->c.vol, x->omegan};
  x->timed("observables", [&] {
    launch_observables(a, d.RR, dipI, d.masses, d.uk, Upot, virt, virialc, d.obs_part, x->d_rec,
                       d.clk, x->stream);
  });
}

// Enqueues one PIMD_VV step (MD.f90:35-110) on the context's stream: the MD
// step and the record slot come from the device clock, which the step
// advances at its end, so the same launches serve every step (and can be
// captured once).  Records Upot, virial, virialc, uk and the observables.
// Does not synchronise.  noise: NULL or the normals of this step,
// [2][Natoms][P][3].
static int md_step_vv(pimd_ctx *x, const double *noise, uint64_t seed) {
  const int P = x->c.P, na = x->natoms;
  const size_t n = (size_t)P * na * 3;
  const pimd_md_params &md = x->md;
  const double delt = md.delt_fs / 1000.0, delt2 = delt / 2.0;
  const int L = 3 * na * P;
  DevBufs &d = x->d;
  hipStream_t s = x->stream;
  // the half kicks ride on the kernel that next reads the momenta: the free
  // ring-polymer step and the second Langevin half step
  const bool lang = md.bead_thermostat == 1;
  auto thermostat = [&](int half, const double *kick) {
    if (!lang) return;
    const double *nz = nullptr;
    if (noise) {
      (void)hipMemcpyAsync(d.noise, noise + (size_t)half * n, sizeof(double) * n,
                           hipMemcpyHostToDevice, s);
      nz = d.noise;
    }
    x->timed("langevin", [&] {
      launch_langevin(P, na, d, d.PP, nz, seed, 0, x->p.mass_o, x->p.mass_h, s, d.clk, half, kick,
                      delt2);
    });
  };
  thermostat(0, nullptr);
  if (md.global_nh) {
    x->timed("nose_hoover", [&] { launch_nose_hoover(d.uk, x->d_vxi, x->d_s, md.tau, delt2, L, md.temp * P, s); });
    x->timed("scale", [&] { launch_scale((int)n, d.PP, x->d_s, s); });
  }
  if (P > 1) {
    x->timed("nm_evolve", [&] { launch_nm_evolve(P, na, d, d.RR, d.PP, s, d.dRR, delt2); });
  } else {
    x->timed("kick", [&] { launch_kick((int)n, d.PP, d.dRR, delt2, s); });
    x->timed("free_flight", [&] { launch_free_flight((int)n, d.RR, d.PP, d.masses, delt, s); });
  }
  x->timed("centroid_pbc", [&] { launch_centroid_pbc(P, na, x->p.box[0], 1.0 / x->p.box[0], d.RR, d.RRc, s); });
  int rc = run_forces(x, d.RR, d.RRc, d.Upot, d.dRR, d.virt, d.virialc, d.dipI, d.dipE, d.chg,
                      d.iters_out);
  if (rc != PIMD_OK) return rc;
  md_kinetic(x);
  md_observables(x, d.dipI, d.Upot, d.virt, d.virialc);  // also the step record
  if (lang)
    thermostat(1, d.dRR);
  else
    x->timed("kick", [&] { launch_kick((int)n, d.PP, d.dRR, delt2, s); });
  md_kinetic(x);
  if (md.global_nh) {
    x->timed("nose_hoover", [&] { launch_nose_hoover(d.uk, x->d_vxi, x->d_s, md.tau, delt2, L, md.temp * P, s); });
    x->timed("scale", [&] { launch_scale((int)n, d.PP, x->d_s, s); });
  }
  launch_advance_clock(d.clk, s, d.uk, x->d_rec);  // with the record's kinetic energy
  HIPCHK(hipGetLastError());
  return PIMD_OK;
}

// Enqueues one contracted_MD step (MD.f90:276-449; clock as md_step_vv): the
// monomer PES moves every bead in intra_timesteps inner steps of the ring
// polymer, the intermolecular forces of the centroid act on every bead once
// per step.  Records sum of bead energies, virial, centroid virial and uk.
// The fast gradient is kept from one step to the next.
static int md_step_contracted(pimd_ctx *x, const double *noise, uint64_t seed) {
  const int P = x->c.P, na = x->natoms, N = x->c.N;
  const size_t n = (size_t)P * na * 3;
  const pimd_md_params &md = x->md;
  const int intra = md.intra_timesteps;
  const double delt = md.delt_fs / 1000.0, delt2 = delt / 2.0;
  const double deltfast = delt / intra, delt2fast = deltfast / 2.0;
  const int L = 3 * na * P;
  DevBufs &d = x->d;
  pimd_ctx *cen = x->cen;
  hipStream_t s = x->stream;
  const bool lang = md.bead_thermostat == 1;
  auto thermostat = [&](int half, const double *kick) {
    if (!lang) return;
    const double *nz = nullptr;
    if (noise) {
      (void)hipMemcpyAsync(d.noise, noise + (size_t)half * n, sizeof(double) * n,
                           hipMemcpyHostToDevice, s);
      nz = d.noise;
    }
    x->timed("langevin", [&] {
      launch_langevin(P, na, d, d.PP, nz, seed, 0, x->p.mass_o, x->p.mass_h, s, d.clk, half, kick,
                      delt2);
    });
  };
  auto centroid = [&] {
    x->timed("centroid_pbc",
             [&] { launch_centroid_pbc(P, na, x->p.box[0], 1.0 / x->p.box[0], d.RR, d.RRc, s); });
  };
  thermostat(0, nullptr);
  if (md.global_nh) {
    x->timed("nose_hoover", [&] { launch_nose_hoover(d.uk, x->d_vxi, x->d_s, md.tau, delt2, L, md.temp * P, s); });
    x->timed("scale", [&] { launch_scale((int)n, d.PP, x->d_s, s); });
  }
  centroid();  // reads and moves coordinates only: the kicks below commute with it
  for (int ti = 1; ti <= intra; ++ti) {
    if (P > 1) {
      // the kicks before each free ring-polymer step ride on it: the outer
      // half kick and the first fast one, then the previous inner step's
      // closing fast kick and this step's opening one (the same gradient)
      x->timed("nm_evolve", [&] {
        if (ti == 1)
          launch_nm_evolve(P, na, d, d.RR, d.PP, s, d.dRR, delt2, d.dRRfast, delt2fast);
        else
          launch_nm_evolve(P, na, d, d.RR, d.PP, s, d.dRRfast, delt2fast, d.dRRfast, delt2fast);
      });
    } else {
      if (ti == 1) x->timed("kick", [&] { launch_kick((int)n, d.PP, d.dRR, delt2, s); });
      else x->timed("kick_fast", [&] { launch_kick((int)n, d.PP, d.dRRfast, delt2fast, s); });
      x->timed("kick_fast", [&] { launch_kick((int)n, d.PP, d.dRRfast, delt2fast, s); });
      x->timed("free_flight", [&] { launch_free_flight((int)n, d.RR, d.PP, d.masses, deltfast, s); });
    }
    x->timed("pes_fast", [&] { launch_pes_fast(x->c, d, d.RR, d.RRc, d.dRRfast, d.mol, ti == intra, s); });
  }
  x->timed("kick_fast", [&] { launch_kick((int)n, d.PP, d.dRRfast, delt2fast, s); });
  centroid();
  int rc = PIMD_OK;
  DevBufs &dc = cen->d;
  x->timed("centroid_forces", [&] {
    rc = run_forces(cen, d.RRc, d.RRc, dc.Upot, dc.dRR, dc.virt, dc.virialc, dc.dipI, dc.dipE,
                    dc.chg, dc.iters_out);
  });
  if (rc != PIMD_OK) return rc;
  x->timed("broadcast", [&] { launch_broadcast_beads(P, na * 3, dc.dRR, d.dRR, s); });
  md_kinetic(x);
  launch_contract_record(P, N, d.mol, dc.Upot, dc.virt, dc.virialc, d.uk, x->d_rec, d.clk, s);
  md_observables(x, nullptr, nullptr, nullptr, nullptr);  // no per-bead dipoles here
  if (lang)
    thermostat(1, d.dRR);  // with the outer half kick
  else
    x->timed("kick", [&] { launch_kick((int)n, d.PP, d.dRR, delt2, s); });
  md_kinetic(x);
  if (md.global_nh) {
    x->timed("nose_hoover", [&] { launch_nose_hoover(d.uk, x->d_vxi, x->d_s, md.tau, delt2, L, md.temp * P, s); });
    x->timed("scale", [&] { launch_scale((int)n, d.PP, x->d_s, s); });
  }
  launch_advance_clock(d.clk, s, d.uk, x->d_rec);  // with the record's kinetic energy
  HIPCHK(hipGetLastError());
  return PIMD_OK;
}

static int md_step_eager(pimd_ctx *x, const double *noise, uint64_t seed) {
  return x->md.intra_timesteps > 0 ? md_step_contracted(x, noise, seed)
                                   : md_step_vv(x, noise, seed);
}

// A step with device normals can run as one replay of a captured graph of
// md_step_eager, rebuilt when the launch arguments it bakes in change (seed,
// launch shapes, parameters, buffers).  Host noise, polling and per-kernel
// timing keep the eager launches.
static bool graph_eligible(const pimd_ctx *x, const double *noise) {
  const int on = env_int("PIMD_MD_GRAPH", 1, 0, 1);  // read per step: a test switches it
  const SysConst &c = x->c;
  const bool cen_ok = !x->cen || (!x->cen->c.poll && !x->cen->prof);
  return on && !noise && !x->prof && !c.poll && cen_ok;
}

static int md_graph_build(pimd_ctx *x, uint64_t seed) {
  if (x->graph) (void)hipGraphExecDestroy(x->graph);
  x->graph = nullptr;
  x->graph_ok = false;
  hipStream_t s = x->stream;
  // PIMD_GRAPH_BRANCHES=1: force calls inside the step split into two graph
  // branches where their real- and reciprocal-space kernels are independent.
  // Off by default: on MI355X it gained 4 % on one FP64 system, nothing in
  // the mixed form, and slowed batches of replicas by a third.
  const bool br = env_int("PIMD_GRAPH_BRANCHES", 0, 0, 1) != 0;
  for (pimd_ctx *y : {x, x->cen}) {
    if (!y || !(br || y->c.overlap)) continue;
    const int err = ensure_stream2(y);
    if (err != PIMD_OK) return err;
  }
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  x->branches = br;
  if (x->cen) x->cen->branches = br;
  const int rc = md_step_eager(x, nullptr, seed);
  x->branches = false;
  if (x->cen) x->cen->branches = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (rc != PIMD_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(PIMD_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
  const hipError_t e2 = hipGraphInstantiate(&x->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e2 != hipSuccess) {
    x->graph = nullptr;
    return fail(PIMD_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e2));
  }
  x->graph_ok = true;
  x->graph_seed = seed;
  return PIMD_OK;
}

static int md_step(pimd_ctx *x, const double *noise, uint64_t seed) {
  if (x->graph_off || !graph_eligible(x, noise)) return md_step_eager(x, noise, seed);
  if (!x->graph_ok || x->graph_seed != seed) {
    if (md_graph_build(x, seed) != PIMD_OK) {
      // capture is not available here: launch eagerly from now on (an error
      // of the step itself shows again in the eager launch)
      x->graph_off = true;
      (void)hipGetLastError();
      return md_step_eager(x, noise, seed);
    }
  }
  HIPCHK(hipGraphLaunch(x->graph, x->stream));
  return PIMD_OK;
}

// Copies nsteps records to the host arrays (any may be NULL); the stream must
// be synchronised.
static int md_copy_records(pimd_ctx *x, int nsteps, double *Upot, double *virial, double *virialc,
                           double *uk) {
  x->rec_steps = std::max(nsteps, 0);
  if (nsteps <= 0) return PIMD_OK;
  std::vector<double> rec((size_t)nsteps * MD_REC);
  HIPCHK(hipMemcpy(rec.data(), x->d_rec, sizeof(double) * nsteps * MD_REC, hipMemcpyDeviceToHost));
  for (int st = 0; st < nsteps; ++st) {
    if (Upot) Upot[st] = rec[st * MD_REC];
    if (virial) virial[st] = rec[st * MD_REC + 1];
    if (virialc) virialc[st] = rec[st * MD_REC + 2];
    if (uk) uk[st] = rec[st * MD_REC + 3];
  }
  return PIMD_OK;
}

extern "C" int pimd_md_observables(pimd_ctx *x, int nsteps, double *obs) {
  if (!x || !obs) return fail(PIMD_EINVAL, "null argument");
  if (nsteps < 0 || nsteps > x->rec_steps)
    return fail(PIMD_EINVAL, "the last run recorded %d steps, asked for %d", x->rec_steps, nsteps);
  if (nsteps == 0) return PIMD_OK;
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  std::vector<double> rec((size_t)nsteps * MD_REC);
  HIPCHK(hipMemcpy(rec.data(), x->d_rec, sizeof(double) * rec.size(), hipMemcpyDeviceToHost));
  for (int st = 0; st < nsteps; ++st)
    for (int k = 0; k < PIMD_NOBS; ++k) obs[(size_t)st * PIMD_NOBS + k] = rec[(size_t)st * MD_REC + 4 + k];
  return PIMD_OK;
}

extern "C" int pimd_md_run(pimd_ctx *x, int t0, int nsteps, const double *noise, uint64_t seed,
                           double *Upot, double *virial, double *virialc, double *uk) {
  int rc = md_check(x);
  if (rc != PIMD_OK) return rc;
  if (nsteps < 0) return fail(PIMD_EINVAL, "nsteps must be >= 0");
  HIPCHK(hipSetDevice(x->p.device));
  if ((rc = md_records(x, nsteps)) != PIMD_OK) return rc;
  const size_t n = (size_t)x->c.P * x->natoms * 3;
  md_begin(x, t0);
  for (int st = 0; st < nsteps; ++st)
    if ((rc = md_step(x, noise ? noise + (size_t)st * 2 * n : nullptr, seed)) != PIMD_OK)
      return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  return md_copy_records(x, nsteps, Upot, virial, virialc, uk);
}

// ---------------------------------------------------------------------------
// Batches of independent systems (replicas) run side by side: each context
// keeps its own stream, the host enqueues step st of every replica before
// step st + 1 of any, and small replicas fill the device together.
// ---------------------------------------------------------------------------
struct pimd_batch {
  std::vector<pimd_ctx *> ctx;
};

// Launch shapes of one context: lanes per molecule of the pair kernels and of
// the LDS-staged dipole kernel, column groups of the field kernels and sites
// per structure-factor chunk.
struct Shape {
  int g_qq, g_dd, g_fin, fd_cg, sf_sites;
};
constexpr int NSHAPE = 5;

static Shape get_shape(const pimd_ctx *x) {
  return Shape{x->c.g_qq, x->c.g_dd, x->c.g_fin, x->c.fd_cg, x->c.sf_sites};
}
static void set_shape(pimd_ctx *x, const Shape &sh) {
  invalidate_graphs(x);
  x->c.g_qq = sh.g_qq;
  x->c.g_dd = sh.g_dd;
  x->c.g_fin = sh.g_fin;
  x->c.fd_cg = sh.fd_cg;
  x->c.sf_sites = sh.sf_sites;
}

// Mean time of one all-beads force evaluation on the loaded coordinates (ms),
// writing the gradient to scratch so the MD state is untouched.
static int time_forces(pimd_ctx *x, double *scratch, hipEvent_t e0, hipEvent_t e1, int reps,
                       double &ms) {
  DevBufs &d = x->d;
  hipStream_t s = x->stream;
  launch_set_clock(d.clk, 1, 0, s);
  int rc = run_forces(x, d.RR, d.RRc, d.Upot, scratch, d.virt, d.virialc, d.dipI, d.dipE,
                      d.chg, d.iters_out);
  if (rc != PIMD_OK) return rc;
  HIPCHK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r)
    if ((rc = run_forces(x, d.RR, d.RRc, d.Upot, scratch, d.virt, d.virialc, d.dipI, d.dipE,
                         d.chg, d.iters_out)) != PIMD_OK)
      return rc;
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float f = 0.f;
  HIPCHK(hipEventElapsedTime(&f, e0, e1));
  ms = (double)f / reps;
  return PIMD_OK;
}

// Coordinate search over the launch shapes of one context, timing force
// evaluations of its current coordinates.  Results do not depend on the
// shapes beyond summation order.  The dipole predictor history is saved and
// restored around the trials.
static int autotune_ctx(pimd_ctx *x, double *best_ms) {
  HIPCHK(hipSetDevice(x->p.device));
  invalidate_graphs(x);  // the shapes of a contracted context's centroid change below
  if (x->cen) {
    // contracted MD: the force call of a step is the centroid's
    const size_t nc = (size_t)x->natoms * 3;
    HIPCHK(hipMemcpyAsync(x->cen->d.RR, x->d.RRc, sizeof(double) * nc, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(x->cen->d.RRc, x->d.RRc, sizeof(double) * nc, hipMemcpyDeviceToDevice, x->stream));
    x = x->cen;
  }
  DevBufs &d = x->d;
  const SysConst &c = x->c;
  const size_t nh = (size_t)c.P * 4 * c.N * c.nd * 3, ng = (size_t)c.P * x->natoms * 3;
  double *hist_save = nullptr, *scratch = nullptr;
  int rc = dalloc(&hist_save, nh);
  if (rc == PIMD_OK) rc = dalloc(&scratch, ng);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == PIMD_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(PIMD_EDEVICE, "hipEventCreate failed");
  const bool prof = x->prof;
  x->prof = false;
  if (rc == PIMD_OK &&
      hipMemcpyAsync(hist_save, d.hist, sizeof(double) * nh, hipMemcpyDeviceToDevice, x->stream) !=
          hipSuccess)
    rc = fail(PIMD_EDEVICE, "hipMemcpyAsync failed");
  Shape best = get_shape(x);
  double tbest = 0.0;
  if (rc == PIMD_OK) rc = time_forces(x, scratch, e0, e1, 2, tbest);
  // candidate values per knob, in the order of Shape
  // (0: no further candidate)
  const int cand[NSHAPE][5] = {{4, 8, 16, 32, 0}, {8, 16, 32, 64, 0}, {4, 8, 16, 32, 0},
                               {4, 8, 16, 0, 0}, {4, 8, 16, 32, 64}};
  for (int knob = 0; knob < NSHAPE && rc == PIMD_OK; ++knob) {
    const int *vals = cand[knob];
    for (int v = 0; v < 5 && vals[v] != 0 && rc == PIMD_OK; ++v) {
      Shape sh = best;
      int *fields[NSHAPE] = {&sh.g_qq, &sh.g_dd, &sh.g_fin, &sh.fd_cg, &sh.sf_sites};
      int &field = *fields[knob];
      if (field == vals[v]) continue;
      field = vals[v];
      set_shape(x, sh);
      double t = 0.0;
      rc = time_forces(x, scratch, e0, e1, 2, t);
      if (rc == PIMD_OK && t < tbest) {
        tbest = t;
        best = sh;
      }
    }
  }
  set_shape(x, best);
  x->prof = prof;
  if (hist_save) {
    (void)hipMemcpyAsync(d.hist, hist_save, sizeof(double) * nh, hipMemcpyDeviceToDevice, x->stream);
    (void)hipStreamSynchronize(x->stream);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (hist_save) (void)hipFree(hist_save);
  if (scratch) (void)hipFree(scratch);
  if (best_ms) *best_ms = tbest;
  return rc;
}

extern "C" int pimd_batch_create(pimd_ctx *const *ctxs, int n, pimd_batch **out) {
  if (!ctxs || !out || n < 1) return fail(PIMD_EINVAL, "need at least one context");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return fail(PIMD_EINVAL, "context %d is null", i);
    if (ctxs[i]->p.device != ctxs[0]->p.device)
      return fail(PIMD_EINVAL, "all contexts of a batch must be on one device");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return fail(PIMD_EINVAL, "context %d appears twice", i);
  }
  pimd_batch *b = new pimd_batch();
  b->ctx.assign(ctxs, ctxs + n);
  *out = b;
  return PIMD_OK;
}

extern "C" int pimd_batch_destroy(pimd_batch *b) {
  delete b;
  return PIMD_OK;
}

extern "C" int pimd_batch_autotune(pimd_batch *b, double *ms) {
  if (!b) return fail(PIMD_EINVAL, "null argument");
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    pimd_ctx *x = b->ctx[i];
    if (!x->state_ready) return fail(PIMD_ESTATE, "replica %zu: pimd_state_set has not been called", i);
    int rc = autotune_ctx(x, ms ? ms + i : nullptr);
    if (rc != PIMD_OK) return rc;
  }
  return PIMD_OK;
}

extern "C" int pimd_batch_shape(pimd_batch *b, int i, int *shape) {
  if (!b || !shape || i < 0 || i >= (int)b->ctx.size()) return fail(PIMD_EINVAL, "bad argument");
  const pimd_ctx *x = b->ctx[i];
  const Shape sh = get_shape(x->cen ? x->cen : x);  // contracted: the centroid's force call
  shape[0] = sh.g_qq;
  shape[1] = sh.g_dd;
  shape[2] = sh.g_fin;
  shape[3] = sh.fd_cg;
  shape[4] = sh.sf_sites;
  return PIMD_OK;
}

extern "C" int pimd_batch_md_run(pimd_batch *b, int t0, int nsteps, const uint64_t *seeds,
                                 double *Upot, double *uk) {
  if (!b || !seeds) return fail(PIMD_EINVAL, "null argument");
  if (nsteps < 0) return fail(PIMD_EINVAL, "nsteps must be >= 0");
  const int n = (int)b->ctx.size();
  int rc;
  HIPCHK(hipSetDevice(b->ctx[0]->p.device));
  for (int r = 0; r < n; ++r) {
    if ((rc = md_check(b->ctx[r])) != PIMD_OK) return rc;
    if ((rc = md_records(b->ctx[r], nsteps)) != PIMD_OK) return rc;
  }
  for (int r = 0; r < n; ++r) md_begin(b->ctx[r], t0);
  for (int st = 0; st < nsteps; ++st)
    for (int r = 0; r < n; ++r)
      if ((rc = md_step(b->ctx[r], nullptr, seeds[r])) != PIMD_OK) return rc;
  for (int r = 0; r < n; ++r) HIPCHK(hipStreamSynchronize(b->ctx[r]->stream));
  for (int r = 0; r < n; ++r)
    if ((rc = md_copy_records(b->ctx[r], nsteps, Upot ? Upot + (size_t)r * nsteps : nullptr,
                              nullptr, nullptr, uk ? uk + (size_t)r * nsteps : nullptr)) != PIMD_OK)
      return rc;
  return PIMD_OK;
}

// ---------------------------------------------------------------------------
// PES parameters and their adjoint (for fitting the monomer surface).
// ---------------------------------------------------------------------------
extern "C" int pimd_pes_params_get(pimd_ctx *x, double *theta) {
  if (!x || !theta) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipMemcpy(theta, x->d.pes, sizeof(double) * PES_NPARAM, hipMemcpyDeviceToHost));
  return PIMD_OK;
}

extern "C" int pimd_pes_params_set(pimd_ctx *x, const double *theta) {
  if (!x || !theta) return fail(PIMD_EINVAL, "null argument");
  for (int j = 0; j < PES_NPARAM; ++j)
    if (!std::isfinite(theta[j])) return fail(PIMD_EINVAL, "PES parameter %d is not finite", j);
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(x->d.pes, theta, sizeof(double) * PES_NPARAM, hipMemcpyHostToDevice));
  return PIMD_OK;
}

extern "C" int pimd_pes_param_grad(pimd_ctx *x, const double *RRt, const double *ew,
                                   const double *gw, double *grad) {
  if (!x || !RRt || !ew || !gw || !grad) return fail(PIMD_EINVAL, "null argument");
  if (x->p.pot_model >= PIMD_QSPCFW)
    return fail(PIMD_EINVAL, "the PS monomer surface belongs to the TTM models (pot_model 2, 3)");
  HIPCHK(hipSetDevice(x->p.device));
  const SysConst &c = x->c;
  DevBufs &d = x->d;
  hipStream_t s = x->stream;
  const size_t n = (size_t)c.P * x->natoms * 3;
  const int nb = pes_adjoint_blocks(c);
  double *dR = d.adj, *dG = d.adj + n, *dE = d.adj + 2 * n, *part = dE + c.P;
  double *dgrad = part + (size_t)nb * PES_NPARAM;
  HIPCHK(hipMemcpyAsync(dR, RRt, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dG, gw, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dE, ew, sizeof(double) * c.P, hipMemcpyHostToDevice, s));
  x->timed("pes_adjoint", [&] { launch_pes_adjoint(c, d, dR, dE, dG, part, nb, dgrad, s); });
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(grad, dgrad, sizeof(double) * PES_NPARAM, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return PIMD_OK;
}

extern "C" int pimd_ttm_params_get(pimd_ctx *x, double *theta) {
  if (!x || !theta) return fail(PIMD_EINVAL, "null argument");
  std::copy(x->ttm, x->ttm + TTM_NPARAM, theta);
  return PIMD_OK;
}

extern "C" int pimd_ttm_params_set(pimd_ctx *x, const double *theta) {
  if (!x || !theta) return fail(PIMD_EINVAL, "null argument");
  for (int i = 0; i < TTM_NPARAM; ++i)
    if (!std::isfinite(theta[i])) return fail(PIMD_EINVAL, "theta[%d] is not finite", i);
  for (int i : {TP_POLFACO, TP_POLFACH, TP_POLFACM})
    if (!(theta[i] > 0.0)) return fail(PIMD_EINVAL, "polfac parameters must be > 0 (theta[%d])", i);
  for (int i : {TP_POLARO, TP_POLARH, TP_POLARM, TP_ADD, TP_ACC})
    if (theta[i] < 0.0) return fail(PIMD_EINVAL, "polarisabilities and Thole factors must be >= 0 (theta[%d])", i);
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipStreamSynchronize(x->stream));  // launches already queued keep the old values
  for (pimd_ctx *y : {x, x->cen}) {
    if (!y) continue;
    std::copy(theta, theta + TTM_NPARAM, y->ttm);
    apply_ttm_params(y);
  }
  invalidate_graphs(x);  // captured launches hold the model constants
  return PIMD_OK;
}

extern "C" int pimd_ttm_param_grad(pimd_ctx *x, const double *RRt, const double *ew,
                                   const double *gw, double *grad, double *E, double *dEdtheta) {
  if (!x || !RRt || !ew || !gw || !grad) return fail(PIMD_EINVAL, "null argument");
  if (x->p.contraction)
    return fail(PIMD_EINVAL, "parameter derivatives need a context with contraction = 0");
  if (x->p.pot_model >= PIMD_QSPCFW)
    return fail(PIMD_EINVAL, "TTM parameter derivatives need a TTM model (pot_model 2, 3)");
  HIPCHK(hipSetDevice(x->p.device));
  const SysConst &c = x->c;
  DevBufs &d = x->d;
  hipStream_t s = x->stream;
  const size_t P = (size_t)c.P, n = P * x->natoms * 3, nout = P * TTM_NPARAM * 4;
  const size_t ws = tangent_ws_duals(c) * P * TTM_NPARAM * 2, cfg = tangent_cfg_doubles(c) * P;
  const size_t bytes = sizeof(double) * (ws + cfg + 2 * n + P + nout);
  if (bytes > ((size_t)32 << 30))
    return fail(PIMD_EINVAL, "parameter derivatives of %zu configurations need %.1f GB of workspace; "
                "split them into smaller contexts", P, bytes / 1e9);
  if (bytes > d.tan_bytes) {
    if (d.tan) (void)hipFree(d.tan);
    d.tan = nullptr;
    d.tan_bytes = 0;
    int rc = dalloc((char **)&d.tan, bytes);
    if (rc != PIMD_OK) return rc;
    d.tan_bytes = bytes;
  }
  TanLaunch L;
  L.ws = d.tan;  // dual workspace first: 16-byte aligned
  L.cfg = d.tan + ws;
  double *dRR = L.cfg + cfg, *dgw = dRR + n, *dew = dgw + n;
  L.out = dew + P;
  L.RR = dRR;
  L.gw = dgw;
  L.ew = dew;
  L.pes = d.pes;
  L.kv = d.kv;
  L.khkl = d.khkl;
  L.dG = nullptr;
  std::copy(x->ttm, x->ttm + TTM_NPARAM, L.theta);
  const double Rc = x->p.rc, pre = TWOPI * (double)(c.N * c.N) / c.vol;
  L.lrc_coef[0] = pre / (9.0 * std::pow(Rc, 9));
  L.lrc_coef[1] = pre / (7.0 * std::pow(Rc, 7));
  L.lrc_coef[2] = pre / (3.0 * (Rc * Rc * Rc));
  HIPCHK(hipMemcpyAsync(dRR, RRt, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dgw, gw, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dew, ew, sizeof(double) * P, hipMemcpyHostToDevice, s));
  x->timed("ttm_tangent", [&] { launch_ttm_tangent(c, L, s); });
  HIPCHK(hipGetLastError());
  std::vector<double> out(nout);
  HIPCHK(hipMemcpyAsync(out.data(), L.out, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int j = 0; j < TTM_NPARAM; ++j) {
    double g = 0.0;
    for (size_t b = 0; b < P; ++b) g += out[(b * TTM_NPARAM + j) * 4 + 2];
    grad[j] = g;
  }
  for (size_t b = 0; b < P; ++b) {
    if (E) E[b] = out[(b * TTM_NPARAM) * 4];
    if (dEdtheta)
      for (int j = 0; j < TTM_NPARAM; ++j) dEdtheta[b * TTM_NPARAM + j] = out[(b * TTM_NPARAM + j) * 4 + 1];
  }
  return PIMD_OK;
}

extern "C" int pimd_state_forces(pimd_ctx *x, int t) {
  int rc = md_check(x);
  if (rc != PIMD_OK) return rc;
  HIPCHK(hipSetDevice(x->p.device));
  DevBufs &d = x->d;
  launch_set_clock(d.clk, t, 0, x->stream);
  return run_forces(x, d.RR, d.RRc, d.Upot, d.dRR, d.virt, d.virialc, d.dipI, d.dipE, d.chg,
                    d.iters_out);
}

__global__ void k_copy_probe(const double4 *__restrict__ a, double4 *__restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

extern "C" int pimd_profile_copy(pimd_ctx *x, size_t bytes, double *ms) {
  if (!x || bytes < 32) return fail(PIMD_EINVAL, "need a context and at least 32 bytes");
  HIPCHK(hipSetDevice(x->p.device));
  const size_t n = bytes / 32;
  double4 *a = nullptr, *b = nullptr;
  int rc = dalloc(&a, n);
  if (rc == PIMD_OK) rc = dalloc(&b, n);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == PIMD_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(PIMD_EDEVICE, "hipEventCreate failed");
  if (rc == PIMD_OK) {
    hipStream_t s = x->stream;
    (void)hipMemsetAsync(a, 0, n * 32, s);
    (void)hipEventRecord(e0, s);
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, s, a, b, n);
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) rc = fail(PIMD_EDEVICE, "copy probe failed");
    float f = 0.f;
    (void)hipEventElapsedTime(&f, e0, e1);
    if (ms) *ms = f;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

extern "C" {

int pimd_profile_enable(pimd_ctx *x, int on) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  x->prof = on != 0;
  return PIMD_OK;
}

int pimd_profile_count(pimd_ctx *x) {
  if (!x) return 0;
  x->resolve();
  return (int)x->names.size();
}

int pimd_profile_get(pimd_ctx *x, int i, char *name, size_t len, long *launches, double *total_ms) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  x->resolve();
  if (i < 0 || i >= (int)x->names.size()) return fail(PIMD_EINVAL, "profile index out of range");
  if (name && len) snprintf(name, len, "%s", x->names[i].c_str());
  if (launches) *launches = x->launches[i];
  if (total_ms) *total_ms = x->total_ms[i];
  return PIMD_OK;
}

int pimd_profile_reset(pimd_ctx *x) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  x->resolve();
  for (auto &v : x->launches) v = 0;
  for (auto &v : x->total_ms) v = 0.0;
  return PIMD_OK;
}

int pimd_sync(pimd_ctx *x) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(x->stream));
  return PIMD_OK;
}

}  // extern "C"
