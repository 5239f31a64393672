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
#include <mutex>
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
  void *arena = nullptr;  // the device block that holds the buffers of d
  hipStream_t stream = nullptr;
  bool stream_borrowed = false;  // the stream belongs to the context that owns this one
  // contracted MD: the centroid's intermolecular forces (one bead, no
  // monomer PES), computed on this context's stream
  pimd_ctx *cen = nullptr;
  // this context as a replica: its device record (in the arena), the
  // ring-polymer constants of its MD step, and the group of one replica
  // that every launch of a call on the context covers
  Rep *rep_dev = nullptr;
  MDConst m{};
  Group self;
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
  // the force call of the all-beads entry points, captured once and
  // replayed while the model stays the same
  hipGraphExec_t fgraph = nullptr;
  bool fgraph_ok = false, fgraph_off = false;
  double omegan = 0.0;    // ring frequency kB T P / hbar of the MD temperature
  double *d_rec = nullptr;  // per-step records of md runs
  int rec_cap = 0;
  int rec_steps = 0;        // steps recorded by the last run
  int *h_nactive = nullptr;
  // pinned staging of the host-array force call: coordinates in, all outputs
  // out (made on the first call)
  double *h_stage = nullptr;
  bool fab_pending = false;  // a host-array force call started, not finished
  // a parameter-derivative sweep started and not yet finished: its pinned
  // inputs and outputs, and what finishing it needs
  double *h_tan = nullptr;
  size_t h_tan_bytes = 0;
  int tan_nparam = 0;  // parameters pimd_ttm_adjoint_finish returns
  struct TanPending {
    bool on = false;
    int mode = 0, ndir = 0;
    size_t nout = 0, nadj = 0;
    std::vector<double> ew;
    double *out = nullptr;  // in h_tan
  } tanp;
  // profiling
  bool prof = false;
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
  // a counter among the timed entries (no time): k more of n
  void count(const char *n, long k) {
    if (prof) launches[kid(n)] += k;
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
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, n * sizeof(T));
  if (e != hipSuccess) return fail(PIMD_ENOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T),
                                   hipGetErrorString(e));
  *p = (T *)q;  // a C cast: T may carry the kernels' address space
  return PIMD_OK;
}

void free_all(pimd_ctx *x) {
  DevBufs &d = x->d;
  if (x->cen) {
    free_all(x->cen);
    delete x->cen;
    x->cen = nullptr;
  }
  // one block holds the buffers of layout_system; these are allocated apart
  void *ptrs[] = {x->arena, d.dRRfast, d.mol, d.tan};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  x->arena = nullptr;
  d = DevBufs{};
  if (x->h_nactive) (void)hipHostFree(x->h_nactive);
  if (x->h_stage) (void)hipHostFree(x->h_stage);
  if (x->h_tan) (void)hipHostFree(x->h_tan);
  x->h_tan = nullptr;
  x->h_tan_bytes = 0;
  for (auto &q : x->pending) {
    (void)hipEventDestroy(q.a);
    (void)hipEventDestroy(q.b);
  }
  for (auto e : x->pool) (void)hipEventDestroy(e);
  if (x->graph) (void)hipGraphExecDestroy(x->graph);
  if (x->fgraph) (void)hipGraphExecDestroy(x->fgraph);
  if (x->stream && !x->stream_borrowed) (void)hipStreamDestroy(x->stream);
  if (x->d_rec) (void)hipFree(x->d_rec);
}

// The context's replica record (model constants, buffers, MD constants) onto
// the device, after any change to them; waits for the context's queued work,
// which reads the record.
int sync_rep(pimd_ctx *x) {
  Rep h;
  h.c = x->c;
  h.b = x->d;
  h.m = x->m;
  x->self.hc.assign(1, x->c);
  x->self.hm.assign(1, x->m);
  x->self.d = x->rep_dev;
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpyAsync(x->rep_dev, &h, sizeof(Rep), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return PIMD_OK;
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

// Host-side tables of a system: k-vectors, kz-columns and their segments.
struct HostTables {
  std::vector<double4> kv;
  std::vector<int4> kh, cols, segs;
  std::vector<int> colk, segk;
};

// The k-vector tables, kz-columns and segments and the default launch shapes
// of a set-up system (no device work).
void plan_system(pimd_ctx *x, HostTables &t) {
  SysConst &c = x->c;
  const size_t nk = c.nk;
  // k-vectors and kz-columns
  t.kv.resize(nk);
  t.kh.resize(nk);
  std::vector<int4> &cols = t.cols;
  std::vector<int> &colk = t.colk;
  cols.clear();
  colk.clear();
  for (size_t k = 0; k < nk; ++k) {
    t.kv[k] = make_double4(x->kvec[3 * k], x->kvec[3 * k + 1], x->kvec[3 * k + 2], x->p4vAk[k]);
    t.kh[k] = make_int4(x->hkl[3 * k], x->hkl[3 * k + 1], x->hkl[3 * k + 2], 0);
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
  std::vector<int4> &segs = t.segs;
  std::vector<int> &segk = t.segk;
  segs.clear();
  segk.clear();
  {
    // 4 kz per segment: twice the lanes of 8 for the same sites (512 H2O x 32
    // beads: 1.273 -> 1.251 ms per mixed step, 1.592 -> 1.586 FP64; the
    // 32-water, 32-bead Rc 4.9 A system gains more)
    const int seg = env_int("PIMD_SFAC_SEG", 4, 1, 32);
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
  // One bead (the centroid call of contracted MD): 0, the charge and final
  // pair passes from global memory with lanes by neighbour count (64 at
  // 512 H2O, Rc = L/2), twice the workgroups of the LDS form's 32 lanes
  // (contracted step 1.05 -> 1.00 ms mixed, 1.12 -> 1.04 ms FP64).
  const int dflt_qq = c.P == 1 ? 0 : pow2_clamp(131072 / std::max(mb, 1L), 8, 32);
  // The dipole pairs' screening factors stay fixed through the SOR solve:
  // kept from the first iteration up to PIMD_DD_CACHE_MB (default 4096) per
  // context, recomputed every iteration beyond (512 H2O x 32 beads, dipole
  // pass per step: FP64 0.356 -> 0.251 ms, mixed 0.238 -> 0.209 ms; TTM2.1-F
  // 128 x 32: 0.438 -> 0.282 ms, 0.316 -> 0.264 ms)
  const double cache_mb = (double)dd_cache_elems(c) * (c.mixed ? 8 : 16) / (1 << 20);
  c.dd_cache = (c.model == PIMD_TTM21F || c.model == PIMD_TTM3F) &&
               env_int("PIMD_DD_CACHE", 1, 0, 1) &&
               cache_mb <= env_int("PIMD_DD_CACHE_MB", 4096, 0, 1 << 20);
  // lanes per molecule of the dipole pass: 16 from 16384 molecule-beads on;
  // an FP32 pass that computes its factors runs lighter lanes, twice as many
  // (512 H2O x 32 beads: 1.685 -> 1.657 ms per force call); TTM2.1-F's three
  // dipole sites at most 32 (128 x 32: 1.50 -> 1.37 ms FP64), 16 with the
  // factors kept (0.312 -> 0.282 ms per step)
  const int dflt_dd = pow2_clamp((c.mixed && !c.dd_cache ? 524288 : 262144) / std::max(mb, 1L), 16,
                                 c.model == 2 ? (c.dd_cache ? 16 : 32) : 64);
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
  c.overlap = 0;
  // the charge pair pass evaluates each molecule pair once (not with the
  // cell-list neighbour search, whose lists are not in partner order):
  // 0 never, 1 where its blocks fill the device, 2 always
  c.qq_half = ((c.model == PIMD_TTM21F || c.model == PIMD_TTM3F) && !c.cells)
                  ? env_int("PIMD_QQ_HALF", 1, 0, 2)
                  : 0;
  c.fin_half = env_int("PIMD_FIN_HALF", 1, 0, 1);
  if (c.fd_cg != 0 && c.fd_cg != 4 && c.fd_cg != 8 && c.fd_cg != 16) c.fd_cg = 0;
}

// Device buffers of a context laid out in one block, each 256-byte aligned.
// With base == nullptr only the size is counted.
struct Arena {
  char *base = nullptr;
  size_t off = 0;
  template <class T>
  void take(T **p, size_t n) {
    off = (off + 255) & ~(size_t)255;
    if (base) *p = (T *)(base + off);  // a C cast: T may carry the kernels' address space
    off += std::max<size_t>(n, 1) * sizeof(T);
  }
};

// full: every buffer of a context (force call, ring polymer, MD, adjoint);
// otherwise only what one force call on caller-provided coordinates uses.
void layout_system(pimd_ctx *x, Arena &a, bool full) {
  SysConst &c = x->c;
  DevBufs &d = x->d;
  const size_t P = c.P, N = c.N, S = 4 * N, na = 3 * N, nk = c.nk, nd = c.nd;
  a.take(&d.RR, P * na * 3);
  a.take(&d.RRc, na * 3);
  a.take(&d.site, P * S);
  a.take(&d.Rcs, S * 3);
  a.take(&d.grdq, P * N * 27);
  a.take(&d.phi, P * S);
  a.take(&d.Efq, P * S * 3);
  a.take(&d.kE, P * S * 3);
  a.take(&d.kphi, P * S);
  a.take(&d.dR, P * S * 3);
  a.take(&d.dip, P * N * nd * 3);
  a.take(&d.dip_prev, P * N * nd * 3);
  a.take(&d.Efd, P * N * nd * 3);
  a.take(&d.diff2, P * N * nd);
  a.take(&d.tally, P * N * T_COUNT);
  a.take(&d.hist, P * 4 * N * nd * 3);
  a.take(&d.nbr, P * N * (size_t)c.nbr_cap);
  a.take(&d.nbrc, P * N);
  a.take(&d.nbrn, P * N);
  a.take(&d.nbrh, P * N * 4);
  if (c.qq_half) a.take(&d.qqslab, qq_slab_elems(c));
  a.take(&d.tab, P * (size_t)c.T * S);
  a.take(&d.tabd, P * (size_t)c.T * N * nd);
  if (c.mixed) a.take(&d.tabdf, P * (size_t)c.T * N * nd);
  if (c.dd_cache) {
    const size_t n = dd_cache_elems(c);
    a.take(&d.ddts, c.mixed ? (n + 1) / 2 : n);
    d.ddtsf = (decltype(d.ddtsf))d.ddts;
  }
  a.take(&d.Sq, P * nk * 2);
  a.take(&d.Sm, P * nk * 2);
  a.take(&d.Sv, P * nk * 6);
  a.take(&d.virk, P * ((nk + 255) / 256) * 18);
  a.take(&d.part, P * (size_t)c.sf_chunks_max * nk * 6);
  a.take(&d.active, P);
  a.take(&d.iters, P);
  a.take(&d.nactive, (size_t)c.maxiter + 1);
  a.take(&d.kv, nk);
  a.take(&d.khkl, nk);
  a.take(&d.red, nk);  // A_k
  a.take(&d.cols, c.ncol);
  a.take(&d.colk, c.ncol);
  a.take(&d.segs, c.nseg);
  a.take(&d.segk, c.nseg);
  a.take(&d.Upot, P);
  a.take(&d.virt, P * 9);
  a.take(&d.virialc, P);
  a.take(&d.terms, P * 5);
  a.take(&d.dRR, P * na * 3);
  a.take(&d.dipI, P * N * 3);
  a.take(&d.dipE, P * N * 3);
  a.take(&d.chg, P * na);
  a.take(&d.iters_out, P);
  a.take(&d.radius, P * N);
  if (c.cells) {
    a.take(&d.cell_start, P * ((size_t)c.ncell[0] * c.ncell[1] * c.ncell[2] + 1));
    a.take(&d.cell_mol, P * N);
  }
  a.take(&d.pes, PES_NALL);
  a.take(&d.clk_own, 2);
  a.take(&d.flags, 1);
  a.take(&x->rep_dev, 1);
  if (!full) return;
  a.take(&d.C, P * P);
  a.take(&d.A, 4 * P);
  a.take(&d.msf, P);
  a.take(&d.c1, P);
  a.take(&d.c2, P);
  a.take(&d.masses, na);
  a.take(&d.PP, P * na * 3);
  a.take(&d.noise, 2 * na * P * 3);
  a.take(&d.uk, 1);
  a.take(&d.ukpart, 256);
  a.take(&d.adj, 2 * P * na * 3 + P + (size_t)pes_adjoint_blocks(c) * PES_NPARAM + PES_NPARAM);
  a.take(&d.obs_part, 64 * 5);
  a.take(&d.vxi_beads, na * P * 3 * 2);
  a.take(&d.nh_tau, P);
  a.take(&d.nh_vxi, 2);
  a.take(&d.nh_s, 1);
}

// Copies the host tables and the reference's PES parameters into the laid-out
// buffers on stream s and sets the step clock to zero.
int upload_system(pimd_ctx *x, const HostTables &t, hipStream_t s) {
  const SysConst &c = x->c;
  DevBufs &d = x->d;
  const size_t nk = c.nk;
  double theta[PES_NALL];
  pes_default_params(theta);
  HIPCHK(hipMemcpyAsync(d.pes, theta, sizeof theta, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.kv, t.kv.data(), sizeof(double4) * nk, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.khkl, t.kh.data(), sizeof(int4) * nk, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.red, x->Ak.data(), sizeof(double) * nk, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.cols, t.cols.data(), sizeof(int4) * t.cols.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.colk, t.colk.data(), sizeof(int) * t.colk.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.segs, t.segs.data(), sizeof(int4) * t.segs.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.segk, t.segk.data(), sizeof(int) * t.segk.size(), hipMemcpyHostToDevice, s));
  d.clk = d.clk_own;
  HIPCHK(hipMemsetAsync(d.clk_own, 0, 2 * sizeof(int), s));
  HIPCHK(hipMemsetAsync(d.flags, 0, sizeof(int), s));
  HIPCHK(hipGetLastError());
  // the host tables are pageable: done before they go out of scope
  HIPCHK(hipStreamSynchronize(s));
  return PIMD_OK;
}

int alloc_system(pimd_ctx *x) {
  SysConst &c = x->c;
  DevBufs &d = x->d;
  HostTables t;
  plan_system(x, t);
  Arena a;
  layout_system(x, a, true);
  int rc;
  if ((rc = dalloc((char **)&x->arena, a.off)) != PIMD_OK) return rc;
  a.base = static_cast<char *>(x->arena);
  a.off = 0;
  layout_system(x, a, true);
  const size_t P = c.P, N = c.N, na = 3 * N, nd = c.nd;
  if ((rc = upload_system(x, t, x->stream)) != PIMD_OK) return rc;
  HIPCHK(hipMemset(d.hist, 0, sizeof(double) * P * 4 * N * nd * 3));
  std::vector<double> masses(na);
  for (size_t i = 0; i < na; ++i) masses[i] = (i % 3 == 0) ? x->p.mass_o : x->p.mass_h;
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

// Per-kernel timing hook: px (may be NULL) records each launch's time when
// its profiling is on.
template <class F>
void timed(pimd_ctx *px, const char *name, F &&f) {
  if (px) px->timed(name, f);
  else f();
}

// One pot_ttm evaluation per bead of every replica of the group
// (pot_ttm.f90:3-630; pot_spc.f90:1-277 for the fixed-charge models), each
// from its coordinates RR, RRc into its output buffers.  The MD step t of the
// call (dipole predictor, history slot) comes from each replica's device
// clock, set by the caller.  px: the context whose profiling and SOR polling
// apply (a group of that one replica), or NULL.
int run_forces(const Group &g, hipStream_t s, pimd_ctx *px) {
  const int n = g.n();
  bool any_ttm = false, all_fold = true, any_mixed = false;
  int maxiter = 0;
  for (const SysConst &c : g.hc) {
    if (c.model > 3) continue;
    any_ttm = true;
    maxiter = std::max(maxiter, c.maxiter);
    all_fold = all_fold && pair_dd_tests(c);
    any_mixed = any_mixed || c.mixed;
  }
  const bool prof = px && px->prof;
  const bool poll = px && (px->c.poll || prof) && n == 1;
  timed(px, "monomer", [&] {
    launch_monomer(g, s);
    launch_spc_monomer(g, s);
  });
  timed(px, "neighbors", [&] { launch_neighbors(g, s); });
  timed(px, "pair_qq", [&] {
    launch_pair_qq(g, s);
    launch_spc_pair(g, s);
  });
  // the reciprocal-space charge pass (it writes kE / kphi, added by their consumers)
  timed(px, "ewald_tables", [&] { launch_tables(g, s); });
  timed(px, "sfac_q", [&] { launch_sfac(g, SF_Q, s, false, 1); });
  timed(px, "sfac_q_reduce", [&] { launch_sfac(g, SF_Q, s, false, 2); });
  timed(px, "field_q", [&] { launch_field(g, FD_Q, 0, s); });
  if (any_ttm) {
    timed(px, "post_qq", [&] { launch_post_qq(g, s); });
    // the convergence test of iteration it - 1 runs at the start of the
    // dipole pair kernel of iteration it, where every replica's dipole
    // kernel can hold it; the last iteration's test stays a launch
    const bool fold = all_fold && !poll;
    for (int it = 1; it <= maxiter; ++it) {
      // with timing on: the beads this iteration works on, for per-launch work counts
      if (poll && prof) px->count("sor_bead_iterations", it == 1 ? px->c.P : px->h_nactive[it - 1]);
      timed(px, "pair_dd", [&] { launch_pair_dd(g, false, s, fold && it > 1 ? it - 1 : 0, false, it); });
      timed(px, "sfac_d", [&] { launch_sfac(g, SF_D, s, false, 1); });
      timed(px, "sfac_d_reduce", [&] { launch_sfac(g, SF_D, s, false, 2); });
      timed(px, "field_d_sor", [&] { launch_field(g, FD_D, it, s); });
      if (!fold || it == maxiter) timed(px, "converge", [&] { launch_converge(g, it, s); });
      // No host round trip: converged beads are switched off on the device
      // and every kernel of the loop skips them, so the remaining iterations
      // of a converged system are near-empty launches.  (With per-kernel
      // timing on, poll so only live iterations are launched and timed.)
      if (poll && it < maxiter) {
        HIPCHK(hipMemcpyAsync(px->h_nactive + it, px->d.nactive + it, sizeof(int),
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (px->h_nactive[it] == 0) break;
      }
    }
    if (any_mixed) {
      // Iterative refinement of the mixed form: redo each bead's last SOR
      // step in FP64 from the dipoles it started from, both halves of the
      // dipole field (the iterations took them in FP32).
      timed(px, "refine_dd", [&] {
        launch_refine_prep(g, s);
        launch_pair_dd(g, true, s, 0, true);
        launch_sfac(g, SF_D, s, true);
        launch_field(g, FD_D, -1, s, true);
      });
    }
    timed(px, "pair_final", [&] { launch_pair_final(g, s); });
    timed(px, "sfac_v", [&] { launch_sfac(g, SF_V, s, false, 1); });
    timed(px, "sfac_v_reduce", [&] { launch_sfac(g, SF_V, s, false, 2); });
    timed(px, "field_qd", [&] { launch_field(g, FD_QD, 0, s); });
  }
  timed(px, "virial_k", [&] { launch_virk(g, s); });
  timed(px, "finalize", [&] {
    launch_finalize(g, s);
    launch_spc_finalize(g, s);
  });
  timed(px, "reduce_bead", [&] { launch_reduce_bead(g, s); });
  HIPCHK(hipGetLastError());
  return PIMD_OK;
}

int run_forces(pimd_ctx *x) { return run_forces(x->self, x->stream, x); }

// run_forces of a context as one replay of a captured graph, captured again
// after invalidate_graphs.  Per-kernel timing and polling keep eager
// launches, as does a device where capture fails.  PIMD_FORCE_GRAPH=0 turns
// replay off.
int force_call(pimd_ctx *x) {
  const SysConst &c = x->c;
  const bool on = env_int("PIMD_FORCE_GRAPH", 1, 0, 1) != 0;
  if (!on || x->fgraph_off || x->prof || c.poll) return run_forces(x);
  if (!x->fgraph_ok) {
    if (x->fgraph) (void)hipGraphExecDestroy(x->fgraph);
    x->fgraph = nullptr;
    hipStream_t s = x->stream;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = run_forces(x);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &gr);
    if (rc != PIMD_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    hipError_t e2 = e;
    if (e == hipSuccess) e2 = hipGraphInstantiate(&x->fgraph, gr, nullptr, nullptr, 0);
    if (gr) (void)hipGraphDestroy(gr);
    if (e2 != hipSuccess) {
      // capture is not available here: launch eagerly from now on
      x->fgraph = nullptr;
      x->fgraph_off = true;
      (void)hipGetLastError();
      return run_forces(x);
    }
    x->fgraph_ok = true;
  }
  HIPCHK(hipGraphLaunch(x->fgraph, x->stream));
  return PIMD_OK;
}
}  // namespace

namespace {
// Checks run parameters and fills the reference's derived defaults
// (InputOutput.f90:118-125, input_file_reader.f90:104-105); makes the
// device current.
int check_params(pimd_params &p) {
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
  return PIMD_OK;
}

}  // namespace

namespace {
// Non-finite coordinates in host arrays: refused before anything runs.
int check_finite(const double *a, size_t n, const char *what) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(a[i])) return fail(PIMD_EINVAL, "%s[%zu] is not finite", what, i);
  return PIMD_OK;
}

// The device flag of non-finite coordinates in force calls since the last
// check (the monomer kernels set it), read and cleared; the context's stream
// must be synchronised.
int take_flags(pimd_ctx *x) {
  for (pimd_ctx *y : {x, x->cen}) {
    if (!y || !y->d.flags) continue;
    int f = 0;
    HIPCHK(hipMemcpy(&f, y->d.flags, sizeof(int), hipMemcpyDeviceToHost));
    if (f) {
      HIPCHK(hipMemset(y->d.flags, 0, sizeof(int)));
      return fail(PIMD_EINVAL, "non-finite coordinates entered a force call");
    }
  }
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
  int rc0 = check_params(p);
  if (rc0 != PIMD_OK) return rc0;
  pimd_ctx *x = new pimd_ctx();
  x->p = p;
  x->natoms = 3 * p.nwaters;
  int rc = setup_system(x);
  // one stream per context (the second is made when first needed): each
  // context of a batch then gets a hardware queue of its own
  if (rc == PIMD_OK && hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(PIMD_EDEVICE, "hipStreamCreate failed");
  if (rc == PIMD_OK) {
    x->c.g_nolds = pick_group(x->c);
    rc = alloc_system(x);
  }
  if (rc == PIMD_OK) {
    x->m.massO = p.mass_o;
    x->m.massH = p.mass_h;
    x->m.box0 = p.box[0];
    x->m.boxi0 = 1.0 / p.box[0];
    rc = sync_rep(x);
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
  hipStream_t s = x->stream;
  const size_t P = x->c.P, na = x->natoms, N = x->c.N;
  // the call works on the context's own coordinate and output buffers (those
  // of its MD state); the caller's arrays are copied in and out on the stream
  if (RR != d.RR) HIPCHK(hipMemcpyAsync(d.RR, RR, sizeof(double) * P * na * 3, hipMemcpyDeviceToDevice, s));
  if (RRc != d.RRc) HIPCHK(hipMemcpyAsync(d.RRc, RRc, sizeof(double) * na * 3, hipMemcpyDeviceToDevice, s));
  launch_set_clock(x->self, t, 0, s);
  int rc = force_call(x);
  if (rc != PIMD_OK) return rc;
  struct Out {
    void *dst;
    const void *src;
    size_t bytes;
  } outs[] = {{Upot, d.Upot, sizeof(double) * P},
              {dRR, d.dRR, sizeof(double) * P * na * 3},
              {virt, d.virt, sizeof(double) * P * 9},
              {virialc, d.virialc, sizeof(double) * P},
              {dipI, d.dipI, sizeof(double) * P * N * 3},
              {dipE, d.dipE, sizeof(double) * P * N * 3},
              {chg, d.chg, sizeof(double) * P * na},
              {iters, d.iters_out, sizeof(int) * P}};
  for (const Out &o : outs)
    if (o.dst && o.dst != o.src) HIPCHK(hipMemcpyAsync(o.dst, o.src, o.bytes, hipMemcpyDeviceToDevice, s));
  return PIMD_OK;
}

// The host-array force call in two halves.  fab_start stages the
// coordinates in pinned memory, queues the call and the copies of the outputs
// (those set in want, bit i for output i of fab_outs) into the staging block
// and returns; fab_finish waits, copies the wanted outputs to the caller and
// reports non-finite input.
struct FabOut {
  void *dst;
  const void *src;
  size_t off, bytes;
};
static void fab_outs(pimd_ctx *x, FabOut *o, void *const dst[8]) {
  const size_t P = x->c.P, na = x->natoms, N = x->c.N, nR = P * na * 3, nc = na * 3;
  const size_t o_U = nR + nc, o_dR = o_U + P, o_vt = o_dR + nR, o_vc = o_vt + P * 9,
               o_dI = o_vc + P, o_dE = o_dI + P * N * 3, o_ch = o_dE + P * N * 3, o_it = o_ch + P * na;
  DevBufs &d = x->d;
  const FabOut all[8] = {{dst[0], d.Upot, o_U, sizeof(double) * P},
                         {dst[1], d.dRR, o_dR, sizeof(double) * nR},
                         {dst[2], d.virt, o_vt, sizeof(double) * P * 9},
                         {dst[3], d.virialc, o_vc, sizeof(double) * P},
                         {dst[4], d.dipI, o_dI, sizeof(double) * P * N * 3},
                         {dst[5], d.dipE, o_dE, sizeof(double) * P * N * 3},
                         {dst[6], d.chg, o_ch, sizeof(double) * P * na},
                         {dst[7], d.iters_out, o_it, sizeof(int) * P}};
  for (int i = 0; i < 8; ++i) o[i] = all[i];
}

static int fab_start(pimd_ctx *x, const double *RRt, const double *RRc, int t, unsigned want) {
  if (!x || !RRt || !RRc) return fail(PIMD_EINVAL, "null argument");
  if (x->fab_pending) return fail(PIMD_EINVAL, "a started force call is still to be finished");
  HIPCHK(hipSetDevice(x->p.device));
  const size_t P = x->c.P, na = x->natoms, N = x->c.N;
  int rc0;
  if ((rc0 = check_finite(RRt, P * na * 3, "RRt")) != PIMD_OK ||
      (rc0 = check_finite(RRc, na * 3, "RRc")) != PIMD_OK)
    return rc0;
  DevBufs &d = x->d;
  hipStream_t s = x->stream;
  // the caller's arrays go through a pinned staging block: DMA copies at
  // full link speed instead of the runtime's pageable path
  const size_t nR = P * na * 3, nc = na * 3;
  const size_t n_stage = 2 * nR + nc + P * 11 + 2 * P * N * 3 + P * na + P;
  if (!x->h_stage)
    HIPCHK(hipHostMalloc((void **)&x->h_stage, sizeof(double) * n_stage, hipHostMallocDefault));
  double *h = x->h_stage;
  std::memcpy(h, RRt, sizeof(double) * nR);
  std::memcpy(h + nR, RRc, sizeof(double) * nc);
  HIPCHK(hipMemcpyAsync(d.RR, h, sizeof(double) * nR, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.RRc, h + nR, sizeof(double) * nc, hipMemcpyHostToDevice, s));
  launch_set_clock(x->self, t, 0, s);
  int rc = force_call(x);
  if (rc != PIMD_OK) return rc;
  FabOut o[8];
  void *const none[8] = {};
  fab_outs(x, o, none);
  for (int i = 0; i < 8; ++i)
    if (want >> i & 1) HIPCHK(hipMemcpyAsync(h + o[i].off, o[i].src, o[i].bytes, hipMemcpyDeviceToHost, s));
  x->fab_pending = true;
  return PIMD_OK;
}

static int fab_finish(pimd_ctx *x, void *const dst[8]) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  if (!x->fab_pending) return fail(PIMD_EINVAL, "no force call was started");
  HIPCHK(hipSetDevice(x->p.device));
  x->fab_pending = false;
  HIPCHK(hipStreamSynchronize(x->stream));
  FabOut o[8];
  fab_outs(x, o, dst);
  for (int i = 0; i < 8; ++i)
    if (o[i].dst) std::memcpy(o[i].dst, x->h_stage + o[i].off, o[i].bytes);
  return take_flags(x);
}

int pimd_forces_all_beads(pimd_ctx *x, const double *RRt, const double *RRc, int t,
                          double *Upott, double *dRRt, double *virt, double *virialct,
                          double *dipI, double *dipE, double *chg, int *iters) {
  void *const dst[8] = {Upott, dRRt, virt, virialct, dipI, dipE, chg, iters};
  unsigned want = 0;
  for (int i = 0; i < 8; ++i)
    if (dst[i]) want |= 1u << i;
  const int rc = fab_start(x, RRt, RRc, t, want);
  return rc != PIMD_OK ? rc : fab_finish(x, dst);
}

int pimd_forces_all_beads_start(pimd_ctx *x, const double *RRt, const double *RRc, int t) {
  return fab_start(x, RRt, RRc, t, 0xffu);
}

int pimd_forces_all_beads_finish(pimd_ctx *x, double *Upott, double *dRRt, double *virt,
                                 double *virialct, double *dipI, double *dipE, double *chg,
                                 int *iters) {
  void *const dst[8] = {Upott, dRRt, virt, virialct, dipI, dipE, chg, iters};
  return fab_finish(x, dst);
}

int pimd_energy_terms(pimd_ctx *x, double *terms) {
  if (!x || !terms) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipMemcpyAsync(terms, x->d.terms, sizeof(double) * x->c.P * 5, hipMemcpyDeviceToHost,
                        x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return PIMD_OK;
}

// calc_monomer_dip_moments (force_calc.f90:205-264) for every bead: RRt
// [P][Natoms][3] in, dip_momIt [P][N][3] out, through the context's
// coordinate staging.
int pimd_monomer_dipoles(pimd_ctx *x, const double *RRt, double *dip) {
  if (!x || !RRt || !dip) return fail(PIMD_EINVAL, "null argument");
  int rc = check_finite(RRt, (size_t)x->c.P * x->natoms * 3, "RRt");
  if (rc != PIMD_OK) return rc;
  HIPCHK(hipSetDevice(x->p.device));
  const size_t P = x->c.P, N = x->c.N;
  // two intermediates of the force call, rewritten by every force call:
  // kE [P][4N][3] takes the coordinates, kphi [P][4N] the dipoles
  HIPCHK(hipMemcpyAsync(x->d.kE, RRt, sizeof(double) * P * 3 * N * 3, hipMemcpyHostToDevice,
                        x->stream));
  launch_monomer_dipoles(x->c, x->d.pes, x->d.kE, x->d.kphi, x->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dip, x->d.kphi, sizeof(double) * P * N * 3, hipMemcpyDeviceToHost,
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
  x->timed("nm_evolve", [&] { launch_nm_evolve(x->self, s, KICK_NONE, KICK_NONE, true); });
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
    launch_langevin(x->self, s, true, noise ? 1 : 0, 0, seed, counter, 0, KICK_NONE);
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
  if (md->bead_thermostat < 0 || md->bead_thermostat > 2)
    return fail(PIMD_EINVAL, "bead_thermostat must be 0 (none), 1 (Langevin) or 2 (Nose-Hoover)");
  if (md->bead_thermostat == 2 && md->centroid_thermostat && !(md->tau_centroid > 0.0))
    return fail(PIMD_EINVAL, "the centroid's Nose-Hoover chain needs tau_centroid > 0");
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
  x->d.cen_Upot = x->d.cen_virt = x->d.cen_virialc = x->d.cen_dRR = nullptr;
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
    // its coordinates are this context's centroids (potential(RRc, RRc, ...),
    // MD.f90:398) and its step the step of this context's clock
    cen->d.RR = x->d.RRc;
    cen->d.RRc = x->d.RRc;
    cen->d.clk = x->d.clk;
    std::copy(x->ttm, x->ttm + TTM_NPARAM, cen->ttm);  // parameters set on this context
    apply_ttm_params(cen);
    if ((rc = sync_rep(cen)) != PIMD_OK) return rc;
    x->d.cen_Upot = cen->d.Upot;
    x->d.cen_virt = cen->d.virt;
    x->d.cen_virialc = cen->d.virialc;
    x->d.cen_dRR = cen->d.dRR;
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
  if (md->bead_thermostat == 2) {
    // bead_thermostat (main_stuff.f90:145-171): tau_centroid for the centroid
    // mode, 1/omega_j for the others; chain velocities start at zero
    // (InputOutput.f90:402-403)
    std::vector<double> tau(P);
    for (int j = 0; j < P; ++j) tau[j] = j == 0 ? md->tau_centroid : 1.0 / x->om[j];
    HIPCHK(hipMemcpy(x->d.nh_tau, tau.data(), sizeof(double) * P, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(x->d.vxi_beads, 0, sizeof(double) * (size_t)x->natoms * P * 3 * 2));
  }
  const double vxi0[2] = {1.0, 1.0}, one = 1.0;  // vxi_global = 1 (InputOutput.f90:399)
  HIPCHK(hipMemcpy(x->d.nh_vxi, vxi0, sizeof vxi0, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(x->d.nh_s, &one, sizeof one, hipMemcpyHostToDevice));
  // the replica's MD constants, read by the ring-polymer and thermostat kernels
  MDConst &m = x->m;
  m.delt = delt;
  m.delt2 = delt2;
  m.deltf = md->intra_timesteps > 0 ? delt / md->intra_timesteps : delt;
  m.delt2f = m.deltf / 2.0;
  m.tempP = P * md->temp;
  m.omegan = omegan;
  m.tau = md->tau;
  m.bead_tt = md->bead_thermostat;
  m.centroid = md->centroid_thermostat ? 1 : 0;
  m.global_nh = md->global_nh ? 1 : 0;
  m.cmd = x->cmd ? 1 : 0;
  if ((rc = sync_rep(x)) != PIMD_OK) return rc;
  x->md_ready = true;
  return PIMD_OK;
}

int pimd_state_set(pimd_ctx *x, const double *RRt, const double *PPt, const double *dRRt) {
  if (!x || !RRt || !PPt) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipSetDevice(x->p.device));
  const size_t n = (size_t)x->c.P * x->natoms * 3;
  int rc0;
  if ((rc0 = check_finite(RRt, n, "RRt")) != PIMD_OK ||
      (rc0 = check_finite(PPt, n, "PPt")) != PIMD_OK ||
      (dRRt && (rc0 = check_finite(dRRt, n, "dRRt")) != PIMD_OK))
    return rc0;
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

// Which replicas of a group have the bead thermostats, the global chain,
// one bead; used to leave out launches no replica needs.
struct GroupFlags {
  bool langevin = false, bead_nh = false, global_nh = false, p1 = false, multi = false,
       no_bead_tt = false;
};
static GroupFlags group_flags(const Group &g) {
  GroupFlags f;
  for (int i = 0; i < g.n(); ++i) {
    f.langevin |= g.hm[i].bead_tt == 1;
    f.bead_nh |= g.hm[i].bead_tt == 2;
    f.no_bead_tt |= g.hm[i].bead_tt == 0;
    f.global_nh |= g.hm[i].global_nh != 0;
    f.p1 |= g.hc[i].P == 1;
    f.multi |= g.hc[i].P > 1;
  }
  return f;
}

// bead_thermostat (main_stuff.f90:132-188) of an MD step: the NM-space
// Langevin thermostat (half 0 / 1 of the step) and the NM-space Nose-Hoover
// chains of the replicas that have them; with kick = KICK_DRR the half kick
// -MASSCON dRRt delt2 comes first, and replicas without a bead thermostat
// take it as a kick of their own.  load_noise(half): loads host normals into
// the noise buffer first (one replica, eager steps), or NULL.
template <class LoadNoise>
static void bead_thermostat(const Group &g, const GroupFlags &f, hipStream_t s, pimd_ctx *px,
                            int half, int kick, LoadNoise load_noise) {
  const bool host = load_noise(half);
  if (f.langevin)
    timed(px, "langevin",
          [&] { launch_langevin(g, s, false, host ? 1 : 0, 1, 0, 0, half, kick); });
  if (f.bead_nh) timed(px, "bead_nh", [&] { launch_bead_nh(g, kick, s); });
  if (kick != KICK_NONE && f.no_bead_tt)
    timed(px, "kick", [&] { launch_kick(g, kick, SEL_NO_BEAD_TT, s); });
}

// Enqueues one PIMD_VV step (MD.f90:35-110) of every replica of g on stream
// s: the MD step and the record slot come from each replica's device clock,
// which the step advances at its end, so the same launches serve every step
// (and can be captured once).  Records Upot, virial, virialc, uk and the
// observables.  Does not synchronise.
template <class LoadNoise>
static int step_vv(const Group &g, hipStream_t s, pimd_ctx *px, LoadNoise load_noise) {
  const GroupFlags f = group_flags(g);
  bead_thermostat(g, f, s, px, 0, KICK_NONE, load_noise);
  if (f.global_nh) timed(px, "nose_hoover", [&] { launch_global_nh(g, s); });
  // the first half kick rides on the free ring-polymer step
  if (f.multi) timed(px, "nm_evolve", [&] { launch_nm_evolve(g, s, KICK_DRR); });
  if (f.p1) {
    timed(px, "kick", [&] { launch_kick(g, KICK_DRR, SEL_P1, s); });
    timed(px, "free_flight", [&] { launch_free_flight(g, false, s); });
  }
  timed(px, "centroid_pbc", [&] { launch_centroid_pbc(g, s); });
  int rc = run_forces(g, s, px);
  if (rc != PIMD_OK) return rc;
  timed(px, "kinetic", [&] { launch_kinetic(g, s); });
  timed(px, "observables", [&] { launch_observables(g, true, true, s); });  // also the step record
  // the second half kick rides on the second bead-thermostat half step
  bead_thermostat(g, f, s, px, 1, KICK_DRR, load_noise);
  timed(px, "kinetic", [&] { launch_kinetic(g, s); });
  if (f.global_nh) timed(px, "nose_hoover", [&] { launch_global_nh(g, s); });
  launch_advance_clock(g, true, s);  // with the record's kinetic energy
  HIPCHK(hipGetLastError());
  return PIMD_OK;
}

// Enqueues one contracted_MD step (MD.f90:276-449; clock as step_vv) of the
// replicas of g, all with intra inner steps, whose centroids form the group
// cen: the monomer PES moves every bead in intra inner steps of the ring
// polymer, the intermolecular forces of the centroid act on every bead once
// per step.  Records sum of bead energies, virial, centroid virial and uk.
// The fast gradient is kept from one step to the next.
template <class LoadNoise>
static int step_contracted(const Group &g, const Group &cen, int intra, hipStream_t s,
                           pimd_ctx *px, LoadNoise load_noise) {
  const GroupFlags f = group_flags(g);
  bead_thermostat(g, f, s, px, 0, KICK_NONE, load_noise);
  if (f.global_nh) timed(px, "nose_hoover", [&] { launch_global_nh(g, s); });
  // reads and moves coordinates only: the kicks below commute with it
  timed(px, "centroid_pbc", [&] { launch_centroid_pbc(g, s); });
  for (int ti = 1; ti <= intra; ++ti) {
    // the kicks before each free ring-polymer step ride on it: the outer
    // half kick and the first fast one, then the previous inner step's
    // closing fast kick and this step's opening one (the same gradient)
    if (f.multi)
      timed(px, "nm_evolve",
            [&] { launch_nm_evolve(g, s, ti == 1 ? KICK_DRR : KICK_FAST, KICK_FAST); });
    if (f.p1) {
      timed(px, "kick", [&] { launch_kick(g, ti == 1 ? KICK_DRR : KICK_FAST, SEL_P1, s); });
      timed(px, "kick_fast", [&] { launch_kick(g, KICK_FAST, SEL_P1, s); });
      timed(px, "free_flight", [&] { launch_free_flight(g, true, s); });
    }
    timed(px, "pes_fast", [&] { launch_pes_fast(g, ti == intra, s); });
  }
  timed(px, "kick_fast", [&] { launch_kick(g, KICK_FAST, SEL_ALL, s); });
  timed(px, "centroid_pbc", [&] { launch_centroid_pbc(g, s); });
  int rc = PIMD_OK;
  timed(px, "centroid_forces", [&] { rc = run_forces(cen, s, nullptr); });
  if (rc != PIMD_OK) return rc;
  timed(px, "broadcast", [&] { launch_broadcast_beads(g, s); });
  timed(px, "kinetic", [&] { launch_kinetic(g, s); });
  launch_contract_record(g, s);
  // the bead dipoles of contracted_MD: monomer dipoles of every bead
  // (calc_monomer_dip_moments after the centroid call, MD.f90:407)
  timed(px, "monomer_dipoles", [&] { launch_monomer_dipoles(g, s); });
  timed(px, "observables", [&] { launch_observables(g, true, false, s); });
  bead_thermostat(g, f, s, px, 1, KICK_DRR, load_noise);  // with the outer half kick
  timed(px, "kinetic", [&] { launch_kinetic(g, s); });
  if (f.global_nh) timed(px, "nose_hoover", [&] { launch_global_nh(g, s); });
  launch_advance_clock(g, true, s);  // with the record's kinetic energy
  HIPCHK(hipGetLastError());
  return PIMD_OK;
}

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
  HIPCHK(hipStreamSynchronize(x->stream));
  if (x->d_rec) (void)hipFree(x->d_rec);
  x->d_rec = nullptr;
  x->rec_cap = 0;
  int rc = dalloc(&x->d_rec, (size_t)nsteps * MD_REC);
  if (rc != PIMD_OK) return rc;
  x->rec_cap = nsteps;
  x->m.rec = x->d_rec;
  return sync_rep(x);
}

// One MD step of a context.  noise: NULL or the normals of this step,
// [2][Natoms][P][3] (host), copied into the noise buffer before each half.
static int md_step_eager(pimd_ctx *x, const double *noise) {
  const size_t n = (size_t)x->c.P * x->natoms * 3;
  auto load = [&](int half) {
    if (!noise) return false;
    (void)hipMemcpyAsync(x->d.noise, noise + (size_t)half * n, sizeof(double) * n,
                         hipMemcpyHostToDevice, x->stream);
    return true;
  };
  if (x->md.intra_timesteps > 0)
    return step_contracted(x->self, x->cen->self, x->md.intra_timesteps, x->stream, x, load);
  return step_vv(x->self, x->stream, x, load);
}

// A step with device normals can run as one replay of a captured graph of
// md_step_eager, rebuilt after invalidate_graphs (launch shapes, parameters,
// buffers).  Host noise, polling and per-kernel timing keep the eager
// launches.
static bool graph_eligible(const pimd_ctx *x, const double *noise) {
  const int on = env_int("PIMD_MD_GRAPH", 1, 0, 1);  // read per step: a test switches it
  const SysConst &c = x->c;
  const bool cen_ok = !x->cen || (!x->cen->c.poll && !x->cen->prof);
  return on && !noise && !x->prof && !c.poll && cen_ok;
}

static int md_graph_build(pimd_ctx *x) {
  if (x->graph) (void)hipGraphExecDestroy(x->graph);
  x->graph = nullptr;
  x->graph_ok = false;
  hipStream_t s = x->stream;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = md_step_eager(x, nullptr);
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
  return PIMD_OK;
}

static int md_step(pimd_ctx *x, const double *noise) {
  if (x->graph_off || !graph_eligible(x, noise)) return md_step_eager(x, noise);
  if (!x->graph_ok) {
    if (md_graph_build(x) != PIMD_OK) {
      // capture is not available here: launch eagerly from now on (an error
      // of the step itself shows again in the eager launch)
      x->graph_off = true;
      (void)hipGetLastError();
      return md_step_eager(x, noise);
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
  if (x->m.seed != seed) {  // the device normals' key, read from the replica record
    x->m.seed = seed;
    if ((rc = sync_rep(x)) != PIMD_OK) return rc;
  }
  const size_t n = (size_t)x->c.P * x->natoms * 3;
  // the clock, and the kinetic energy of the loaded state (needed by the
  // first thermostat call)
  launch_set_clock(x->self, t0, 0, x->stream);
  x->timed("kinetic", [&] { launch_kinetic(x->self, x->stream); });
  for (int st = 0; st < nsteps; ++st)
    if ((rc = md_step(x, noise ? noise + (size_t)st * 2 * n : nullptr)) != PIMD_OK) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  if ((rc = md_copy_records(x, nsteps, Upot, virial, virialc, uk)) != PIMD_OK) return rc;
  return take_flags(x);
}

// ---------------------------------------------------------------------------
// Batches of independent systems (replicas).  Every kernel of a batch step
// covers all replicas at once (one grid per kernel variant, see launch_over):
// the plain PIMD_VV replicas form one group, the contracted ones one group
// per intra_timesteps with their centroids as another.  The whole step of
// the batch is captured once as a graph and replayed.
// ---------------------------------------------------------------------------
struct pimd_batch {
  std::vector<pimd_ctx *> ctx;
  hipStream_t stream = nullptr;
  Rep *table = nullptr;  // device replica records of every group, group by group
  int table_cap = 0;
  // plain PIMD members, one group per force class (model and arithmetic)
  std::vector<Group> vv;
  // one stream per group beyond the first, so that the groups' kernel chains
  // run side by side (parallel branches of the step's graph)
  std::vector<hipStream_t> lanes;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  struct Con {
    int intra;
    Group outer, cen;
  };
  std::vector<Con> con;
  hipGraphExec_t graph = nullptr;
  bool graph_ok = false;
  std::vector<unsigned char> sig;  // what the captured launches depend on
};

static Rep host_rep(const pimd_ctx *x) {
  Rep h;
  h.c = x->c;
  h.b = x->d;
  h.m = x->m;
  return h;
}

// The batch's groups and device replica table from the contexts' current
// records (uploaded on the batch stream); a change to anything a captured
// launch depends on (model constants, launch shapes, bead counts, MD
// options) drops the captured step.
static int batch_tables(pimd_batch *b) {
  std::vector<Rep> host;
  std::vector<unsigned char> sig;
  auto add_sig = [&](const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    sig.insert(sig.end(), c, c + n);
  };
  auto push = [&](Group &g, const pimd_ctx *x) {
    const Rep h = host_rep(x);
    host.push_back(h);
    g.hc.push_back(h.c);
    g.hm.push_back(h.m);
    add_sig(&h.c, sizeof(SysConst));
    const int flags[5] = {h.m.bead_tt, h.m.centroid, h.m.global_nh, h.m.cmd, x->md.intra_timesteps};
    add_sig(flags, sizeof flags);
  };
  b->vv.clear();
  b->con.clear();
  std::vector<std::pair<int, int>> keys;
  for (pimd_ctx *x : b->ctx) {
    if (x->md.intra_timesteps != 0) continue;
    const std::pair<int, int> k{x->c.model, x->c.mixed};
    if (std::find(keys.begin(), keys.end(), k) == keys.end()) keys.push_back(k);
  }
  // members per lane: a force class; while there are fewer lanes than
  // PIMD_BATCH_LANES (default 4, the hardware queues of a process), the
  // costliest class is dealt into two (members sorted by molecule-beads,
  // dealt in turn).  64 uneven replicas, FP64: 106.3 ms per step on two
  // lanes, 102.5 on four.
  std::vector<std::vector<pimd_ctx *>> lanes;
  for (const auto &k : keys) {
    lanes.emplace_back();
    for (pimd_ctx *x : b->ctx)
      if (x->md.intra_timesteps == 0 && x->c.model == k.first && x->c.mixed == k.second)
        lanes.back().push_back(x);
  }
  const int want = env_int("PIMD_BATCH_LANES", 4, 0, 8);
  auto cost = [](const std::vector<pimd_ctx *> &v) {
    long w = 0;
    for (const pimd_ctx *x : v) w += (long)x->c.P * x->c.N;
    return w;
  };
  while ((int)lanes.size() < want) {
    size_t big = 0;
    for (size_t i = 1; i < lanes.size(); ++i)
      if (cost(lanes[i]) > cost(lanes[big])) big = i;
    if (lanes[big].size() < 2) break;
    std::vector<pimd_ctx *> v = lanes[big];
    std::stable_sort(v.begin(), v.end(), [](const pimd_ctx *a, const pimd_ctx *c2) {
      return (long)a->c.P * a->c.N > (long)c2->c.P * c2->c.N;
    });
    std::vector<pimd_ctx *> h0, h1;
    for (size_t i = 0; i < v.size(); ++i) ((i % 2 == 0) ? h0 : h1).push_back(v[i]);
    lanes[big] = h0;
    lanes.push_back(h1);
  }
  for (const auto &l : lanes) {
    b->vv.emplace_back();
    for (pimd_ctx *x : l) push(b->vv.back(), x);
  }
  std::vector<int> intras;
  for (pimd_ctx *x : b->ctx)
    if (x->md.intra_timesteps > 0 &&
        std::find(intras.begin(), intras.end(), x->md.intra_timesteps) == intras.end())
      intras.push_back(x->md.intra_timesteps);
  for (int in : intras) {
    pimd_batch::Con cn;
    cn.intra = in;
    for (pimd_ctx *x : b->ctx)
      if (x->md.intra_timesteps == in) push(cn.outer, x);
    for (pimd_ctx *x : b->ctx)
      if (x->md.intra_timesteps == in) push(cn.cen, x->cen);
    b->con.push_back(cn);
  }
  const int n = (int)host.size();
  if (n > b->table_cap) {
    if (b->table) (void)hipFree(b->table);
    b->table = nullptr;
    b->table_cap = 0;
    int rc = dalloc(&b->table, (size_t)n);
    if (rc != PIMD_OK) return rc;
    b->table_cap = n;
    b->graph_ok = false;
  }
  int off = 0;
  for (Group &g : b->vv) {
    g.d = b->table + off;
    off += g.n();
  }
  for (auto &cn : b->con) {
    cn.outer.d = b->table + off;
    off += cn.outer.n();
    cn.cen.d = b->table + off;
    off += cn.cen.n();
  }
  HIPCHK(hipMemcpyAsync(b->table, host.data(), sizeof(Rep) * n, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  // streams and events of the side lanes (default priority)
  const size_t nl = b->vv.size() + b->con.size();
  if (!b->fork && nl > 1 && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    return fail(PIMD_EDEVICE, "hipEventCreate failed");
  while (nl > 1 && b->lanes.size() < nl - 1) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
      return fail(PIMD_EDEVICE, "hipStreamCreate failed");
    b->lanes.push_back(st);
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
      return fail(PIMD_EDEVICE, "hipEventCreate failed");
    b->joins.push_back(ev);
  }
  if (sig != b->sig) {
    b->sig.swap(sig);
    b->graph_ok = false;
  }
  return PIMD_OK;
}

// One MD step of every member: each group (a force class of plain PIMD
// members, or the contracted members of one inner-step count) on a lane of
// its own, the first on the batch's stream, the others forked from it and
// joined back.
static int batch_step(pimd_batch *b) {
  auto no_noise = [](int) { return false; };
  const size_t nl = b->vv.size() + b->con.size();
  if (nl > 1) {
    HIPCHK(hipEventRecord(b->fork, b->stream));
    for (size_t i = 0; i + 1 < nl; ++i) HIPCHK(hipStreamWaitEvent(b->lanes[i], b->fork, 0));
  }
  int rc = PIMD_OK;
  size_t lane = 0;
  auto stream_of = [&](size_t i) { return i == 0 ? b->stream : b->lanes[i - 1]; };
  for (Group &g : b->vv)
    if ((rc = step_vv(g, stream_of(lane++), nullptr, no_noise)) != PIMD_OK) return rc;
  for (auto &cn : b->con)
    if ((rc = step_contracted(cn.outer, cn.cen, cn.intra, stream_of(lane++), nullptr, no_noise)) !=
        PIMD_OK)
      return rc;
  for (size_t i = 0; i + 1 < nl; ++i) {
    HIPCHK(hipEventRecord(b->joins[i], b->lanes[i]));
    HIPCHK(hipStreamWaitEvent(b->stream, b->joins[i], 0));
  }
  return PIMD_OK;
}

// The batch step as one replay of a captured graph (captured again when the
// tables' signature changes); eager launches where capture fails or with
// PIMD_MD_GRAPH=0.
static int batch_md_step(pimd_batch *b) {
  if (env_int("PIMD_MD_GRAPH", 1, 0, 1) == 0) return batch_step(b);
  if (!b->graph_ok) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_step(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != PIMD_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    hipError_t e2 = e;
    if (e == hipSuccess) e2 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      b->graph = nullptr;
      (void)hipGetLastError();
      return batch_step(b);
    }
    b->graph_ok = true;
  }
  HIPCHK(hipGraphLaunch(b->graph, b->stream));
  return PIMD_OK;
}

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
static int set_shape(pimd_ctx *x, const Shape &sh) {
  invalidate_graphs(x);
  x->c.g_qq = sh.g_qq;
  x->c.g_dd = sh.g_dd;
  x->c.g_fin = sh.g_fin;
  x->c.fd_cg = sh.fd_cg;
  x->c.sf_sites = sh.sf_sites;
  return sync_rep(x);
}

// The contexts whose force calls a batch step makes (the replicas of plain
// PIMD, the centroids of contracted MD), in classes that share launch
// shapes: one model and arithmetic each.
static std::vector<std::vector<pimd_ctx *>> force_classes(const pimd_batch *b) {
  std::vector<std::vector<pimd_ctx *>> cls;
  std::vector<std::pair<int, int>> keys;
  for (pimd_ctx *x : b->ctx) {
    pimd_ctx *f = x->cen ? x->cen : x;
    const std::pair<int, int> k{f->c.model, f->c.mixed};
    auto it = std::find(keys.begin(), keys.end(), k);
    if (it == keys.end()) {
      keys.push_back(k);
      cls.push_back({f});
    } else {
      cls[it - keys.begin()].push_back(f);
    }
  }
  return cls;
}

// A class's members take the launch shapes of its largest member (in
// molecule-beads), so that each kernel of a batch step is one launch per class.
static int unify_shapes(pimd_batch *b) {
  for (auto &cl : force_classes(b)) {
    pimd_ctx *big = cl[0];
    for (pimd_ctx *f : cl)
      if ((long)f->c.P * f->c.N > (long)big->c.P * big->c.N) big = f;
    const Shape sh = get_shape(big);
    for (pimd_ctx *f : cl) {
      const int rc = set_shape(f, sh);
      if (rc != PIMD_OK) return rc;
    }
  }
  return PIMD_OK;
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
  HIPCHK(hipSetDevice(ctxs[0]->p.device));
  pimd_batch *b = new pimd_batch();
  b->ctx.assign(ctxs, ctxs + n);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(PIMD_EDEVICE, "hipStreamCreate failed");
  }
  // PIMD_BATCH_UNIFY=0 keeps every member's own launch shapes
  int rc = env_int("PIMD_BATCH_UNIFY", 1, 0, 1) ? unify_shapes(b) : PIMD_OK;
  if (rc != PIMD_OK) {
    (void)pimd_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PIMD_OK;
}

extern "C" int pimd_batch_destroy(pimd_batch *b) {
  if (!b) return PIMD_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (hipStream_t st : b->lanes) (void)hipStreamSynchronize(st);
  for (hipStream_t st : b->lanes) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : b->joins) (void)hipEventDestroy(ev);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->table) (void)hipFree(b->table);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PIMD_OK;
}

// Mean time (ms) of one force call of every member of a class, all in one
// group, on the members' loaded coordinates.
static int time_class(pimd_batch *b, const std::vector<pimd_ctx *> &cl, Rep *tab,
                      hipEvent_t e0, hipEvent_t e1, int reps, double &ms) {
  Group g;
  std::vector<Rep> host;
  for (pimd_ctx *f : cl) {
    host.push_back(host_rep(f));
    g.hc.push_back(f->c);
    g.hm.push_back(f->m);
  }
  g.d = tab;
  HIPCHK(hipMemcpyAsync(tab, host.data(), sizeof(Rep) * host.size(), hipMemcpyHostToDevice, b->stream));
  for (pimd_ctx *f : cl) launch_set_clock(f->self, 1, 0, b->stream);
  int rc = run_forces(g, b->stream, nullptr);
  if (rc != PIMD_OK) return rc;
  HIPCHK(hipEventRecord(e0, b->stream));
  for (int r = 0; r < reps; ++r)
    if ((rc = run_forces(g, b->stream, nullptr)) != PIMD_OK) return rc;
  HIPCHK(hipEventRecord(e1, b->stream));
  HIPCHK(hipEventSynchronize(e1));
  float f = 0.f;
  HIPCHK(hipEventElapsedTime(&f, e0, e1));
  ms = (double)f / reps;
  return PIMD_OK;
}

// Coordinate search over the launch shapes of each class, timing the class's
// force calls together on the members' current coordinates.  Results do not
// depend on the shapes beyond summation order.  What the force calls
// overwrite of the MD state (the gradient, the dipole predictor history, the
// centroid of contracted MD) is saved and restored around the trials.
extern "C" int pimd_batch_autotune(pimd_batch *b, double *ms) {
  if (!b) return fail(PIMD_EINVAL, "null argument");
  for (size_t i = 0; i < b->ctx.size(); ++i)
    if (!b->ctx[i]->state_ready)
      return fail(PIMD_ESTATE, "replica %zu: pimd_state_set has not been called", i);
  HIPCHK(hipSetDevice(b->ctx[0]->p.device));
  // saved state: gradient and history of the plain replicas
  size_t nsave = 0;
  for (pimd_ctx *x : b->ctx) {
    const size_t P = x->c.P, N = x->c.N;
    nsave += P * x->natoms * 3 + P * 4 * N * x->c.nd * 3;
    HIPCHK(hipStreamSynchronize(x->stream));
  }
  double *save = nullptr;
  Rep *tab = nullptr;
  int rc = dalloc(&save, nsave);
  if (rc == PIMD_OK) rc = dalloc(&tab, b->ctx.size());
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == PIMD_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(PIMD_EDEVICE, "hipEventCreate failed");
  auto each_state = [&](bool to_save) {
    size_t o = 0;
    for (pimd_ctx *x : b->ctx) {
      const size_t ng = (size_t)x->c.P * x->natoms * 3, nh = (size_t)x->c.P * 4 * x->c.N * x->c.nd * 3;
      double *a = save + o, *h = save + o + ng;
      if (to_save) {
        (void)hipMemcpyAsync(a, x->d.dRR, sizeof(double) * ng, hipMemcpyDeviceToDevice, b->stream);
        (void)hipMemcpyAsync(h, x->d.hist, sizeof(double) * nh, hipMemcpyDeviceToDevice, b->stream);
      } else {
        (void)hipMemcpyAsync(x->d.dRR, a, sizeof(double) * ng, hipMemcpyDeviceToDevice, b->stream);
        (void)hipMemcpyAsync(x->d.hist, h, sizeof(double) * nh, hipMemcpyDeviceToDevice, b->stream);
      }
      o += ng + nh;
    }
  };
  if (rc == PIMD_OK) each_state(true);
  // candidate values per knob, in the order of Shape
  const int cand[NSHAPE][5] = {{4, 8, 16, 32, 0}, {8, 16, 32, 64, 0}, {4, 8, 16, 32, 0},
                               {4, 8, 16, 0, 0}, {4, 8, 16, 32, 64}};
  const auto classes = force_classes(b);
  std::map<const pimd_ctx *, double> best_ms;
  for (const auto &cl : classes) {
    if (rc != PIMD_OK) break;
    Shape best = get_shape(cl[0]);
    double tbest = 0.0;
    rc = time_class(b, cl, tab, e0, e1, 2, tbest);
    for (int knob = 0; knob < NSHAPE && rc == PIMD_OK; ++knob) {
      for (int v = 0; v < 5 && cand[knob][v] != 0 && rc == PIMD_OK; ++v) {
        Shape sh = best;
        int *fields[NSHAPE] = {&sh.g_qq, &sh.g_dd, &sh.g_fin, &sh.fd_cg, &sh.sf_sites};
        if (*fields[knob] == cand[knob][v]) continue;
        *fields[knob] = cand[knob][v];
        for (pimd_ctx *f : cl)
          if ((rc = set_shape(f, sh)) != PIMD_OK) break;
        double t = 0.0;
        if (rc == PIMD_OK) rc = time_class(b, cl, tab, e0, e1, 2, t);
        if (rc == PIMD_OK && t < tbest) {
          tbest = t;
          best = sh;
        }
      }
    }
    for (pimd_ctx *f : cl) {
      if (rc == PIMD_OK) rc = set_shape(f, best);
      best_ms[f] = tbest;
    }
  }
  if (ms)
    for (size_t i = 0; i < b->ctx.size(); ++i) {
      const pimd_ctx *f = b->ctx[i]->cen ? b->ctx[i]->cen : b->ctx[i];
      ms[i] = best_ms.count(f) ? best_ms[f] : 0.0;
    }
  if (save) {
    each_state(false);
    (void)hipStreamSynchronize(b->stream);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (save) (void)hipFree(save);
  if (tab) (void)hipFree(tab);
  b->graph_ok = false;
  return rc;
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
    pimd_ctx *x = b->ctx[r];
    if ((rc = md_check(x)) != PIMD_OK) return rc;
    if ((rc = md_records(x, nsteps)) != PIMD_OK) return rc;
    HIPCHK(hipStreamSynchronize(x->stream));
    x->m.seed = seeds[r];  // uploaded with the batch's replica records
  }
  if ((rc = batch_tables(b)) != PIMD_OK) return rc;
  // the clocks, and the kinetic energies of the loaded states (needed by the
  // first thermostat call)
  for (Group *g = b->vv.data(); g != b->vv.data() + b->vv.size(); ++g) {
    if (!g->n()) continue;
    launch_set_clock(*g, t0, 0, b->stream);
    launch_kinetic(*g, b->stream);
  }
  for (auto &cn : b->con) {
    launch_set_clock(cn.outer, t0, 0, b->stream);
    launch_kinetic(cn.outer, b->stream);
  }
  for (int st = 0; st < nsteps; ++st)
    if ((rc = batch_md_step(b)) != PIMD_OK) return rc;
  HIPCHK(hipStreamSynchronize(b->stream));
  for (int r = 0; r < n; ++r)
    if ((rc = md_copy_records(b->ctx[r], nsteps, Upot ? Upot + (size_t)r * nsteps : nullptr,
                              nullptr, nullptr, uk ? uk + (size_t)r * nsteps : nullptr)) != PIMD_OK ||
        (rc = take_flags(b->ctx[r])) != PIMD_OK)
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
  bool ok = true;
  x->timed("pes_adjoint", [&] { ok = launch_pes_adjoint(c, d, dR, dE, dG, part, nb, dgrad, s); });
  if (!ok) return fail(PIMD_EDEVICE, "copying the PES exponent table to the device failed");
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
    const int rc = sync_rep(y);
    if (rc != PIMD_OK) return rc;
  }
  invalidate_graphs(x);  // captured launches hold the model constants
  return PIMD_OK;
}

// Tangent sweeps along ndir directions (mode 0 the TTM parameters, 1 the DMS
// coefficients, 2 gammaM), or with mode 3 the adjoint of all TTM_ADJ_NPARAM
// parameters in one sweep (ndir = TTM_ADJ_NPARAM outputs): the loss
// gradient, and optionally each configuration's energy and its derivatives.
// A parameter-derivative sweep in two halves: tangent_start queues the
// inputs (through pinned host memory), the sweep and the copy of its outputs
// on the context's stream and returns; tangent_finish waits for them and
// forms the gradient.  Sweeps of different contexts run side by side.
static int tangent_start(pimd_ctx *x, const double *RRt, const double *ew, const double *gw,
                         int mode, int ndir) {
  if (!x || !RRt || !ew || !gw) return fail(PIMD_EINVAL, "null argument");
  if (x->tanp.on) return fail(PIMD_EINVAL, "a parameter-derivative sweep is still to be finished");
  if (x->p.contraction)
    return fail(PIMD_EINVAL, "parameter derivatives need a context with contraction = 0");
  if (x->p.pot_model >= PIMD_QSPCFW)
    return fail(PIMD_EINVAL, "TTM parameter derivatives need a TTM model (pot_model 2, 3)");
  HIPCHK(hipSetDevice(x->p.device));
  const SysConst &c = x->c;
  DevBufs &d = x->d;
  hipStream_t s = x->stream;
  const bool adj = mode == 3;
  const int nsweep = adj ? 1 : ndir;  // workgroups per configuration
  const size_t P = (size_t)c.P, n = P * x->natoms * 3, nout = P * nsweep * 4;
  const size_t nadj = adj ? P * TTM_ADJ_NPARAM * 2 : 0;
  const size_t ws = tangent_ws_duals(c) * P * nsweep * 2, cfg = tangent_cfg_doubles(c) * P;
  const size_t bytes = sizeof(double) * (ws + cfg + 2 * n + P + nout + nadj);
  if (bytes > ((size_t)32 << 30))
    return fail(PIMD_EINVAL, "parameter derivatives of %zu configurations need %.1f GB of workspace; "
                "split them into smaller contexts", P, bytes / 1e9);
  if (bytes > d.tan_bytes) {
    if (d.tan) (void)hipFree(d.tan);
    d.tan = nullptr;
    d.tan_bytes = 0;
    char *t = nullptr;
    int rc = dalloc(&t, bytes);
    if (rc != PIMD_OK) return rc;
    d.tan = (decltype(d.tan))t;
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
  L.dms_coef = mode == 1;
  L.gamma_m = mode == 2;
  L.adjoint = adj;
  L.adj = L.out + nout;
  L.ndir = nsweep;
  std::copy(x->ttm, x->ttm + TTM_NPARAM, L.theta);
  const double Rc = x->p.rc, pre = TWOPI * (double)(c.N * c.N) / c.vol;
  L.lrc_coef[0] = pre / (9.0 * std::pow(Rc, 9));
  L.lrc_coef[1] = pre / (7.0 * std::pow(Rc, 7));
  L.lrc_coef[2] = pre / (3.0 * (Rc * Rc * Rc));
  // pinned: the inputs (the caller's arrays are free once this returns), then
  // the outputs
  const size_t hbytes = sizeof(double) * (2 * n + P + nout + nadj);
  if (hbytes > x->h_tan_bytes) {
    if (x->h_tan) (void)hipHostFree(x->h_tan);
    x->h_tan = nullptr;
    x->h_tan_bytes = 0;
    HIPCHK(hipHostMalloc((void **)&x->h_tan, hbytes, hipHostMallocDefault));
    x->h_tan_bytes = hbytes;
  }
  double *hin = x->h_tan;
  std::copy(RRt, RRt + n, hin);
  std::copy(gw, gw + n, hin + n);
  std::copy(ew, ew + P, hin + 2 * n);
  HIPCHK(hipMemcpyAsync(dRR, hin, sizeof(double) * (2 * n + P), hipMemcpyHostToDevice, s));
  static const char *const names[4] = {"ttm_tangent", "dms_tangent", "gamma_m_tangent",
                                       "ttm_adjoint"};
  x->timed(names[mode], [&] { launch_ttm_tangent(c, L, s); });
  HIPCHK(hipGetLastError());
  double *hout = hin + 2 * n + P;
  HIPCHK(hipMemcpyAsync(hout, L.out, sizeof(double) * (nout + nadj), hipMemcpyDeviceToHost, s));
  auto &tp = x->tanp;
  tp.on = true;
  tp.mode = mode;
  tp.ndir = ndir;
  tp.nout = nout;
  tp.nadj = nadj;
  tp.ew.assign(ew, ew + P);
  tp.out = hout;
  return PIMD_OK;
}

static int tangent_finish(pimd_ctx *x, double *grad, double *E, double *dEdtheta) {
  if (!x || !grad) return fail(PIMD_EINVAL, "null argument");
  auto &tp = x->tanp;
  if (!tp.on) return fail(PIMD_EINVAL, "no parameter-derivative sweep was started");
  HIPCHK(hipSetDevice(x->p.device));
  tp.on = false;
  HIPCHK(hipStreamSynchronize(x->stream));
  const size_t P = (size_t)x->c.P, nout = tp.nout;
  const int ndir = tp.ndir;
  const bool adj = tp.mode == 3;
  const double *ew = tp.ew.data();
  const double *out = tp.out;
  if (adj) {
    // grad_j = sum_b ew_b dE_b/dtheta_j + gw . d(dE_b/dr)/dtheta_j
    const double *A = out + nout;
    for (int j = 0; j < ndir; ++j) {
      double g = 0.0;
      for (size_t b = 0; b < P; ++b)
        g += ew[b] * A[(b * ndir + j) * 2] + A[(b * ndir + j) * 2 + 1];
      grad[j] = g;
    }
    for (size_t b = 0; b < P; ++b) {
      if (E) E[b] = out[b * 4];
      if (dEdtheta)
        for (int j = 0; j < ndir; ++j) dEdtheta[b * ndir + j] = A[(b * ndir + j) * 2];
    }
    return PIMD_OK;
  }
  for (int j = 0; j < ndir; ++j) {
    double g = 0.0;
    for (size_t b = 0; b < P; ++b) g += out[(b * ndir + j) * 4 + 2];
    grad[j] = g;
  }
  for (size_t b = 0; b < P; ++b) {
    if (E) E[b] = out[(b * ndir) * 4];
    if (dEdtheta)
      for (int j = 0; j < ndir; ++j) dEdtheta[b * ndir + j] = out[(b * ndir + j) * 4 + 1];
  }
  return PIMD_OK;
}

static int tangent_grad(pimd_ctx *x, const double *RRt, const double *ew, const double *gw,
                        int mode, int ndir, double *grad, double *E, double *dEdtheta) {
  if (!grad) return fail(PIMD_EINVAL, "null argument");
  const int rc = tangent_start(x, RRt, ew, gw, mode, ndir);
  return rc != PIMD_OK ? rc : tangent_finish(x, grad, E, dEdtheta);
}

extern "C" int pimd_ttm_param_grad(pimd_ctx *x, const double *RRt, const double *ew,
                                   const double *gw, double *grad, double *E, double *dEdtheta) {
  return tangent_grad(x, RRt, ew, gw, 0, TTM_NPARAM, grad, E, dEdtheta);
}

extern "C" int pimd_dms_coef_grad(pimd_ctx *x, const double *RRt, const double *ew,
                                  const double *gw, double *grad, double *E, double *dEdc) {
  return tangent_grad(x, RRt, ew, gw, 1, DMS_NCOEF, grad, E, dEdc);
}

extern "C" int pimd_gamma_m_grad(pimd_ctx *x, const double *RRt, const double *ew,
                                 const double *gw, double *grad, double *E, double *dEdg) {
  return tangent_grad(x, RRt, ew, gw, 2, 1, grad, E, dEdg);
}

extern "C" int pimd_ttm_adjoint(pimd_ctx *x, const double *RRt, const double *ew, const double *gw,
                                int nparam, double *grad, double *E, double *dEdtheta) {
  if (nparam != TTM_NPARAM && nparam != TTM_NPARAM + DMS_NCOEF && nparam != TTM_ADJ_NPARAM)
    return fail(PIMD_EINVAL, "nparam must be %d, %d or %d", TTM_NPARAM, TTM_NPARAM + DMS_NCOEF,
                TTM_ADJ_NPARAM);
  if (!x || !grad) return fail(PIMD_EINVAL, "null argument");
  const size_t P = (size_t)x->c.P;
  std::vector<double> g(TTM_ADJ_NPARAM), dE(dEdtheta ? P * TTM_ADJ_NPARAM : 0);
  int rc = tangent_grad(x, RRt, ew, gw, 3, TTM_ADJ_NPARAM, g.data(), E,
                        dEdtheta ? dE.data() : nullptr);
  if (rc != PIMD_OK) return rc;
  std::copy(g.begin(), g.begin() + nparam, grad);
  if (dEdtheta)
    for (size_t b = 0; b < P; ++b)
      std::copy(dE.begin() + b * TTM_ADJ_NPARAM, dE.begin() + b * TTM_ADJ_NPARAM + nparam,
                dEdtheta + b * nparam);
  if (nparam < TTM_ADJ_NPARAM) return PIMD_OK;
  // gammaM moves the M sites, so its energy derivative is a position
  // derivative of the energy as computed, which the reference's gradient
  // matches only to ~1e-5: one tangent sweep along it instead
  std::vector<double> dEg(P);
  double gg = 0.0;
  if ((rc = tangent_grad(x, RRt, ew, gw, 2, 1, &gg, nullptr, dEg.data())) != PIMD_OK) return rc;
  grad[TTM_ADJ_NPARAM - 1] = gg;
  if (dEdtheta)
    for (size_t b = 0; b < P; ++b) dEdtheta[b * TTM_ADJ_NPARAM + TTM_ADJ_NPARAM - 1] = dEg[b];
  return PIMD_OK;
}

extern "C" int pimd_ttm_adjoint_start(pimd_ctx *x, const double *RRt, const double *ew,
                                      const double *gw, int nparam) {
  if (nparam != TTM_NPARAM && nparam != TTM_NPARAM + DMS_NCOEF)
    return fail(PIMD_EINVAL, "nparam must be %d or %d (gammaM: pimd_ttm_adjoint)", TTM_NPARAM,
                TTM_NPARAM + DMS_NCOEF);
  if (!x) return fail(PIMD_EINVAL, "null argument");
  const int rc = tangent_start(x, RRt, ew, gw, 3, TTM_ADJ_NPARAM);
  if (rc == PIMD_OK) x->tan_nparam = nparam;
  return rc;
}

extern "C" int pimd_ttm_adjoint_finish(pimd_ctx *x, double *grad, double *E, double *dEdtheta) {
  if (!x || !grad) return fail(PIMD_EINVAL, "null argument");
  if (!x->tanp.on || x->tanp.mode != 3)
    return fail(PIMD_EINVAL, "no adjoint sweep was started (pimd_ttm_adjoint_start)");
  const size_t P = (size_t)x->c.P;
  const int nparam = x->tan_nparam;
  std::vector<double> g(TTM_ADJ_NPARAM), dE(dEdtheta ? P * TTM_ADJ_NPARAM : 0);
  const int rc = tangent_finish(x, g.data(), E, dEdtheta ? dE.data() : nullptr);
  if (rc != PIMD_OK) return rc;
  std::copy(g.begin(), g.begin() + nparam, grad);
  if (dEdtheta)
    for (size_t b = 0; b < P; ++b)
      std::copy(dE.begin() + b * TTM_ADJ_NPARAM, dE.begin() + b * TTM_ADJ_NPARAM + nparam,
                dEdtheta + b * nparam);
  return PIMD_OK;
}

extern "C" int pimd_gamma_m_get(pimd_ctx *x, double *gamma_m) {
  if (!x || !gamma_m) return fail(PIMD_EINVAL, "null argument");
  *gamma_m = x->c.gammaM;
  return PIMD_OK;
}

extern "C" int pimd_gamma_m_set(pimd_ctx *x, double gamma_m) {
  if (!x) return fail(PIMD_EINVAL, "null argument");
  if (x->p.pot_model >= PIMD_QSPCFW)
    return fail(PIMD_EINVAL, "gammaM is a parameter of the TTM models (pot_model 2, 3)");
  if (!(gamma_m >= 0.0 && gamma_m < 1.0))
    return fail(PIMD_EINVAL, "gammaM must lie in [0, 1) (the M charge is q_O / (1 - gammaM))");
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipStreamSynchronize(x->stream));  // launches already queued keep the old value
  for (pimd_ctx *y : {x, x->cen}) {
    if (!y) continue;
    y->c.gammaM = gamma_m;
    const int rc = sync_rep(y);
    if (rc != PIMD_OK) return rc;
  }
  invalidate_graphs(x);  // captured launches hold the model constants
  return PIMD_OK;
}

extern "C" int pimd_dms_coefs_get(pimd_ctx *x, double *coef) {
  if (!x || !coef) return fail(PIMD_EINVAL, "null argument");
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(coef, x->d.pes + PES_DMSC, sizeof(double) * DMS_NCOEF, hipMemcpyDeviceToHost));
  return PIMD_OK;
}

extern "C" int pimd_dms_coefs_set(pimd_ctx *x, const double *coef) {
  if (!x || !coef) return fail(PIMD_EINVAL, "null argument");
  for (int j = 0; j < DMS_NCOEF; ++j)
    if (!std::isfinite(coef[j])) return fail(PIMD_EINVAL, "DMS coefficient %d is not finite", j);
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(x->d.pes + PES_DMSC, coef, sizeof(double) * DMS_NCOEF, hipMemcpyHostToDevice));
  // the centroid context of contracted MD charges its sites with them too
  if (x->cen)
    HIPCHK(hipMemcpy(x->cen->d.pes + PES_DMSC, coef, sizeof(double) * DMS_NCOEF,
                     hipMemcpyHostToDevice));
  return PIMD_OK;
}

extern "C" int pimd_state_forces(pimd_ctx *x, int t) {
  int rc = md_check(x);
  if (rc != PIMD_OK) return rc;
  HIPCHK(hipSetDevice(x->p.device));
  launch_set_clock(x->self, t, 0, x->stream);
  return run_forces(x);
}

// ---------------------------------------------------------------------------
// A force call that allocates nothing (pimd_forces_all_beads_noalloc): every
// intermediate array of the call lives in one static device workspace per
// GPU, set up at the first call on that device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) char scratch[PIMD_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[MAX_DEVICES] = {};

// The device address of this device's workspace (the device must be
// current); the symbol's own name is only a host-side handle.
int scratch_on_device(int dev, char **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= MAX_DEVICES) return fail(PIMD_EINVAL, "device %d out of range", dev);
  if (!g_scratch_ptr[dev]) {
    char *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(PIMD_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return PIMD_OK;
}
}  // namespace

extern "C" int pimd_forces_all_beads_noalloc(const pimd_params *params, const double *RR,
                                             const double *RRc, int t, double *Upot, double *dRR,
                                             double *virt, double *virialc, double *dipI,
                                             double *dipE, double *chg, int *iters, void *stream) {
  if (!params || !RR || !RRc) return fail(PIMD_EINVAL, "null argument");
  pimd_params p = *params;
  int rc = check_params(p);
  if (rc != PIMD_OK) return rc;
  if (p.guess_initdip)
    return fail(PIMD_EINVAL, "the dipole predictor keeps a history between calls; "
                "pimd_forces_all_beads_noalloc needs guess_initdip = 0");
  char *base = nullptr;
  if ((rc = scratch_on_device(p.device, &base)) != PIMD_OK) return rc;
  // a host-side context whose device buffers are carved from the workspace
  pimd_ctx x;
  x.p = p;
  x.natoms = 3 * p.nwaters;
  x.stream = static_cast<hipStream_t>(stream);
  if ((rc = setup_system(&x)) != PIMD_OK) return rc;
  HostTables tabs;
  plan_system(&x, tabs);
  // one stream and no host polling: nothing here creates streams or pinned
  // memory; the dipole pass recomputes its pair screening factors every SOR
  // iteration (the same values as the cache would hold)
  x.c.poll = 0;
  x.c.overlap = 0;
  x.c.dd_cache = 0;
  // partial sums for the chunks the structure factors use (the same
  // chunking as a context's, so the results are the same bits)
  x.c.sf_chunks_max = sfac_chunks_used(x.c);
  Arena a;
  layout_system(&x, a, false);
  if (a.off > sizeof(scratch))
    return fail(PIMD_EINVAL, "%d waters x %d beads need %zu bytes of device workspace, more than "
                "the %zu bytes of pimd_forces_all_beads_noalloc", p.nwaters, p.nbeads, a.off,
                sizeof(scratch));
  a.base = base;
  a.off = 0;
  layout_system(&x, a, false);
  if ((rc = upload_system(&x, tabs, x.stream)) != PIMD_OK) return rc;
  x.c.g_nolds = pick_group(x.c);
  x.m.massO = p.mass_o;
  x.m.massH = p.mass_h;
  if ((rc = sync_rep(&x)) != PIMD_OK) return rc;  // the replica record, in the workspace too
  DevBufs &d = x.d;
  hipStream_t s = x.stream;
  const size_t P = p.nbeads, N = p.nwaters, na = 3 * N;
  HIPCHK(hipMemcpyAsync(d.RR, RR, sizeof(double) * P * na * 3, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(d.RRc, RRc, sizeof(double) * na * 3, hipMemcpyDeviceToDevice, s));
  launch_set_clock(x.self, t, 0, s);
  rc = run_forces(&x);
  struct Out {
    void *dst;
    const void *src;
    size_t bytes;
  } outs[] = {{Upot, d.Upot, sizeof(double) * P},
              {dRR, d.dRR, sizeof(double) * P * na * 3},
              {virt, d.virt, sizeof(double) * P * 9},
              {virialc, d.virialc, sizeof(double) * P},
              {dipI, d.dipI, sizeof(double) * P * N * 3},
              {dipE, d.dipE, sizeof(double) * P * N * 3},
              {chg, d.chg, sizeof(double) * P * na},
              {iters, d.iters_out, sizeof(int) * P}};
  for (const Out &o : outs)
    if (rc == PIMD_OK && o.dst)
      HIPCHK(hipMemcpyAsync(o.dst, o.src, o.bytes, hipMemcpyDeviceToDevice, s));
  x.d = DevBufs{};  // workspace addresses, not allocations of x
  return rc;
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
  HIPCHK(hipSetDevice(x->p.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  return take_flags(x);
}

}  // extern "C"

int pimd::xcd_numbering() {
  static const int v = env_int("PIMD_XCD_MAP", 1, 0, 1);
  return v;
}
