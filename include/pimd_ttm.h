/*
 * pimd_ttm.h - C ABI of the MI355X (gfx950) TTM2.1-F / TTM3-F force path and
 * ring-polymer propagation.
 *
 * All P beads of one system are evaluated together on one GPU.  Arrays keep
 * the reference's Fortran column-major layouts read as C row-major:
 *   RRt(3,Natoms,Nbeads)   -> double [P][Natoms][3]   (atoms in O,H,H order)
 *   virt(3,3)              -> double [col][row]
 *   dip_momIt(3,Nwaters,P) -> double [P][Nwaters][3]
 * Units are the reference's: Angstrom, kcal/mol, amu*Angstrom/ps, ps.
 * dRRt is the gradient dU/dr (not the force), as in the reference.
 *
 * Every entry point returns PIMD_OK (0) or an error code; the message of the
 * last failure on the calling thread is available from pimd_last_error().
 * A context is not thread-safe: calls on one context must be serialised.
 */
#ifndef PIMD_TTM_H
#define PIMD_TTM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIMD_OK 0
#define PIMD_EINVAL 1   /* invalid argument or configuration */
#define PIMD_EDEVICE 2  /* HIP runtime / device failure */
#define PIMD_ENOMEM 3   /* device allocation failed */
#define PIMD_ESTATE 4   /* call out of order (e.g. nm_evolve before nm_init) */

#define PIMD_TTM21F 2
#define PIMD_TTM3F 3
#define PIMD_QSPCFW 4   /* q-SPC/Fw, fixed charges (pot_spc.f90:69-106)       */
#define PIMD_SPCF 5     /* SPC/F, fixed charges (pot_spc.f90:107-137)         */

#define PIMD_FP64 0     /* plain double precision, the reference's arithmetic */
#define PIMD_MIXED 1    /* real-space pair terms and the real-space dipole field
                           of the SOR solve in FP32 (accumulated in FP64); the
                           last SOR step of every bead redone in FP64; Ewald
                           k-space, energies, virials and ring-polymer steps
                           in FP64.  No counterpart in the reference.         */

typedef struct pimd_ctx pimd_ctx;

/* Run parameters.  Sentinels follow the reference's derived defaults
 * (InputOutput.f90:118-125): rc <= 0 -> min(box)/2; rc1 < 0 -> 0.8*rc with
 * 0.8 a single-precision literal. */
typedef struct {
  int pot_model;      /* PIMD_TTM21F or PIMD_TTM3F (potential.f90:84-113), or
                         PIMD_QSPCFW or PIMD_SPCF (potential.f90:121-145)    */
  int nwaters;        /* Natoms = 3*nwaters                                  */
  int nbeads;         /* P                                                   */
  int polar_maxiter;  /* input_file_reader.f90:79, default 15                */
  int guess_initdip;  /* 4-point dipole predictor (pot_ttm.f90:281-297); the
                         history is kept per bead, as in a run with one MPI
                         rank per bead                                        */
  int precision;      /* PIMD_FP64 or PIMD_MIXED                             */
  int device;         /* HIP device ordinal                                  */
  int contraction;    /* 1: pot_ttm's CONTRACTION branch, no monomer PES energy
                         or gradient (pot_ttm.f90:92-132); 0 otherwise       */
  double box[3];      /* Angstrom                                            */
  double rc;          /* real-space cut-off (Angstrom)                       */
  double rc1;         /* vdW switch start (TTM3-F shift, potential.f90:107)  */
  double eps_ewald;   /* Ewald accuracy, sets alpha and |k| cut (ewald.f90:312) */
  double polar_sor;   /* SOR mixing (default real(4) 0.7)                    */
  double polar_eps;   /* dipole convergence, Debye RMS (pot_ttm.f90:306,381) */
  double mass_o;      /* amu, used by the Langevin thermostat                */
  double mass_h;
} pimd_params;

/* ---- context ---------------------------------------------------------- */

/* Replaces init_pot + ewald_set(.true.) (potential.f90:43-199,
 * ewald.f90:293-378) and the per-rank allocations. */
int pimd_ctx_create(const pimd_params *params, pimd_ctx **out);
int pimd_ctx_destroy(pimd_ctx *ctx);
/* Copies the message of the last failed call on this thread (NUL-terminated). */
int pimd_last_error(char *buf, size_t len);

/* Derived set-up values: out[0]=aewald, out[1]=Uvdw_lrc, out[2]=shiftA,
 * out[3]=shiftB, out[4..7]=polfac(1:4), out[8]=nkvec, out[9]=hmax,
 * out[10]=kmax, out[11]=lmax, out[12]=1 when the neighbour search uses a
 * cell list (13 doubles). */
int pimd_ctx_info(const pimd_ctx *ctx, double *out);

/* ---- forces: force_calc.f90:26-106 full_bead_forces ------------------- */

/* One pot_ttm evaluation (pot_ttm.f90:3-630) for each of the P beads, all in
 * one set of launches.  Host arrays:
 *   RRt[P][Natoms][3], RRc[Natoms][3]           in
 *   Upott[P], dRRt[P][Natoms][3], virt[P][3][3] (column-major per bead),
 *   virialct[P], dip_momIt[P][N][3], dip_momEt[P][N][3], chg[P][Natoms],
 *   iters[P] (SOR iterations run; negative when polar_maxiter was reached
 *   without convergence)                         out, any may be NULL
 * t is the MD step index passed to pot_ttm (drives the dipole predictor).
 * Non-finite coordinates are refused with PIMD_EINVAL before anything runs. */
int pimd_forces_all_beads(pimd_ctx *ctx, const double *RRt, const double *RRc, int t,
                          double *Upott, double *dRRt, double *virt, double *virialct,
                          double *dip_momIt, double *dip_momEt, double *chg, int *iters);
/* pimd_forces_all_beads in two halves: _start checks and copies the inputs
 * (the arrays are free once it returns) and queues the call and the copy of
 * every output on the context's stream; _finish waits and fills the outputs
 * given (any may be NULL), the same values as pimd_forces_all_beads.  Calls
 * started on several contexts run side by side.  No counterpart in the
 * reference (its ranks each call pot_ttm and wait).  PIMD_EINVAL: _start
 * while a started call of the context is unfinished, _finish with none
 * started. */
int pimd_forces_all_beads_start(pimd_ctx *ctx, const double *RRt, const double *RRc, int t);
int pimd_forces_all_beads_finish(pimd_ctx *ctx, double *Upott, double *dRRt, double *virt,
                                 double *virialct, double *dip_momIt, double *dip_momEt,
                                 double *chg, int *iters);

/* Same computation on device-resident arrays (same shapes), enqueued on the
 * context's stream.  Outputs may be NULL.  Does not synchronise: non-finite
 * coordinates (seen by the kernels) are reported as PIMD_EINVAL by the next
 * pimd_sync, pimd_forces_all_beads or pimd_md_run of the context. */
int pimd_forces_all_beads_dev(pimd_ctx *ctx, const double *d_RRt, const double *d_RRc, int t,
                              double *d_Upott, double *d_dRRt, double *d_virt,
                              double *d_virialct, double *d_dip_momIt, double *d_dip_momEt,
                              double *d_chg, int *d_iters);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pimd_forces_all_beads_dev without a context.  The
 * parameters are checked and set up as pimd_ctx_create does; every
 * intermediate array of the call lives in one static device workspace of
 * PIMD_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, and the
 * call makes no device or pinned allocation.  Systems whose arrays exceed the
 * workspace are refused with PIMD_EINVAL (128 waters x 32 beads fit at
 * eps_ewald 1e-6 with Rc = L/2; 512 x 32 does not).  guess_initdip must be 0
 * (the predictor history would have to persist between calls).  Device
 * arrays as in pimd_forces_all_beads_dev; outputs may be NULL.  stream: a
 * hipStream_t, NULL for the default stream; the kernels are enqueued there
 * and the call returns without waiting for them.  Every call on one GPU
 * shares that GPU's workspace: calls that overlap on one GPU (on different
 * streams or host threads) race. */
#define PIMD_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int pimd_forces_all_beads_noalloc(const pimd_params *params, const double *d_RRt,
                                  const double *d_RRc, int t, double *d_Upott, double *d_dRRt,
                                  double *d_virt, double *d_virialct, double *d_dip_momIt,
                                  double *d_dip_momEt, double *d_chg, int *d_iters, void *stream);

/* calc_monomer_dip_moments (force_calc.f90:205-264): the monomer dipole
 * moment of every molecule of every bead from its dipole-moment-surface
 * charges.  RRt [P][Natoms][3] host; dip_momIt [P][Nwaters][3] host.  Does
 * not change the results of the last force call. */
int pimd_monomer_dipoles(pimd_ctx *ctx, const double *RRt, double *dip_momIt);

/* Energy terms of the last force call, per bead: [P][5] =
 * Umon, Uvdw, Uelec, Uind, Uvdw_lrc (system_mod.f90:30 module variables). */
int pimd_energy_terms(pimd_ctx *ctx, double *terms);

/* ---- ring polymer: NormalModes.f90 / Langevin.f90 / Nose_Hoover.f90 --- */

/* InitNormalModes(Nbeads, omegan, delta, setNMfreq) (NormalModes.f90:68-218);
 * delta in ps. */
int pimd_nm_init(pimd_ctx *ctx, double omegan, double delta, double setNMfreq);
/* Module data C[mode][bead], A11..A22, omegalist, MassScaleFactor (any NULL). */
int pimd_nm_get(pimd_ctx *ctx, double *C, double *A11, double *A12, double *A21, double *A22,
                double *omegalist, double *mass_scale);
/* EvolveRing for every atom at once (replaces the 3N calls of MD.f90:52-58).
 * RRt, PPt [P][Natoms][3] updated in place; masses[Natoms] in amu. */
int pimd_nm_evolve(pimd_ctx *ctx, double *RRt, double *PPt, const double *masses);

/* Init_Langevin_NM(delt2, CENTROIDTHERMOSTAT, tau_bead, Nbeads, temp)
 * (Langevin.f90:68-98); temp is the bead temperature P*T as the reference
 * passes it (InputOutput.f90:405). */
int pimd_langevin_init(pimd_ctx *ctx, double delt2, int centroid_thermostat, double tau_bead,
                       double temp);
/* Langevin_NM(PPt) (Langevin.f90:30-64), PPt [P][Natoms][3] in place.
 * noise: [Natoms][P][3] standard normals in the reference's consumption
 * order, or NULL to draw them on the device (counter-based, seed/counter). */
int pimd_langevin(pimd_ctx *ctx, double *PPt, const double *noise, uint64_t seed,
                  uint64_t counter);

/* Nose_Hoover(s, uk, 2, vxi, tau, delt2, L, temp) (Nose_Hoover.f90:33-112),
 * chain length 2, on the host (scalar). */
int pimd_nose_hoover(double *s, double *uk, double *vxi, double tau, double delt2, int L,
                     double temp);

/* ---- device-resident PIMD velocity-Verlet: MD.f90:35-110 ------------- */

typedef struct {
  double delt_fs;        /* time step (fs)                                  */
  double temp;           /* K                                                */
  int bead_thermostat;   /* 0 none, 1 Langevin (NM, PILE, Langevin.f90:30-98),
                            2 Nose-Hoover chains in normal-mode space
                            (main_stuff.f90:145-171)                          */
  int centroid_thermostat; /* thermostat the centroid mode too              */
  double tau_centroid;   /* ps                                               */
  int global_nh;         /* global Nose-Hoover chain (MD.f90:43-46)          */
  int intra_timesteps;   /* 0: PIMD_VV (MD.f90:35-110); n >= 1: contracted_MD
                            (MD.f90:276-449), monomer forces on the beads in n
                            inner steps of delt/n, TTM forces of the centroid
                            once per step                                     */
  double tau;            /* ps, global thermostat                            */
  double setNMfreq;      /* cm^-1, 0 = RPMD                                  */
} pimd_md_params;

/* Configures the integrator: runs pimd_nm_init/pimd_langevin_init with the
 * reference's derived values (omegan = kB*T*P/hbar, delt2 = delt/2). */
int pimd_md_init(pimd_ctx *ctx, const pimd_md_params *md);
/* Loads RRt, PPt, dRRt [P][Natoms][3] into device state (dRRt may be NULL ->
 * zero, as the reference starts).  Non-finite values: PIMD_EINVAL.  A run
 * whose coordinates stop being finite (pimd_md_run, pimd_batch_md_run)
 * returns PIMD_EINVAL at its end. */
int pimd_state_set(pimd_ctx *ctx, const double *RRt, const double *PPt, const double *dRRt);
/* Copies the device state out (any may be NULL); RRc [Natoms][3]. */
int pimd_state_get(pimd_ctx *ctx, double *RRt, double *PPt, double *dRRt, double *RRc);
/* One force evaluation (full_bead_forces, force_calc.f90:26-106) on the
 * device-resident state coordinates; the gradient replaces the state's. */
int pimd_state_forces(pimd_ctx *ctx, int t);
/* Runs nsteps PIMD_VV steps (contracted_MD steps when intra_timesteps >= 1)
 * on the device starting at MD step t0.
 * noise: NULL (device RNG keyed by seed) or [nsteps][2][Natoms][P][3]
 * standard normals for the two Langevin calls of each step.
 * Per step outputs (any may be NULL): Upot[nsteps] = sum over beads,
 * virial[nsteps], virialc[nsteps] (force_calc.f90:94-97), uk[nsteps].
 * Contracted: Upot sums the monomer energies of every bead and P times the
 * centroid's energy, virial = monomer virial + trace of the centroid's,
 * virialc = monomer centroid virial / P + the centroid's (MD.f90:391-418).
 * The fast gradient starts at zero at pimd_md_init and carries over between
 * runs. */
int pimd_md_run(pimd_ctx *ctx, int t0, int nsteps, const double *noise, uint64_t seed,
                double *Upot, double *virial, double *virialc, double *uk);

/* Per-step observables of the last pimd_md_run / pimd_batch_md_run, as
 * write_out computes them after each step's force call
 * (InputOutput.f90:556-621, estimators.f90:30-96): obs[nsteps][PIMD_NOBS] =
 * sys_temp (K); quantum virial energy (kcal/mol per molecule) and pressure
 * (bar); simple (primitive) energy and pressure; total molecular dipole
 * x, y, z and the sum of molecular dipole magnitudes (Debye, each molecule's
 * dipole averaged over beads; NaN in contracted MD, which has no per-bead
 * dipoles). */
#define PIMD_NOBS 9
int pimd_md_observables(pimd_ctx *ctx, int nsteps, double *obs);

/* ---- batches of independent systems ------------------------------------
 * No counterpart in the reference, which runs one system per job.  A batch
 * runs several contexts (replicas of any size, model and bead count, each
 * already set up with pimd_md_init and pimd_state_set) side by side on one
 * device, each on its own stream. */
typedef struct pimd_batch pimd_batch;
int pimd_batch_create(pimd_ctx *const *ctxs, int n, pimd_batch **out);
int pimd_batch_destroy(pimd_batch *batch);
/* Picks each replica's launch shapes (lanes per molecule of the charge pair
 * kernel, of the dipole kernel and of the final pair kernel, column groups
 * of the field kernels, sites per structure-factor chunk) by timing force
 * evaluations of its loaded coordinates; ms[n] (may be NULL) receives the
 * best time per replica.  The MD state, including the dipole predictor
 * history, is left as it was. */
int pimd_batch_autotune(pimd_batch *batch, double *ms);
/* The five launch-shape values of replica i (int[5]), in the order above. */
int pimd_batch_shape(pimd_batch *batch, int i, int *shape);
/* nsteps PIMD_VV steps of every replica from MD step t0 (device normals,
 * seeds[n]); Upot[n][nsteps] and uk[n][nsteps] (may be NULL). */
int pimd_batch_md_run(pimd_batch *batch, int t0, int nsteps, const uint64_t *seeds, double *Upot,
                      double *uk);

/* ---- PES parameters and their adjoint ------------------------------------
 * No counterpart in the reference, which fits the monomer surface offline
 * (PES_fitter/fit_energy.m).  theta[PIMD_PES_NPARAM]: the 245 combined
 * Partridge-Schwenke coefficients c5z (nasa.f90:45), then deohA, alphaoh,
 * roh, phh1A, phh2 (nasa.f90:46-57).  A context starts with the reference's
 * values; later force calls use the values set. */
#define PIMD_PES_NPARAM 250
int pimd_pes_params_get(pimd_ctx *ctx, double *theta);
int pimd_pes_params_set(pimd_ctx *ctx, const double *theta);
/* Gradient of a loss over the P configurations held as beads,
 *   L = sum_b (E_b - E*_b)^2 + sum_{b,atoms} |dE_b/dr - (dE_b/dr)*|^2,
 * with respect to theta, given the configurations RRt[P][Natoms][3], the
 * weights ew[P] = 2 (E_b - E*_b) and gw[P][Natoms][3] = 2 (dE_b/dr - target):
 *   grad[j] = sum_b ew_b dE_b/dtheta_j + sum gw . d(dE_b/dr)/dtheta_j. */
int pimd_pes_param_grad(pimd_ctx *ctx, const double *RRt, const double *ew, const double *gw,
                        double *grad);

/* ---- TTM model parameters and their derivatives ------------------------
 * No counterpart in the reference, whose TTM2.1-F / TTM3-F parameters are
 * constants (pot_mod.f90:7-26, selected in potential.f90:84-113).
 * theta[PIMD_TTM_NPARAM]: vdwA, vdwB, vdwC, vdwD, vdwE, polarO, polarH,
 * polarM, polfacO, polfacH, polfacM, aDD, aCCaCD, dms_param1, dms_param2,
 * dms_param3.  A context starts with its model's values; later force calls
 * use the values set.  The M-site fraction gammaM is not among them. */
#define PIMD_TTM_NPARAM 16
int pimd_ttm_params_get(pimd_ctx *ctx, double *theta);
int pimd_ttm_params_set(pimd_ctx *ctx, const double *theta);
/* The loss gradient of pimd_pes_param_grad with respect to the theta above,
 * every configuration evaluated as a first MD step (no dipole predictor):
 *   grad[j] = sum_b ew_b dE_b/dtheta_j + sum gw . d(dE_b/dr)/dtheta_j.
 * E[P] and dEdtheta[P][PIMD_TTM_NPARAM] (either may be NULL) receive the
 * energies and their derivatives. */
int pimd_ttm_param_grad(pimd_ctx *ctx, const double *RRt, const double *ew, const double *gw,
                        double *grad, double *E, double *dEdtheta);

/* The 84 coefficients coefD of the dipole moment surface that charges the
 * sites (nasa_mod.f90 values; dms_nasa, nasa.f90:185-242).  No counterpart
 * in the reference, where they are constants.  A context starts with the
 * reference's values; later force calls use the ones set. */
#define PIMD_DMS_NCOEF 84
int pimd_dms_coefs_get(pimd_ctx *ctx, double *coef);
int pimd_dms_coefs_set(pimd_ctx *ctx, const double *coef);
/* pimd_ttm_param_grad with respect to the 84 DMS coefficients: grad[84],
 * and optionally E[P] and dEdc[P][84]. */
int pimd_dms_coef_grad(pimd_ctx *ctx, const double *RRt, const double *ew, const double *gw,
                       double *grad, double *E, double *dEdc);

/* The M-site fraction gammaM (pot_mod.f90:12,24): the M site sits at
 * r_O + gammaM (r_OH1 + r_OH2) / 2 and carries q_O / (1 - gammaM), the H
 * sites gain gammaM / (2 (1 - gammaM)) (q_H1 + q_H2) (pot_ttm.f90:74-79,
 * 134-137).  No counterpart in the reference, where it is a constant.  A
 * context starts with its model's value; later force calls use the one set,
 * which must lie in [0, 1). */
int pimd_gamma_m_get(pimd_ctx *ctx, double *gamma_m);
int pimd_gamma_m_set(pimd_ctx *ctx, double gamma_m);
/* pimd_ttm_param_grad with respect to gammaM, the M sites moving with it:
 * grad[1], and optionally E[P] and dEdg[P]. */
int pimd_gamma_m_grad(pimd_ctx *ctx, const double *RRt, const double *ew, const double *gw,
                      double *grad, double *E, double *dEdg);

/* The adjoint of the same loss with respect to the first nparam of the
 * PIMD_TTM_ADJ_NPARAM parameters: the 16 TTM parameters, the 84 DMS
 * coefficients, then gammaM (the orders of pimd_ttm_params_get and
 * pimd_dms_coefs_get); nparam is 16, 100 or 101.  One sweep per
 * configuration for all of the first 100: the induction energy is stationary
 * in the converged dipoles, so dE/dtheta is the explicit derivative at fixed
 * dipoles and the SOR solve needs no separate adjoint solve; the gradient
 * term gw . d(dE/dr)/dtheta is that derivative's rate of change along the
 * atoms' direction gw, carried through the force path as a dual number with
 * the screening functions' derivatives taken from the chain the reference's
 * gradient uses.  gammaM (nparam 101) takes one tangent sweep more.  Exact up
 * to the SOR convergence (the tangents above differentiate the iterations as
 * executed): use polar_eps of 1e-10 or below for fitting.
 * grad[nparam]; E[P] and dEdtheta[P][nparam] may be NULL. */
#define PIMD_TTM_ADJ_NPARAM (PIMD_TTM_NPARAM + PIMD_DMS_NCOEF + 1)
int pimd_ttm_adjoint(pimd_ctx *ctx, const double *RRt, const double *ew, const double *gw,
                     int nparam, double *grad, double *E, double *dEdtheta);
/* pimd_ttm_adjoint in two halves, for nparam 16 or 100: _start copies the
 * inputs (the arrays are free once it returns) and queues the sweep on the
 * context's stream; _finish waits for it and returns what pimd_ttm_adjoint
 * returns, bit for bit.  Sweeps started on several contexts run side by side
 * (a fit's configuration groups).  No counterpart in the reference, which has
 * no parameter derivatives.  PIMD_EINVAL: _start while a sweep of the context
 * is unfinished, _finish with none started. */
int pimd_ttm_adjoint_start(pimd_ctx *ctx, const double *RRt, const double *ew, const double *gw,
                           int nparam);
int pimd_ttm_adjoint_finish(pimd_ctx *ctx, double *grad, double *E, double *dEdtheta);

/* ---- instrumentation --------------------------------------------------- */

/* Enables per-kernel HIP-event timing on the context stream. */
int pimd_profile_enable(pimd_ctx *ctx, int on);
/* Number of timed kernels, then per kernel: name, launch count, total ms. */
int pimd_profile_count(pimd_ctx *ctx);
int pimd_profile_get(pimd_ctx *ctx, int i, char *name, size_t len, long *launches,
                     double *total_ms);
int pimd_profile_reset(pimd_ctx *ctx);
/* A plain device-to-device copy of `bytes` on the context stream, a kernel
 * of known HBM traffic (bytes read, bytes written) for calibrating
 * counter-based byte counts; ms receives its duration. */
int pimd_profile_copy(pimd_ctx *ctx, size_t bytes, double *ms);
/* Blocks until all work queued on the context stream is done; PIMD_EINVAL
 * if a force call since the last check saw non-finite coordinates. */
int pimd_sync(pimd_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* PIMD_TTM_H */
