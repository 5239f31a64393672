"""Benchmark: PIMD of 512 H2O, P = 32 beads, TTM3-F on one MI355X.

A step is one full PIMD velocity-Verlet step (MD.f90:35-110): NM-space
Langevin thermostat, half kicks, free ring-polymer propagation of all atoms,
centroid wrap and the TTM3-F force evaluation of all 32 beads (monomer PES,
real-space Ewald with Thole damping, induced-dipole SOR solve, reciprocal
space), all device resident.  Inputs are synthetic (seeded water box at
0.997 g/cm^3 with free ring-polymer bead offsets at 300 K).

Both precision forms are timed: FP64 (the reference's arithmetic) and mixed
(FP32 real-space pair terms with FP64 accumulation, FP64 refinement of the
dipoles, FP64 k-space and ring-polymer steps).  Prints one JSON line.
value = molecule-bead force evaluations per second of the mixed form (whole
job); the FP64 form, the ratio and the mixed form's error against FP64 are
reported beside it, with ns_per_day.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pimd-f90_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# 64 independent replicas: waters from {32..512}, beads from {8..64}, T in
# [250, 350] K, models alternating TTM2.1-F / TTM3-F, Rc = L/2, eps 1e-8
REPLICAS64 = "replicas64_mixed_sizes"
# 128 five-water clusters (O-O 2.6-3.2 A, 40 A box) + 128 periodic 32-water
# boxes (Rc 4.9 A), eps 1e-8; targets from the FP64 model at parameters
# perturbed by 2 %; one step = loss and its gradient over the PES and TTM
# parameters
FIT256 = "fit256_clusters_and_boxes"

WORKLOADS = {
    # name: (pot_model, nwaters, nbeads, rc, eps_ewald)
    "h2o512_p32_ttm3f": (3, 512, 32, -1.0, 1e-6),
    "h2o128_p32_ttm21f": (2, 128, 32, -1.0, 1e-6),
    "h2o32_p32_ttm3f_rc49": (3, 32, 32, 4.9, 1e-8),
    # the fixed-charge q-SPC/Fw model of pot_spc on the same system
    "h2o512_p32_qspcfw": (4, 512, 32, -1.0, 1e-6),
}
MODEL_NAMES = {2: "TTM2.1-F", 3: "TTM3-F", 4: "q-SPC/Fw", 5: "SPC/F"}


def byte_model(kernel, c, fp32):
    """Algorithmic HBM bytes per launch of the kernels whose bound is memory:
    the SOR dipole pass with its pair screening factors kept from the first
    iteration (PIMD_DD_CACHE, on by default) reads, per neighbour-list entry
    of every live bead, the partner index (4 bytes) and the (ts1, ts2) of
    each dipole-site pair (2 x 4 or 2 x 8 bytes); per dipole site it stages
    the position and dipole (48 bytes) and writes the field (24 bytes)."""
    if kernel != "pair_dd" or os.environ.get("PIMD_DD_CACHE", "1") == "0":
        return None
    N, P, nnb, nd = c["N"], c["P"], c["nnb"], c["nd"]
    return P * N * nnb * (4 + nd * nd * 2 * (4 if fp32 else 8)) + P * N * nd * (48 + 24)


def flop_model(kernel, c):
    """Algorithmic FP64/FP32 operations per launch: what the algorithm needs,
    not what the kernel issues.  A fused multiply-add counts two, a
    transcendental, square root or division one.  Pair terms are counted once
    per unordered molecule pair within the cutoff (the kernels evaluate each
    pair from both molecules, which is not credited); reciprocal-space terms
    once per (site, k) with one complex product to advance the phase along
    the kz column:
      structure factors  q:  phase 6 + accumulate 4                     = 10
                         d:  phase 6 + k.mu 2 + accumulate 4 + setup 2   = 14
                         v:  phase 6 + 3 components x 4                  = 18
      fields             q:  phase 6 + field 6 + potential 3             = 15
                         d:  phase 6 + field 6                           = 12
                         qd: phase 6 + charge 7 + dipole 7 + column 3    = 23
      pairs  qq: O-O vdW 30 + 9 charge-site pairs x 71 (vector and
                 minimum image 21, damped erfc-screened kernels 40, both
                 sides' fields 10); fixed charges 30 + 9 x 52
             dd: per dipole-site pair: vector 21, smeared tensor 45, both
                 sides' T.mu 18                                          = 84
             final: 5 site pairs (TTM3-F) x 111."""
    N, P, nk, nnb, nd = c["N"], c["P"], c["nk"], c["nnb"], c["nd"]
    pairs = P * N * nnb / 2.0  # unordered molecule pairs within the cutoff
    if c.get("model", 3) >= 4 and kernel == "pair_qq":
        return pairs * (30 + 9 * 52)
    table = {
        "field_d_sor": P * N * nd * nk * 12,
        "sfac_d": P * N * nd * nk * 14,
        "pair_dd": pairs * nd * nd * 84,
        "pair_qq": pairs * (30 + 9 * 71),
        "pair_final": pairs * 5 * 111,
        "field_q": P * N * 3 * nk * 15,
        "field_qd": P * N * 3 * nk * 23,
        "sfac_q": P * N * 3 * nk * 10,
        "sfac_v": P * N * nd * nk * 18,
    }
    v = table.get(kernel)
    return None if v is None else int(v)


def measured_valu(form, key):
    """Issued VALU floating-point operations per launch of a kernel from the
    committed rocprofv3 SQ_INSTS_VALU_* summary (64 lanes per instruction,
    an FMA two), or None."""
    path = os.path.join(ROOT, "profiles", "valu_counters_h2o512_p32_ttm3f.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == form and cells[2] == key:
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def build_system(workload, seed=2024):
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    model, N, P, rc, eps = WORKLOADS[workload]
    RRt, PPt, RRc, L = make_system(N, P, seed=seed)
    # guess_initdip: the reference's default (input_file_reader.f90:83), the
    # 4-point dipole predictor from the previous steps
    p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=rc, eps_ewald=eps,
                  mass_o=15.995, mass_h=1.008, guess_initdip=True).resolved()
    return p, RRt, PPt, RRc


def host_cpu():
    """The host CPU's model name and the logical CPUs this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        ncpu = os.cpu_count()
    return {"model": model, "logical_cpus": ncpu}


def cpu_baseline(p, RRt, RRc, budget_s=15.0):
    """The reference Fortran (oracle/_ref) on a bounded sample of beads of the
    same system, one core; falls back to the C restatement."""
    try:
        import refrun
        kind = "reference" if refrun.available() else None
    except Exception:
        kind = None
    from pimd_amd.params import TTMParams
    one = TTMParams(**{**p.__dict__, "nbeads": 1})
    t0 = time.perf_counter()
    if kind == "reference":
        refrun.forces(one, RRt[:1], RRc)
    else:
        import oracle
        kind = "port"
        sysm = oracle.System(one)
        sysm.forces(RRt[:1], RRc)
    t1 = time.perf_counter()
    per_bead = t1 - t0
    nb = int(max(1, min(p.nbeads, budget_s // max(per_bead, 1e-3))))
    q = TTMParams(**{**p.__dict__, "nbeads": nb})
    t0 = time.perf_counter()
    if kind == "reference":
        refrun.forces(q, RRt[:nb], RRc)
    else:
        oracle.System(q).forces(RRt[:nb], RRc)
    dt = time.perf_counter() - t0
    return {"value": p.nwaters * nb / dt, "unit": "molecule-bead force evals/s", "cores": 1,
            "kind": kind, "host": host_cpu(),
            "sample": f"{'pot_spc' if p.pot_model >= 4 else 'pot_ttm'} on {nb} of the {p.nbeads} beads of the same system "
                      f"({'reference Fortran, amdflang -O2' if kind == 'reference' else 'C restatement, gcc -O2'}), "
                      f"serial, {dt:.1f} s"}


PEAK_FP32_TFLOPS = 157.3
SOR_KERNELS = ("pair_dd", "sfac_d", "field_d_sor")  # launched once per SOR iteration
# FP32 arithmetic in the mixed form: the real-space pair passes and the SOR
# iterations' reciprocal dipole pass
FP32_KERNELS = ("pair_qq", "pair_dd", "pair_final", "sfac_d", "field_d_sor")


def run_form(p, RRt, PPt, RRc, args):
    """Times args.steps PIMD steps of one precision form; returns its figures."""
    from pimd_amd.ttm import PIMDIntegrator, TTMForceField, load_library
    ff = TTMForceField(p)
    info = ff.info()
    if args.copy_mib > 0:
        ff.profile_copy(args.copy_mib << 20)
    integ = PIMDIntegrator(ff, delt_fs=0.25, temp=300.0, bead_thermostat=1,
                           centroid_thermostat=True, tau_centroid=0.1)
    # start from consistent forces (the reference's first step kicks with zero)
    f0 = ff.full_bead_forces(RRt, RRc)
    integ.set_state(RRt, PPt, f0["dRRt"])
    if args.warmup > 0:
        integ.run(args.warmup, t0=1, seed=99)
    L = load_library()
    L.pimd_sync(ff.handle)
    t0 = time.perf_counter()
    rec = integ.run(args.steps, t0=1 + args.warmup, seed=99)
    L.pimd_sync(ff.handle)
    elapsed = time.perf_counter() - t0
    ms = 1000.0 * elapsed / args.steps
    out = {"ms_per_step": ms, "value": p.nwaters * p.nbeads / (ms / 1000.0),
           "ns_per_day": (1000.0 / ms) * 0.25e-6 * 86400.0, "f0": f0, "rec": rec,
           "info": info, "kernels": {}, "roofline": None}
    out["force_only"] = force_only(ff, integ, p, args.steps)
    if args.no_profile:
        return out
    ff.profile(True)
    ff.profile_reset()
    integ.run(args.steps, t0=1 + args.warmup + args.steps, seed=99)
    kernels = ff.profile_report()
    ff.profile(False)
    out["kernels"] = kernels
    nnb = (4.0 / 3.0) * np.pi * p.rc ** 3 / (p.box[0] * p.box[1] * p.box[2]) * p.nwaters
    c = {"N": p.nwaters, "P": p.nbeads, "nk": info["nkvec"], "nnb": min(nnb, p.nwaters - 1),
         "nd": {2: 3, 3: 1}.get(p.pot_model, 0), "model": p.pot_model}
    # beads a SOR launch works on, on average (converged beads leave the
    # remaining iterations' launches early)
    bead_iters = kernels.pop("sor_bead_iterations", (0, 0.0))[0]
    n_sor = kernels.get("field_d_sor", (0, 0.0))[0]
    active = bead_iters / n_sor if n_sor else float(p.nbeads)
    out["sor_active_beads_per_launch"] = round(active, 2)
    tot = {k: v[1] for k, v in kernels.items()}
    dom = max(tot, key=tot.get)
    cnt, tms = kernels[dom]
    fl = flop_model(dom, {**c, "P": active} if dom in SOR_KERNELS else c)
    if fl is not None and cnt > 0:
        fp32 = p.precision == 1 and dom in FP32_KERNELS
        peak = PEAK_FP32_TFLOPS if fp32 else PEAK_FP64_TFLOPS
        avg_s = tms / cnt / 1000.0
        ach = fl / avg_s / 1e12
        form = "mixed" if p.precision == 1 else "fp64"
        traffic = measured_traffic(args.workload, form, dom)
        issued = measured_valu(form, dom) if args.workload == "h2o512_p32_ttm3f" else None
        out["roofline"] = {"bound": "valu_fp32" if fp32 else "valu_fp64",
                           "achieved": round(ach, 3), "peak": peak,
                           "unit": "TFLOP/s", "frac": round(ach / peak, 4), "traffic": traffic,
                           "kernel": dom, "avg_launch_ms": round(tms / cnt, 4),
                           "flops_per_launch": fl,
                           "issued_flops_per_launch": issued,
                           "issued_frac": (round(issued / avg_s / 1e12 / peak, 4)
                                           if issued else None),
                           "traffic_source": "profiles/hbm_traffic_h2o512_p32_ttm3f.md",
                           "issued_source": "profiles/valu_counters_h2o512_p32_ttm3f.md"}
        nb = byte_model(dom, {**c, "P": active} if dom in SOR_KERNELS else c, fp32)
        if nb is not None:
            # a memory-bound kernel: its roofline is HBM bandwidth, the flop
            # figures above stay as a second view
            gbs = nb / avg_s / 1e9
            r = out["roofline"]
            r["valu"] = {k: r.pop(k) for k in ("bound", "achieved", "peak", "unit", "frac")}
            out["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                               "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                               "traffic": traffic, "bytes_per_launch": int(nb), **r}
    ff.close()
    return out


def run_batched(p, nrep, args, seed=2024):
    """nrep replicas of the workload (seeds seed, seed + 1, ...) stepped side
    by side as one batch (pimd_batch_md_run: each kernel one launch over all
    the replicas, one captured graph per batch step); whole-job
    molecule-bead force evaluations per second."""
    from pimd_amd.systems import make_system
    from pimd_amd.ttm import PIMDIntegrator, ReplicaBatch, TTMForceField
    reps = []
    for r in range(nrep):
        RRt, PPt, RRc, L = make_system(p.nwaters, p.nbeads, seed=seed + r)
        ff = TTMForceField(p)
        integ = PIMDIntegrator(ff, delt_fs=0.25, temp=300.0, bead_thermostat=1,
                               centroid_thermostat=True, tau_centroid=0.1)
        integ.set_state(RRt, PPt, ff.full_bead_forces(RRt, RRc)["dRRt"])
        reps.append((ff, integ))
    batch = ReplicaBatch([i for _, i in reps])
    seeds = [99 + r for r in range(nrep)]
    batch.run(args.warmup, seeds, t0=1)
    t0 = time.perf_counter()
    batch.run(args.steps, seeds, t0=1 + args.warmup)
    s = (time.perf_counter() - t0) / args.steps
    batch.close()
    for ff, _ in reps:
        ff.close()
    return {"replicas": nrep, "ms_per_step": round(1000 * s, 4),
            "value": round(nrep * p.nwaters * p.nbeads / s, 1)}


def time_md(p, args, delt_fs=0.25, **md):
    """Whole PIMD steps of a variant of the workload (another model, or
    contracted MD with md['intra_timesteps']), from the same synthetic system:
    ms per step, molecule-bead force evaluations (or, for contracted MD,
    molecule-bead steps) per second, ns/day at delt_fs."""
    from pimd_amd.systems import make_system
    from pimd_amd.ttm import PIMDIntegrator, TTMForceField, load_library
    RRt, PPt, RRc, L = make_system(p.nwaters, p.nbeads, seed=2024)
    ff = TTMForceField(p)
    integ = PIMDIntegrator(ff, delt_fs=delt_fs, temp=300.0, bead_thermostat=1,
                           centroid_thermostat=True, tau_centroid=0.1, **md)
    integ.set_state(RRt, PPt, ff.full_bead_forces(RRt, RRc)["dRRt"])
    integ.run(args.warmup, t0=1, seed=99)
    lib = load_library()
    lib.pimd_sync(ff.handle)
    t0 = time.perf_counter()
    integ.run(args.steps, t0=1 + args.warmup, seed=99)
    lib.pimd_sync(ff.handle)
    ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    ff.close()
    return {"ms_per_step": round(ms, 4), "value": round(p.nwaters * p.nbeads / (ms / 1000.0), 1),
            "ns_per_day": round((1000.0 / ms) * delt_fs * 1e-6 * 86400.0, 3), "dt_fs": delt_fs}


def force_only(ff, integ, p, steps):
    """The all-beads force call alone (full_bead_forces equivalent) on the
    device-resident coordinates of the last MD step, dipole predictor off
    (every call starts the SOR solve from alpha E_q)."""
    from pimd_amd.ttm import load_library
    L = load_library()
    integ.state_forces(1)
    L.pimd_sync(ff.handle)
    t0 = time.perf_counter()
    for _ in range(steps):
        integ.state_forces(1)
    L.pimd_sync(ff.handle)
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    return {"ms_per_call": round(ms, 4), "value": round(p.nwaters * p.nbeads / (ms / 1000.0), 1),
            "unit": "molecule-bead force evals/s", "predictor": "off"}


def cpu_parallel_baseline(p, RRt, RRc, threads=16, budget_s=10.0):
    """The C restatement over beads in parallel threads (the reference's
    MPI-over-beads decomposition on the host cores), on as many beads of the
    same system as fit the budget."""
    import concurrent.futures
    import oracle
    from pimd_amd.params import TTMParams
    threads = max(1, min(threads, os.cpu_count() or 1, p.nbeads))
    one = TTMParams(**{**p.__dict__, "nbeads": 1})
    systems = [oracle.System(one) for _ in range(threads)]
    t0 = time.perf_counter()
    systems[0].forces(RRt[:1], RRc)
    per_bead = time.perf_counter() - t0
    nb = int(min(p.nbeads, max(threads, threads * budget_s // max(per_bead, 1e-3))))

    def work(i):
        sysm = systems[i % threads]
        for b in range(i, nb, threads):
            sysm.forces(RRt[b:b + 1], RRc)
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(threads)))
    dt = time.perf_counter() - t0
    return {"value": p.nwaters * nb / dt, "unit": "molecule-bead force evals/s", "cores": threads,
            "kind": "port", "host": host_cpu(),
            "sample": f"C restatement (gcc -O2) on {nb} of the {p.nbeads} beads, one bead per "
                      f"thread at a time, {threads} threads, {dt:.1f} s"}


def measured_traffic(workload, form, key):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_h2o512_p32_ttm3f.md")
    if workload != "h2o512_p32_ttm3f" or not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] == form and cells[2] == key:
            try:
                return float(cells[5])
            except ValueError:
                return None
    return None


def spring_energy(RRt, masses, P, temp):
    """Ring-polymer spring energy sum_b sum_i m_i omega_n^2 |r_b - r_(b+1)|^2 / 2 in
    kcal/mol, omega_n = kB T P / hbar (the frequency the normal-mode propagator
    uses, NormalModes.f90:125-153)."""
    from pimd_amd.systems import HBAR, KB_AMUA2PS2PERK, MASSCON
    wn = KB_AMUA2PS2PERK * temp * P / HBAR
    d = RRt - np.roll(RRt, -1, axis=0)
    return float(0.5 * wn * wn * np.sum(masses[None, :, None] * d * d) / MASSCON)


def accuracy_report(p, RRt, PPt, RRc, nsteps=2000, every=100):
    """The mixed form against FP64 over a trajectory: (1) an FP64 Langevin
    trajectory of nsteps steps whose configurations are replayed every `every`
    steps through both forms (identical coordinates, predictor off), and (2)
    the drift of the ring-polymer Hamiltonian in NVE runs of both forms from
    the same state."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import MASSCON, atom_masses
    from pimd_amd.ttm import PIMDIntegrator, TTMForceField
    q64 = TTMParams(**{**p.__dict__, "precision": 0})
    qmx = TTMParams(**{**p.__dict__, "precision": 1})
    ff = TTMForceField(q64)
    ev64, evmx = TTMForceField(q64), TTMForceField(qmx)
    integ = PIMDIntegrator(ff, delt_fs=0.25, temp=300.0, bead_thermostat=1,
                           centroid_thermostat=True, tau_centroid=0.1)
    f0 = ff.full_bead_forces(RRt, RRc)
    integ.set_state(RRt, PPt, f0["dRRt"])
    frms, fmax, erel = [], [], []
    t = 1
    for _ in range(nsteps // every):
        integ.run(every, t0=t, seed=7)
        t += every
        st = integ.get_state()
        a = ev64.full_bead_forces(st["RRt"], st["RRc"])
        m = evmx.full_bead_forces(st["RRt"], st["RRc"])
        d = m["dRRt"] - a["dRRt"]
        frms.append(float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(a["dRRt"] ** 2))))
        fmax.append(float(np.max(np.abs(d)) / np.max(np.abs(a["dRRt"]))))
        # per molecule: a bead's total energy (monomer strain minus cohesion) can pass near zero
        erel.append(float(np.max(np.abs(m["Upott"] - a["Upott"])) / p.nwaters))
    masses = atom_masses(p.nwaters, p.mass_o, p.mass_h)
    drift = {}
    for name, q in (("fp64", q64), ("mixed", qmx)):
        f = TTMForceField(q)
        it = PIMDIntegrator(f, delt_fs=0.25, temp=300.0, bead_thermostat=0)
        g = f.full_bead_forces(RRt, RRc)
        it.set_state(RRt, PPt, g["dRRt"])
        H, times = [], []
        chunk = 50
        for c in range(nsteps // chunk):
            rec = it.run(chunk, t0=1 + c * chunk, seed=0)
            st = it.get_state()
            H.append(rec["uk"][-1] / MASSCON + spring_energy(st["RRt"], masses, p.nbeads, 300.0)
                     + rec["Upot"][-1])
            times.append((c + 1) * chunk * 0.25e-3)
        H = np.array(H)
        slope = float(np.polyfit(np.array(times), H, 1)[0])
        drift[name] = {"kcal_mol_per_ps": round(slope, 5),
                       "per_molecule_bead_kcal_mol_per_ps": round(slope / (p.nwaters * p.nbeads), 9),
                       "H_first": round(float(H[0]), 4), "H_last": round(float(H[-1]), 4),
                       "H_rms_fluct": round(float(np.std(H)), 5)}
    return {"replayed_configurations": len(frms), "trajectory_steps": nsteps,
            "force_rms_rel_max": max(frms), "force_max_rel_max": max(fmax),
            "energy_abs_max_kcal_mol_per_molecule": max(erel),
            "nve_drift": drift,
            "note": "FP64 Langevin trajectory replayed through both forms on identical "
                    "coordinates; NVE runs from the same state, H = kinetic + springs + "
                    "sum of bead potentials"}


def _kern(res, steps):
    return {k: round(v[1] / max(steps, 1), 4) for k, v in res["kernels"].items()}


def replica_specs(n=64, seed=3000):
    rng = np.random.default_rng(seed)
    specs = []
    for r in range(n):
        specs.append({"model": 2 if r % 2 == 0 else 3,
                      "nwaters": int(rng.choice([32, 64, 128, 256, 512])),
                      "nbeads": int(rng.choice([8, 16, 32, 64])),
                      "temp": float(rng.uniform(250.0, 350.0)), "seed": seed + r})
    return specs


def replica_params(sp, precision=0):
    """Model parameters of one replica of replica_specs()."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    L = make_system(sp["nwaters"], 1, seed=sp["seed"], spaced=True)[3]
    return TTMParams(pot_model=sp["model"], nwaters=sp["nwaters"], nbeads=sp["nbeads"],
                     box=(L, L, L), rc=-1.0, eps_ewald=1e-8, mass_o=15.995, mass_h=1.008,
                     guess_initdip=True, precision=precision)


def make_replica(sp, precision=0):
    """One replica of replica_specs() loaded with its initial state and
    gradient: (integrator, force field, RRt, PPt, dRRt)."""
    from pimd_amd.systems import make_system
    from pimd_amd.ttm import PIMDIntegrator, TTMForceField
    RRt, PPt, RRc, _ = make_system(sp["nwaters"], sp["nbeads"], temp=sp["temp"], seed=sp["seed"],
                                   spaced=True)
    ff = TTMForceField(replica_params(sp, precision))
    integ = PIMDIntegrator(ff, delt_fs=0.25, temp=sp["temp"], bead_thermostat=1,
                           centroid_thermostat=True, tau_centroid=0.1)
    f0 = ff.full_bead_forces(RRt, RRc)
    integ.set_state(RRt, PPt, f0["dRRt"])
    return integ, ff, RRt, PPt, f0["dRRt"]


def run_replicas(args):
    """Aggregate throughput of 64 uneven replicas: run one after another with
    the default launch shapes, then side by side as one batch, untuned and
    autotuned.  Every run starts each replica from the same state."""
    from pimd_amd.ttm import ReplicaBatch, load_library
    specs = replica_specs()
    reps = [make_replica(sp, 1 if args.precision == "mixed" else 0) for sp in specs]
    work = sum(sp["nwaters"] * sp["nbeads"] for sp in specs)
    seeds = [sp["seed"] for sp in specs]
    L = load_library()

    def reset():
        for integ, ff, RRt, PPt, g in reps:
            integ.set_state(RRt, PPt, g)

    def sequential():
        reset()
        for (integ, *_), sd in zip(reps, seeds):
            integ.run(args.warmup, t0=1, seed=sd)
        t0 = time.perf_counter()
        for (integ, *_), sd in zip(reps, seeds):
            integ.run(args.steps, t0=1 + args.warmup, seed=sd)
        return (time.perf_counter() - t0) / args.steps

    def batched(batch):
        reset()
        batch.run(args.warmup, seeds, t0=1)
        for integ, ff, *_ in reps:
            L.pimd_sync(ff.handle)
        t0 = time.perf_counter()
        batch.run(args.steps, seeds, t0=1 + args.warmup)
        return (time.perf_counter() - t0) / args.steps

    s_seq = sequential()
    batch = ReplicaBatch([r[0] for r in reps])
    s_bat = batched(batch)
    reset()
    tune_ms = batch.autotune()
    shapes = batch.shapes()
    s_tuned = batched(batch)
    return {
        "metric": "PIMD ns/day + molecule-bead force evals/s, 512 H2O P=32 TTM3-F, mixed vs FP64",
        "value": round(work / s_tuned, 1), "unit": "molecule-bead force evals/s (sum over replicas)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000 * s_tuned, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32+fp64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": REPLICAS64, "replicas": len(specs),
                   "molecule_beads_per_step": work, "eps_ewald": 1e-8, "rc": "L/2",
                   "precision": args.precision if args.precision != "both" else "fp64"},
        "sequential": {"value": round(work / s_seq, 1), "ms_per_step": round(1000 * s_seq, 3)},
        "batched_untuned": {"value": round(work / s_bat, 1), "ms_per_step": round(1000 * s_bat, 3)},
        "batched_autotuned": {"value": round(work / s_tuned, 1),
                              "ms_per_step": round(1000 * s_tuned, 3)},
        "batch_over_sequential": round(s_seq / s_tuned, 3),
        "autotuned_shapes": {"fields": ["qq_group", "dd_group", "final_group", "field_cg",
                                        "sfac_sites"],
                             "per_replica": shapes[:8], "force_ms": [round(x, 4) for x in tune_ms[:8]]},
        "replicas": specs[:8],
    }


def run_fit(args):
    """Throughput of the fitting gradient: per step the forward energies and
    gradients of 256 configurations, the PES adjoint and the TTM tangents."""
    from pimd_amd.fit import FitGroup, loss_and_grad
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import distorted_box, water_cluster
    from pimd_amd.ttm import TTMForceField
    import oracle
    ncl = nbx = 128
    cl = np.array([water_cluster(5, seed=6000 + i) for i in range(ncl)])
    bx = []
    for i in range(nbx):
        r, L = distorted_box(32, seed=6500 + i)
        bx.append(r)
    bx = np.array(bx)
    common = dict(pot_model=3, eps_ewald=1e-8, polar_eps=1e-12, polar_maxiter=100)
    # contexts hold up to 64 configurations (beads) each.  The box groups'
    # adjoint sweeps (16 ms each) dominate; their contexts are made first so
    # that their streams land on different hardware queues (the runtime
    # spreads streams over its 4 queues in creation order; the last two
    # made share one) and their sweeps overlap in loss_and_grad
    pcl = TTMParams(nwaters=5, nbeads=64, box=(40.0, 40.0, 40.0), **common)
    pbx = TTMParams(nwaters=32, nbeads=64, box=(L, L, L), rc=4.9, **common)
    sets = [(pbx, bx[:64]), (pbx, bx[64:]), (pcl, cl[:64]), (pcl, cl[64:])]
    ffs = [TTMForceField(q) for q, _ in sets]
    theta0 = np.concatenate([ffs[0].pes_params(), ffs[0].ttm_params()])
    rng = np.random.default_rng(6000)
    theta_t = theta0 * (1.0 + 0.02 * rng.standard_normal(theta0.size))
    groups = []
    for ff, R in zip(ffs, [R for _, R in sets]):
        ff.set_pes_params(theta_t[:250])
        ff.set_ttm_params(theta_t[250:])
        r = ff.full_bead_forces(R, R[0])
        groups.append(FitGroup(ff, R, r["Upott"], r["dRRt"]))
    for _ in range(args.warmup):
        loss_and_grad(groups, theta0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        L0, g = loss_and_grad(groups, theta0)
    s = (time.perf_counter() - t0) / max(args.steps, 1)
    # the same step with the TTM part from its 16 tangent sweeps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, gt = loss_and_grad(groups, theta0, method="tangent")
    s_tan = (time.perf_counter() - t0) / max(args.steps, 1)
    # the parts of one step
    parts = {}
    for name, fn in (("forward", lambda grp: grp.ff.full_bead_forces(grp.RRt, grp.RRc)),
                     ("pes_adjoint", lambda grp: grp.ff.pes_param_grad(
                         grp.RRt, np.ones(grp.RRt.shape[0]), np.ones_like(grp.RRt))),
                     ("ttm_tangent", lambda grp: grp.ff.ttm_param_grad(
                         grp.RRt, np.ones(grp.RRt.shape[0]), np.ones_like(grp.RRt))),
                     # not part of the step: the 84 DMS coefficients' tangents
                     ("dms_tangent_84", lambda grp: grp.ff.dms_coef_grad(
                         grp.RRt, np.ones(grp.RRt.shape[0]), np.ones_like(grp.RRt))),
                     # one adjoint sweep over the 16 TTM parameters, 84 DMS
                     # coefficients and gammaM
                     ("ttm_adjoint_101", lambda grp: grp.ff.ttm_adjoint(
                         grp.RRt, np.ones(grp.RRt.shape[0]), np.ones_like(grp.RRt)))):
        t1 = time.perf_counter()
        for grp in groups:
            fn(grp)
        parts[name] = round(1000 * (time.perf_counter() - t1), 3)
    # CPU: one forward pass of the oracle (the C restatement) over a bounded
    # sample; a finite-difference gradient needs 2 x 266 of them
    cpu = None
    if not args.no_cpu:
        t1 = time.perf_counter()
        nsamp = 0
        for grp, n in ((groups[0], 16), (groups[2], 4)):
            q = TTMParams(**{**grp.ff.params.__dict__, "nbeads": n})
            o = oracle.System(q)
            o.forces(grp.RRt[:n], grp.RRt[0], t=1)
            nsamp += n
        tc = time.perf_counter() - t1
        cpu = {"value": round(nsamp / tc, 2), "unit": "configuration forward evaluations/s",
               "cores": 1, "kind": "port",
               "sample": "oracle forward pass of 16 clusters + 4 boxes; a central-difference "
                         "gradient over the 266 parameters costs 532 such passes per "
                         "configuration"}
    return {
        "metric": "PIMD ns/day + molecule-bead force evals/s, 512 H2O P=32 TTM3-F, mixed vs FP64",
        "value": round((ncl + nbx) / s, 1),
        "unit": "configurations/s (energy+force loss and its gradient over 266 parameters)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000 * s, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": FIT256, "clusters": ncl, "boxes": nbx, "parameters": int(theta0.size),
                   "pes_parameters": 250, "ttm_parameters": 16, "eps_ewald": 1e-8,
                   "polar_eps": 1e-12},
        "gradient": "adjoint (pimd_ttm_adjoint) for the TTM part, PES adjoint",
        "tangent_step": {"ms_per_step": round(1000 * s_tan, 3),
                         "grad_rel_diff": float(np.max(np.abs(gt - g)) / np.max(np.abs(gt)))},
        "loss": L0, "grad_norm": float(np.linalg.norm(g)),
        "parts_ms": parts,
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="h2o512_p32_ttm3f",
                    choices=sorted(WORKLOADS) + [REPLICAS64, FIT256])
    ap.add_argument("--precision", default="both", choices=("fp64", "mixed", "both"),
                    help="form(s) to time; with both, value is the mixed form")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--batch", type=int, default=3,
                    help="replicas of the workload in the batched measurement (1 skips it)")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=2000,
                    help="length of the mixed-vs-FP64 trajectory report (0 skips it)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB (k_copy_probe, a kernel "
                         "of known HBM traffic, to scale rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the force path "
              "shards by beads or replicas, which this build does not launch.", file=sys.stderr)
        sys.exit(2)
    if args.workload == REPLICAS64:
        print(json.dumps(run_replicas(args)), flush=True)
        return
    if args.workload == FIT256:
        print(json.dumps(run_fit(args)), flush=True)
        return

    from pimd_amd.params import TTMParams
    p, RRt, PPt, RRc = build_system(args.workload)
    forms = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res = {}
    for f in forms:
        q = TTMParams(**{**p.__dict__, "precision": 1 if f == "mixed" else 0})
        res[f] = run_form(q, RRt, PPt, RRc, args)
    main_form = forms[-1]
    r = res[main_form]

    cpu = None if args.no_cpu else cpu_baseline(p, RRt, RRc)
    cpu_par = None if args.no_cpu else cpu_parallel_baseline(p, RRt, RRc)
    out = {
        "metric": "PIMD ns/day + molecule-bead force evals/s, 512 H2O P=32 TTM3-F, mixed vs FP64",
        "value": round(r["value"], 1),
        "unit": "molecule-bead force evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4), "ns_per_day": round(r["ns_per_day"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64" if main_form == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "nwaters": p.nwaters, "nbeads": p.nbeads,
                   "model": MODEL_NAMES[p.pot_model], "precision": main_form,
                   "rc": round(p.rc, 4), "eps_ewald": p.eps_ewald, "nkvec": r["info"]["nkvec"],
                   "dt_fs": 0.25, "thermostat": "NM Langevin (PILE), 300 K",
                   "polar_maxiter": p.polar_maxiter, "polar_eps": p.polar_eps,
                   "guess_initdip": p.guess_initdip},
        "roofline": r["roofline"],
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_par,
        "force_only": r["force_only"],
        "kernels_ms_per_step": _kern(r, args.steps),
        "sor_active_beads_per_launch": r.get("sor_active_beads_per_launch"),
        "sor_iterations_per_step": (round(r["kernels"]["field_d_sor"][0] / max(args.steps, 1), 2)
                                    if "field_d_sor" in r["kernels"] else None),
    }
    if len(forms) == 2 and args.accuracy_steps > 0:
        out["mixed_accuracy"] = accuracy_report(p, RRt, PPt, RRc, nsteps=args.accuracy_steps)
    if len(forms) == 2:
        a, m = res["fp64"], res["mixed"]
        df = m["f0"]["dRRt"] - a["f0"]["dRRt"]
        out["fp64"] = {"value": round(a["value"], 1), "ms_per_step": round(a["ms_per_step"], 4),
                       "ns_per_day": round(a["ns_per_day"], 3), "roofline": a["roofline"],
                       "force_only": a["force_only"],
                       "kernels_ms_per_step": _kern(a, args.steps)}
        out["mixed_over_fp64"] = round(m["value"] / a["value"], 3)
        if args.batch > 1:
            # several replicas of the same workload side by side: the batched
            # forms against the plain (one system) FP64 form
            bat = {}
            for f in forms:
                q = TTMParams(**{**p.__dict__, "precision": 1 if f == "mixed" else 0})
                bat[f] = run_batched(q, args.batch, args)
            bat["mixed_over_plain_fp64"] = round(bat["mixed"]["value"] / a["value"], 3)
            bat["fp64_over_plain_fp64"] = round(bat["fp64"]["value"] / a["value"], 3)
            out["batched"] = bat
        # the other built rows on the same system: contracted MD (the
        # reference's default intra_timesteps 10, inner step 0.25 fs) and the
        # fixed-charge q-SPC/Fw model
        extra = {"contracted_md": {"intra_timesteps": 10, "outer_dt_fs": 2.5},
                 "qspcfw": {"model": "q-SPC/Fw"}}
        for f in forms:
            q = TTMParams(**{**p.__dict__, "precision": 1 if f == "mixed" else 0})
            extra["contracted_md"][f] = time_md(q, args, delt_fs=2.5, intra_timesteps=10)
            qs = TTMParams(**{**q.__dict__, "pot_model": 4, "guess_initdip": False})
            extra["qspcfw"][f] = time_md(qs, args)
        extra["contracted_md"]["note"] = ("value = molecule-bead steps/s (one centroid TTM call and "
                                          "10 monomer passes per step)")
        out["other_rows"] = extra
        out["mixed_error_vs_fp64"] = {
            "force_rms_rel": float(np.sqrt(np.mean(df ** 2)) / np.sqrt(np.mean(a["f0"]["dRRt"] ** 2))),
            "force_max_rel": float(np.max(np.abs(df)) / np.max(np.abs(a["f0"]["dRRt"]))),
            "energy_rel_max": float(np.max(np.abs(m["f0"]["Upott"] - a["f0"]["Upott"])
                                           / np.abs(a["f0"]["Upott"]))),
            "trajectory_upot_rel_max": float(np.max(np.abs(m["rec"]["Upot"] - a["rec"]["Upot"])
                                                    / np.abs(a["rec"]["Upot"]))),
            "note": "first evaluation on the start configuration, and the summed potential "
                    "over the timed steps of both forms run from the same state and noise"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
