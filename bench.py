"""Benchmark: PIMD of 512 H2O, P = 32 beads, TTM3-F on one MI355X.

A step is one full PIMD velocity-Verlet step (MD.f90:35-110): NM-space
Langevin thermostat, half kicks, free ring-polymer propagation of all atoms,
centroid wrap and the TTM3-F force evaluation of all 32 beads (monomer PES,
real-space Ewald with Thole damping, induced-dipole SOR solve, reciprocal
space), all device resident.  Inputs are synthetic (seeded water box at
0.997 g/cm^3 with free ring-polymer bead offsets at 300 K).

Prints one JSON line.  value = molecule-bead force evaluations per second
(whole job); ns_per_day is reported beside it.
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

WORKLOADS = {
    # name: (pot_model, nwaters, nbeads, rc, eps_ewald)
    "h2o512_p32_ttm3f": (3, 512, 32, -1.0, 1e-6),
    "h2o128_p32_ttm21f": (2, 128, 32, -1.0, 1e-6),
    "h2o32_p32_ttm3f_rc49": (3, 32, 32, 4.9, 1e-8),
}


def flop_model(kernel, c):
    """Algorithmic FP64 operations per launch (transcendentals, sqrt and
    divisions count one each), from the operation counts of each kernel's
    inner loop; see the kernel-table of the design notes."""
    N, P, nk, nnb, nd = c["N"], c["P"], c["nk"], c["nnb"], c["nd"]
    pairs = P * N * nnb  # ordered molecule pairs (each side evaluated once)
    if kernel == "field_d_sor":   # per (dipole site, k): phase step 6 + field 8
        return P * N * nd * nk * 14
    if kernel == "sfac_d":        # per (dipole site, k): phase 12 + k.mu 5 + acc 4
        return P * N * nd * nk * 21
    if kernel == "pair_dd":       # per ordered pair and dipole-site pair: vectors 21, smear2 ~45, T.mu 18
        return pairs * nd * nd * 84
    if kernel == "pair_qq":       # vdW ~30 + per site pair: vectors 21, smear01 ~40 (+Thole), field 10
        return pairs * (30 + 9 * 71)
    if kernel == "pair_final":    # 5 site pairs (TTM3-F): smear2/3 ~50, tensors ~40, vectors 21
        return pairs * 5 * 111
    if kernel == "field_q":
        return P * N * 3 * nk * 18
    if kernel == "field_qd":
        return P * N * 3 * nk * 24
    if kernel == "sfac_q":
        return P * N * 3 * nk * 16
    if kernel == "sfac_v":
        return P * N * nd * nk * 24
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
            "kind": kind,
            "sample": f"pot_ttm on {nb} of the {p.nbeads} beads of the same system "
                      f"({'reference Fortran, amdflang -O2' if kind == 'reference' else 'C restatement, gcc -O2'}), "
                      f"serial, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="h2o512_p32_ttm3f", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the force path "
              "shards by beads or replicas, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    from pimd_amd.ttm import PIMDIntegrator, TTMForceField
    p, RRt, PPt, RRc = build_system(args.workload)
    ff = TTMForceField(p)
    info = ff.info()
    integ = PIMDIntegrator(ff, delt_fs=0.25, temp=300.0, bead_thermostat=1,
                           centroid_thermostat=True, tau_centroid=0.1)
    integ.set_state(RRt, PPt, None)
    # start from consistent forces (the reference's first step kicks with zero)
    f0 = ff.full_bead_forces(RRt, RRc)
    integ.set_state(RRt, PPt, f0["dRRt"])
    if args.warmup > 0:
        integ.run(args.warmup, t0=1, seed=99)
    from pimd_amd.ttm import load_library
    L = load_library()
    L.pimd_sync(ff.handle)
    t0 = time.perf_counter()
    rec = integ.run(args.steps, t0=1 + args.warmup, seed=99)
    L.pimd_sync(ff.handle)
    elapsed = time.perf_counter() - t0
    ms = 1000.0 * elapsed / args.steps
    evals = p.nwaters * p.nbeads / (ms / 1000.0)
    ns_day = (1000.0 / ms) * 0.25e-6 * 86400.0

    roof = None
    kernels = {}
    if not args.no_profile:
        ff.profile(True)
        ff.profile_reset()
        integ.run(args.steps, t0=1 + args.warmup + args.steps, seed=99)
        kernels = ff.profile_report()
        ff.profile(False)
        nnb = (4.0 / 3.0) * np.pi * p.rc ** 3 / (p.box[0] * p.box[1] * p.box[2]) * p.nwaters
        c = {"N": p.nwaters, "P": p.nbeads, "nk": info["nkvec"], "nnb": min(nnb, p.nwaters - 1),
             "nd": 1 if p.pot_model == 3 else 3}
        tot = {k: v[1] for k, v in kernels.items()}
        dom = max(tot, key=tot.get)
        cnt, tms = kernels[dom]
        fl = flop_model(dom, c)
        if fl is not None and cnt > 0:
            avg_s = tms / cnt / 1000.0
            ach = fl / avg_s / 1e12
            roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(ach / PEAK_FP64_TFLOPS, 4), "traffic": None,
                    "kernel": dom, "avg_launch_ms": round(tms / cnt, 4),
                    "flops_per_launch": fl}

    cpu = None if args.no_cpu else cpu_baseline(p, RRt, RRc)
    out = {
        "metric": "PIMD ns/day + molecule-bead force evals/s, 512 H2O P=32 TTM3-F, mixed vs FP64",
        "value": round(evals, 1),
        "unit": "molecule-bead force evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "ns_per_day": round(ns_day, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "nwaters": p.nwaters, "nbeads": p.nbeads,
                   "model": "TTM3-F" if p.pot_model == 3 else "TTM2.1-F", "precision": "fp64",
                   "rc": round(p.rc, 4), "eps_ewald": p.eps_ewald, "nkvec": info["nkvec"],
                   "dt_fs": 0.25, "thermostat": "NM Langevin (PILE), 300 K",
                   "polar_maxiter": p.polar_maxiter, "polar_eps": p.polar_eps,
                   "guess_initdip": p.guess_initdip},
        "roofline": roof,
        "cpu_baseline": cpu,
        "energy_per_step_last": float(rec["Upot"][-1]) if args.steps else None,
        "kernels_ms_per_step": {k: round(v[1] / max(args.steps, 1), 4) for k, v in kernels.items()},
        "sor_iterations_per_step": (round(kernels["field_d_sor"][0] / max(args.steps, 1), 2)
                                    if "field_d_sor" in kernels else None),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
