"""Generates the golden vectors in tests/golden/*.npz from the reference.

Each fixture holds seeded synthetic inputs and the outputs of the reference's
own Fortran routines, built from /root/reference by oracle/Makefile into
oracle/_ref/ref_driver (see oracle/ref_driver.f90).  Run from the repo root:

    make -C oracle ref && python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "pimd-f90_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import refrun  # noqa: E402
from pimd_amd.params import TTMParams  # noqa: E402
from pimd_amd.systems import make_system, KB_AMUA2PS2PERK, HBAR  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def params_dict(p):
    return {f"p_{k}": np.asarray(v) for k, v in p.__dict__.items()}


def save(name, **arrs):
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrs)
    print("wrote", name)


# Force cases: (name, model, N, P, lattice, box, rc, eps, seed, extra)
FORCE_CASES = [
    # 32 H2O, 0.997 g/cm^3, Rc 4.9, eps 1e-8 (the CPU-runnable benchmark system)
    ("forces_ttm3f_n32_p4_rc49", 3, 32, 4, None, None, 4.9, 1e-8, 11, {}),
    ("forces_ttm21f_n32_p3_rc49", 2, 32, 3, None, None, 4.9, 1e-8, 12, {}),
    # default cut-off L/2 and eps 1e-6
    ("forces_ttm3f_n32_p2_rchalf", 3, 32, 2, None, None, -1.0, 1e-6, 13, {}),
    ("forces_ttm21f_n32_p2_rchalf", 2, 32, 2, None, None, -1.0, 1e-6, 14, {}),
    # 64 H2O, single bead, larger box
    ("forces_ttm3f_n64_p1_rchalf", 3, 64, 1, None, None, -1.0, 1e-6, 15, {}),
    # non-cubic box (lattice 4x2x2 in a 12x8x8 box): per-axis k ranges and images
    ("forces_ttm3f_n16_p2_ortho", 3, 16, 2, (4, 2, 2), (12.0, 8.0, 8.0), 3.9, 1e-7, 16, {}),
    # SOR cap reached: tight tolerance with 4 iterations
    ("forces_ttm3f_n32_p2_maxiter4", 3, 32, 2, None, None, 4.9, 1e-8, 17,
     {"polar_maxiter": 4, "polar_eps": 1e-12}),
    ("forces_ttm21f_n27_p2_eps1e5", 2, 27, 2, (3, 3, 3), None, -1.0, 1e-5, 18,
     {"polar_maxiter": 40, "polar_eps": 1e-9}),
    # the fixed-charge models of pot_spc: q-SPC/Fw (4) and SPC/F (5)
    ("forces_qspcfw_n32_p2_rc49", 4, 32, 2, None, None, 4.9, 1e-8, 19, {}),
    ("forces_spcf_n32_p2_rchalf", 5, 32, 2, None, None, -1.0, 1e-6, 20, {}),
    ("forces_qspcfw_n16_p2_ortho", 4, 16, 2, (4, 2, 2), (12.0, 8.0, 8.0), 3.9, 1e-7, 22, {}),
    ("forces_spcf_n64_p1_rchalf", 5, 64, 1, None, None, -1.0, 1e-6, 23, {}),
]


def force_fixtures(only=None):
    for name, model, N, P, lat, box, rc, eps, seed, extra in FORCE_CASES:
        if only and name not in only:
            continue
        if box is None:
            RRt, PPt, RRc, L = make_system(N, P, seed=seed, lattice=lat)
            bx = (L, L, L)
        else:
            RRt, PPt, RRc, _ = make_system(N, P, seed=seed, lattice=lat, box=box[0])
            # rescale y, z of a cubic construction into the orthorhombic box
            sc = np.array(box) / box[0]
            cen = RRc.reshape(N, 3, 3)
            oc = cen[:, :1, :]
            RRt = (RRt.reshape(P, N, 3, 3) - oc[None]) + oc[None] * sc
            RRt = RRt.reshape(P, 3 * N, 3)
            RRc = RRt.mean(axis=0)
            bx = box
        p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=bx, rc=rc, eps_ewald=eps,
                      **extra).resolved()
        # move one molecule of bead 0 to a periodic image outside the box: the
        # centroid wrap leaves beads there (main_stuff.f90:100-106)
        RRt = RRt.copy()
        RRt[0, 0:3] += np.array([bx[0], 0.0, -bx[2]])
        RRc = RRt.mean(axis=0)
        f = refrun.forces(p, RRt, RRc, t=1)
        save(name, RRt=RRt, RRc=RRc, t=np.asarray(1), En=f["En"], Umon=f["Umon"], Uvdw=f["Uvdw"],
             Uelec=f["Uelec"], Uind=f["Uind"], dRR=f["dRR"], virt_cm=f["virt"],
             virialc=f["virialc"], dipI=f["dipI"], dipE=f["dipE"], chg=f["chg"],
             iters=f["iters"], **params_dict(p))


def predictor_fixture():
    """guess_initdip: per-bead predictor history over 7 MD steps (t > 4 uses it)."""
    N, P = 16, 2
    RRt, PPt, RRc, L = make_system(N, P, seed=21, lattice=(4, 2, 2))
    p = TTMParams(pot_model=3, nwaters=N, nbeads=P, box=(L, L, L), rc=-1.0, eps_ewald=1e-6,
                  guess_initdip=True).resolved()
    dR0 = np.zeros_like(RRt)
    r = refrun.trajectory(p, RRt, PPt, dR0, nsteps=7, temp=300.0, delt_fs=0.25,
                          bead_thermostat=0, per_bead_hist=True)
    save("traj_ttm3f_predictor_n16_p2", RRt=RRt, PPt=PPt,
         Upot=np.array([s["Upot"] for s in r["steps"]]), uk=np.array([s["uk"] for s in r["steps"]]),
         obs=np.array([s["obs"] for s in r["steps"]]),
         RRt_end=r["RRt"], PPt_end=r["PPt"], dRRt_end=r["dRRt"], iters=r["iters"],
         **params_dict(p))


def traj_fixture(only=None):
    """PIMD_VV with the NM Langevin thermostat; the normals it consumed are
    recorded.  The last case is centroid MD (setNMfreq 1500 cm^-1: scaled
    normal-mode masses, mode-space kinetic energy) with the global
    Nose-Hoover chain."""
    for name, model, N, P, lat, glob, setnm in [
            ("traj_ttm3f_langevin_n8_p4", 3, 8, 4, (2, 2, 2), False, 0.0),
            ("traj_ttm21f_langevin_nh_n8_p2", 2, 8, 2, (2, 2, 2), True, 0.0),
            ("traj_ttm3f_cmd_nh_n8_p4", 3, 8, 4, (2, 2, 2), True, 1500.0),
            ("traj_qspcfw_langevin_n8_p4", 4, 8, 4, (2, 2, 2), False, 0.0),
            ("traj_spcf_langevin_nh_n8_p2", 5, 8, 2, (2, 2, 2), True, 0.0)]:
        if only and name not in only:
            continue
        RRt, PPt, RRc, L = make_system(N, P, seed=31 + N + P, lattice=lat)
        p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=-1.0,
                      eps_ewald=1e-6, mass_o=15.995, mass_h=1.008).resolved()
        dR0 = np.zeros_like(RRt)
        r = refrun.trajectory(p, RRt, PPt, dR0, nsteps=6, temp=300.0, delt_fs=0.25,
                              bead_thermostat=1, centroid=True, tau_centroid=0.05,
                              global_nh=glob, tau=0.1, per_bead_hist=True, setnm=setnm)
        noise = np.array([[s["noise1"], s["noise2"]] for s in r["steps"]])
        save(name, RRt=RRt, PPt=PPt, noise=noise, setnm=np.asarray(setnm),
             obs=np.array([s["obs"] for s in r["steps"]]),
             Upot=np.array([s["Upot"] for s in r["steps"]]),
             uk=np.array([s["uk"] for s in r["steps"]]), RRt_end=r["RRt"], PPt_end=r["PPt"],
             dRRt_end=r["dRRt"], RRc_end=r["RRc"], iters=r["iters"], global_nh=np.asarray(glob),
             **params_dict(p))


def config0_fixture():
    """The CPU-runnable benchmark system as stated: 32 H2O (0.997 g/cm^3,
    4 x 2 x 4 lattice), TTM3-F, Rc 4.9, eps_ewald 1e-8, P = 32, 100 NVE steps
    of 0.25 fs from a thermal ring polymer at 300 K; per-step energies,
    kinetic energy and write_out observables and the final state."""
    N, P = 32, 32
    RRt, PPt, RRc, L = make_system(N, P, seed=1000, lattice=(4, 2, 4), temp=300.0)
    p = TTMParams(pot_model=3, nwaters=N, nbeads=P, box=(L, L, L), rc=4.9, eps_ewald=1e-8,
                  mass_o=15.995, mass_h=1.008).resolved()
    r = refrun.trajectory(p, RRt, PPt, np.zeros_like(RRt), nsteps=100, temp=300.0, delt_fs=0.25,
                          bead_thermostat=0, per_bead_hist=True)
    save("traj_ttm3f_nve_n32_p32_rc49", RRt=RRt, PPt=PPt,
         obs=np.array([s["obs"] for s in r["steps"]]),
         Upot=np.array([s["Upot"] for s in r["steps"]]),
         uk=np.array([s["uk"] for s in r["steps"]]), RRt_end=r["RRt"], PPt_end=r["PPt"],
         dRRt_end=r["dRRt"], RRc_end=r["RRc"], iters=r["iters"], **params_dict(p))


def bead_nh_fixture():
    """PIMD_VV with the normal-mode-space Nose-Hoover bead thermostat
    (bead_thermostat_type 'Nose-Hoover', main_stuff.f90:145-171): with and
    without the centroid chain, one with the global chain as well."""
    for name, model, N, P, cen, glob, setnm in [
            ("traj_ttm3f_beadnh_n8_p4", 3, 8, 4, True, False, 0.0),
            ("traj_ttm21f_beadnh_nocen_n8_p3", 2, 8, 3, False, True, 0.0),
            ("traj_ttm3f_beadnh_cmd_n8_p4", 3, 8, 4, True, False, 1500.0)]:
        RRt, PPt, RRc, L = make_system(N, P, seed=60 + N + P, lattice=(2, 2, 2))
        p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=-1.0,
                      eps_ewald=1e-6, mass_o=15.995, mass_h=1.008).resolved()
        r = refrun.trajectory(p, RRt, PPt, np.zeros_like(RRt), nsteps=6, temp=300.0,
                              delt_fs=0.25, bead_thermostat=2, centroid=cen, tau_centroid=0.05,
                              global_nh=glob, tau=0.1, per_bead_hist=True, setnm=setnm)
        save(name, RRt=RRt, PPt=PPt, bead_tt=np.asarray(2), centroid=np.asarray(cen),
             setnm=np.asarray(setnm), obs=np.array([s["obs"] for s in r["steps"]]),
             Upot=np.array([s["Upot"] for s in r["steps"]]),
             uk=np.array([s["uk"] for s in r["steps"]]), RRt_end=r["RRt"], PPt_end=r["PPt"],
             dRRt_end=r["dRRt"], RRc_end=r["RRc"], iters=r["iters"], global_nh=np.asarray(glob),
             vxi_beads=r["vxi_beads"], **params_dict(p))


def contracted_fixture():
    """contracted_MD: monomer forces in inner steps, centroid TTM forces once per step."""
    for name, model, N, P, intra, lat in [("contracted_ttm3f_n8_p4_intra3", 3, 8, 4, 3, (2, 2, 2)),
                                          ("contracted_ttm21f_n8_p1_intra4", 2, 8, 1, 4, (2, 2, 2))]:
        RRt, PPt, RRc, L = make_system(N, P, seed=70 + N + P, lattice=lat)
        p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=-1.0,
                      eps_ewald=1e-6, mass_o=15.995, mass_h=1.008).resolved()
        dR0 = np.zeros_like(RRt)
        r = refrun.contracted_trajectory(p, RRt, PPt, dR0, nsteps=5, intra=intra, temp=300.0,
                                         delt_fs=0.5, bead_thermostat=1, centroid=True,
                                         tau_centroid=0.05, seed=900 + P)
        noise = np.array([[s["noise1"], s["noise2"]] for s in r["steps"]])
        rec = np.array([s["rec"] for s in r["steps"]])
        save(name, RRt=RRt, PPt=PPt, noise=noise, intra=np.asarray(intra),
             Upot=rec[:, 0], virial=rec[:, 1], virialc=rec[:, 2],
             uk=np.array([s["uk"] for s in r["steps"]]), RRt_end=r["RRt"], PPt_end=r["PPt"],
             dRRt_end=r["dRRt"], RRc_end=r["RRc"], iters=r["iters"], **params_dict(p))


DROPIN_INPUT = """old
!Input file of the drop-in program test (the reference's positional format)

BASIC INFO
'dropin_n32_p4.img'    ! input filename - either a centroid .xyz or a full bead configuration
'dropin'    ! run name - this will be appended to all the output files
4                    ! Number of beads
0                    ! number of steps to equilibrate
10                   ! number of steps to run
.25                  ! timestep (delt) (fs)

OUTPUT OPTIONS
.f.                      ! centroid coordinates
.f.                      ! centroid momenta
.f.                      ! dipoles
.f.                      ! polarization dipoles
.f.                      ! total dipole
.f.                      ! coordinates for all images
.f.                      ! dipoles for all images
.f.                      ! box size running average
.t.                      ! Temp/Press (.t. = to file, .f. = to terminal)
.t.                      ! avg. radius of gyration (in Temp/Press file)
.f.                      ! output dielectric constant (running average in Temp/Press file)
.f.                      ! write out charges in seperate file?
.t.                      ! write final configuration of all beads to seperate file?
20                       ! Total dipole output frequency
1                        ! Temp/Press output frequency
2000                     ! all images output frequency
10                       ! Output frequency for everything else

MD OPTIONS
3                    ! Model: 2=ttm21f 3=ttm3f 4= qspcfw 5=spcf 6=SIESTA
'siesta','32h2o', 1                 ! sys_label for SIESTA runs, # SIESTA processors
4.9d0 3.92d0 1.d-8           ! realspace cutoff (Ang), start of VdW force switch (Ang) (TTM3F only), eps for aewald
15  0.7  1.d-5 .f.  .f. ! params for dipole iterative procedure: Max iterations, pol factor, accuracy, guess initial?, debug output
.f.                     ! generate initial velocities?
.f.                     ! Global thermostat? t/f
.f.                     ! Thermostat non-centroid modes ?
.f.                     ! Thermostat centroid mode ?
'none'                  ! type of bead thermostat - 'Nose-Hoover', 'Langevin' or 'none'
.5                      ! tau for global thermostat (ps)
.1                      ! tau for centroid thermostat (ps)
2              ! global Nose-Hoover chain length
2              ! bead Nose-Hoover chain length
300      ! temperature (Kelvin)
.f.          ! Berendson barostat ?
.2       ! tau for barostat (ps)
1            ! reference pressure (bar)
.f.          ! pressure coupling during equilibration only?

MORE RUN OPTIONS
0            ! frequency (inverse cm) to set NMs to ( or enter 0 for none/RPMD)
{ptype}      ! type of PIMD run - "full", "contracted", "monomerPIMD"
3            ! ratio of slow timestep / fast timestep for contraction scheme
15.994       ! mass of Oxygen (au)
1.008        ! mass of Hyrdrogen
.false.      ! restart? (append to previous files)
"""


def parse_temppress(path):
    """The '@' lines of a run's Temp/Press file (InputOutput.f90:633-670):
    time, temperature, running simple pressure, simple energy, its running
    mean, running potential energy per molecule, mean r_O and r_H, mean
    molecular dipole."""
    rows = []
    with open(path) as f:
        for ln in f:
            if ln.startswith("@"):
                rows.append([float(x) for x in ln[1:].split()])
    return np.array(rows)


def dropin_fixture():
    """The reference program (PIMD.f90, built unmodified into
    oracle/_ref/PIMD_ref.x) run from a committed input and bead image for 10
    NVE steps, full and contracted (3 inner steps): its Temp/Press lines and
    the final bead image it writes.  The GPU test runs the drop-in programs on
    the same files."""
    import shutil
    import subprocess
    import tempfile
    d = os.path.join(OUT, "inputs")
    from pimd_amd import runinput
    from pimd_amd.params import f32
    RRt, PPt, RRc, L = make_system(32, 4, seed=1100, lattice=(4, 2, 4), temp=300.0)
    runinput.write_image(os.path.join(d, "dropin_n32_p4.img"), RRt, PPt, (L, L, L), 0, 0.25e-3,
                         f32(15.994), f32(1.008))
    out = {}
    for ptype in ("full", "contracted"):
        with open(os.path.join(d, f"dropin_{ptype}.inp"), "w") as f:
            f.write(DROPIN_INPUT.replace("{ptype}", ptype))
        with tempfile.TemporaryDirectory() as tmp:
            shutil.copy(os.path.join(d, f"dropin_{ptype}.inp"), tmp)
            shutil.copy(os.path.join(d, "dropin_n32_p4.img"), tmp)
            subprocess.run([refrun.PIMD_BIN, f"dropin_{ptype}.inp"], cwd=tmp, check=True,
                           stdout=subprocess.DEVNULL)
            out[f"tp_{ptype}"] = parse_temppress(os.path.join(tmp, "out_dropin_TempPress.dat"))
            with open(os.path.join(tmp, "out_dropin_finalchk.img")) as f:
                out[f"final_{ptype}"] = np.array(f.read())
    save("dropin_ttm3f_n32_p4", **out)


def io_fixtures():
    """The reference's input-file reader on its own example inputs and on a
    crafted one (tests/golden/inputs/quirks.inp), and its checkpoint writer and
    reader on a seeded bead configuration."""
    import json
    import shutil
    import subprocess
    from pimd_amd.params import f32
    d = os.path.join(OUT, "inputs")
    os.makedirs(d, exist_ok=True)
    for src in ("input32.inp", "input_new_64.inp", "test_input.inp",
                "input_files/input4nm.inp", "input_files/inputdimer.inp"):
        dst = os.path.join(d, os.path.basename(src))
        shutil.copyfile(os.path.join(refrun.REF, src), dst)
    for f in sorted(os.listdir(d)):
        if not f.endswith(".inp"):
            continue
        r = refrun.parse_input(os.path.join(d, f))
        with open(os.path.join(d, f[:-4] + ".json"), "w") as fh:
            json.dump(r, fh, indent=1, sort_keys=True)
        print("wrote inputs/" + f[:-4] + ".json")
    # checkpoint: the reference's save_checkpoint text and load_checkpoint result
    RRt, PPt, RRc, L = make_system(8, 3, seed=81, lattice=(2, 2, 2))
    RRt = RRt + np.array([4.0, -3.1, 0.2])       # centroids outside the box
    mO, mH = f32(15.994), f32(1.008)
    tmp = os.path.join(d, "_tmp.bin")
    with open(tmp, "wb") as fh:
        fh.write(np.array([24, 3], np.int32).tobytes())
        fh.write(np.array([mO, mH, L, L, L, 0.5, 0.25e-3, 40.0]).tobytes())
        fh.write(np.ascontiguousarray(RRt).tobytes())
        fh.write(np.ascontiguousarray(PPt).tobytes())
    img = os.path.join(d, "beads_n8_p3.img")
    subprocess.run([refrun.IMG_BIN, "write", tmp, img], check=True)
    out = os.path.join(d, "_out.bin")
    subprocess.run([refrun.IMG_BIN, "read", tmp, img, out], check=True)
    o = np.fromfile(out)
    n = 3 * 24 * 3
    save("inputs/beads_n8_p3", RRt=RRt, PPt=PPt, box=np.array([L, L, L]), t=np.asarray(40),
         delt_ps=np.asarray(0.25e-3), init_time=np.asarray(0.5), mass_o=np.asarray(mO),
         mass_h=np.asarray(mH), RRt_loaded=o[:n].reshape(3, 24, 3),
         PPt_loaded=o[n:2 * n].reshape(3, 24, 3), RRc_loaded=o[2 * n:].reshape(24, 3))
    os.remove(tmp)
    os.remove(out)


def ring_fixtures():
    for P, setnm in [(8, 0.0), (7, 0.0), (6, 1500.0)]:
        N = 4
        RRt, PPt, RRc, L = make_system(N, P, seed=40 + P, lattice=(2, 2, 1), box=7.0)
        p = TTMParams(pot_model=3, nwaters=N, nbeads=P, box=(L, L, L), mass_o=15.995,
                      mass_h=1.008).resolved()
        r = refrun.nm(p, RRt, PPt, 300.0, 0.25, setnm)
        r["RRt_out"], r["PPt_out"] = r.pop("RRt"), r.pop("PPt")
        save(f"nm_p{P}_setnm{int(setnm)}", RRt=RRt, PPt=PPt, temp=np.asarray(300.0),
             delt_fs=np.asarray(0.25), setnm=np.asarray(setnm), **r, **params_dict(p))
    for P, cen in [(8, True), (5, False)]:
        N = 4
        RRt, PPt, RRc, L = make_system(N, P, seed=50 + P, lattice=(2, 2, 1), box=7.0)
        p = TTMParams(pot_model=3, nwaters=N, nbeads=P, box=(L, L, L), mass_o=15.995,
                      mass_h=1.008).resolved()
        r = refrun.langevin(p, PPt, 300.0, 0.25, tau_centroid=0.05, centroid=cen, seed=777 + P)
        save(f"langevin_p{P}_centroid{int(cen)}", PPt_in=PPt, temp=np.asarray(300.0),
             delt_fs=np.asarray(0.25), tau_centroid=np.asarray(0.05), centroid=np.asarray(cen),
             **r, **params_dict(p))
    nh = []
    for uk, v, tau, L, T in [(12.5, (0.3, -0.2), 0.1, 300, 3000.0), (800.0, (1.0, 1.0), 0.05, 2880, 9600.0)]:
        r = refrun.nose_hoover(uk, v, tau, 0.25, L, T)
        nh.append([uk, v[0], v[1], tau, L, T, r["s"], r["uk"], r["vxi"][0], r["vxi"][1]])
    save("nose_hoover", table=np.array(nh))


def scalar_fixtures():
    RRt, PPt, RRc, L = make_system(32, 1, seed=60)
    p = TTMParams(pot_model=3, nwaters=32, nbeads=1, box=(L, L, L), rc=4.9, eps_ewald=1e-8).resolved()
    rng = np.random.default_rng(61)
    xs = np.concatenate([np.linspace(0.0, 6.0, 61), rng.uniform(0, 40, 40)])
    d2 = np.concatenate([rng.uniform(0.6, 24.99, 70), rng.uniform(25.0, 40.0, 31)])
    ap = rng.uniform(0.15, 0.45, len(xs))
    f = refrun.funcs(p, xs, d2, ap)
    save("scalar_funcs", xs=xs, drsq=d2, apol=ap, **f, **params_dict(p))
    # monomers: bead geometries of a quantum ring polymer (stretched/bent)
    RRt, PPt, RRc, L = make_system(8, 8, seed=62, lattice=(2, 2, 2), box=6.3)
    r1s = RRt.reshape(-1, 3, 3)
    for model in (3, 2):
        q = TTMParams(pot_model=model, nwaters=8, nbeads=8, box=(L, L, L)).resolved()
        m = refrun.monomer(q, r1s)
        save(f"monomer_model{model}", r1s=r1s, e1=m["e1"], dr1=m["dr1"], q3=m["q3"],
             dq3_f=m["dq3"], **params_dict(q))
    for name, (N, box, rc, eps) in {"setup_n32_rc49": (32, None, 4.9, 1e-8),
                                    "setup_n128_rchalf": (128, None, -1.0, 1e-6)}.items():
        Lb = make_system(N, 1, seed=1)[3] if box is None else box
        q = TTMParams(pot_model=3, nwaters=N, nbeads=1, box=(Lb, Lb, Lb), rc=rc,
                      eps_ewald=eps).resolved()
        s = refrun.setup(q)
        save(name, **{k: np.asarray(v) for k, v in s.items()}, **params_dict(q))


if __name__ == "__main__":
    if not refrun.available():
        sys.exit("build the reference first: make -C oracle ref")
    if len(sys.argv) > 1:   # named fixture groups or force and trajectory fixtures only
        groups = {"config0": config0_fixture, "bead_nh": bead_nh_fixture,
                  "dropin": dropin_fixture}
        for a in sys.argv[1:]:
            if a in groups:
                groups[a]()
        force_fixtures(set(sys.argv[1:]))
        traj_fixture(set(sys.argv[1:]))
        sys.exit(0)
    force_fixtures()
    predictor_fixture()
    traj_fixture()
    config0_fixture()
    bead_nh_fixture()
    contracted_fixture()
    io_fixtures()
    dropin_fixture()
    ring_fixtures()
    scalar_fixtures()
