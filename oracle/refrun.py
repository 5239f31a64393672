"""Runs the reference Fortran routines (oracle/_ref/ref_driver) on given inputs.

Test infrastructure only: imported by tests/ and the CPU-baseline leg of
bench.py, never by the product.  The binary is built by oracle/Makefile from
the reference's own sources; see ref_driver.f90 for the modes.
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "_ref", "ref_driver")
PIMD_BIN = os.path.join(HERE, "_ref", "PIMD_ref.x")   # the reference program, unmodified
REF = "/root/reference"   # the reference sources (fixture generation only)


def available() -> bool:
    return os.path.exists(BIN)


def _header(p, mode, t=1, nsteps=0, centroid=0, seed=1, nfunc=0, bead_tt=0, glob_t=0,
            nh_L=0, per_bead_hist=0, delt_fs=0.25, temp=300.0, tau_centroid=0.1, setnm=0.0,
            tau=0.1, uk=0.0, vxi=(0.0, 0.0), intra=0):
    ih = np.zeros(16, dtype=np.int32)
    ih[:16] = [mode, p.pot_model, p.nwaters, p.nbeads, p.polar_maxiter, int(p.guess_initdip), t,
               nsteps, centroid, seed, nfunc, bead_tt, glob_t, nh_L, per_bead_hist, intra]
    dh = np.zeros(32, dtype=np.float64)
    dh[0:3] = p.box
    dh[3:18] = [p.rc, p.rc1, p.eps_ewald, p.polar_sor, p.polar_eps, p.mass_o, p.mass_h, delt_fs,
                temp, tau_centroid, setnm, tau, uk, vxi[0], vxi[1]]
    return ih.tobytes() + dh.tobytes()


def _run(blob: bytes):
    if not available():
        raise RuntimeError("reference binary missing: run `make -C oracle ref`")
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(blob)
        r = subprocess.run([BIN, fin, fout], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"ref_driver failed: {r.stdout}\n{r.stderr}")
        out = np.fromfile(fout, dtype=np.float64)
    iters = [int(m) for m in re.findall(r"#of iterations =\s*(\d+)", r.stdout)]
    return out, iters, r.stdout


class _Reader:
    def __init__(self, a):
        self.a, self.i = a, 0

    def take(self, *shape):
        n = int(np.prod(shape)) if shape else 1
        v = self.a[self.i:self.i + n]
        self.i += n
        return v.reshape(shape) if shape else float(v[0])


def forces(p, RRt, RRc, t=1, per_bead_hist=False):
    """pot_ttm for every bead (pot_ttm.f90:3-630), serial bead order."""
    N, na, P = p.nwaters, p.natoms, p.nbeads
    blob = _header(p, 1, t=t, per_bead_hist=int(per_bead_hist))
    blob += np.ascontiguousarray(RRt, dtype=np.float64).tobytes()
    blob += np.ascontiguousarray(RRc, dtype=np.float64).tobytes()
    out, iters, _ = _run(blob)
    r = _Reader(out)
    res = {k: [] for k in ("En", "Umon", "Uvdw", "Uelec", "Uind", "virialc", "virt", "dRR",
                           "dipI", "dipE", "chg")}
    for _ in range(P):
        for k in ("En", "Umon", "Uvdw", "Uelec", "Uind", "virialc"):
            res[k].append(r.take())
        res["virt"].append(r.take(3, 3))   # stored column-major: virt[col][row]
        res["dRR"].append(r.take(na, 3))
        res["dipI"].append(r.take(N, 3))
        res["dipE"].append(r.take(N, 3))
        res["chg"].append(r.take(na))
    res = {k: np.array(v) for k, v in res.items()}
    # the SPC models have no SOR loop and report no iterations
    res["iters"] = np.array(iters) if iters else np.zeros(P, dtype=int)
    return res


def setup(p):
    """init_pot + ewald_set (potential.f90:43-199, ewald.f90:293-378)."""
    out, _, _ = _run(_header(p, 7))
    r = _Reader(out)
    res = {}
    for k in ("aewald", "hmax", "kmax", "lmax", "nkvec", "Uvdw_lrc"):
        res[k] = r.take()
    res["polfac"] = r.take(4)
    for k in ("shiftA", "shiftB", "const_ts1", "const_ts2", "const_ts3", "volume"):
        res[k] = r.take()
    nk = int(res["nkvec"])
    res["kvec"] = r.take(nk, 3)
    res["Ak"] = r.take(nk)
    res["hkl"] = r.take(nk, 3).astype(int)
    return res


def funcs(p, xs, drsq, apol):
    """erfc, gammq2_3, PBCsmear01/2/3, self1/2 at the given arguments."""
    xs, drsq, apol = (np.asarray(a, dtype=np.float64) for a in (xs, drsq, apol))
    blob = _header(p, 6, nfunc=len(xs)) + xs.tobytes() + drsq.tobytes() + apol.tobytes()
    out, _, _ = _run(blob)
    a = out.reshape(len(xs), 13)
    names = ["erfc", "gammq", "s01_ts0", "s01_ts1", "s2_ts1", "s2_ts2", "s3_ts1", "s3_ts2",
             "s3_ts3", "self1_ts0", "self1_ts1", "self2_ts1", "self2_ts2"]
    return {n: a[:, i] for i, n in enumerate(names)}


def monomer(p, r1s):
    """pot_nasa + dms_nasa for monomers r1s (n, 3 atoms, 3)."""
    r1s = np.ascontiguousarray(r1s, dtype=np.float64)
    out, _, _ = _run(_header(p, 8, nfunc=len(r1s)) + r1s.tobytes())
    a = out.reshape(len(r1s), 1 + 9 + 3 + 27)
    return {"e1": a[:, 0], "dr1": a[:, 1:10].reshape(-1, 3, 3), "q3": a[:, 10:13],
            "dq3": a[:, 13:].reshape(-1, 3, 3, 3)}  # dr1[atom][xyz], dq3[k][j][i] = dq3(i,j,k)


def nm(p, RRt, PPt, temp, delt_fs, setnm=0.0):
    """InitNormalModes + EvolveRing for every atom (NormalModes.f90:34-218)."""
    P, na = p.nbeads, p.natoms
    blob = _header(p, 2, temp=temp, delt_fs=delt_fs, setnm=setnm)
    blob += np.ascontiguousarray(RRt).tobytes() + np.ascontiguousarray(PPt).tobytes()
    out, _, _ = _run(blob)
    r = _Reader(out)
    res = {"C": r.take(P, P).T.copy()}  # C[mode][bead]
    for k in ("A11", "A12", "A21", "A22", "omegalist", "MassScaleFactor"):
        res[k] = r.take(P)
    res["RRt"] = r.take(P, na, 3)
    res["PPt"] = r.take(P, na, 3)
    return res


def langevin(p, PPt, temp, delt_fs, tau_centroid=0.1, centroid=False, seed=12345, setnm=0.0):
    """Init_Langevin_NM + one Langevin_NM call; returns the normals it consumed."""
    P, na = p.nbeads, p.natoms
    blob = _header(p, 3, temp=temp, delt_fs=delt_fs, tau_centroid=tau_centroid,
                   centroid=int(centroid), seed=seed, setnm=setnm)
    blob += np.ascontiguousarray(PPt).tobytes()
    out, _, _ = _run(blob)
    r = _Reader(out)
    return {"c1": r.take(P), "c2": r.take(P), "noise": r.take(na, P, 3), "PPt": r.take(P, na, 3)}


def nose_hoover(uk, vxi, tau, delt_fs, L, temp):
    """One Nose_Hoover half-step call (Nose_Hoover.f90:33-112)."""
    from types import SimpleNamespace
    p = SimpleNamespace(pot_model=3, nwaters=1, nbeads=1, polar_maxiter=1, guess_initdip=False,
                        box=(10.0, 10.0, 10.0), rc=5.0, rc1=4.0, eps_ewald=1e-6, polar_sor=0.7,
                        polar_eps=1e-5, mass_o=16.0, mass_h=1.0)
    out, _, _ = _run(_header(p, 4, delt_fs=delt_fs, temp=temp, tau=tau, uk=uk, vxi=vxi, nh_L=L))
    return {"s": out[0], "uk": out[1], "vxi": out[2:4].copy()}


def trajectory(p, RRt, PPt, dRRt, nsteps, temp, delt_fs, bead_thermostat=1, centroid=False,
               tau_centroid=0.1, global_nh=False, tau=0.1, seed=4242, per_bead_hist=True,
               setnm=0.0):
    """PIMD_VV steps (MD.f90:35-110) with forces from pot_ttm, bead by bead.

    bead_thermostat: 0 none, 1 Langevin (normals recorded), 2 NM Nose-Hoover.
    """
    P, na = p.nbeads, p.natoms
    blob = _header(p, 5, nsteps=nsteps, temp=temp, delt_fs=delt_fs, bead_tt=bead_thermostat,
                   centroid=int(centroid), tau_centroid=tau_centroid, glob_t=int(global_nh),
                   tau=tau, seed=seed, per_bead_hist=int(per_bead_hist), setnm=setnm)
    for a in (RRt, PPt, dRRt):
        blob += np.ascontiguousarray(a, dtype=np.float64).tobytes()
    out, iters, _ = _run(blob)
    r = _Reader(out)
    steps = []
    for _ in range(nsteps):
        st = {}
        if bead_thermostat == 1:
            st["noise1"] = r.take(na, P, 3)
        st["Upot"] = np.array([r.take(3) for _ in range(P)])  # En, virial, virialc per bead
        # sys_temp, virial energy and pressure, simple energy and pressure,
        # total dipole (3), sum of dipole magnitudes
        st["obs"] = r.take(9)
        if bead_thermostat == 1:
            st["noise2"] = r.take(na, P, 3)
        st["uk"] = r.take()
        steps.append(st)
    res = {"steps": steps, "RRt": r.take(P, na, 3), "PPt": r.take(P, na, 3),
           "dRRt": r.take(P, na, 3), "RRc": r.take(na, 3), "iters": np.array(iters)}
    if global_nh:
        res["vxi_global"] = r.take(2)
    if bead_thermostat == 2:
        # vxi_beads(2, Natoms, Nbeads, 3) column-major, as [atom][bead][component][chain]
        res["vxi_beads"] = r.take(3, P, na, 2).transpose(2, 1, 0, 3).copy()
    return res


def contracted_trajectory(p, RRt, PPt, dRRt, nsteps, intra, temp, delt_fs, bead_thermostat=1,
                          centroid=False, tau_centroid=0.1, seed=4242):
    """contracted_MD steps (MD.f90:276-449): per step the recorded normals of
    the two thermostat calls, (sum of bead energies, virial, centroid virial)
    and the kinetic energy; then the final state."""
    P, na = p.nbeads, p.natoms
    blob = _header(p, 9, nsteps=nsteps, temp=temp, delt_fs=delt_fs, bead_tt=bead_thermostat,
                   centroid=int(centroid), tau_centroid=tau_centroid, seed=seed, intra=intra)
    for a in (RRt, PPt, dRRt):
        blob += np.ascontiguousarray(a, dtype=np.float64).tobytes()
    out, iters, _ = _run(blob)
    r = _Reader(out)
    steps = []
    for _ in range(nsteps):
        st = {}
        if bead_thermostat == 1:
            st["noise1"] = r.take(na, P, 3)
        st["rec"] = r.take(3)
        if bead_thermostat == 1:
            st["noise2"] = r.take(na, P, 3)
        st["uk"] = r.take()
        steps.append(st)
    return {"steps": steps, "RRt": r.take(P, na, 3), "PPt": r.take(P, na, 3),
            "dRRt": r.take(P, na, 3), "RRc": r.take(na, 3), "iters": np.array(iters)}


INPUT_BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "ref_input")
IMG_BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "ref_img")


def parse_input(path):
    """The reference's read_inputfile on an input file: {key: value} of every
    variable it sets, or {"error": message} when it stops on a missing key."""
    r = subprocess.run([INPUT_BIN, os.path.abspath(path)], capture_output=True, text=True,
                       cwd=os.path.dirname(os.path.abspath(path)))
    lines = r.stdout.splitlines()
    if "#parsed" not in lines:
        return {"error": r.stdout.strip()}
    out = {"messages": [ln.strip() for ln in lines[:lines.index("#parsed")]]}
    for ln in lines[lines.index("#parsed") + 1:]:
        k, _, v = ln.partition(" ")
        out[k] = v
    return out

