"""The C restatement (oracle) against the reference's own outputs (golden vectors).

The fixtures were produced by the reference Fortran routines built from source
(tests/golden/gen_golden.py).  The restatement follows the reference's
statement order, so agreement is expected to rounding (<= 1e-13 relative).
"""
import dataclasses

import numpy as np
import pytest

from conftest import golden_names, load_golden, params_from
from helpers import max_abs, ref_iters, rel_err
from pimd_amd.systems import HBAR, KB_AMUA2PS2PERK, MASSCON, atom_masses

TIGHT = 1e-13


@pytest.mark.parametrize("name", golden_names("forces_"))
def test_forces_vs_reference(oracle_lib, name):
    g = load_golden(name)
    p = params_from(g)
    o = oracle_lib.System(p).forces(g["RRt"], g["RRc"], t=int(g["t"]))
    assert rel_err(o["En"], g["En"]) < TIGHT
    for i, k in enumerate(("Umon", "Uvdw", "Uelec", "Uind")):
        assert rel_err(o["terms"][:, i], g[k]) < TIGHT, k
    assert max_abs(o["dRR"], g["dRR"]) < 1e-11
    assert rel_err(o["virt"], g["virt_cm"]) < TIGHT
    assert rel_err(o["virialc"], g["virialc"]) < TIGHT
    assert max_abs(o["dipI"], g["dipI"]) < 1e-11
    assert max_abs(o["dipE"], g["dipE"]) < 1e-11
    assert max_abs(o["chg"], g["chg"]) < 1e-12
    assert np.array_equal(ref_iters(o["iters"], p.polar_maxiter), ref_iters(g["iters"], p.polar_maxiter))


def test_scalar_funcs_vs_reference(oracle_lib):
    g = load_golden("scalar_funcs")
    s = oracle_lib.System(params_from(g))
    erfc = np.array([oracle_lib.erfc(x) for x in g["xs"]])
    assert rel_err(erfc, g["erfc"]) < TIGHT
    gq = np.array([oracle_lib.gammq23(x) for x in g["xs"]])
    assert rel_err(gq, g["gammq"]) < TIGHT
    sm = np.array([s.smear(a, b) for a, b in zip(g["drsq"], g["apol"])])
    names = ["s01_ts0", "s01_ts1", "s2_ts1", "s2_ts2", "s3_ts1", "s3_ts2", "s3_ts3",
             "self1_ts0", "self1_ts1", "self2_ts1", "self2_ts2"]
    for i, n in enumerate(names):
        assert rel_err(sm[:, i], g[n]) < TIGHT, n


@pytest.mark.parametrize("model", [3, 2])
def test_monomer_vs_reference(oracle_lib, model):
    g = load_golden(f"monomer_model{model}")
    p = params_from(g)
    dms = (0.5, 0.9578, 0.012) if model == 3 else (0.0, 0.0, 0.0)
    res = [oracle_lib.monomer(r, p.box, dms) for r in g["r1s"]]
    assert rel_err([r[0] for r in res], g["e1"]) < TIGHT
    assert rel_err([r[1] for r in res], g["dr1"]) < TIGHT
    assert rel_err([r[2] for r in res], g["q3"]) < TIGHT
    # reference gradq(i,j,k) was written column-major: fixture [k][j][i]
    dq = np.transpose(g["dq3_f"], (0, 3, 2, 1))
    assert rel_err([r[3] for r in res], dq) < TIGHT


@pytest.mark.parametrize("name", golden_names("setup_"))
def test_setup_vs_reference(oracle_lib, name):
    g = load_golden(name)
    s = oracle_lib.System(params_from(g))
    c = s.consts()
    kv, Ak, hkl = s.kvecs()
    assert s.nk == int(g["nkvec"])
    assert np.array_equal(hkl, g["hkl"])
    assert rel_err(kv, g["kvec"]) < TIGHT and rel_err(Ak, g["Ak"]) < TIGHT
    for k in ("aewald", "Uvdw_lrc", "shiftA", "shiftB", "const_ts1", "const_ts2", "const_ts3"):
        assert abs(c[k] - float(g[k])) <= 1e-14 * max(1.0, abs(float(g[k]))), k
    assert rel_err(c["polfac"], g["polfac"]) < TIGHT


@pytest.mark.parametrize("name", golden_names("nm_"))
def test_normal_modes_vs_reference(oracle_lib, name):
    g = load_golden(name)
    p = params_from(g)
    P = p.nbeads
    omegan = KB_AMUA2PS2PERK * float(g["temp"]) * P / HBAR
    nm = oracle_lib.nm_init(P, omegan, float(g["delt_fs"]) / 1000.0, float(g["setnm"]))
    for k in ("C", "A11", "A12", "A21", "A22", "omegalist", "MassScaleFactor"):
        assert rel_err(nm[k], g[k]) < TIGHT, k
    R, Pm = oracle_lib.nm_evolve(nm, atom_masses(p.nwaters, p.mass_o, p.mass_h), g["RRt"], g["PPt"])
    assert rel_err(R, g["RRt_out"]) < TIGHT and rel_err(Pm, g["PPt_out"]) < TIGHT


@pytest.mark.parametrize("name", golden_names("langevin_"))
def test_langevin_vs_reference(oracle_lib, name):
    g = load_golden(name)
    p = params_from(g)
    P = p.nbeads
    T, dt = float(g["temp"]), float(g["delt_fs"]) / 1000.0
    nm = oracle_lib.nm_init(P, KB_AMUA2PS2PERK * T * P / HBAR, dt)
    c1, c2 = oracle_lib.langevin_init(P, dt / 2, bool(g["centroid"]), float(g["tau_centroid"]),
                                      P * T, nm["omegalist"])
    assert rel_err(c1, g["c1"]) < TIGHT and rel_err(c2, g["c2"]) < TIGHT
    out = oracle_lib.langevin(nm, c1, c2, p.mass_o, p.mass_h, g["noise"], g["PPt_in"])
    assert rel_err(out, g["PPt"]) < 1e-12


def test_nose_hoover_vs_reference(oracle_lib):
    for row in load_golden("nose_hoover")["table"]:
        uk, v0, v1, tau, L, T, s, uk1, w0, w1 = row
        r = oracle_lib.nose_hoover(uk, (v0, v1), tau, 0.25 / 2000, int(L), T)
        assert abs(r["s"] - s) < 1e-15 and abs(r["uk"] - uk1) < 1e-12 * abs(uk1)
        assert max_abs(r["vxi"], [w0, w1]) < 1e-14


TEMPFACTOR, PRESSCON2, KB_KCAL, DEBYE, CHARGECON = (0.80181484527, 69477.0, 0.0019872041,
                                                   4.8033324, 18.22234397655801030455)


def observables(p, R, dipI, Upot, virial, virialc, uk, omegan, masses):
    """write_out's per-step observables (InputOutput.f90:556-621) from the
    reference's estimators (estimators.f90:30-96), restated in numpy."""
    P, na, N = p.nbeads, p.natoms, p.nwaters
    vol = p.box[0] * p.box[1] * p.box[2]
    temp = TEMPFACTOR * uk / (na * P * P)
    ke = 1.5 * na * KB_KCAL * temp
    qe = (ke + 0.5 * (virial - virialc) / P + Upot / P) / N
    qp = PRESSCON2 * (1 / (3 * vol)) * (2 * ke - virialc / P) / N
    d = R - np.roll(R, 1, axis=0)       # bead k against bead k-1, bead 0 against P-1
    k0 = np.sum(masses[None, :, None] * d * d)
    kes = 1.5 * na * P * KB_KCAL * temp * P
    k0 = 0.5 * k0 * omegan ** 2 / 418.4
    se = (kes - k0 + Upot) / (N * P)
    sp = PRESSCON2 * (1 / (3 * vol)) * (2 * (kes - k0) - virial) / N
    mol = (dipI * DEBYE / CHARGECON).sum(axis=0) / P
    return np.array([temp, qe, qp, se, sp, *mol.sum(axis=0), np.sum(np.linalg.norm(mol, axis=1))])


def nh_chain2(u, v0, v1, tau, dt2, L, temp):
    """Nose_Hoover with a chain of length 2 (Nose_Hoover.f90:33-112),
    elementwise over arrays: returns (s, u, v0, v1)."""
    d4, d8 = dt2 / 2.0, dt2 / 4.0
    kT = KB_AMUA2PS2PERK * temp
    M1 = L * kT * (tau * tau)
    M2 = M1
    v1 = v1 + (M1 * (v0 * v0) - kT) * d4 / M2
    v0 = v0 * np.exp(-v1 * d8)
    v0 = v0 + (2 * u - L * kT) * d4 / M1
    v0 = v0 * np.exp(-v1 * d8)
    s = np.exp(-v0 * dt2)
    u = u * s * s
    v0 = v0 * np.exp(-v1 * d8)
    v0 = v0 + (2 * u - L * kT) * d4 / M1
    v0 = v0 * np.exp(-v1 * d8)
    v1 = v1 + (M1 * (v0 * v0) - kT) * d4 / M2
    return s, u, v0, v1


def bead_nose_hoover(nm, Pm, vxi, masses, centroid, tau_centroid, dt2, temp_beads):
    """bead_thermostat's Nose-Hoover branch (main_stuff.f90:145-171): real2NM
    of every atom's momenta, one chain per (atom, mode, component) with L = 1,
    the centroid only with CENTROIDTHERMOSTAT, NM2real.  vxi [Natoms][P][3][2]
    is updated in place."""
    C, om, msf = nm["C"], nm["omegalist"], nm["MassScaleFactor"]
    P = C.shape[0]
    q = np.einsum("jk,kac->ajc", C, Pm)           # [atom][mode][component]
    u = 0.5 * (1.0 / masses)[:, None, None] * q ** 2
    u[:, 1:, :] = u[:, 1:, :] / msf[None, 1:, None]
    tau = np.where(np.arange(P) == 0, tau_centroid, 1.0 / np.where(om > 0, om, 1.0))
    s, _, v0, v1 = nh_chain2(u, vxi[..., 0], vxi[..., 1], tau[None, :, None], dt2, 1.0,
                             temp_beads)
    live = np.ones(P, dtype=bool)
    live[0] = centroid
    s = np.where(live[None, :, None], s, 1.0)
    vxi[..., 0] = np.where(live[None, :, None], v0, vxi[..., 0])
    vxi[..., 1] = np.where(live[None, :, None], v1, vxi[..., 1])
    return np.einsum("kj,akc->jac", C, q * s)


def oracle_pimd_vv(oracle_lib, p, RRt, PPt, dRRt, nsteps, temp, delt_fs, noise=None,
                   centroid=True, tau_centroid=0.05, global_nh=False, tau=0.1, setnm=0.0,
                   bead_nh=False):
    """PIMD_VV (MD.f90:35-110) composed from the oracle's pieces, with the
    per-step observables of write_out.  Bead thermostat: Langevin with the
    given normals, or with bead_nh the normal-mode Nose-Hoover chains."""
    P, na = p.nbeads, p.natoms
    dt = delt_fs / 1000.0
    dt2 = dt / 2
    omegan = KB_AMUA2PS2PERK * temp * P / HBAR
    nm = oracle_lib.nm_init(P, omegan, dt, setnm)
    c1, c2 = oracle_lib.langevin_init(P, dt2, centroid, tau_centroid, P * temp, nm["omegalist"])
    masses = atom_masses(p.nwaters, p.mass_o, p.mass_h)
    sys_ = oracle_lib.System(p)
    R, Pm, dR = RRt.copy(), PPt.copy(), dRRt.copy()
    vxi = np.array([1.0, 1.0])
    L = 3 * na * P

    def uk_of(Pm):
        if setnm != 0.0 and P > 1:   # calc_uk in mode space (main_stuff.f90:366-385)
            q = np.einsum("jk,kac->jac", nm["C"], Pm)
            return 0.5 * np.sum((1.0 / masses)[None, :, None] * q ** 2
                                / nm["MassScaleFactor"][:, None, None])
        return 0.5 * np.sum((1.0 / masses)[None, :, None] * Pm ** 2)

    uk = uk_of(Pm)
    vxi_b = np.zeros((na, P, 3, 2))

    def thermostat(Pm, half):
        if noise is not None:
            return oracle_lib.langevin(nm, c1, c2, p.mass_o, p.mass_h, noise[st, half], Pm)
        if bead_nh:
            return bead_nose_hoover(nm, Pm, vxi_b, masses, centroid, tau_centroid, dt2, P * temp)
        return Pm

    out = []
    for st in range(nsteps):
        Pm = thermostat(Pm, 0)
        if global_nh:
            r = oracle_lib.nose_hoover(uk, vxi, tau, dt2, L, temp * P)
            vxi, uk = r["vxi"], r["uk"]
            Pm = Pm * r["s"]
        Pm = Pm - MASSCON * dR * dt2
        R, Pm = oracle_lib.nm_evolve(nm, masses, R, Pm)
        Rc = R.sum(axis=0) / P
        sh = p.box[0] * np.round(Rc / p.box[0])
        Rc = Rc - sh
        R = R - sh[None]
        f = sys_.forces(R, Rc, t=st + 1)
        dR = f["dRR"]
        obs = observables(p, R, f["dipI"], f["En"].sum(), np.trace(f["virt"], axis1=1, axis2=2).sum(),
                          f["virialc"].sum() / P, uk_of(Pm), omegan, masses)
        Pm = Pm - MASSCON * dR * dt2
        Pm = thermostat(Pm, 1)
        uk = uk_of(Pm)
        if global_nh:
            r = oracle_lib.nose_hoover(uk, vxi, tau, dt2, L, temp * P)
            vxi, uk = r["vxi"], r["uk"]
            Pm = Pm * r["s"]
        out.append((f["En"].sum(), uk, *obs))
    return R, Pm, dR, np.array(out), vxi_b


# the 100-step 32-bead run of the CPU-runnable system: its first steps here
# (the whole run takes minutes on the oracle; the GPU tests run all of it)
ORACLE_STEPS = {"traj_ttm3f_nve_n32_p32_rc49": 20}


@pytest.mark.parametrize("name", golden_names("traj_"))
def test_trajectory_vs_reference(oracle_lib, name):
    g = load_golden(name)
    p = params_from(g)
    noise = g.get("noise")
    glob = bool(g.get("global_nh", False))
    bead_nh = int(g.get("bead_tt", 0)) == 2
    full = g["Upot"].shape[0]
    nsteps = ORACLE_STEPS.get(name, full)
    R, Pm, dR, rec, vxi_b = oracle_pimd_vv(
        oracle_lib, p, g["RRt"], g["PPt"], np.zeros_like(g["RRt"]), nsteps, 300.0, 0.25,
        noise=noise, global_nh=glob, setnm=float(g.get("setnm", 0.0)), bead_nh=bead_nh,
        centroid=bool(g.get("centroid", True)))
    assert rel_err(rec[:, 0], g["Upot"][:nsteps, :, 0].sum(axis=1)) < 1e-12
    assert rel_err(rec[:, 1], g["uk"][:nsteps]) < 1e-12
    for i in range(9):
        assert rel_err(rec[:, 2 + i], g["obs"][:nsteps, i]) < 1e-11, i
    if nsteps < full:
        return
    assert rel_err(R, g["RRt_end"]) < 1e-12
    assert rel_err(Pm, g["PPt_end"]) < 1e-11
    assert max_abs(dR, g["dRRt_end"]) < 1e-10
    if bead_nh:
        assert rel_err(vxi_b, g["vxi_beads"]) < 1e-11


def test_bead_nose_hoover_restatement_is_pinned(oracle_lib):
    """A plausible slip in the bead chains (the centroid chain without the
    centroid thermostat, or unscaled mode masses) fails the reference
    trajectory that the test above matches."""
    g = load_golden("traj_ttm3f_beadnh_cmd_n8_p4")
    p = params_from(g)
    args = (oracle_lib, p, g["RRt"], g["PPt"], np.zeros_like(g["RRt"]), 6, 300.0, 0.25)
    good = oracle_pimd_vv(*args, setnm=1500.0, bead_nh=True)[1]
    assert rel_err(good, g["PPt_end"]) < 1e-11
    plain = oracle_pimd_vv(*args, setnm=1500.0, bead_nh=False)[1]
    assert rel_err(plain, g["PPt_end"]) > 1e-6


def _pes_all(oracle_lib, p, R):
    """pot_nasa of every monomer of every bead: energies [P][N], gradients [P][Natoms][3]."""
    box = np.asarray(p.box, dtype=np.float64)
    e = np.zeros((R.shape[0], p.nwaters))
    d = np.zeros_like(R)
    for b in range(R.shape[0]):
        for w in range(p.nwaters):
            e[b, w], d[b, 3 * w:3 * w + 3], _, _ = oracle_lib.monomer(R[b, 3 * w:3 * w + 3], box)
    return e, d


def _oh_virial(p, R, d):
    """Sum over molecules of roh1 . dr1(H1) + roh2 . dr1(H2) with minimum image."""
    box = np.asarray(p.box)
    v = 0.0
    for w in range(p.nwaters):
        for h in (1, 2):
            r = R[3 * w + h] - R[3 * w]
            r = r - box * np.round(r / box)
            v += float(r @ d[3 * w + h])
    return v


def oracle_contracted(oracle_lib, p, RRt, PPt, nsteps, intra, temp, delt_fs, noise,
                      tau_centroid=0.05):
    """contracted_MD (MD.f90:276-449) composed from the oracle's pieces.  The
    centroid's CONTRACTION forces (pot_ttm.f90:92-132) are the full TTM forces
    of the centroid minus its monomer PES terms, which is what the branch
    leaves out."""
    P, na = p.nbeads, p.natoms
    dt = delt_fs / 1000.0
    dt2, dtf = dt / 2, dt / intra
    dt2f = dtf / 2
    nm = oracle_lib.nm_init(P, KB_AMUA2PS2PERK * temp * P / HBAR, dtf)
    c1, c2 = oracle_lib.langevin_init(P, dt2, True, tau_centroid, P * temp, nm["omegalist"])
    masses = atom_masses(p.nwaters, p.mass_o, p.mass_h)
    cen = oracle_lib.System(dataclasses.replace(p, nbeads=1))
    R, Pm = RRt.copy(), PPt.copy()
    dR, dRf = np.zeros_like(R), np.zeros_like(R)

    def centroid(R):
        Rc = R.sum(axis=0) / P
        sh = p.box[0] * np.round(Rc / p.box[0])
        return R - sh[None], Rc - sh

    out = []
    for st in range(nsteps):
        Pm = oracle_lib.langevin(nm, c1, c2, p.mass_o, p.mass_h, noise[st, 0], Pm)
        Pm = Pm - MASSCON * dR * dt2
        R, Rc = centroid(R)
        for ti in range(1, intra + 1):
            Pm = Pm - MASSCON * dRf * dt2f
            if P > 1:
                R, Pm = oracle_lib.nm_evolve(nm, masses, R, Pm)
            else:
                R = R + (1.0 / masses)[None, :, None] * Pm * dtf
            e1, dRf = _pes_all(oracle_lib, p, R)
            if ti == intra:
                umon = e1.sum()
                vmon = sum(_oh_virial(p, R[b], dRf[b]) for b in range(P))
                vcmon = sum(_oh_virial(p, Rc, dRf[b]) for b in range(P))
            Pm = Pm - MASSCON * dRf * dt2f
        R, Rc = centroid(R)
        f = cen.forces(Rc[None], Rc, t=st + 1)
        ec, dc = _pes_all(oracle_lib, p, Rc[None])
        vc_mon = _oh_virial(p, Rc, dc[0])
        ucen = f["En"][0] - ec.sum()
        dR = np.broadcast_to(f["dRR"][0] - dc[0], R.shape).copy()
        vir = vmon + np.trace(f["virt"][0]) - vc_mon
        virc = vcmon / P + f["virialc"][0] - vc_mon
        Pm = Pm - MASSCON * dR * dt2
        Pm = oracle_lib.langevin(nm, c1, c2, p.mass_o, p.mass_h, noise[st, 1], Pm)
        uk = 0.5 * np.sum((1.0 / masses)[None, :, None] * Pm ** 2)
        out.append((umon + P * ucen, vir, virc, uk))
    return R, Pm, dR, Rc, np.array(out)


@pytest.mark.parametrize("name", golden_names("contracted_"))
def test_contracted_vs_reference(oracle_lib, name):
    g = load_golden(name)
    p = params_from(g)
    R, Pm, dR, Rc, rec = oracle_contracted(oracle_lib, p, g["RRt"], g["PPt"], g["Upot"].shape[0],
                                           int(g["intra"]), 300.0, 0.5, g["noise"])
    for i, k in enumerate(("Upot", "virial", "virialc", "uk")):
        assert rel_err(rec[:, i], g[k]) < 1e-11, k
    assert rel_err(R, g["RRt_end"]) < 1e-12
    assert rel_err(Pm, g["PPt_end"]) < 1e-11
    assert max_abs(dR, g["dRRt_end"]) < 1e-10
    assert max_abs(Rc, g["RRc_end"]) < 1e-12
