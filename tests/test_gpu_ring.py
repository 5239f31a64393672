"""GPU parity of the ring-polymer propagation and of whole PIMD_VV steps.

Normal modes, the Langevin thermostat (with the normals the reference
consumed, recorded by its driver) and short trajectories are compared with the
reference's outputs held in tests/golden.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden, params_from
from helpers import max_abs, rel_err
from pimd_amd.systems import CHARGECON, DEBYE, HBAR, KB_AMUA2PS2PERK, atom_masses

pytestmark = pytest.mark.gpu


def _ff(p):
    from pimd_amd.ttm import TTMForceField
    return TTMForceField(p)


@pytest.mark.parametrize("name", golden_names("nm_"))
def test_normal_modes_vs_reference(name):
    from pimd_amd.ttm import RingPolymer
    g = load_golden(name)
    p = params_from(g)
    P = p.nbeads
    print("case", name, flush=True)
    rp = RingPolymer(_ff(p))
    rp.init_normal_modes(KB_AMUA2PS2PERK * float(g["temp"]) * P / HBAR,
                         float(g["delt_fs"]) / 1000.0, float(g["setnm"]))
    md = rp.module_data()
    for k in ("C", "A11", "A12", "A21", "A22", "omegalist", "MassScaleFactor"):
        assert rel_err(md[k], g[k]) < 1e-14, k
    R, Pm = rp.evolve_ring(g["RRt"], g["PPt"], atom_masses(p.nwaters, p.mass_o, p.mass_h))
    assert rel_err(R, g["RRt_out"]) < 1e-12 and rel_err(Pm, g["PPt_out"]) < 1e-12


@pytest.mark.parametrize("name", golden_names("langevin_"))
def test_langevin_vs_reference(name):
    from pimd_amd.ttm import RingPolymer
    g = load_golden(name)
    p = params_from(g)
    P = p.nbeads
    T, dt = float(g["temp"]), float(g["delt_fs"]) / 1000.0
    print("case", name, flush=True)
    rp = RingPolymer(_ff(p))
    rp.init_normal_modes(KB_AMUA2PS2PERK * T * P / HBAR, dt)
    rp.init_langevin_nm(dt / 2, bool(g["centroid"]), float(g["tau_centroid"]), P * T)
    out = rp.langevin_nm(g["PPt_in"], noise=g["noise"])
    assert rel_err(out, g["PPt"]) < 1e-12


def test_langevin_device_noise_statistics():
    """Device normals (Philox4x32-10 + Box-Muller): from p = 0 a thermostat call
    leaves p~_j = c2_j sqrt(m) xi in mode space, so the mode variances must be
    c2_j^2 m and the samples uncorrelated with zero mean."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    from pimd_amd.ttm import RingPolymer
    N, P, T = 64, 32, 300.0
    _, _, _, L = make_system(N, 1, seed=1)
    p = TTMParams(pot_model=3, nwaters=N, nbeads=P, box=(L, L, L)).resolved()
    rp = RingPolymer(_ff(p))
    rp.init_normal_modes(KB_AMUA2PS2PERK * T * P / HBAR, 0.25e-3)
    rp.init_langevin_nm(0.125e-3, True, 0.1, P * T)
    md = rp.module_data()
    out = rp.langevin_nm(np.zeros((P, 3 * N, 3)), seed=1234, counter=7)
    m = atom_masses(N, p.mass_o, p.mass_h)
    modes = np.einsum("jk,kan->jan", md["C"], out) / np.sqrt(m)[None, :, None]
    c1 = np.exp(-0.125e-3 * np.where(np.arange(P) == 0, 10.0, 2 * md["omegalist"]))
    c2 = np.sqrt(KB_AMUA2PS2PERK * P * T * (1 - c1 ** 2))
    z = modes / c2[:, None, None]
    n = z.size
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(z.var() - 1.0) < 0.05
    again = rp.langevin_nm(np.zeros((P, 3 * N, 3)), seed=1234, counter=7)
    assert np.array_equal(again, out)
    other = rp.langevin_nm(np.zeros((P, 3 * N, 3)), seed=1234, counter=8)
    assert not np.allclose(other, out)


@pytest.mark.parametrize("name", golden_names("traj_"))
def test_trajectory_vs_reference(name):
    from pimd_amd.ttm import PIMDIntegrator
    g = load_golden(name)
    p = params_from(g)
    noise = g.get("noise")
    glob = bool(g.get("global_nh", False))
    nsteps = g["Upot"].shape[0]
    print("case", name, flush=True)
    ff = _ff(p)
    setnm = float(g.get("setnm", 0.0))
    # Langevin with the reference's recorded normals, the NM-space Nose-Hoover
    # chains (bead_tt 2), or none
    bead_tt = int(g["bead_tt"]) if "bead_tt" in g else (1 if noise is not None else 0)
    integ = PIMDIntegrator(ff, delt_fs=0.25, temp=300.0, bead_thermostat=bead_tt,
                           centroid_thermostat=bool(g.get("centroid", True)), tau_centroid=0.05,
                           global_nh=glob, tau=0.1, setNMfreq=setnm)
    integ.set_state(g["RRt"], g["PPt"], None)
    rec = integ.run(nsteps, t0=1, noise=noise)
    obs = integ.observables(nsteps)
    st = integ.get_state()
    errs = {"Upot": rel_err(rec["Upot"], g["Upot"][:, :, 0].sum(axis=1)),
            "uk": rel_err(rec["uk"], g["uk"]), "R": rel_err(st["RRt"], g["RRt_end"]),
            "P": rel_err(st["PPt"], g["PPt_end"]), "dR": max_abs(st["dRRt"], g["dRRt_end"])}
    print(name, errs, flush=True)
    assert errs["Upot"] < 1e-10 and errs["uk"] < 1e-10
    assert errs["R"] < 1e-12 and errs["P"] < 1e-10 and errs["dR"] < 1e-8
    # write_out's observables (temperature, virial and simple estimators, dipoles)
    oerr = [rel_err(obs[:, i], g["obs"][:, i]) for i in range(obs.shape[1])]
    print(name, "observables rel err", np.max(oerr), flush=True)
    assert max(oerr[:5]) < 1e-10 and max(oerr[5:]) < 1e-8
    if "RRc_end" in g:
        assert rel_err(st["RRc"], g["RRc_end"]) < 1e-12


@pytest.mark.parametrize("name", golden_names("contracted_"))
def test_contracted_vs_reference(name):
    """contracted_MD (MD.f90:276-449) against the reference's own step loop:
    monomer forces on every bead in intra inner steps, the centroid's TTM
    forces (CONTRACTION branch, pot_ttm.f90:92-132) broadcast to every bead,
    driven by the normals the reference's thermostat drew."""
    from pimd_amd.ttm import PIMDIntegrator
    g = load_golden(name)
    p = params_from(g)
    nsteps = g["Upot"].shape[0]
    print("case", name, flush=True)
    ff = _ff(p)
    integ = PIMDIntegrator(ff, delt_fs=0.5, temp=300.0, bead_thermostat=1,
                           centroid_thermostat=True, tau_centroid=0.05,
                           intra_timesteps=int(g["intra"]))
    integ.set_state(g["RRt"], g["PPt"], None)
    rec = integ.run(nsteps, t0=1, noise=g["noise"])
    st = integ.get_state()
    errs = {k: rel_err(rec[k], g[k]) for k in ("Upot", "virial", "virialc", "uk")}
    errs.update({"R": rel_err(st["RRt"], g["RRt_end"]), "P": rel_err(st["PPt"], g["PPt_end"]),
                 "dR": max_abs(st["dRRt"], g["dRRt_end"]), "Rc": max_abs(st["RRc"], g["RRc_end"])})
    print(name, errs, flush=True)
    assert errs["Upot"] < 1e-10 and errs["uk"] < 1e-10
    assert errs["virial"] < 1e-9 and errs["virialc"] < 1e-9
    assert errs["R"] < 1e-12 and errs["P"] < 1e-10 and errs["dR"] < 1e-8 and errs["Rc"] < 1e-12
    # the step's dipole observables: the monomer dipoles of every bead
    # (calc_monomer_dip_moments after the centroid call, MD.f90:407) averaged
    # over beads per molecule, in Debye, at the last step's coordinates
    obs = integ.observables(nsteps)
    mono = ff.monomer_dipoles(st["RRt"])
    v = mono.mean(axis=0) * (DEBYE / CHARGECON)
    want = np.concatenate([v.sum(axis=0), [np.sum(np.linalg.norm(v, axis=1))]])
    assert rel_err(obs[-1, 5:], want) < 1e-12
    assert np.all(np.isfinite(obs))
    # a second run continues from the state, the fast gradient included
    rec2 = integ.run(1, t0=nsteps + 1)
    assert np.all(np.isfinite(rec2["Upot"]))


@pytest.mark.parametrize("path", ["nm_fma", "langevin_mfma"])
def test_other_transform_kernels_vs_reference(path, monkeypatch):
    """The normal-mode step on FMA slabs (PIMD_NM_MFMA=0) and the Langevin
    step on the FP64 matrix cores (PIMD_LANGEVIN_MFMA=1), the forms not
    taken by default, against the reference's outputs."""
    if path == "nm_fma":
        monkeypatch.setenv("PIMD_NM_MFMA", "0")
        for name in golden_names("nm_"):
            test_normal_modes_vs_reference(name)
    else:
        monkeypatch.setenv("PIMD_LANGEVIN_MFMA", "1")
        for name in golden_names("langevin_"):
            test_langevin_vs_reference(name)
