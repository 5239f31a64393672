"""The reference program with its force path swapped for the GPU one, run
end to end from the reference's own input format.

pimd-f90_amd/fortran/_dropin holds two builds of the reference program
(PIMD.f90 and all its units, compiled from the reference sources by
`make -C pimd-f90_amd/fortran dropin`): PIMD_gpu.x with force_calc.f90
replaced by force_calc_gpu.f90 (every bead in one launch) and
PIMD_gpu_potttm.x with pot_ttm.f90 replaced by pot_ttm_gpu.f90 (one bead
per call).  Each runs the committed input (positional format) from the
committed bead image for 10 NVE steps, as full PIMD (PIMD_VV) and as
contracted PIMD (contracted_MD, 3 inner steps, whose centroid call needs the
CONTRACTION branch of pot_ttm.f90:92-132).  Its Temp/Press lines and final
bead image must equal those the unmodified reference program printed
(tests/golden/dropin_ttm3f_n32_p4.npz) to the last printed digit.  The
same input and image also go through the Python host side (runinput into
the device integrator), checked against the same outputs.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_golden

pytestmark = pytest.mark.gpu

DROPIN = os.path.join(ROOT, "pimd-f90_amd", "fortran", "_dropin")
INPUTS = os.path.join(GOLDEN, "inputs")
# one unit of the last printed digit of each Temp/Press column
# (InputOutput.f90:633-670): time f10.3, temperature f10.2, pressure f10.2,
# energies f12.3, radii f6.4, dipole f8.4
TP_ULP = np.array([1e-3, 1e-2, 1e-2, 1e-3, 1e-3, 1e-3, 1e-4, 1e-4, 1e-4])


def temppress(path):
    with open(path) as f:
        return np.array([[float(x) for x in ln[1:].split()] for ln in f if ln.startswith("@")])


def image_numbers(text):
    """Header values and the coordinate / velocity columns of a bead image."""
    vals = []
    for ln in text.splitlines():
        for tok in (ln[2:] if ln[:1].isalpha() else ln).split():
            vals.append(float(tok))
    return np.array(vals)


@pytest.mark.parametrize("ptype", ["full", "contracted"])
@pytest.mark.parametrize("exe", ["PIMD_gpu.x", "PIMD_gpu_potttm.x"])
def test_dropin_program_matches_reference_run(exe, ptype, tmp_path):
    path = os.path.join(DROPIN, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: build it with `make -C pimd-f90_amd/fortran dropin`")
    g = load_golden("dropin_ttm3f_n32_p4")
    for f in (f"dropin_{ptype}.inp", "dropin_n32_p4.img"):
        shutil.copy(os.path.join(INPUTS, f), tmp_path)
    print("case", exe, ptype, flush=True)
    r = subprocess.run([path, f"dropin_{ptype}.inp"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    tp = temppress(tmp_path / "out_dropin_TempPress.dat")
    ref = g[f"tp_{ptype}"]
    assert tp.shape == ref.shape
    units = np.abs(tp - ref) / TP_ULP[None, :]
    print(exe, ptype, "Temp/Press: largest difference", float(units.max()),
          "units of the last printed digit", flush=True)
    assert units.max() <= 1.0 + 1e-6
    with open(tmp_path / "out_dropin_finalchk.img") as f:
        mine = image_numbers(f.read())
    theirs = image_numbers(str(g[f"final_{ptype}"]))
    assert mine.shape == theirs.shape
    d = float(np.max(np.abs(mine - theirs)))
    print(exe, ptype, "final image: largest difference", d, flush=True)
    assert d <= 1e-6 * (1 + 1e-6)      # one unit of f12.6


@pytest.mark.parametrize("ptype", ["full", "contracted"])
def test_python_run_from_reference_input(ptype, tmp_path):
    """The same committed input and bead image through the Python host side:
    runinput.read_input, read_image (the .img restart path,
    InputOutput.f90:1030-1074) and run_setup into TTMForceField and
    PIMDIntegrator, 10 NVE steps on the device.  The temperatures and the
    final bead image written by runinput.write_image must equal the
    unmodified reference program's to the last printed digit."""
    from pimd_amd import runinput
    from pimd_amd.ttm import PIMDIntegrator, TTMForceField
    g = load_golden("dropin_ttm3f_n32_p4")
    inp = runinput.read_input(os.path.join(INPUTS, f"dropin_{ptype}.inp"), warn=lambda m: None)
    R, Pm, Rc, box, t_init, _ = runinput.read_image(os.path.join(INPUTS, "dropin_n32_p4.img"),
                                                   inp["massO"], inp["massH"], inp["Nbeads"])
    p, md = runinput.run_setup(inp, box, R.shape[1])
    n = inp["eq_timesteps"] + inp["run_timesteps"]
    print("case python", ptype, "steps", n, flush=True)
    ff = TTMForceField(p)
    integ = PIMDIntegrator(ff, **md)
    integ.set_state(R, Pm, None)
    integ.run(n, t0=1)
    obs = integ.observables(n)
    st = integ.get_state()
    ff.close()
    temp = np.round(obs[:, 0], 2)
    ref = g[f"tp_{ptype}"]
    dt = float(np.max(np.abs(temp - ref[:, 1])))
    print("python", ptype, "temperature: largest difference", dt, flush=True)
    assert dt <= TP_ULP[1] * (1 + 1e-6)
    # the final checkpoint is written after the step loop (PIMD.f90:74), whose
    # index then stands one past its bound (InputOutput.f90:870-872)
    out = tmp_path / "final.img"
    runinput.write_image(str(out), st["RRt"], st["PPt"], box, n + 1, inp["delt"] * 1e-3,
                         p.mass_o, p.mass_h, init_time=t_init)
    mine = image_numbers(out.read_text())
    theirs = image_numbers(str(g[f"final_{ptype}"]))
    assert mine.shape == theirs.shape
    d = float(np.max(np.abs(mine - theirs)))
    print("python", ptype, "final image: largest difference", d, flush=True)
    assert d <= 1e-6 * (1 + 1e-6)      # one unit of f12.6


@pytest.mark.parametrize("model", [2, 3, 4])
def test_monomer_dipoles_vs_oracle_charges(model):
    """pimd_monomer_dipoles (force_calc.f90:205-264) on every bead against the
    oracle's dipole-moment-surface charges (pinned by the reference's monomer
    fixtures) put through the formula of force_calc.f90:221-255: the M-site
    share folded into the H charges, H positions relative to the M site for
    TTM3-F and q-SPC/Fw (gamma_M 0.46 and 0), to the O for TTM2.1-F."""
    import oracle
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import CHARGECON, make_system
    from pimd_amd.ttm import TTMForceField
    N, P = 16, 3
    RRt, _, RRc, L = make_system(N, P, seed=31)
    p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L)).resolved()
    got = TTMForceField(p).monomer_dipoles(RRt)
    gM = {2: 0.426706882, 3: 0.46, 4: 0.0}[model]
    dms = {2: (0.0, 0.0, 0.0), 3: (0.5, 0.9578, 0.012), 4: (0.0, 0.0, 0.0)}[model]
    box = np.array([L, L, L])
    want = np.zeros((P, N, 3))
    tq = 0.5 * gM / (1.0 - gM)
    for b in range(P):
        for w in range(N):
            r = RRt[b, 3 * w:3 * w + 3]
            q = oracle.monomer(r, box, dms)[2] * CHARGECON
            qh1, qh2 = q[1] + tq * (q[1] + q[2]), q[2] + tq * (q[1] + q[2])
            o1 = r[1] - r[0]
            o1 -= box * np.rint(o1 / box)
            o2 = r[2] - r[0]
            o2 -= box * np.rint(o2 / box)
            h1, h2 = o1, o2
            if model in (3, 4):
                m = 0.5 * gM * (o1 + o2) + r[0]
                h1 = r[1] - m
                h1 -= box * np.rint(h1 / box)
                h2 = r[2] - m
                h2 -= box * np.rint(h2 / box)
            want[b, w] = qh1 * h1 + qh2 * h2
    err = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    print("model", model, "monomer dipoles rel err", err, flush=True)
    assert err < 1e-12
