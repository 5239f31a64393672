"""The Fortran host side (ISO_C_BINDING modules over the C ABI).

pimd-f90_amd/fortran/_build/pimd_gpu_driver is a Fortran program that calls
the GPU path through the pimd_gpu module with the reference's array layouts
(RRt(3,Natoms,Nbeads), virt(3,3), noise(3,Nbeads,Natoms)); its outputs are
compared with the reference outputs held in tests/golden.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, golden_names, load_golden, params_from
from helpers import E_REL, F_ABS, max_abs, rel_err

DRIVER = os.path.join(ROOT, "pimd-f90_amd", "fortran", "_build", "pimd_gpu_driver")


def _header(p, mode, t=1, centroid=0, delt_fs=0.25, temp=300.0, tau_centroid=0.1, setnm=0.0):
    ih = np.zeros(16, dtype=np.int32)
    ih[:9] = [mode, p.pot_model, p.nwaters, p.nbeads, p.polar_maxiter, int(p.guess_initdip), t,
              0, centroid]
    dh = np.zeros(32, dtype=np.float64)
    dh[0:3] = p.box
    dh[3:14] = [p.rc, p.rc1, p.eps_ewald, p.polar_sor, p.polar_eps, p.mass_o, p.mass_h,
                delt_fs, temp, tau_centroid, setnm]
    return ih.tobytes() + dh.tobytes()


def _run(blob, check=True):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built: run __graft_entry__.build()")
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(blob)
        r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        out = np.fromfile(fout, dtype=np.float64) if os.path.exists(fout) else None
    if check:
        assert r.returncode == 0, r.stdout + r.stderr
    return r, out


def _blob_forces(g, p):
    return (_header(p, 1, t=int(g["t"])) + np.ascontiguousarray(g["RRt"]).tobytes()
            + np.ascontiguousarray(g["RRc"]).tobytes())


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_driver_fails_loudly_without_gpu():
    """No CPU fallback: without a GPU the program stops with the library's error."""
    g = load_golden("forces_ttm3f_n16_p2_ortho")
    r, _ = _run(_blob_forces(g, params_from(g)), check=False)
    assert r.returncode != 0
    assert "pimd_gpu" in (r.stdout + r.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_names("forces_"))
def test_fortran_forces_vs_reference(name):
    g = load_golden(name)
    p = params_from(g)
    N, na, P = p.nwaters, 3 * p.nwaters, p.nbeads
    print("case", name, flush=True)
    _, out = _run(_blob_forces(g, p))
    rec = 6 + 9 + 3 * na + 6 * N + na
    a = out.reshape(P, rec)
    o = 0

    def take(n):
        nonlocal o
        v = a[:, o:o + n]
        o += n
        return v
    En, Umon, Uvdw, Uelec, Uind, vc = (take(1)[:, 0] for _ in range(6))
    virt = take(9).reshape(P, 3, 3)
    dRR = take(3 * na).reshape(P, na, 3)
    dipI = take(3 * N).reshape(P, N, 3)
    dipE = take(3 * N).reshape(P, N, 3)
    chg = take(na)
    assert rel_err(En, g["En"]) < E_REL
    for k, v in (("Umon", Umon), ("Uvdw", Uvdw), ("Uelec", Uelec), ("Uind", Uind)):
        assert rel_err(v, g[k]) < E_REL, k
    assert rel_err(vc, g["virialc"]) < 1e-10
    assert rel_err(virt, g["virt_cm"]) < 1e-10
    assert max_abs(dRR, g["dRR"]) < F_ABS
    assert max_abs(dipI, g["dipI"]) < 1e-8 and max_abs(dipE, g["dipE"]) < 1e-8
    assert max_abs(chg, g["chg"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_names("nm_"))
def test_fortran_normal_modes_vs_reference(name):
    g = load_golden(name)
    p = params_from(g)
    P, na = p.nbeads, 3 * p.nwaters
    print("case", name, flush=True)
    blob = (_header(p, 2, temp=float(g["temp"]), delt_fs=float(g["delt_fs"]),
                    setnm=float(g["setnm"]))
            + np.ascontiguousarray(g["RRt"]).tobytes() + np.ascontiguousarray(g["PPt"]).tobytes())
    _, out = _run(blob)
    C = out[:P * P].reshape(P, P).T          # written as the reference's C(mode, bead)
    o = P * P
    for k in ("A11", "A12", "A21", "A22", "omegalist", "MassScaleFactor"):
        assert rel_err(out[o:o + P], g[k]) < 1e-14, k
        o += P
    assert rel_err(C, g["C"]) < 1e-14
    R = out[o:o + P * na * 3].reshape(P, na, 3)
    Pm = out[o + P * na * 3:].reshape(P, na, 3)
    assert rel_err(R, g["RRt_out"]) < 1e-12 and rel_err(Pm, g["PPt_out"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_names("langevin_"))
def test_fortran_langevin_vs_reference(name):
    g = load_golden(name)
    p = params_from(g)
    P, na = p.nbeads, 3 * p.nwaters
    print("case", name, flush=True)
    blob = (_header(p, 3, centroid=int(g["centroid"]), temp=float(g["temp"]),
                    delt_fs=float(g["delt_fs"]), tau_centroid=float(g["tau_centroid"]))
            + np.ascontiguousarray(g["PPt_in"]).tobytes()
            + np.ascontiguousarray(g["noise"]).tobytes())
    _, out = _run(blob)
    assert rel_err(out.reshape(P, na, 3), g["PPt"]) < 1e-12
