"""GPU parity of the all-beads force path against the reference.

Fixture cases hold the reference's own outputs; larger systems are checked
against the C restatement (itself pinned to the reference bit-for-bit in
test_oracle_golden.py); the full 512-water, 32-bead system is also checked
through properties that do not depend on size.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden, params_from
from helpers import E_REL, F_ABS, max_abs, ref_iters, rel_err

pytestmark = pytest.mark.gpu


def _ff(p):
    from pimd_amd.ttm import TTMForceField
    return TTMForceField(p)


def _report(tag, **errs):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


def check_against(r, ref, p, tag, iters=True):
    errs = {"En": rel_err(r["Upott"], ref["En"]), "dRR": max_abs(r["dRRt"], ref["dRR"]),
            "virt": rel_err(r["virt_cm"], ref["virt"]), "virc": rel_err(r["virialct"], ref["virialc"]),
            "dipI": max_abs(r["dip_momIt"], ref["dipI"]), "dipE": max_abs(r["dip_momEt"], ref["dipE"]),
            "chg": max_abs(r["chg"], ref["chg"])}
    _report(tag, **errs)
    assert errs["En"] < E_REL
    assert errs["dRR"] < F_ABS
    assert errs["virt"] < 1e-10 and errs["virc"] < 1e-10
    assert errs["dipI"] < 1e-8 and errs["dipE"] < 1e-8
    assert errs["chg"] < 1e-12
    if iters:
        assert np.array_equal(ref_iters(r["iters"], p.polar_maxiter),
                              ref_iters(ref["iters"], p.polar_maxiter))


@pytest.mark.parametrize("name", golden_names("forces_"))
def test_forces_vs_reference_fixture(name):
    g = load_golden(name)
    p = params_from(g)
    print("case", name, flush=True)
    ff = _ff(p)
    r = ff.full_bead_forces(g["RRt"], g["RRc"], t=int(g["t"]))
    ref = {"En": g["En"], "dRR": g["dRR"], "virt": g["virt_cm"], "virialc": g["virialc"],
           "dipI": g["dipI"], "dipE": g["dipE"], "chg": g["chg"], "iters": g["iters"]}
    check_against(r, ref, p, name)
    terms = ff.energy_terms()
    for i, k in enumerate(("Umon", "Uvdw", "Uelec", "Uind")):
        assert rel_err(terms[:, i], g[k]) < E_REL, k
    ff.close()


def _oracle_forces(p, RRt, RRc, beads):
    import oracle
    from pimd_amd.params import TTMParams
    q = TTMParams(**{**p.__dict__, "nbeads": len(beads)})
    o = oracle.System(q).forces(RRt[beads], RRc)
    return {"En": o["En"], "dRR": o["dRR"], "virt": o["virt"], "virialc": o["virialc"],
            "dipI": o["dipI"], "dipE": o["dipE"], "chg": o["chg"], "iters": o["iters"]}


def _select(r, beads):
    return {k: (v[beads] if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[0] > max(beads)
                else v) for k, v in r.items()}


@pytest.mark.parametrize("model,N,P,rc,eps", [(2, 128, 4, -1.0, 1e-6), (3, 128, 3, -1.0, 1e-8),
                                              (3, 256, 2, 9.0, 1e-6), (4, 128, 4, -1.0, 1e-8),
                                              (5, 256, 2, 9.0, 1e-6), (3, 256, 2, 6.0, 1e-6),
                                              (2, 512, 1, 7.0, 1e-6), (4, 256, 2, 6.0, 1e-6),
                                              (3, 512, 1, 6.0, 1e-6)])
def test_forces_vs_oracle_medium(model, N, P, rc, eps, monkeypatch):
    """PIMD_CELLS=2: the cases with Rc = 6 and 7 A (three cells per axis)
    search neighbours through the cell list, the others over all molecules."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    RRt, _, RRc, L = make_system(N, P, seed=100 + N + model)
    p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=rc,
                  eps_ewald=eps).resolved()
    print("case", model, N, P, rc, eps, flush=True)
    monkeypatch.setenv("PIMD_CELLS", "2")
    ff = _ff(p)
    assert ff.info()["cells"] == (rc in (6.0, 7.0))
    r = ff.full_bead_forces(RRt, RRc)
    beads = list(range(P))
    check_against(r, _oracle_forces(p, RRt, RRc, beads), p, f"ttm{model} N{N} P{P}")


@pytest.mark.parametrize("model,N,P", [(2, 128, 4), (3, 128, 3)])
def test_two_sided_charge_pass_vs_oracle(model, N, P, monkeypatch):
    """The charge pair pass that takes every pair from both sides
    (PIMD_QQ_HALF=0), against the oracle, and against the once-per-pair pass
    (PIMD_QQ_HALF=2) on the same input: the two differ by rounding only."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    RRt, _, RRc, L = make_system(N, P, seed=300 + N + model)
    p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=-1.0,
                  eps_ewald=1e-6).resolved()
    print("case", model, N, P, flush=True)
    out = {}
    for mode in ("0", "2"):
        monkeypatch.setenv("PIMD_QQ_HALF", mode)
        ff = _ff(p)
        out[mode] = ff.full_bead_forces(RRt, RRc)
        ff.close()
    check_against(out["0"], _oracle_forces(p, RRt, RRc, list(range(P))), p, f"ttm{model} two-sided")
    assert rel_err(out["0"]["Upott"], out["2"]["Upott"]) < 1e-12
    assert max_abs(out["0"]["dRRt"], out["2"]["dRRt"]) < 1e-9
    assert rel_err(out["0"]["virialct"], out["2"]["virialct"]) < 1e-11


@pytest.fixture(scope="module")
def big_system():
    """512 H2O, P = 32, TTM3-F, box 24.86 A, Rc = L/2 (the benchmark system)."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    RRt, PPt, RRc, L = make_system(512, 32, seed=2002)
    p = TTMParams(pot_model=3, nwaters=512, nbeads=32, box=(L, L, L), rc=-1.0,
                  eps_ewald=1e-6).resolved()
    ff = _ff(p)
    r = ff.full_bead_forces(RRt, RRc)
    return p, ff, RRt, RRc, r


def test_full_size_matches_oracle_on_two_beads(big_system):
    p, ff, RRt, RRc, r = big_system
    beads = [0, 17]
    ref = _oracle_forces(p, RRt, RRc, beads)
    check_against(_select(r, beads), ref, p, "N512 P32 beads 0,17")


def test_full_size_deterministic_and_bead_order(big_system):
    p, ff, RRt, RRc, r = big_system
    r2 = ff.full_bead_forces(RRt, RRc)
    for k in ("Upott", "dRRt", "virt_cm", "virialct", "dip_momIt"):
        assert np.array_equal(r[k], r2[k]), k
    perm = np.arange(p.nbeads)[::-1].copy()
    r3 = ff.full_bead_forces(RRt[perm], RRc)
    for k in ("Upott", "dRRt", "virt_cm", "virialct", "dip_momIt", "iters"):
        assert np.array_equal(r[k][perm], r3[k]), k


def test_full_size_periodic_image_invariance(big_system):
    p, ff, RRt, RRc, r = big_system
    L = p.box[0]
    shifted = RRt.copy()
    shifted[:, 0:3, 0] += L        # one molecule moved to a periodic image in every bead
    shifted[5, :, 1] -= L          # every atom of bead 5 by -L along y
    rs = ff.full_bead_forces(shifted, RRc)
    assert rel_err(rs["Upott"], r["Upott"]) < E_REL
    assert max_abs(rs["dRRt"], r["dRRt"]) < F_ABS
    assert max_abs(rs["dip_momIt"], r["dip_momIt"]) < 1e-8


def test_gradient_sums_to_small_net_force(big_system):
    """Every term of the path is pairwise antisymmetric or sums to zero per
    molecule (PS gradient, DMS chain rule, M-site weights (1-g)+g/2+g/2, Ewald
    k-space).  The one exception is the reciprocal charge-dipole term, which
    pairs the dipole structure factor of the last SOR iterate with the final
    dipoles (ewald.f90:180-192); it leaves a net gradient of the order of the
    last dipole update, far below the gradient scale."""
    p, ff, RRt, RRc, r = big_system
    net = np.abs(r["dRRt"].sum(axis=1)).max()
    scale = np.abs(r["dRRt"]).max()
    print("net gradient", net, "max component", scale, flush=True)
    assert net < 1e-5 * scale


def test_fixed_charge_models_refuse_ttm_only_paths():
    """Contraction (its fast forces are the TTM models' PS surface) and the
    TTM / PES parameter derivatives are refused for q-SPC/Fw and SPC/F; the
    force call itself works on one bead as on many."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    from pimd_amd.ttm import PimdError, TTMForceField
    RRt, _, RRc, L = make_system(16, 1, seed=3, lattice=(4, 2, 2))
    for model in (4, 5):
        kw = dict(pot_model=model, nwaters=16, nbeads=1, box=(L, L, L))
        with pytest.raises(PimdError, match="contraction"):
            TTMForceField(TTMParams(**kw, contraction=True))
        ff = TTMForceField(TTMParams(**kw))
        r = ff.full_bead_forces(RRt, RRc)
        assert np.all(np.isfinite(r["dRRt"])) and np.all(r["iters"] == 0)
        assert np.all(r["virialct"] == 0.0) and np.all(r["dip_momEt"] == 0.0)
        na = 48
        with pytest.raises(PimdError, match="TTM"):
            ff.ttm_param_grad(RRt, np.zeros(1), np.zeros((1, na, 3)))
        with pytest.raises(PimdError, match="TTM"):
            ff.pes_param_grad(RRt, np.zeros(1), np.zeros((1, na, 3)))
        ff.close()


@pytest.mark.parametrize("model,N,P,rc", [(3, 256, 3, 6.0), (2, 512, 2, 7.0), (5, 256, 2, 6.0),
                                          (3, 1024, 1, 6.0)])
def test_cell_list_matches_full_search(model, N, P, rc, monkeypatch):
    """With at least three cells of width >= Rc per axis the neighbour search
    can go through a cell list (by default from 108 cells on, PIMD_CELLS=2
    from 27); the forces equal those of the full search (PIMD_CELLS=0) up to
    the order of the pair sums."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    RRt, _, RRc, L = make_system(N, P, seed=300 + N + model)
    p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=rc,
                  eps_ewald=1e-6).resolved()
    print("case cells", model, N, P, rc, flush=True)
    ncell = int(np.floor(L / rc)) ** 3
    ff = _ff(p)
    assert ff.info()["cells"] == (ncell >= 108)
    ff.close()
    monkeypatch.setenv("PIMD_CELLS", "2")
    ff = _ff(p)
    assert ff.info()["cells"] == 1
    rc_ = ff.full_bead_forces(RRt, RRc)
    ff.close()
    monkeypatch.setenv("PIMD_CELLS", "0")
    ff = _ff(p)
    assert ff.info()["cells"] == 0
    rf = ff.full_bead_forces(RRt, RRc)
    ff.close()
    _report(f"cells vs full ttm{model} N{N}", En=rel_err(rc_["Upott"], rf["Upott"]),
            dRR=max_abs(rc_["dRRt"], rf["dRRt"]))
    assert rel_err(rc_["Upott"], rf["Upott"]) < 1e-12
    assert max_abs(rc_["dRRt"], rf["dRRt"]) < 1e-10
    assert np.array_equal(rc_["iters"], rf["iters"])


def test_force_call_in_two_halves():
    """pimd_forces_all_beads_start / _finish on two contexts, both started
    before either is finished, give pimd_forces_all_beads's outputs bit for
    bit; a second start before the finish and a finish with none started are
    refused."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    from pimd_amd.ttm import PimdError
    cases = []
    for model, N, P in ((3, 64, 4), (2, 32, 2)):
        RRt, _, RRc, L = make_system(N, P, seed=400 + N)
        p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=-1.0,
                      eps_ewald=1e-6).resolved()
        ff = _ff(p)
        cases.append((ff, RRt, RRc, ff.full_bead_forces(RRt, RRc, t=3)))
    print("case two halves", flush=True)
    for ff, RRt, RRc, _ in cases:
        ff.full_bead_forces_start(RRt, RRc, t=3)
    for ff, RRt, RRc, ref in cases:
        got = ff.full_bead_forces_finish()
        for k in ref:
            assert np.array_equal(got[k], ref[k]), k
    ff, RRt, RRc, ref = cases[0]
    with pytest.raises(PimdError):
        ff.full_bead_forces_finish()
    ff.full_bead_forces_start(RRt, RRc, t=3)
    with pytest.raises(PimdError):
        ff.full_bead_forces_start(RRt, RRc, t=3)
    with pytest.raises(PimdError):
        ff.full_bead_forces(RRt, RRc, t=3)
    assert np.array_equal(ff.full_bead_forces_finish()["dRRt"], ref["dRRt"])
    for c in cases:
        c[0].close()
