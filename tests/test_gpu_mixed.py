"""The mixed-precision form (PIMD_MIXED) against the FP64 oracle.

Real-space pair terms and the real-space dipole field run in FP32 with FP64
accumulation, each bead's last SOR step is redone in FP64, and everything else
is FP64.  The bar is a force (gradient) RMS error of at most 1e-5 relative to
the FP64 result; energies, virials and dipoles are reported and held to
bounds of the same order.  The FP64 GPU form is itself pinned to the reference
at 1e-10 (test_gpu_forces.py), so it stands in for the oracle on all beads;
two beads are also compared with the oracle directly.
"""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

F_RMS = 1e-5      # gradient RMS error / gradient RMS
E_MIX = 1e-5      # relative energy error per bead
DIP_MIX = 1e-5    # dipole error relative to the largest dipole component


def _ff(p):
    from pimd_amd.ttm import TTMForceField
    return TTMForceField(p)


def rms_rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _compare(rm, ref, tag):
    errs = {"F_rms": rms_rel(rm["dRRt"], ref["dRR"]),
            "F_max": float(np.max(np.abs(rm["dRRt"] - ref["dRR"])) / np.max(np.abs(ref["dRR"]))),
            "E": float(np.max(np.abs(rm["Upott"] - ref["En"]) / np.abs(ref["En"]))),
            "virt": rel_err(rm["virt_cm"], ref["virt"]),
            "dip": rel_err(rm["dip_momIt"], ref["dipI"])}
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert errs["F_rms"] < F_RMS
    assert errs["E"] < E_MIX
    assert errs["virt"] < 1e-5
    assert errs["dip"] < DIP_MIX
    return errs


def _as_ref(r):
    return {"En": r["Upott"], "dRR": r["dRRt"], "virt": r["virt_cm"], "dipI": r["dip_momIt"]}


@pytest.fixture(scope="module")
def big_pair():
    """512 H2O, 32 beads, TTM3-F, Rc = L/2, eps 1e-6 in both forms."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    RRt, PPt, RRc, L = make_system(512, 32, seed=2002)
    kw = dict(pot_model=3, nwaters=512, nbeads=32, box=(L, L, L), rc=-1.0, eps_ewald=1e-6)
    p64 = TTMParams(**kw).resolved()
    pmx = TTMParams(**kw, precision=1).resolved()
    r64 = _ff(p64).full_bead_forces(RRt, RRc)
    ffm = _ff(pmx)
    rmx = ffm.full_bead_forces(RRt, RRc)
    return p64, pmx, ffm, RRt, RRc, r64, rmx


def test_mixed_full_size_vs_fp64_all_beads(big_pair):
    p64, pmx, ffm, RRt, RRc, r64, rmx = big_pair
    _compare(rmx, _as_ref(r64), "mixed vs fp64, 512x32")
    # the FP32 field does not change the SOR path: iteration counts agree
    # with the FP64 form to within one step
    assert np.max(np.abs(np.abs(rmx["iters"]) - np.abs(r64["iters"]))) <= 1


def test_mixed_full_size_vs_oracle_two_beads(big_pair):
    import oracle
    from pimd_amd.params import TTMParams
    p64, pmx, ffm, RRt, RRc, r64, rmx = big_pair
    beads = [3, 29]
    q = TTMParams(**{**p64.__dict__, "nbeads": len(beads)})
    o = oracle.System(q).forces(RRt[beads], RRc)
    sel = {k: v[beads] for k, v in rmx.items() if isinstance(v, np.ndarray) and v.ndim >= 1
           and v.shape[0] == p64.nbeads}
    _compare(sel, {"En": o["En"], "dRR": o["dRR"], "virt": o["virt"], "dipI": o["dipI"]},
             "mixed vs oracle, beads 3,29")


def test_mixed_deterministic(big_pair):
    p64, pmx, ffm, RRt, RRc, r64, rmx = big_pair
    again = ffm.full_bead_forces(RRt, RRc)
    for k in ("Upott", "dRRt", "virt_cm", "dip_momIt"):
        assert np.array_equal(again[k], rmx[k]), k


@pytest.mark.parametrize("model,N,P,rc,eps", [(2, 128, 8, -1.0, 1e-6), (3, 32, 8, 4.9, 1e-8),
                                              (4, 256, 4, -1.0, 1e-6), (5, 128, 4, -1.0, 1e-8)])
def test_mixed_vs_fp64_other_systems(model, N, P, rc, eps):
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    RRt, _, RRc, L = make_system(N, P, seed=500 + N + model)
    kw = dict(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=rc, eps_ewald=eps)
    print("case", model, N, P, rc, eps, flush=True)
    r64 = _ff(TTMParams(**kw).resolved()).full_bead_forces(RRt, RRc)
    rmx = _ff(TTMParams(**kw, precision=1).resolved()).full_bead_forces(RRt, RRc)
    _compare(rmx, _as_ref(r64), f"mixed vs fp64 ttm{model} N{N} P{P}")


@pytest.mark.parametrize("model,N,P,half", [(2, 128, 4, "0"), (3, 128, 3, "0"),
                                            (2, 128, 1, "1"), (3, 256, 1, "1")])
def test_mixed_two_sided_pair_passes(model, N, P, half, monkeypatch):
    """The FP32 charge and final pair passes that take every pair from both
    sides: the LDS form (PIMD_QQ_HALF=0) and, with one bead, the
    global-memory form that one-bead contexts use; mixed against FP64 at the
    same bar."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    monkeypatch.setenv("PIMD_QQ_HALF", half)
    RRt, _, RRc, L = make_system(N, P, seed=700 + N + model)
    kw = dict(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=-1.0, eps_ewald=1e-6)
    print("case two-sided", model, N, P, half, flush=True)
    r64 = _ff(TTMParams(**kw).resolved()).full_bead_forces(RRt, RRc)
    rmx = _ff(TTMParams(**kw, precision=1).resolved()).full_bead_forces(RRt, RRc)
    _compare(rmx, _as_ref(r64), f"two-sided mixed vs fp64 ttm{model} N{N} P{P}")
