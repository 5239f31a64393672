"""MD steps replayed from a captured HIP graph against the same steps launched
one kernel at a time (PIMD_MD_GRAPH=0): every replay, with host reads and
synchronisations between the runs, must give the eager records, observables
and state bit for bit (plain, centroid-MD and contracted steps, a fixed-charge
q-SPC/Fw system, and a batch)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _integrator(model, N, P, seed, **md):
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    from pimd_amd.ttm import PIMDIntegrator, TTMForceField
    RRt, PPt, RRc, L = make_system(N, P, seed=seed)
    p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), rc=-1.0, eps_ewald=1e-6,
                  guess_initdip=True)
    ff = TTMForceField(p)
    md.setdefault("bead_thermostat", 1)
    integ = PIMDIntegrator(ff, delt_fs=0.25, temp=300.0, centroid_thermostat=True,
                           tau_centroid=0.1, **md)
    integ.set_state(RRt, PPt, ff.full_bead_forces(RRt, RRc)["dRRt"])
    return integ


def _runs(integ, graph: bool, chunks=(3, 2, 4)):
    os.environ["PIMD_MD_GRAPH"] = "1" if graph else "0"
    out, t0 = [], 1
    try:
        for n in chunks:
            rec = integ.run(n, t0=t0, seed=1234)
            obs = integ.observables(n)
            st = integ.get_state()          # synchronises and reads back between replays
            out.append((rec, obs, st))
            t0 += n
    finally:
        os.environ.pop("PIMD_MD_GRAPH", None)
    return out


@pytest.mark.parametrize("case", ["ttm3f", "ttm21f_cmd_nh", "contracted", "qspcfw",
                                  "ttm3f_bead_nh", "contracted_bead_nh"])
def test_graph_replay_matches_eager(case):
    """*_bead_nh: the normal-mode Nose-Hoover chains on the beads in place of
    the Langevin thermostat."""
    kw = {"ttm3f": (3, 64, 8, {}),
          "ttm21f_cmd_nh": (2, 32, 4, {"setNMfreq": 1500.0, "global_nh": True}),
          "contracted": (3, 32, 4, {"intra_timesteps": 3}),
          "qspcfw": (4, 64, 8, {}),
          "ttm3f_bead_nh": (3, 64, 8, {"bead_thermostat": 2}),
          "contracted_bead_nh": (3, 32, 4, {"intra_timesteps": 3, "bead_thermostat": 2})}[case]
    model, N, P, md = kw
    print("case", case, flush=True)
    a = _runs(_integrator(model, N, P, 77, **md), graph=False)
    b = _runs(_integrator(model, N, P, 77, **md), graph=True)
    for (ra, oa, sa), (rb, ob, sb) in zip(a, b):
        for k in ("Upot", "virial", "virialc", "uk"):
            assert np.array_equal(ra[k], rb[k]), k
        assert np.array_equal(oa, ob, equal_nan=True)
        for k in ("RRt", "PPt", "dRRt", "RRc"):
            assert np.array_equal(sa[k], sb[k]), k


# a batch of unlike replicas: Langevin and Nose-Hoover beads, CMD with the
# global chains, a fixed-charge model and contracted MD
BATCH = [(3, 32, 4, {}), (2, 32, 8, {"bead_thermostat": 2}),
         (3, 64, 4, {"setNMfreq": 1500.0, "global_nh": True}), (4, 32, 4, {}),
         (3, 32, 4, {"intra_timesteps": 3})]


def test_batch_graph_matches_eager():
    from pimd_amd.ttm import ReplicaBatch
    res = []
    seeds = [5, 6, 7, 8, 9]
    for graph in (False, True):
        os.environ["PIMD_MD_GRAPH"] = "1" if graph else "0"
        try:
            integs = [_integrator(m, N, P, 90 + r, **dict(md)) for r, (m, N, P, md) in enumerate(BATCH)]
            batch = ReplicaBatch(integs)
            r1 = batch.run(3, seeds, t0=1)
            r2 = batch.run(2, seeds, t0=4)
            res.append((r1, r2, [i.get_state()["RRt"] for i in integs]))
            batch.close()
        finally:
            os.environ.pop("PIMD_MD_GRAPH", None)
    (a1, a2, sa), (b1, b2, sb) = res
    for k in ("Upot", "uk"):
        assert np.array_equal(a1[k], b1[k]) and np.array_equal(a2[k], b2[k]), k
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)


def test_force_graph_matches_eager():
    """The all-beads force call replayed from a captured graph (the default)
    against eager launches (PIMD_FORCE_GRAPH=0), bit for bit: calls at several
    MD steps on moving coordinates (the dipole predictor reads the step and
    its history), then after a change of model parameters, for TTM3-F,
    TTM2.1-F and q-SPC/Fw."""
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    from pimd_amd.ttm import TTMForceField
    RRt, PPt, RRc, L = make_system(64, 8, seed=5)
    res = {}
    for graph in ("0", "1"):
        os.environ["PIMD_FORCE_GRAPH"] = graph
        try:
            for model in (3, 2, 4):
                print("case force graph", graph, "model", model, flush=True)
                p = TTMParams(pot_model=model, nwaters=64, nbeads=8, box=(L, L, L), rc=-1.0,
                              eps_ewald=1e-6, guess_initdip=True)
                ff = TTMForceField(p)
                out = []
                for t in range(1, 8):
                    R = RRt + 0.002 * t
                    out.append(ff.full_bead_forces(R, R.mean(axis=0), t=t))
                th = ff.ttm_params()
                th[2] *= 1.01
                ff.set_ttm_params(th)
                out.append(ff.full_bead_forces(RRt, RRc, t=8))
                ff.close()
                res[(graph, model)] = out
        finally:
            os.environ.pop("PIMD_FORCE_GRAPH", None)
    for model in (3, 2, 4):
        for ra, rb in zip(res[("0", model)], res[("1", model)]):
            for k in ("Upott", "dRRt", "virt_cm", "virialct", "dip_momIt", "chg", "iters"):
                assert np.array_equal(ra[k], rb[k]), (model, k)
