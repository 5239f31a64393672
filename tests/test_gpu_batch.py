"""Batches of independent replicas stepped as one (pimd_batch_*).

The batch is the 64-replica benchmark set (bench.replica_specs(): TTM2.1-F
and TTM3-F alternating, 32-512 waters, 8-64 beads, 250-350 K, Rc = L/2,
eps 1e-8, dipole predictor on).  Every kernel of a batch step is one launch
over all the replicas of a model, so each replica must keep exactly the
trajectory it has when run alone with the same launch shapes: the records
and final state of every replica are compared bit for bit with solo runs of
contexts whose shapes were unified by a batch of the same systems.  Two
replicas are also followed by the oracle on two beads after the first steps
(no predictor yet at t <= 4), and the autotuned batch is held to the FP64
parity bar (its shapes only change summation order).
"""
import os
import sys

import numpy as np
import pytest

from helpers import E_REL, F_ABS, max_abs, rel_err

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

NSTEPS = (3, 3)        # two batch calls: t = 1..3, then t = 4..6 (predictor at t > 4)


def _specs():
    import bench
    return bench.replica_specs()


def _make_all(specs):
    import bench
    return [bench.make_replica(sp) for sp in specs]


def _close(reps):
    for r in reps:
        r[1].close()


def _run_batch(specs):
    """Records of the two batch calls, the states after the first call and
    after the second, from one batch of fresh replicas."""
    from pimd_amd.ttm import ReplicaBatch
    reps = _make_all(specs)
    seeds = [sp["seed"] for sp in specs]
    batch = ReplicaBatch([r[0] for r in reps])
    shapes = batch.shapes()
    r1 = batch.run(NSTEPS[0], seeds, t0=1)
    st1 = [r[0].get_state() for r in reps]
    r2 = batch.run(NSTEPS[1], seeds, t0=1 + NSTEPS[0])
    st2 = [r[0].get_state() for r in reps]
    batch.close()
    _close(reps)
    return shapes, (r1, r2), st1, st2


@pytest.fixture(scope="module")
def batch_run():
    specs = _specs()
    print("case 64-replica batch,", sum(s["nwaters"] * s["nbeads"] for s in specs),
          "molecule-beads", flush=True)
    return specs, _run_batch(specs)


def test_batch_matches_each_replica_run_alone(batch_run):
    from pimd_amd.ttm import ReplicaBatch
    specs, (shapes, (r1, r2), _, st2) = batch_run
    reps = _make_all(specs)
    # the same shape unification as the batch above, then every replica alone
    b = ReplicaBatch([r[0] for r in reps])
    assert b.shapes() == shapes
    b.close()
    for i, (r, sp) in enumerate(zip(reps, specs)):
        a1 = r[0].run(NSTEPS[0], t0=1, seed=sp["seed"])
        a2 = r[0].run(NSTEPS[1], t0=1 + NSTEPS[0], seed=sp["seed"])
        st = r[0].get_state()
        for k in ("Upot", "uk"):
            assert np.array_equal(a1[k], r1[k][i]) and np.array_equal(a2[k], r2[k][i]), (i, sp, k)
        for k in ("RRt", "PPt", "dRRt", "RRc"):
            assert np.array_equal(st[k], st2[i][k]), (i, sp, k)
        r[1].close()
    print("64 replicas bit-identical to their solo runs over", sum(NSTEPS), "steps", flush=True)


def test_batch_replicas_vs_oracle(batch_run):
    """Gradient of two beads of the two smallest replicas after the first
    batch call (t = 3: the predictor starts at t = 5) against the oracle on
    the replica's own coordinates; the batch's energy of the last step against
    the oracle's sum over all beads."""
    import bench
    import oracle
    specs, (_, (r1, _), st1, _) = batch_run
    order = sorted(range(len(specs)), key=lambda i: specs[i]["nwaters"] * specs[i]["nbeads"])
    for i in order[:2]:
        sp, st = specs[i], st1[i]
        beads = [0, sp["nbeads"] - 1]
        o = oracle.System(bench.replica_params(sp).resolved()).forces(st["RRt"], st["RRc"], t=NSTEPS[0])
        e = {"dRR": max_abs(st["dRRt"][beads], o["dRR"][beads]),
             "Upot": rel_err(r1["Upot"][i, -1], o["En"].sum())}
        print("replica", i, sp, " ".join(f"{k}={v:.2e}" for k, v in e.items()), flush=True)
        assert e["dRR"] < F_ABS and e["Upot"] < E_REL


def test_autotuned_batch_matches_to_parity(batch_run):
    from pimd_amd.ttm import ReplicaBatch
    specs, (_, (r1, _), _, _) = batch_run
    reps = _make_all(specs)
    batch = ReplicaBatch([r[0] for r in reps])
    ms = batch.autotune()
    print("autotuned shapes", batch.shapes()[:4], "ms", ms[:4], flush=True)
    assert np.all(ms > 0)
    r = batch.run(NSTEPS[0], [sp["seed"] for sp in specs], t0=1)
    batch.close()
    _close(reps)
    assert rel_err(r["Upot"], r1["Upot"]) < 1e-10
