"""Batches of independent replicas run side by side (pimd_batch_*).

A batch must give each replica exactly the trajectory it has when run alone:
bit for bit with the default launch shapes, and to the FP64 parity bar after
the autotuner has changed them (the shapes only change summation order).
"""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

# (model, waters, beads, T): TTM3-F, TTM2.1-F and the fixed-charge q-SPC/Fw
# and SPC/F models side by side
REPLICAS = [(3, 32, 8, 250.0), (2, 64, 4, 300.0), (3, 128, 16, 350.0), (4, 48, 8, 300.0),
            (5, 32, 2, 280.0)]


def _make(model, N, P, T, seed):
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import make_system
    from pimd_amd.ttm import PIMDIntegrator, TTMForceField
    RRt, PPt, RRc, L = make_system(N, P, temp=T, seed=seed)
    p = TTMParams(pot_model=model, nwaters=N, nbeads=P, box=(L, L, L), guess_initdip=True)
    ff = TTMForceField(p)
    integ = PIMDIntegrator(ff, delt_fs=0.25, temp=T, bead_thermostat=1, centroid_thermostat=True)
    f0 = ff.full_bead_forces(RRt, RRc)
    integ.set_state(RRt, PPt, f0["dRRt"])
    return integ


def _fresh():
    return [_make(m, N, P, T, 900 + i) for i, (m, N, P, T) in enumerate(REPLICAS)]


SEEDS = [11, 22, 33, 44, 55]
NSTEPS = 6


@pytest.fixture(scope="module")
def alone():
    out = []
    for integ, sd in zip(_fresh(), SEEDS):
        out.append(integ.run(NSTEPS, t0=1, seed=sd)["Upot"])
    return np.array(out)


def test_batch_matches_replicas_run_alone(alone):
    from pimd_amd.ttm import ReplicaBatch
    batch = ReplicaBatch(_fresh())
    r = batch.run(NSTEPS, SEEDS, t0=1)
    print("batch Upot", r["Upot"][:, -1], flush=True)
    assert np.array_equal(r["Upot"], alone)


def test_autotuned_batch_matches_to_parity(alone):
    from pimd_amd.ttm import ReplicaBatch
    batch = ReplicaBatch(_fresh())
    ms = batch.autotune()
    print("autotuned shapes", batch.shapes(), "ms", ms, flush=True)
    assert np.all(ms > 0)
    r = batch.run(NSTEPS, SEEDS, t0=1)
    assert rel_err(r["Upot"], alone) < 1e-10
