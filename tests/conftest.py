import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "pimd-f90_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(ROOT, "tests", "golden")

# The library takes the once-per-pair charge pass only where its blocks fill
# the device; the tests take it wherever it applies (test_gpu_forces also
# runs the two-sided pass, PIMD_QQ_HALF=0, against the oracle).
os.environ.setdefault("PIMD_QQ_HALF", "2")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950)")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def golden_names(prefix):
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.startswith(prefix) and f.endswith(".npz"))


def params_from(g):
    from pimd_amd.params import TTMParams
    kw = {}
    for k, v in g.items():
        if k.startswith("p_"):
            key = k[2:]
            kw[key] = tuple(v.tolist()) if key == "box" else v.item()
    kw["guess_initdip"] = bool(kw["guess_initdip"])
    return TTMParams(**kw)


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    try:
        oracle.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    return oracle
