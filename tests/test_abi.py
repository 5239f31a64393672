"""The C-ABI library loads and exports every entry point include/pimd_ttm.h
declares; host-only entry points and argument checks work without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_golden

HEADER = os.path.join(ROOT, "include", "pimd_ttm.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*int\s+(pimd_\w+)\s*\(", txt, re.M)))


@pytest.fixture(scope="module")
def lib():
    from pimd_amd import ttm
    if not os.path.exists(ttm.LIB_PATH):
        pytest.fail(f"{ttm.LIB_PATH} missing: run `make -C pimd-f90_amd`")
    return ttm.load_library()


def test_exports_every_declared_symbol(lib):
    from pimd_amd import ttm
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(ttm.EXPORTS) == syms


def test_nose_hoover_host_entry_matches_reference(lib):
    from pimd_amd.ttm import nose_hoover
    for row in load_golden("nose_hoover")["table"]:
        uk, v0, v1, tau, L, T, s, uk1, w0, w1 = row
        r = nose_hoover(uk, (v0, v1), tau, 0.25 / 2000, int(L), T)
        assert abs(r["s"] - s) < 1e-15 and abs(r["uk"] - uk1) < 1e-12 * abs(uk1)
        assert np.max(np.abs(r["vxi"] - [w0, w1])) < 1e-14


def test_invalid_parameters_are_rejected(lib):
    from pimd_amd import ttm
    from pimd_amd.params import TTMParams
    bad = [TTMParams(pot_model=4, nwaters=8, nbeads=2, box=(8, 8, 8)),
           TTMParams(pot_model=3, nwaters=8, nbeads=0, box=(8, 8, 8)),
           TTMParams(pot_model=3, nwaters=8, nbeads=2, box=(8, 8, 8), rc=4.5),
           TTMParams(pot_model=3, nwaters=8, nbeads=2, box=(8, 8, 8), eps_ewald=2.0)]
    for p in bad:
        with pytest.raises(ttm.PimdError):
            ttm.TTMForceField(p)


def test_null_and_order_errors(lib):
    assert lib.pimd_ctx_create(None, None) == 1
    buf = C.create_string_buffer(256)
    assert lib.pimd_last_error(buf, 256) == 0 and b"null" in buf.value
    assert lib.pimd_nm_get(None, *([None] * 7)) == 1
