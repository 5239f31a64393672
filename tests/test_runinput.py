"""The run inputs against the reference's own input-file reader and checkpoint
writer/reader (fixtures in tests/golden/inputs/, made by running the
reference's read_inputfile, save_checkpoint and load_checkpoint).

The example inputs are the reference's own (input32.inp, inputdimer.inp:
positional old style; input_new_64.inp, test_input.inp, input4nm.inp: key and
value); quirks.inp exercises the reader's quirks: the unreachable massH key,
DIELECTRICOUT read into nothing, an unknown key, a tab-indented key, a
multi-item value, d exponents and unquoted strings.
"""
import json
import os

import numpy as np
import pytest

from pimd_amd import runinput

HERE = os.path.join(os.path.dirname(__file__), "golden", "inputs")
# inputs with the reference reader's record beside them (the drop-in program
# inputs in the same folder have none)
CASES = sorted(f[:-4] for f in os.listdir(HERE)
               if f.endswith(".inp") and os.path.exists(os.path.join(HERE, f[:-4] + ".json")))


def _typed(kind, text):
    if kind == "l":
        return text == "T"
    if kind == "i":
        return int(text)
    if kind in ("d", "r"):
        return float(text)
    return text


@pytest.mark.parametrize("name", CASES)
def test_input_file_matches_reference_reader(name):
    ref = json.load(open(os.path.join(HERE, name + ".json")))
    msgs = []
    got = runinput.read_input(os.path.join(HERE, name + ".inp"), warn=msgs.append)
    for key, (kind, _) in runinput.KEYS.items():
        assert got[key] == _typed(kind, ref[key]), key
    # one report per line the reference could not understand
    assert len(msgs) == sum(m.startswith("ERROR") for m in ref["messages"])


def test_quirks_as_the_reference():
    got = runinput.read_input(os.path.join(HERE, "quirks.inp"), warn=lambda m: None)
    assert got["massH"] == np.float32(1.008)       # the massH key is never matched
    assert got["massO"] == 15.995
    assert got["DIELECTRICOUT"] is True           # read into nothing
    assert got["tau"] == np.float32(0.1)          # tab-indented key not matched
    assert got["Rc"] == 4.9 and got["polar_eps"] == 1e-8
    assert got["PIMD_type"] == "contracted" and got["bead_thermostat_type"] == "Langevin"


def test_missing_required_key(tmp_path):
    f = tmp_path / "x.inp"
    f.write_text("Nbeads 2\npot_model 3\n")
    with pytest.raises(runinput.InputError, match="fconfig"):
        runinput.read_input(str(f), warn=lambda m: None)


def test_checkpoint_written_as_the_reference(tmp_path):
    g = np.load(os.path.join(HERE, "beads_n8_p3.npz"))
    out = tmp_path / "x.img"
    runinput.write_image(str(out), g["RRt"], g["PPt"], g["box"], int(g["t"]), float(g["delt_ps"]),
                         float(g["mass_o"]), float(g["mass_h"]), init_time=float(g["init_time"]))
    assert out.read_bytes() == open(os.path.join(HERE, "beads_n8_p3.img"), "rb").read()


def test_checkpoint_read_as_the_reference():
    g = np.load(os.path.join(HERE, "beads_n8_p3.npz"))
    R, Pm, Rc, box, t0, vxi = runinput.read_image(os.path.join(HERE, "beads_n8_p3.img"),
                                                  float(g["mass_o"]), float(g["mass_h"]), 3)
    assert np.array_equal(R, g["RRt_loaded"])
    assert np.array_equal(Pm, g["PPt_loaded"])
    assert np.array_equal(Rc, g["RRc_loaded"])
    assert np.allclose(box, g["box"], atol=5e-7) and t0 == 0.5 + 40 * 0.25e-3
    with pytest.raises(runinput.InputError):
        runinput.read_image(os.path.join(HERE, "beads_n8_p3.img"), 16.0, 1.0, 4)


def test_run_setup_from_reference_input():
    inp = runinput.read_input(os.path.join(HERE, "quirks.inp"), warn=lambda m: None)
    inp["Nbeads"] = 3
    p, md = runinput.run_setup(inp, (9.865, 9.865, 9.865), 96)
    assert p.pot_model == 3 and p.nwaters == 32 and p.nbeads == 3
    assert p.rc == 4.9 and p.rc1 == np.float32(0.8) * 4.9     # rc1 default 0.8 Rc
    assert p.mass_h == np.float32(1.008) and p.polar_sor == 0.6
    assert md["bead_thermostat"] == 1 and md["intra_timesteps"] == 4 and md["delt_fs"] == 0.25
    # the normal-mode Nose-Hoover bead thermostat (main_stuff.f90:145-171)
    inp["bead_thermostat_type"] = "Nose-Hoover"
    assert runinput.run_setup(inp, (9.865, 9.865, 9.865), 96)[1]["bead_thermostat"] == 2
    inp["bead_thermostat_type"] = "Andersen"
    with pytest.raises(runinput.InputError, match="Andersen"):
        runinput.run_setup(inp, (9.865, 9.865, 9.865), 96)
    inp2 = runinput.read_input(os.path.join(HERE, "input_new_64.inp"), warn=lambda m: None)
    with pytest.raises(runinput.InputError, match="pot_model 6"):
        runinput.run_setup(inp2, (12.4, 12.4, 12.4), 192)


def test_run_setup_fixed_charge_models():
    """pot_model 4 (q-SPC/Fw) and 5 (SPC/F) set up as the other models;
    6 (SIESTA) is refused."""
    for model in (4, 5):
        inp = runinput.read_input(os.path.join(HERE, "quirks.inp"), warn=lambda m: None)
        inp["pot_model"] = model
        inp["Nbeads"] = 2
        p, md = runinput.run_setup(inp, (9.865, 9.865, 9.865), 96)
        assert p.pot_model == model and p.nwaters == 32 and p.nbeads == 2
