"""The reference's run inputs: the key/value input file and the bead
configuration images (.img checkpoints), read and written the way the
reference does, so a run prepared for PIMD-F90 starts here unchanged.

read_input   read_inputfile (input_file_reader.f90:30-352): defaults, comment
             and blank lines, one key per line followed by a list-directed
             value, the required keys; a line 'old' switches to the
             positional format of read_old_style_input_file
             (input_file_reader.f90:359-431), in which most of the
             reference's own input files are written.  Kept as the reference behaves: the
             key 'massH' is never matched (the reference's case label is
             ' massH'), 'DIELECTRICOUT' consumes its value without setting
             anything, an unknown key is reported and skipped, and real
             defaults written as single-precision literals keep their
             single-precision values (polar_sor 0.7, tau 0.1, massO 15.994 ...).
read_image   the header read of read_and_initialize_all_nodes
             (InputOutput.f90:77-93) and load_checkpoint
             (InputOutput.f90:1031-1074): velocities on file, momenta in
             memory (p = v m), centroids and the centroid wrap of PBCs
             (main_stuff.f90:107-125).
write_image  save_checkpoint (InputOutput.f90:987-1027).
run_setup    TTMParams and PIMDIntegrator arguments from an input dict and
             the box (InputOutput.f90:118-145: cut-off defaults, fs -> ps).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .params import TTMParams, f32

# key -> (type, default): i integer, d double, r double set from a
# single-precision default literal, l logical, s string (input_file_reader.f90:43-123)
KEYS = {
    "coord_out": ("l", False), "momenta_out": ("l", False), "dip_out": ("l", False),
    "Edip_out": ("l", False), "TD_out": ("l", False), "images_out": ("l", False),
    "IMAGEDIPOLESOUT": ("l", False), "BOXSIZEOUT": ("l", False), "TP_out": ("l", False),
    "ENERGYOUT": ("l", True), "HISTOUT": ("l", True), "hist_out_freq": ("i", 1000),
    "CALCGEOMETRY": ("l", True), "CALCDIFFUSION": ("l", True), "read_method": ("i", 1),
    "CALCIRSPECTRA": ("l", False), "CALCDOS": ("l", True), "CALC_RADIUS_GYRATION": ("l", True),
    "DIELECTRICOUT": ("l", True), "CHARGESOUT": ("l", False), "WRITECHECKPOINTS": ("l", True),
    "td_freq": ("i", 1), "tp_freq": ("i", 10), "ti_freq": ("i", 2000), "t_freq": ("i", 10),
    "checkpoint_freq": ("i", 2000), "siesta_name": ("s", "siesta"), "num_SIESTA_nodes": ("i", 1),
    "Rc": ("d", -1.0), "rc1": ("d", -1.0), "eps_ewald": ("d", 1e-6), "polar_maxiter": ("i", 15),
    "polar_sor": ("r", 0.7), "polar_eps": ("d", 1e-5), "guess_initdip": ("l", True),
    "print_dipiters": ("l", False), "GENVEL": ("l", True), "THERMOSTAT": ("l", False),
    "BEADTHERMOSTAT": ("l", False), "CENTROIDTHERMOSTAT": ("l", False),
    "bead_thermostat_type": ("s", "none"), "tau": ("r", 0.1), "tau_centroid": ("r", 0.1),
    "global_chain_length": ("i", 2), "bead_chain_length": ("i", 2), "temp": ("d", 300.0),
    "BAROSTAT": ("l", False), "tau_P": ("r", 0.2), "press": ("d", 1.0), "PEQUIL": ("l", False),
    "setNMfreq": ("d", 0.0), "PIMD_type": ("s", "full"), "intra_timesteps": ("i", 10),
    "massO": ("r", 15.994), "massH": ("r", 1.008), "RESTART": ("l", False),
    "SIESTA_MON_CALC": ("l", False), "Nbeads": ("i", 1), "fconfig": ("s", "-1"),
    "fsave": ("s", "-1"), "eq_timesteps": ("i", -1), "run_timesteps": ("i", -1),
    "delt": ("d", -1.0), "pot_model": ("i", -1), "sys_label": ("s", "-1"),
    "mon_siesta_name": ("s", "-1"),
}
# keys the reference's select case cannot match or reads into nothing
UNMATCHED = {"massH"}
NO_TARGET = {"DIELECTRICOUT"}
REQUIRED = ("fconfig", "Nbeads", "fsave", "eq_timesteps", "run_timesteps", "delt", "pot_model")
# character lengths of the string variables (main_stuff.f90:36,54,95)
STRLEN = {"bead_thermostat_type": 11, "fconfig": 250, "fsave": 250, "siesta_name": 200,
          "PIMD_type": 200, "mon_siesta_name": 200, "sys_label": 200}


class InputError(ValueError):
    pass


def _first_item(text: str) -> str:
    """The first item of a list-directed read: a quoted string, or the text up
    to the first blank, comma or slash."""
    t = text.lstrip()
    if t[:1] in ("'", '"'):
        q = t[0]
        end = t.find(q, 1)
        return t[1:end] if end > 0 else t[1:]
    out = []
    for ch in t:
        if ch in " ,/\t":
            break
        out.append(ch)
    return "".join(out)


def _logical(item: str) -> bool:
    s = item.lstrip(".").upper()
    if s[:1] == "T":
        return True
    if s[:1] == "F":
        return False
    raise InputError(f"not a logical value: {item!r}")


def _real(item: str) -> float:
    return float(item.replace("d", "e").replace("D", "e"))


def _convert(kind: str, item: str):
    if kind == "i":
        try:
            return int(item)
        except ValueError:
            raise InputError(f"integer value absent: {item!r}") from None
    if kind in ("d", "r"):
        return _real(item)
    if kind == "l":
        return _logical(item)
    return item


def defaults() -> dict:
    out = {}
    for k, (kind, v) in KEYS.items():
        out[k] = f32(v) if kind == "r" else v
    return out


def _tokens(line: str):
    """Items of a list-directed record: quoted strings, or runs up to a blank
    or comma; a slash ends the record's items."""
    out, i, n = [], 0, len(line)
    while i < n:
        ch = line[i]
        if ch in " \t,":
            i += 1
            continue
        if ch == "/":
            out.append(None)
            break
        if ch in ("'", '"'):
            j = line.find(ch, i + 1)
            j = n if j < 0 else j
            out.append(line[i + 1:j])
            i = j + 1
            continue
        j = i
        while j < n and line[j] not in " \t,/":
            j += 1
        out.append(line[i:j])
        i = j
    return out


# read_old_style_input_file (input_file_reader.f90:359-431): one read
# statement per entry; None skips a record
OLD_STYLE = [None, None, None, ("fconfig",), ("fsave",), ("Nbeads",), ("eq_timesteps",),
             ("run_timesteps",), ("delt",), None, None, ("coord_out",), ("momenta_out",),
             ("dip_out",), ("Edip_out",), ("TD_out",), ("images_out",), ("IMAGEDIPOLESOUT",),
             ("BOXSIZEOUT",), ("TP_out",), ("CALC_RADIUS_GYRATION",), ("DIELECTRICOUT",),
             ("CHARGESOUT",), ("WRITECHECKPOINTS",), ("td_freq",), ("tp_freq",), ("ti_freq",),
             ("t_freq",), None, None, ("pot_model",),
             ("siesta_name", "sys_label", "num_SIESTA_nodes"), ("Rc", "rc1", "eps_ewald"),
             ("polar_maxiter", "polar_sor", "polar_eps", "guess_initdip", "print_dipiters"),
             ("GENVEL",), ("THERMOSTAT",), ("BEADTHERMOSTAT",), ("CENTROIDTHERMOSTAT",),
             ("bead_thermostat_type",), ("tau",), ("tau_centroid",), ("global_chain_length",),
             ("bead_chain_length",), ("temp",), ("BAROSTAT",), ("tau_P",), ("press",),
             ("PEQUIL",), None, None, ("setNMfreq",), ("PIMD_type",), ("intra_timesteps",),
             ("massO",), ("massH",), ("RESTART",)]


def _read_old_style(lines, row, vals):
    """Positional reads from record row on; returns the next record."""
    for entry in OLD_STYLE:
        if entry is None:
            row += 1
            continue
        items = []
        while len(items) < len(entry):
            if row >= len(lines):
                raise InputError("end of file inside the old-style input")
            toks = _tokens(lines[row])
            row += 1
            if None in toks:
                items.extend(toks[:toks.index(None)])
                break
            items.extend(toks)
        for key, item in zip(entry, items):
            kind = KEYS[key][0]
            v = _convert(kind, item)
            if kind == "s" and key in STRLEN:
                v = v[:STRLEN[key]]
            vals[key] = v
    vals["checkpoint_freq"] = vals["t_freq"]
    return row


def read_input(path: str, warn=print) -> dict:
    """The values read_inputfile leaves in the reference's variables."""
    vals = defaults()
    with open(path) as f:
        lines = f.read().splitlines()
    row = 0
    while row < len(lines):
        line = lines[row].lstrip(" ")   # adjustl: blanks only, a tab stays
        row += 1
        if line[:1] in ("!", "#", ";") or line.strip() == "":
            continue
        pos = line.find(" ")
        # the line variable is blank-padded, so a lone word is a key
        key, buf = (line[:pos], line[pos + 1:]) if pos >= 0 else (line, "")
        if key in ("old_style", "OLD_STYLE", "OLD", "old"):
            # rewind, skip the first record, read the positional entries
            row = _read_old_style(lines, 1, vals)
            continue
        if key in NO_TARGET:
            continue
        if key not in KEYS or key in UNMATCHED:
            warn(f"ERROR: could not understand the following line: {buf}")
            continue
        kind = KEYS[key][0]
        item = _first_item(buf)
        v = _convert(kind, item)
        if kind == "s" and key in STRLEN:
            v = v[:STRLEN[key]]
        vals[key] = v
    missing = [k for k in REQUIRED if vals[k] in ("-1", -1, -1.0)]
    if missing:
        raise InputError(f"required input variable(s) missing: {', '.join(missing)}")
    return vals


def read_image(path: str, mass_o: float, mass_h: float, nbeads: int | None = None,
               nh_global: bool = False):
    """A bead configuration image: (RRt[P][Natoms][3], PPt[P][Natoms][3],
    RRc[Natoms][3], box[3], init_time, vxi_global or None), the beads moved
    with their centroids into the box as load_checkpoint leaves them."""
    with open(path) as f:
        lines = f.read().splitlines()
    natoms, nb = int(lines[0][0:10]), int(lines[0][10:20])
    if nbeads is not None and nb != nbeads:
        raise InputError(f"the image holds {nb} beads, the input asks for {nbeads}")
    hdr = lines[1]
    init_time = float(hdr[0:12])
    box = np.array([float(hdr[14:26]), float(hdr[27:39]), float(hdr[40:52])])
    nw = natoms // 3
    R = np.zeros((nb, natoms, 3))
    Pm = np.zeros((nb, natoms, 3))
    row = 2
    for w in range(nw):
        for a, m in ((3 * w, mass_o), (3 * w + 1, mass_h), (3 * w + 2, mass_h)):
            for j in range(nb):
                ln = lines[row]
                row += 1
                v = [float(ln[2 + 13 * i + 1: 2 + 13 * (i + 1)]) for i in range(6)]
                R[j, a] = v[:3]
                Pm[j, a] = np.array(v[3:]) * m
    vxi = None
    if nh_global:
        vxi = np.array([float(x.replace("D", "E")) for x in lines[row].split()])
    Rc = R.sum(axis=0) / nb
    shifts = box[0] * np.rint(Rc * (1.0 / box[0]))   # box(1) on every axis
    return R - shifts[None], Pm, Rc - shifts, box, init_time, vxi


def _f12(x: float) -> str:
    s = f"{x:12.6f}"
    return s if len(s) == 12 else "*" * 12


def write_image(path: str, RRt, PPt, box, t: int, delt_ps: float, mass_o: float, mass_h: float,
                init_time: float = 0.0, vxi_global=None):
    """save_checkpoint: header, then per water the O, H1, H2 lines of every
    bead with coordinates and velocities (momentum / mass)."""
    RRt, PPt = np.asarray(RRt), np.asarray(PPt)
    nb, natoms = RRt.shape[:2]
    io, ih = 1.0 / mass_o, 1.0 / mass_h
    out = [f"{natoms:10d}{nb:10d}",
           _f12(t * delt_ps + init_time) + "  " + " ".join(_f12(b) for b in box)]
    for w in range(natoms // 3):
        for a, im, sym in ((3 * w, io, "O "), (3 * w + 1, ih, "H "), (3 * w + 2, ih, "H ")):
            for j in range(nb):
                vals = list(RRt[j, a]) + list(PPt[j, a] * im)
                out.append(sym + "".join(" " + _f12(v) for v in vals))
    if vxi_global is not None:
        out.append(" " + "  ".join(repr(float(v)) for v in vxi_global))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def run_setup(inp: dict, box, natoms: int):
    """(TTMParams, integrator keyword arguments) of a run described by an
    input dict, for pimd_amd.ttm.TTMForceField and PIMDIntegrator."""
    model = inp["pot_model"]
    if model not in (2, 3, 4, 5):
        raise InputError(f"pot_model {model} is not built (TTM2.1-F = 2, TTM3-F = 3, "
                         "q-SPC/Fw = 4 and SPC/F = 5 are)")
    if inp["BAROSTAT"]:
        raise InputError("the barostat is not built")
    bt = inp["bead_thermostat_type"].strip()
    # InputOutput.f90:251-256 admits these three; main_stuff.f90:140,145
    if inp["BEADTHERMOSTAT"] and bt not in ("Langevin", "Nose-Hoover", "none"):
        raise InputError(f"bead thermostat {bt!r} is not one of 'Nose-Hoover', 'Langevin', 'none'")
    ptype = inp["PIMD_type"].strip()
    if ptype not in ("full", "contracted"):
        raise InputError(f"PIMD_type {ptype!r} is not built (full and contracted are)")
    box = tuple(float(b) for b in box)
    rc = inp["Rc"] if inp["Rc"] != -1 else min(box) / 2.0
    rc1 = inp["rc1"] if inp["rc1"] != -1 else f32(0.8) * rc
    p = TTMParams(pot_model=model, nwaters=natoms // 3, nbeads=inp["Nbeads"], box=box, rc=rc,
                  rc1=rc1, eps_ewald=inp["eps_ewald"], polar_maxiter=inp["polar_maxiter"],
                  polar_sor=inp["polar_sor"], polar_eps=inp["polar_eps"],
                  guess_initdip=inp["guess_initdip"], mass_o=inp["massO"], mass_h=inp["massH"])
    md = dict(delt_fs=inp["delt"], temp=inp["temp"],
              bead_thermostat=({"Langevin": 1, "Nose-Hoover": 2}.get(bt, 0)
                               if inp["BEADTHERMOSTAT"] else 0),
              centroid_thermostat=bool(inp["CENTROIDTHERMOSTAT"]),
              tau_centroid=inp["tau_centroid"], global_nh=bool(inp["THERMOSTAT"]), tau=inp["tau"],
              setNMfreq=inp["setNMfreq"],
              intra_timesteps=inp["intra_timesteps"] if ptype == "contracted" else 0)
    return p, md


def replace(p: TTMParams, **kw) -> TTMParams:
    return dataclasses.replace(p, **kw)
