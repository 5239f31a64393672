"""Python host mirror of the reference's force and ring-polymer interface.

Thin ctypes layer over the C ABI in include/pimd_ttm.h (libpimd_ttm.so, built
in-tree by `make -C pimd-f90_amd`).  Names follow the reference:

  TTMForceField.full_bead_forces   force_calc.f90:26-106 (all P beads at once)
  TTMForceField.potential          force_calc.f90:112-170 / pot_ttm.f90:3 (one bead)
  RingPolymer.init_normal_modes    NormalModes.f90:68-218
  RingPolymer.evolve_ring          NormalModes.f90:34-63, every atom in one call
  RingPolymer.init_langevin_nm     Langevin.f90:68-98
  RingPolymer.langevin_nm          Langevin.f90:30-64
  nose_hoover                      Nose_Hoover.f90:33-112
  PIMDIntegrator                   MD.f90:35-110 (PIMD_VV), device resident

There is no CPU fallback: a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .params import TTMParams

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpimd_ttm.so")

# Every symbol include/pimd_ttm.h declares.
EXPORTS = (
    "pimd_ctx_create", "pimd_ctx_destroy", "pimd_last_error", "pimd_ctx_info",
    "pimd_forces_all_beads", "pimd_forces_all_beads_dev", "pimd_energy_terms",
    "pimd_nm_init", "pimd_nm_get", "pimd_nm_evolve", "pimd_langevin_init", "pimd_langevin",
    "pimd_nose_hoover", "pimd_md_init", "pimd_state_set", "pimd_state_get", "pimd_md_run",
    "pimd_profile_enable", "pimd_profile_count", "pimd_profile_get", "pimd_profile_reset",
    "pimd_sync", "pimd_batch_create", "pimd_batch_destroy", "pimd_batch_autotune",
    "pimd_batch_shape", "pimd_batch_md_run", "pimd_pes_params_get", "pimd_pes_params_set",
    "pimd_pes_param_grad", "pimd_state_forces", "pimd_profile_copy", "pimd_ttm_params_get",
    "pimd_ttm_params_set", "pimd_ttm_param_grad", "pimd_md_observables",
    "pimd_forces_all_beads_noalloc", "pimd_monomer_dipoles", "pimd_dms_coefs_get",
    "pimd_dms_coefs_set", "pimd_dms_coef_grad", "pimd_gamma_m_get", "pimd_gamma_m_set",
    "pimd_gamma_m_grad", "pimd_ttm_adjoint", "pimd_ttm_adjoint_start", "pimd_ttm_adjoint_finish",
    "pimd_forces_all_beads_start", "pimd_forces_all_beads_finish",
)

PES_NPARAM = 250   # combined c5z (245), deohA, alphaoh, roh, phh1A, phh2
# vdwA..vdwE, polarO, polarH, polarM, polfacO, polfacH, polfacM, aDD, aCCaCD, dms_param1..3
TTM_PARAM_NAMES = ("vdwA", "vdwB", "vdwC", "vdwD", "vdwE", "polarO", "polarH", "polarM",
                   "polfacO", "polfacH", "polfacM", "aDD", "aCCaCD", "dms_param1", "dms_param2",
                   "dms_param3")
TTM_NPARAM = len(TTM_PARAM_NAMES)
DMS_NCOEF = 84     # coefD of the dipole moment surface (nasa_mod.f90)
# the adjoint's parameters: TTM parameters, DMS coefficients, gammaM
ADJ_NPARAM = TTM_NPARAM + DMS_NCOEF + 1
# per-step observables of write_out (pimd_md_observables)
OBS_NAMES = ("sys_temp", "virial_energy", "virial_press", "simple_energy", "simple_press",
             "dip_x", "dip_y", "dip_z", "dip_mag_sum")


class PimdParams(C.Structure):
    _fields_ = [("pot_model", C.c_int), ("nwaters", C.c_int), ("nbeads", C.c_int),
                ("polar_maxiter", C.c_int), ("guess_initdip", C.c_int), ("precision", C.c_int),
                ("device", C.c_int), ("contraction", C.c_int), ("box", C.c_double * 3),
                ("rc", C.c_double), ("rc1", C.c_double), ("eps_ewald", C.c_double),
                ("polar_sor", C.c_double), ("polar_eps", C.c_double), ("mass_o", C.c_double),
                ("mass_h", C.c_double)]


class MDParams(C.Structure):
    _fields_ = [("delt_fs", C.c_double), ("temp", C.c_double), ("bead_thermostat", C.c_int),
                ("centroid_thermostat", C.c_int), ("tau_centroid", C.c_double),
                ("global_nh", C.c_int), ("intra_timesteps", C.c_int), ("tau", C.c_double),
                ("setNMfreq", C.c_double)]


class PimdError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libpimd_ttm.so and declares the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PimdError(f"{path} is missing: build it with `make -C pimd-f90_amd`")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    sig = {
        "pimd_ctx_create": [C.POINTER(PimdParams), C.POINTER(C.c_void_p)],
        "pimd_ctx_destroy": [vp],
        "pimd_last_error": [C.c_char_p, C.c_size_t],
        "pimd_ctx_info": [vp, dp],
        "pimd_forces_all_beads": [vp, dp, dp, C.c_int, dp, dp, dp, dp, dp, dp, dp, ip],
        "pimd_forces_all_beads_dev": [vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp],
        "pimd_energy_terms": [vp, dp],
        "pimd_monomer_dipoles": [vp, dp, dp],
        "pimd_nm_init": [vp, C.c_double, C.c_double, C.c_double],
        "pimd_nm_get": [vp, dp, dp, dp, dp, dp, dp, dp],
        "pimd_nm_evolve": [vp, dp, dp, dp],
        "pimd_langevin_init": [vp, C.c_double, C.c_int, C.c_double, C.c_double],
        "pimd_langevin": [vp, dp, dp, C.c_uint64, C.c_uint64],
        "pimd_nose_hoover": [dp, dp, dp, C.c_double, C.c_double, C.c_int, C.c_double],
        "pimd_md_init": [vp, C.POINTER(MDParams)],
        "pimd_state_set": [vp, dp, dp, dp],
        "pimd_state_get": [vp, dp, dp, dp, dp],
        "pimd_md_run": [vp, C.c_int, C.c_int, dp, C.c_uint64, dp, dp, dp, dp],
        "pimd_profile_enable": [vp, C.c_int],
        "pimd_profile_count": [vp],
        "pimd_profile_get": [vp, C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_long), dp],
        "pimd_profile_reset": [vp],
        "pimd_sync": [vp],
        "pimd_batch_create": [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)],
        "pimd_batch_destroy": [vp],
        "pimd_batch_autotune": [vp, dp],
        "pimd_batch_shape": [vp, C.c_int, ip],
        "pimd_batch_md_run": [vp, C.c_int, C.c_int, C.POINTER(C.c_uint64), dp, dp],
        "pimd_pes_params_get": [vp, dp],
        "pimd_pes_params_set": [vp, dp],
        "pimd_pes_param_grad": [vp, dp, dp, dp, dp],
        "pimd_ttm_params_get": [vp, dp],
        "pimd_ttm_params_set": [vp, dp],
        "pimd_ttm_param_grad": [vp, dp, dp, dp, dp, dp, dp],
        "pimd_dms_coefs_get": [vp, dp],
        "pimd_dms_coefs_set": [vp, dp],
        "pimd_dms_coef_grad": [vp, dp, dp, dp, dp, dp, dp],
        "pimd_gamma_m_get": [vp, dp],
        "pimd_gamma_m_set": [vp, C.c_double],
        "pimd_gamma_m_grad": [vp, dp, dp, dp, dp, dp, dp],
        "pimd_ttm_adjoint": [vp, dp, dp, dp, C.c_int, dp, dp, dp],
        "pimd_ttm_adjoint_start": [vp, dp, dp, dp, C.c_int],
        "pimd_forces_all_beads_start": [vp, dp, dp, C.c_int],
        "pimd_forces_all_beads_finish": [vp, dp, dp, dp, dp, dp, dp, dp, C.POINTER(C.c_int)],
        "pimd_ttm_adjoint_finish": [vp, dp, dp, dp],
        "pimd_md_observables": [vp, C.c_int, dp],
        "pimd_state_forces": [vp, C.c_int],
        "pimd_profile_copy": [vp, C.c_size_t, dp],
        "pimd_forces_all_beads_noalloc": [C.POINTER(PimdParams), vp, vp, C.c_int, vp, vp, vp, vp,
                                          vp, vp, vp, vp, vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        buf = C.create_string_buffer(512)
        load_library().pimd_last_error(buf, 512)
        raise PimdError(f"pimd error {rc}: {buf.value.decode()}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _arr(x, shape):
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
    return a


def c_params(p: TTMParams, device: int = 0) -> PimdParams:
    """The C ABI's pimd_params of resolved run parameters."""
    return PimdParams(p.pot_model, p.nwaters, p.nbeads, p.polar_maxiter, int(p.guess_initdip),
                      int(p.precision), device, int(p.contraction), (C.c_double * 3)(*p.box), p.rc,
                      p.rc1, p.eps_ewald, p.polar_sor, p.polar_eps, p.mass_o, p.mass_h)


def forces_noalloc(params: TTMParams, d_RRt, d_RRc, t: int = 1, outs=None, stream=None,
                   device: int = 0):
    """pimd_forces_all_beads_noalloc on device addresses (ints): no context and no
    device allocation.  outs: dict of device addresses for Upott, dRRt, virt,
    virialct, dip_momIt, dip_momEt, chg, iters (missing ones stay in the
    library's workspace).  stream: a hipStream_t handle as an int, or None for
    the default stream.  Enqueued only: synchronise before reading."""
    o = outs or {}
    cp = c_params(params.resolved(), device)
    keys = ("Upott", "dRRt", "virt", "virialct", "dip_momIt", "dip_momEt", "chg", "iters")
    _check(load_library().pimd_forces_all_beads_noalloc(
        C.byref(cp), d_RRt, d_RRc, int(t), *[o.get(k) for k in keys], stream))


class TTMForceField:
    """All-beads TTM2.1-F / TTM3-F force evaluation on one GPU."""

    def __init__(self, params: TTMParams, device: int = 0):
        p = params.resolved()
        self.params = p
        L = load_library()
        cp = c_params(p, device)
        h = C.c_void_p()
        _check(L.pimd_ctx_create(C.byref(cp), C.byref(h)))
        self._h = h
        self._single = None

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().pimd_ctx_destroy(self._h)
            self._h = C.c_void_p()
        if self._single is not None:
            self._single.close()
            self._single = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        o = np.zeros(13)
        _check(load_library().pimd_ctx_info(self._h, _ptr(o)))
        return {"aewald": o[0], "Uvdw_lrc": o[1], "shiftA": o[2], "shiftB": o[3],
                "polfac": o[4:8].copy(), "nkvec": int(o[8]), "hmax": int(o[9]),
                "kmax": int(o[10]), "lmax": int(o[11]), "cells": int(o[12])}

    def full_bead_forces(self, RRt, RRc, t: int = 1) -> dict:
        """pot_ttm for every bead; returns Upott, dRRt, virt, virialct, dipoles, chg, iters.

        virt is returned as [P][row][col]; the C ABI's column-major layout is transposed
        here for convenience.  iters < 0 marks beads that reached polar_maxiter.
        """
        p = self.params
        P, na = p.nbeads, p.natoms
        RRt = _arr(RRt, (P, na, 3))
        RRc = _arr(RRc, (na, 3))
        out = self._force_outputs()
        _check(load_library().pimd_forces_all_beads(self._h, _ptr(RRt), _ptr(RRc), int(t),
                                                    *self._force_ptrs(out)))
        return self._force_done(out)

    def full_bead_forces_start(self, RRt, RRc, t: int = 1):
        """full_bead_forces in two halves: this queues the call on the
        context's stream and returns; full_bead_forces_finish waits and
        returns its dict.  Calls of several force fields run side by side."""
        P, na = self.params.nbeads, self.params.natoms
        _check(load_library().pimd_forces_all_beads_start(
            self._h, _ptr(_arr(RRt, (P, na, 3))), _ptr(_arr(RRc, (na, 3))), int(t)))

    def full_bead_forces_finish(self) -> dict:
        out = self._force_outputs()
        _check(load_library().pimd_forces_all_beads_finish(self._h, *self._force_ptrs(out)))
        return self._force_done(out)

    def _force_outputs(self) -> dict:
        P, na, N = self.params.nbeads, self.params.natoms, self.params.nwaters
        return {"Upott": np.zeros(P), "dRRt": np.zeros((P, na, 3)), "virt_cm": np.zeros((P, 3, 3)),
                "virialct": np.zeros(P), "dip_momIt": np.zeros((P, N, 3)),
                "dip_momEt": np.zeros((P, N, 3)), "chg": np.zeros((P, na)),
                "iters": np.zeros(P, dtype=np.int32)}

    @staticmethod
    def _force_ptrs(out: dict):
        return (_ptr(out["Upott"]), _ptr(out["dRRt"]), _ptr(out["virt_cm"]), _ptr(out["virialct"]),
                _ptr(out["dip_momIt"]), _ptr(out["dip_momEt"]), _ptr(out["chg"]),
                out["iters"].ctypes.data_as(C.POINTER(C.c_int)))

    @staticmethod
    def _force_done(out: dict) -> dict:
        out["virt"] = np.transpose(out["virt_cm"], (0, 2, 1)).copy()
        out["virialt"] = np.trace(out["virt"], axis1=1, axis2=2)
        return out

    def profile_copy(self, nbytes: int) -> float:
        """A device copy of nbytes (known HBM traffic, for counter calibration); ms."""
        ms = C.c_double(0.0)
        _check(load_library().pimd_profile_copy(self._h, int(nbytes), C.byref(ms)))
        return ms.value

    def energy_terms(self) -> np.ndarray:
        """[P][5]: Umon, Uvdw, Uelec, Uind, Uvdw_lrc of the last call."""
        t = np.zeros((self.params.nbeads, 5))
        _check(load_library().pimd_energy_terms(self._h, _ptr(t)))
        return t

    def monomer_dipoles(self, RRt) -> np.ndarray:
        """[P][N][3] monomer dipole moments of every bead from the dipole-moment-
        surface charges (calc_monomer_dip_moments, force_calc.f90:205-264)."""
        p = self.params
        R = _arr(RRt, (p.nbeads, 3 * p.nwaters, 3))
        out = np.zeros((p.nbeads, p.nwaters, 3))
        _check(load_library().pimd_monomer_dipoles(self._h, _ptr(R), _ptr(out)))
        return out

    def pes_params(self) -> np.ndarray:
        """The monomer PES parameters in use (combined c5z, deohA, alphaoh, roh, phh1A, phh2)."""
        th = np.zeros(PES_NPARAM)
        _check(load_library().pimd_pes_params_get(self._h, _ptr(th)))
        return th

    def set_pes_params(self, theta):
        th = _arr(theta, (PES_NPARAM,))
        _check(load_library().pimd_pes_params_set(self._h, _ptr(th)))
        if self._single is not None:
            self._single.set_pes_params(th)

    def pes_param_grad(self, RRt, ew, gw) -> np.ndarray:
        """Adjoint of the PES parameters: sum_b ew_b dE_b/dtheta + sum gw . d(dE/dr)/dtheta,
        for configurations RRt[P][Natoms][3] (see pimd_pes_param_grad)."""
        P, na = self.params.nbeads, self.params.natoms
        r, e, g = _arr(RRt, (P, na, 3)), _arr(ew, (P,)), _arr(gw, (P, na, 3))
        out = np.zeros(PES_NPARAM)
        _check(load_library().pimd_pes_param_grad(self._h, _ptr(r), _ptr(e), _ptr(g), _ptr(out)))
        return out

    def ttm_params(self) -> np.ndarray:
        """The TTM model parameters in use, in the order of TTM_PARAM_NAMES."""
        th = np.zeros(TTM_NPARAM)
        _check(load_library().pimd_ttm_params_get(self._h, _ptr(th)))
        return th

    def set_ttm_params(self, theta):
        th = _arr(theta, (TTM_NPARAM,))
        _check(load_library().pimd_ttm_params_set(self._h, _ptr(th)))
        if self._single is not None:
            self._single.set_ttm_params(th)

    def ttm_param_grad(self, RRt, ew, gw):
        """Tangents of the TTM parameters: (grad, E, dE/dtheta) with
        grad = sum_b ew_b dE_b/dtheta + sum gw . d(dE/dr)/dtheta (see pimd_ttm_param_grad)."""
        P, na = self.params.nbeads, self.params.natoms
        r, e, g = _arr(RRt, (P, na, 3)), _arr(ew, (P,)), _arr(gw, (P, na, 3))
        grad, E, dE = np.zeros(TTM_NPARAM), np.zeros(P), np.zeros((P, TTM_NPARAM))
        _check(load_library().pimd_ttm_param_grad(self._h, _ptr(r), _ptr(e), _ptr(g), _ptr(grad),
                                                  _ptr(E), _ptr(dE)))
        return grad, E, dE

    def dms_coefs(self) -> np.ndarray:
        """The 84 dipole-moment-surface coefficients coefD in use."""
        c = np.zeros(DMS_NCOEF)
        _check(load_library().pimd_dms_coefs_get(self._h, _ptr(c)))
        return c

    def set_dms_coefs(self, coef):
        c = _arr(coef, (DMS_NCOEF,))
        _check(load_library().pimd_dms_coefs_set(self._h, _ptr(c)))
        if self._single is not None:
            self._single.set_dms_coefs(c)

    def dms_coef_grad(self, RRt, ew, gw):
        """Tangents of the DMS coefficients: (grad, E, dE/dcoef), as ttm_param_grad."""
        P, na = self.params.nbeads, self.params.natoms
        r, e, g = _arr(RRt, (P, na, 3)), _arr(ew, (P,)), _arr(gw, (P, na, 3))
        grad, E, dE = np.zeros(DMS_NCOEF), np.zeros(P), np.zeros((P, DMS_NCOEF))
        _check(load_library().pimd_dms_coef_grad(self._h, _ptr(r), _ptr(e), _ptr(g), _ptr(grad),
                                                 _ptr(E), _ptr(dE)))
        return grad, E, dE

    def gamma_m(self) -> float:
        """The M-site fraction gammaM in use."""
        g = np.zeros(1)
        _check(load_library().pimd_gamma_m_get(self._h, _ptr(g)))
        return float(g[0])

    def set_gamma_m(self, g: float):
        _check(load_library().pimd_gamma_m_set(self._h, float(g)))
        if self._single is not None:
            self._single.set_gamma_m(g)

    def gamma_m_grad(self, RRt, ew, gw):
        """The tangent along gammaM (the M sites moving with it): (grad, E,
        dE/dgammaM), as ttm_param_grad."""
        P, na = self.params.nbeads, self.params.natoms
        r, e, g = _arr(RRt, (P, na, 3)), _arr(ew, (P,)), _arr(gw, (P, na, 3))
        grad, E, dE = np.zeros(1), np.zeros(P), np.zeros(P)
        _check(load_library().pimd_gamma_m_grad(self._h, _ptr(r), _ptr(e), _ptr(g), _ptr(grad),
                                                _ptr(E), _ptr(dE)))
        return float(grad[0]), E, dE

    def ttm_adjoint(self, RRt, ew, gw, nparam: int = ADJ_NPARAM):
        """The adjoint sweep over the first nparam (16, 100 or 101) of the TTM
        parameters, the DMS coefficients and gammaM (pimd_ttm_adjoint):
        (grad, E, dE/dtheta), the parameters in the order TTM_PARAM_NAMES,
        coefD, gammaM."""
        P, na = self.params.nbeads, self.params.natoms
        r, e, g = _arr(RRt, (P, na, 3)), _arr(ew, (P,)), _arr(gw, (P, na, 3))
        grad, E, dE = np.zeros(nparam), np.zeros(P), np.zeros((P, nparam))
        _check(load_library().pimd_ttm_adjoint(self._h, _ptr(r), _ptr(e), _ptr(g), int(nparam),
                                               _ptr(grad), _ptr(E), _ptr(dE)))
        return grad, E, dE

    def ttm_adjoint_start(self, RRt, ew, gw, nparam: int = TTM_NPARAM):
        """ttm_adjoint in two halves (nparam 16 or 100): this queues the sweep
        on the context's stream and returns; ttm_adjoint_finish waits for it
        and returns (grad, E, dE/dtheta).  Sweeps of several force fields run
        side by side."""
        P, na = self.params.nbeads, self.params.natoms
        r, e, g = _arr(RRt, (P, na, 3)), _arr(ew, (P,)), _arr(gw, (P, na, 3))
        _check(load_library().pimd_ttm_adjoint_start(self._h, _ptr(r), _ptr(e), _ptr(g),
                                                     int(nparam)))
        self._adj_nparam = int(nparam)

    def ttm_adjoint_finish(self):
        P, n = self.params.nbeads, getattr(self, "_adj_nparam", TTM_NPARAM)
        grad, E, dE = np.zeros(n), np.zeros(P), np.zeros((P, n))
        _check(load_library().pimd_ttm_adjoint_finish(self._h, _ptr(grad), _ptr(E), _ptr(dE)))
        return grad, E, dE

    def potential(self, RR, RRc, t: int = 1) -> dict:
        """One bead (force_calc.f90:112-170): the reference's per-rank call."""
        if self._single is None:
            q = TTMParams(**{**self.params.__dict__, "nbeads": 1})
            self._single = TTMForceField(q)
            self._single.set_pes_params(self.pes_params())
            self._single.set_ttm_params(self.ttm_params())
            if self.params.pot_model in (2, 3):
                self._single.set_dms_coefs(self.dms_coefs())
                self._single.set_gamma_m(self.gamma_m())
        r = self._single.full_bead_forces(np.asarray(RR)[None], RRc, t)
        return {k: (v[0] if isinstance(v, np.ndarray) and v.ndim > 0 else v) for k, v in r.items()}

    def profile(self, on: bool = True):
        _check(load_library().pimd_profile_enable(self._h, int(on)))

    def profile_reset(self):
        _check(load_library().pimd_profile_reset(self._h))

    def profile_report(self) -> dict:
        L = load_library()
        n = L.pimd_profile_count(self._h)
        res = {}
        for i in range(n):
            name = C.create_string_buffer(64)
            cnt, ms = C.c_long(), C.c_double()
            _check(L.pimd_profile_get(self._h, i, name, 64, C.byref(cnt), C.byref(ms)))
            res[name.value.decode()] = (cnt.value, ms.value)
        return res


class RingPolymer:
    """NormalModes / Langevin module equivalents bound to a force-field context."""

    def __init__(self, ff: TTMForceField):
        self.ff = ff
        self.P = ff.params.nbeads
        self.natoms = ff.params.natoms

    def init_normal_modes(self, omegan: float, delta: float, setNMfreq: float = 0.0):
        _check(load_library().pimd_nm_init(self.ff.handle, omegan, delta, setNMfreq))

    def module_data(self) -> dict:
        P = self.P
        C_, a = np.zeros((P, P)), [np.zeros(P) for _ in range(6)]
        _check(load_library().pimd_nm_get(self.ff.handle, _ptr(C_), *[_ptr(x) for x in a]))
        return {"C": C_, "A11": a[0], "A12": a[1], "A21": a[2], "A22": a[3], "omegalist": a[4],
                "MassScaleFactor": a[5]}

    def evolve_ring(self, RRt, PPt, masses=None):
        RRt = _arr(RRt, (self.P, self.natoms, 3)).copy()
        PPt = _arr(PPt, (self.P, self.natoms, 3)).copy()
        m = None if masses is None else _arr(masses, (self.natoms,))
        _check(load_library().pimd_nm_evolve(self.ff.handle, _ptr(RRt), _ptr(PPt), _ptr(m)))
        return RRt, PPt

    def init_langevin_nm(self, delt2: float, centroid_thermostat: bool, tau_bead: float,
                         temp_beads: float):
        _check(load_library().pimd_langevin_init(self.ff.handle, delt2, int(centroid_thermostat),
                                                 tau_bead, temp_beads))

    def langevin_nm(self, PPt, noise=None, seed: int = 0, counter: int = 0):
        PPt = _arr(PPt, (self.P, self.natoms, 3)).copy()
        nz = None if noise is None else _arr(noise, (self.natoms, self.P, 3))
        _check(load_library().pimd_langevin(self.ff.handle, _ptr(PPt), _ptr(nz), seed, counter))
        return PPt


def nose_hoover(uk: float, vxi, tau: float, delt2: float, L: int, temp: float) -> dict:
    s, u = C.c_double(0.0), C.c_double(uk)
    v = np.array(vxi, dtype=np.float64)
    _check(load_library().pimd_nose_hoover(C.byref(s), C.byref(u), _ptr(v), tau, delt2, L, temp))
    return {"s": s.value, "uk": u.value, "vxi": v}


class PIMDIntegrator:
    """Device-resident PIMD velocity Verlet (MD.f90:35-110) over a force field."""

    def __init__(self, ff: TTMForceField, delt_fs=0.25, temp=300.0, bead_thermostat=1,
                 centroid_thermostat=False, tau_centroid=0.1, global_nh=False, tau=0.1,
                 setNMfreq=0.0, intra_timesteps=0):
        """intra_timesteps >= 1 runs contracted_MD (MD.f90:276-449) with that many
        inner steps of the monomer forces per outer step."""
        self.ff = ff
        md = MDParams(delt_fs, temp, bead_thermostat, int(centroid_thermostat), tau_centroid,
                      int(global_nh), int(intra_timesteps), tau, setNMfreq)
        _check(load_library().pimd_md_init(ff.handle, C.byref(md)))
        self.P, self.natoms = ff.params.nbeads, ff.params.natoms

    def set_state(self, RRt, PPt, dRRt=None):
        shp = (self.P, self.natoms, 3)
        a, b = _arr(RRt, shp), _arr(PPt, shp)
        c = None if dRRt is None else _arr(dRRt, shp)
        _check(load_library().pimd_state_set(self.ff.handle, _ptr(a), _ptr(b), _ptr(c)))

    def get_state(self) -> dict:
        shp = (self.P, self.natoms, 3)
        o = {"RRt": np.zeros(shp), "PPt": np.zeros(shp), "dRRt": np.zeros(shp),
             "RRc": np.zeros((self.natoms, 3))}
        _check(load_library().pimd_state_get(self.ff.handle, _ptr(o["RRt"]), _ptr(o["PPt"]),
                                             _ptr(o["dRRt"]), _ptr(o["RRc"])))
        return o

    def state_forces(self, t: int = 1):
        """Forces of the device-resident state coordinates (not synchronised)."""
        _check(load_library().pimd_state_forces(self.ff.handle, int(t)))

    def run(self, nsteps: int, t0: int = 1, noise=None, seed: int = 0) -> dict:
        nz = None
        if noise is not None:
            nz = _arr(noise, (nsteps, 2, self.natoms, self.P, 3))
        o = {k: np.zeros(nsteps) for k in ("Upot", "virial", "virialc", "uk")}
        _check(load_library().pimd_md_run(self.ff.handle, int(t0), int(nsteps), _ptr(nz), seed,
                                          _ptr(o["Upot"]), _ptr(o["virial"]), _ptr(o["virialc"]),
                                          _ptr(o["uk"])))
        return o

    def observables(self, nsteps: int) -> np.ndarray:
        """[nsteps][9] observables of the last run's steps, in the order of OBS_NAMES."""
        o = np.zeros((nsteps, len(OBS_NAMES)))
        _check(load_library().pimd_md_observables(self.ff.handle, int(nsteps), _ptr(o)))
        return o


class ReplicaBatch:
    """Independent PIMD systems (replicas of any size, model and bead count) run
    side by side on one GPU, each integrator already loaded with its state.
    No counterpart in the reference, which runs one system per job."""

    def __init__(self, integrators):
        self.integrators = list(integrators)
        handles = (C.c_void_p * len(self.integrators))(*[i.ff.handle.value for i in self.integrators])
        h = C.c_void_p()
        _check(load_library().pimd_batch_create(handles, len(self.integrators), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().pimd_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self) -> np.ndarray:
        """Picks launch shapes per replica; returns the best force time (ms) of each."""
        ms = np.zeros(len(self.integrators))
        _check(load_library().pimd_batch_autotune(self._h, _ptr(ms)))
        return ms

    def shapes(self) -> list:
        out = []
        for i in range(len(self.integrators)):
            sh = (C.c_int * 5)()
            _check(load_library().pimd_batch_shape(self._h, i, sh))
            out.append(tuple(sh))
        return out

    def run(self, nsteps: int, seeds, t0: int = 1) -> dict:
        n = len(self.integrators)
        sd = (C.c_uint64 * n)(*[int(x) for x in seeds])
        o = {"Upot": np.zeros((n, nsteps)), "uk": np.zeros((n, nsteps))}
        _check(load_library().pimd_batch_md_run(self._h, int(t0), int(nsteps), sd,
                                                _ptr(o["Upot"]), _ptr(o["uk"])))
        return o
