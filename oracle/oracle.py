"""ctypes wrapper for the C restatement (oracle/_ref/liboracle.so).

Test infrastructure only (tests/, smoke() and bench.py's CPU-baseline leg).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "liboracle.so")


class OrcParams(C.Structure):
    _fields_ = [("pot_model", C.c_int), ("nwaters", C.c_int), ("nbeads", C.c_int),
                ("polar_maxiter", C.c_int), ("guess_initdip", C.c_int),
                ("box", C.c_double * 3), ("rc", C.c_double), ("rc1", C.c_double),
                ("eps_ewald", C.c_double), ("polar_sor", C.c_double),
                ("polar_eps", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle library missing: run `make -C oracle port`")
        L = C.CDLL(LIB)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_get_ttm_params.argtypes = [C.c_void_p, dp]
        L.orc_set_ttm_params.argtypes = [C.c_void_p, dp]
        L.orc_get_gamma_m.restype = C.c_double
        L.orc_get_gamma_m.argtypes = [C.c_void_p]
        L.orc_set_gamma_m.argtypes = [C.c_void_p, C.c_double]
        L.orc_nkvec.restype = C.c_int
        L.orc_nkvec.argtypes = [C.c_void_p]
        L.orc_consts.argtypes = [C.c_void_p, dp]
        L.orc_kvecs.argtypes = [C.c_void_p, dp, dp, ip]
        L.orc_forces.argtypes = [C.c_void_p, dp, dp, C.c_int, dp, dp, dp, dp, dp, dp, dp, dp,
                                 ip, C.c_void_p]
        L.orc_erfc.restype = C.c_double
        L.orc_erfc.argtypes = [C.c_double]
        L.orc_gammq23.restype = C.c_double
        L.orc_gammq23.argtypes = [C.c_double]
        L.orc_smear.argtypes = [C.c_void_p, C.c_double, C.c_double, dp]
        L.orc_monomer.argtypes = [dp, dp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_double]
        L.orc_pes_default.argtypes = [dp]
        L.orc_pes_monomer.argtypes = [dp, dp, dp, dp, dp]
        L.orc_nm_init.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double] + [dp] * 7
        L.orc_nm_evolve.argtypes = [C.c_int, C.c_int] + [dp] * 9
        L.orc_langevin_init.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                        dp, dp, dp]
        L.orc_langevin.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, C.c_double, C.c_double,
                                   dp, dp]
        L.orc_nose_hoover.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), dp,
                                      C.c_double, C.c_double, C.c_int, C.c_double]
        _lib = L
    return _lib


class System:
    """One TTM system (pot_model, box, cut-offs) for the C restatement."""

    def __init__(self, p):
        self.p = p
        cp = OrcParams(p.pot_model, p.nwaters, p.nbeads, p.polar_maxiter, int(p.guess_initdip),
                       (C.c_double * 3)(*p.box), p.rc, p.rc1, p.eps_ewald, p.polar_sor,
                       p.polar_eps)
        self.h = lib().orc_create(C.byref(cp))
        self.nk = lib().orc_nkvec(self.h)
        self.ndip = 3 * p.nwaters if p.pot_model == 2 else p.nwaters
        self.hist = np.zeros((p.nbeads, 4, self.ndip, 3))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)

    def ttm_params(self):
        """vdwA..vdwE, polarO/H/M, polfacO/H/M, aDD, aCCaCD, dms_param1..3."""
        th = np.zeros(16)
        lib().orc_get_ttm_params(self.h, th)
        return th

    def set_ttm_params(self, theta):
        lib().orc_set_ttm_params(self.h, np.ascontiguousarray(theta, dtype=np.float64))

    def gamma_m(self):
        return lib().orc_get_gamma_m(self.h)

    def set_gamma_m(self, g):
        lib().orc_set_gamma_m(self.h, float(g))

    def consts(self):
        o = np.zeros(11)
        lib().orc_consts(self.h, o)
        return {"aewald": o[0], "Uvdw_lrc": o[1], "shiftA": o[2], "shiftB": o[3],
                "const_ts1": o[4], "const_ts2": o[5], "const_ts3": o[6], "polfac": o[7:11]}

    def kvecs(self):
        kv, Ak, hkl = np.zeros((self.nk, 3)), np.zeros(self.nk), np.zeros((self.nk, 3), np.int32)
        lib().orc_kvecs(self.h, kv, Ak, hkl)
        return kv, Ak, hkl

    def forces(self, RRt, RRc, t=1, use_hist=None):
        p = self.p
        P, na, N = p.nbeads, p.natoms, p.nwaters
        out = {"En": np.zeros(P), "terms": np.zeros((P, 5)), "dRR": np.zeros((P, na, 3)),
               "virt": np.zeros((P, 3, 3)), "virialc": np.zeros(P), "dipI": np.zeros((P, N, 3)),
               "dipE": np.zeros((P, N, 3)), "chg": np.zeros((P, na)),
               "iters": np.zeros(P, np.int32)}
        hist = self.hist if (use_hist if use_hist is not None else p.guess_initdip) else None
        lib().orc_forces(self.h, np.ascontiguousarray(RRt, dtype=np.float64),
                         np.ascontiguousarray(RRc, dtype=np.float64), t, out["En"],
                         out["terms"], out["dRR"], out["virt"], out["virialc"], out["dipI"],
                         out["dipE"], out["chg"], out["iters"],
                         hist.ctypes.data_as(C.c_void_p) if hist is not None else None)
        return out

    def smear(self, drsq, apol):
        o = np.zeros(11)
        lib().orc_smear(self.h, drsq, apol, o)
        return o


def erfc(x):
    return lib().orc_erfc(float(x))


def gammq23(x):
    return lib().orc_gammq23(float(x))


def monomer(r1, box, dms=(0.5, 0.9578, 0.012)):
    e1, dr1, q3, dq3 = np.zeros(1), np.zeros((3, 3)), np.zeros(3), np.zeros((3, 3, 3))
    lib().orc_monomer(np.ascontiguousarray(r1, dtype=np.float64), e1, dr1, q3, dq3,
                      np.asarray(box, dtype=np.float64), *dms)
    return e1[0], dr1, q3, dq3


def nm_init(P, omegan, delta, setnm=0.0):
    C_, a = np.zeros((P, P)), [np.zeros(P) for _ in range(6)]
    lib().orc_nm_init(P, omegan, delta, setnm, C_, *a)
    return {"C": C_, "A11": a[0], "A12": a[1], "A21": a[2], "A22": a[3], "omegalist": a[4],
            "MassScaleFactor": a[5]}


def nm_evolve(nm, masses, RRt, PPt):
    RRt, PPt = np.array(RRt, dtype=np.float64), np.array(PPt, dtype=np.float64)
    P, na = RRt.shape[0], RRt.shape[1]
    lib().orc_nm_evolve(P, na, nm["C"], nm["A11"], nm["A12"], nm["A21"], nm["A22"],
                        nm["MassScaleFactor"], np.asarray(masses, dtype=np.float64), RRt, PPt)
    return RRt, PPt


def langevin_init(P, delt2, centroid, tau_bead, temp_beads, omegalist):
    c1, c2 = np.zeros(P), np.zeros(P)
    lib().orc_langevin_init(P, delt2, int(centroid), tau_bead, temp_beads,
                            np.asarray(omegalist, dtype=np.float64), c1, c2)
    return c1, c2


def langevin(nm, c1, c2, massO, massH, noise, PPt):
    PPt = np.array(PPt, dtype=np.float64)
    P, na = PPt.shape[0], PPt.shape[1]
    lib().orc_langevin(P, na, nm["C"], nm["MassScaleFactor"], c1, c2, massO, massH,
                       np.ascontiguousarray(noise, dtype=np.float64), PPt)
    return PPt


def nose_hoover(uk, vxi, tau, delt2, L, temp):
    s, u, v = C.c_double(0.0), C.c_double(uk), np.array(vxi, dtype=np.float64)
    lib().orc_nose_hoover(C.byref(s), C.byref(u), v, tau, delt2, L, temp)
    return {"s": s.value, "uk": u.value, "vxi": v}


def pes_default():
    """The reference PES parameters (combined c5z, deohA, alphaoh, roh, phh1A, phh2)."""
    th = np.zeros(250)
    lib().orc_pes_default(th)
    return th


def pes_monomer(r1, box, theta):
    """pot_nasa of one monomer r1[atom][xyz] with parameters theta: (e1, dr1[3][3])."""
    e1 = np.zeros(1)
    dr1 = np.zeros((3, 3))
    lib().orc_pes_monomer(np.ascontiguousarray(r1, dtype=np.float64),
                          np.ascontiguousarray(box, dtype=np.float64),
                          np.ascontiguousarray(theta, dtype=np.float64), e1, dr1)
    return float(e1[0]), dr1
