"""Run parameters for one TTM system, with the reference's defaults.

Defaults restate input_file_reader.f90:76-84,104-105 and the derived values of
InputOutput.f90:118-145.  Several reference defaults are single-precision
literals widened to double (input_file_reader.f90:81,104,105 and
InputOutput.f90:124), so they are reproduced through float32 here.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def f32(x: float) -> float:
    """A single-precision literal widened to double, as Fortran does."""
    return float(np.float32(x))


TTM21F = 2
TTM3F = 3
QSPCFW = 4     # q-SPC/Fw, fixed charges (pot_spc.f90:69-106)
SPCF = 5       # SPC/F, fixed charges (pot_spc.f90:107-137)


@dataclass
class TTMParams:
    pot_model: int = TTM3F              # 2 TTM2.1-F, 3 TTM3-F, 4 q-SPC/Fw, 5 SPC/F (potential.f90:84-145)
    nwaters: int = 32
    nbeads: int = 1
    box: tuple = (0.0, 0.0, 0.0)
    rc: float = -1.0                    # -1: min(box)/2 (InputOutput.f90:118-122)
    rc1: float = -1.0                   # -1: 0.8*rc with 0.8 a real(4) literal
    eps_ewald: float = 1e-6
    polar_maxiter: int = 15
    polar_sor: float = field(default_factory=lambda: f32(0.7))
    polar_eps: float = 1e-5
    guess_initdip: bool = False         # reference default is .true.; see the predictor note
    mass_o: float = field(default_factory=lambda: f32(15.994))
    mass_h: float = field(default_factory=lambda: f32(1.008))
    precision: int = 0                  # 0 = FP64 (the reference's arithmetic), 1 = mixed
    contraction: bool = False           # pot_ttm's CONTRACTION branch: no monomer PES

    def resolved(self) -> "TTMParams":
        """Copy with the -1 sentinels replaced by the reference's derived values."""
        p = TTMParams(**self.__dict__)
        box = tuple(float(b) for b in (self.box if np.ndim(self.box) else (self.box,) * 3))
        p.box = box
        if p.rc == -1.0:
            p.rc = min(box) / 2.0
        if p.rc1 == -1.0:
            p.rc1 = f32(0.8) * p.rc
        return p

    @property
    def natoms(self) -> int:
        return 3 * self.nwaters
