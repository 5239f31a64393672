"""Shared comparison helpers for the parity tests."""
import numpy as np

# Parity bar for the FP64 form: 1e-10 relative energy, 1e-8 absolute gradient
# (kcal/mol/A).  Virials and dipoles are held to the same relative/absolute scale.
E_REL = 1e-10
F_ABS = 1e-8


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def ref_iters(it, maxiter):
    """The reference reports maxiter+1 when the SOR loop ran out (pot_ttm.f90:308,594);
    the C ABI reports -maxiter for that case."""
    it = np.asarray(it)
    return np.where(it > maxiter, -maxiter, it)
