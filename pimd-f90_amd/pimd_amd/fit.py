"""Fitting the monomer PES and the TTM parameters to reference energies and forces.

A fitting set is a list of configuration groups; each group is a force field
whose beads hold independent configurations of one size and box (a bead is
one configuration here, not a ring-polymer image), with target energies per
configuration and target gradients per atom.  The loss is

    L(theta) = sum_configs (E - E*)^2 + sum_configs,atoms |dE/dr - (dE/dr)*|^2

over theta: the Partridge-Schwenke parameters (combined c5z, deohA, alphaoh,
roh, phh1A, phh2; PES_NPARAM values), optionally followed by the TTM
parameters (vdW, polarisabilities, Thole factors, DMS corrections;
TTM_NPARAM values, see TTM_PARAM_NAMES), then optionally by the 84
coefficients of the dipole moment surface that charges the sites (DMS_NCOEF)
and then optionally by the M-site fraction gammaM.  Forward values come from
the GPU force path; the gradient from the hand-written derivative kernels
(pimd_pes_param_grad for the PES, pimd_ttm_param_grad for the TTM parameters,
pimd_dms_coef_grad for the DMS coefficients, pimd_gamma_m_grad for gammaM).  PESLoss is a torch.autograd.Function over theta, so any torch
optimiser can drive a fit.
"""
from __future__ import annotations

import numpy as np

from .ttm import DMS_NCOEF, PES_NPARAM, TTM_NPARAM, TTMForceField

NPARAM_ALL = PES_NPARAM + TTM_NPARAM
NPARAM_DMS = NPARAM_ALL + DMS_NCOEF
NPARAM_GM = NPARAM_DMS + 1


def split_theta(theta):
    """(PES part, TTM part, DMS coefficients, gammaM) of a parameter vector
    of any of the four lengths; the parts it does not hold are None."""
    th = np.asarray(theta, dtype=np.float64)
    if th.size == PES_NPARAM:
        return th, None, None, None
    if th.size == NPARAM_ALL:
        return th[:PES_NPARAM], th[PES_NPARAM:], None, None
    if th.size in (NPARAM_DMS, NPARAM_GM):
        gm = float(th[NPARAM_DMS]) if th.size == NPARAM_GM else None
        return th[:PES_NPARAM], th[PES_NPARAM:NPARAM_ALL], th[NPARAM_ALL:NPARAM_DMS], gm
    raise ValueError(f"theta must hold {PES_NPARAM}, {NPARAM_ALL}, {NPARAM_DMS} or {NPARAM_GM} "
                     f"values, got {th.size}")


class FitGroup:
    """P configurations of one size and box evaluated in one all-beads call."""

    def __init__(self, ff: TTMForceField, RRt, E_target, G_target):
        self.ff = ff
        self.RRt = np.ascontiguousarray(RRt, dtype=np.float64)
        self.RRc = self.RRt[0].copy()   # enters only the centroid virial, unused here
        self.E_target = np.asarray(E_target, dtype=np.float64)
        self.G_target = np.ascontiguousarray(G_target, dtype=np.float64)

    def evaluate(self, theta, grad: bool, method: str = "adjoint"):
        """Loss of this group and, with grad, its gradient.  method "adjoint":
        the TTM, DMS and gammaM parts from one adjoint sweep
        (pimd_ttm_adjoint, exact up to the SOR convergence); "tangent": from
        their tangent sweeps (one per parameter, exact for the SOR iterations
        as executed)."""
        return self._evaluate(theta, grad, method)

    def _set_params(self, pes, ttm, dms, gm):
        self.ff.set_pes_params(pes)
        if ttm is not None:
            self.ff.set_ttm_params(ttm)
        if dms is not None:
            self.ff.set_dms_coefs(dms)
        if gm is not None:
            self.ff.set_gamma_m(gm)

    def _residuals(self, r, grad):
        """(loss, PES gradient or None, dE, dG) of a forward result."""
        dE = r["Upott"] - self.E_target
        dG = r["dRRt"] - self.G_target
        loss = float(np.sum(dE * dE) + np.sum(dG * dG))
        g = self.ff.pes_param_grad(self.RRt, 2.0 * dE, 2.0 * dG) if grad else None
        return loss, g, dE, dG

    def _forward(self, pes, ttm, dms, gm, grad):
        """Sets the parameters, runs the forward pass: (loss, PES gradient or
        None, dE, dG)."""
        self._set_params(pes, ttm, dms, gm)
        return self._residuals(self.ff.full_bead_forces(self.RRt, self.RRc), grad)

    def _evaluate(self, theta, grad: bool, method: str):
        pes, ttm, dms, gm = split_theta(theta)
        loss, g, dE, dG = self._forward(pes, ttm, dms, gm, grad)
        if grad:
            if method == "adjoint" and ttm is not None:
                n = TTM_NPARAM + (DMS_NCOEF if dms is not None else 0) + (gm is not None)
                ga = self.ff.ttm_adjoint(self.RRt, 2.0 * dE, 2.0 * dG, n)[0]
                return loss, np.concatenate([g, ga])
            if method not in ("tangent", "adjoint"):
                raise ValueError(f"method must be 'tangent' or 'adjoint', got {method!r}")
            if ttm is not None:
                g = np.concatenate([g, self.ff.ttm_param_grad(self.RRt, 2.0 * dE, 2.0 * dG)[0]])
            if dms is not None:
                g = np.concatenate([g, self.ff.dms_coef_grad(self.RRt, 2.0 * dE, 2.0 * dG)[0]])
            if gm is not None:
                g = np.append(g, self.ff.gamma_m_grad(self.RRt, 2.0 * dE, 2.0 * dG)[0])
        return loss, g


def loss_and_grad(groups, theta, grad: bool = True, method: str = "adjoint"):
    """Loss over every group and, with grad, its gradient dL/dtheta (method:
    see FitGroup.evaluate).  With the adjoint (and no gammaM) every group's
    forward pass is started before any is finished, and every group's
    adjoint sweep before any is finished, so the groups run side by side on
    their force fields' streams."""
    total, g = 0.0, np.zeros(np.asarray(theta).size) if grad else None
    pes, ttm, dms, gm = split_theta(theta)
    if not (grad and method == "adjoint" and ttm is not None and gm is None):
        for grp in groups:
            l, gg = grp.evaluate(theta, grad, method)
            total += l
            if grad:
                g += gg
        return total, g
    n = TTM_NPARAM + (DMS_NCOEF if dms is not None else 0)
    for grp in groups:
        grp._set_params(pes, ttm, dms, gm)
        grp.ff.full_bead_forces_start(grp.RRt, grp.RRc)
    parts = []
    for grp in groups:
        l, gp, dE, dG = grp._residuals(grp.ff.full_bead_forces_finish(), True)
        grp.ff.ttm_adjoint_start(grp.RRt, 2.0 * dE, 2.0 * dG, n)
        parts.append((l, gp))
    for grp, (l, gp) in zip(groups, parts):
        total += l
        g += np.concatenate([gp, grp.ff.ttm_adjoint_finish()[0]])
    return total, g


def _autograd_function():
    import torch

    class PESLoss(torch.autograd.Function):
        """loss = PESLoss.apply(theta, groups): theta a float64 tensor of PES_NPARAM,
        PES_NPARAM + TTM_NPARAM, PES_NPARAM + TTM_NPARAM + DMS_NCOEF values, or
        those and gammaM."""

        @staticmethod
        def forward(ctx, theta, groups, method="adjoint"):
            th = theta.detach().cpu().numpy().astype(np.float64)
            loss, g = loss_and_grad(groups, th, grad=True, method=method)
            ctx.save_for_backward(torch.from_numpy(g).to(theta.device))
            return theta.new_tensor(loss)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None, None

    return PESLoss


def pes_loss(theta, groups, method: str = "adjoint"):
    """Differentiable loss of a torch tensor theta (see PESLoss; method: see
    FitGroup.evaluate)."""
    return _autograd_function().apply(theta, groups, method)
