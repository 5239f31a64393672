"""Fitting the monomer PES to reference energies and forces.

A fitting set is a list of configuration groups; each group is a force field
whose beads hold independent configurations of one size and box (a bead is
one configuration here, not a ring-polymer image), with target energies per
configuration and target gradients per atom.  The loss is

    L(theta) = sum_configs (E - E*)^2 + sum_configs,atoms |dE/dr - (dE/dr)*|^2

over the Partridge-Schwenke parameters theta (combined c5z, deohA, alphaoh,
roh, phh1A, phh2).  Forward values come from the GPU force path; the gradient
comes from the hand-written adjoint kernel (pimd_pes_param_grad).  PESLoss is
a torch.autograd.Function over theta, so any torch optimiser can drive a fit.
"""
from __future__ import annotations

import numpy as np

from .ttm import PES_NPARAM, TTMForceField


class FitGroup:
    """P configurations of one size and box evaluated in one all-beads call."""

    def __init__(self, ff: TTMForceField, RRt, E_target, G_target):
        self.ff = ff
        self.RRt = np.ascontiguousarray(RRt, dtype=np.float64)
        self.RRc = self.RRt[0].copy()   # enters only the centroid virial, unused here
        self.E_target = np.asarray(E_target, dtype=np.float64)
        self.G_target = np.ascontiguousarray(G_target, dtype=np.float64)

    def evaluate(self, theta, grad: bool):
        self.ff.set_pes_params(theta)
        r = self.ff.full_bead_forces(self.RRt, self.RRc)
        dE = r["Upott"] - self.E_target
        dG = r["dRRt"] - self.G_target
        loss = float(np.sum(dE * dE) + np.sum(dG * dG))
        g = self.ff.pes_param_grad(self.RRt, 2.0 * dE, 2.0 * dG) if grad else None
        return loss, g


def loss_and_grad(groups, theta, grad: bool = True):
    """Loss over every group and, with grad, its gradient dL/dtheta."""
    total, g = 0.0, np.zeros(PES_NPARAM) if grad else None
    for grp in groups:
        l, gg = grp.evaluate(theta, grad)
        total += l
        if grad:
            g += gg
    return total, g


def _autograd_function():
    import torch

    class PESLoss(torch.autograd.Function):
        """loss = PESLoss.apply(theta, groups): theta a float64 tensor of 250 values."""

        @staticmethod
        def forward(ctx, theta, groups):
            th = theta.detach().cpu().numpy().astype(np.float64)
            loss, g = loss_and_grad(groups, th, grad=True)
            ctx.save_for_backward(torch.from_numpy(g).to(theta.device))
            return theta.new_tensor(loss)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None

    return PESLoss


def pes_loss(theta, groups):
    """Differentiable loss of a torch tensor theta (see PESLoss)."""
    return _autograd_function().apply(theta, groups)
