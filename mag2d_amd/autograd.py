"""torch.autograd over the library's adjoint (m2d_adjoint_energy_2d).

The mean kinetic energy (eV) of the live particles after a self-consistent
run (Pic::advance with a solve every step, pic.cpp:330-358; optionally with
MCC collisions) is a differentiable function of the voltages of groups of
electrode nodes.  ``ElectrodeEnergy.apply(V, run)`` returns it as a 0-d
tensor; its backward is the library's reverse sweep, so ``J.backward()``
fills ``V.grad`` with dJ/dV.  Every evaluation starts from the same initial
state: the particle store uploaded afresh, u cleared (the solves warm-start
from it), the voltages set on the device.  Collisions keep their random
numbers between evaluations (they are keyed by seed, stream, slot and
step), so nearby voltages see the same events and finite differences of J
agree with the gradient.
"""
from __future__ import annotations

import numpy as np
import torch

from . import abi


class ElectrodeRun:
    """One self-consistent run, set up for repeated evaluation.

    base_voltage: nx*nz node voltages of the grid (Fields::grid.voltage);
    group: nx*nz group index of every FIXED node that moves with a group
    voltage, -1 elsewhere: node k gets base_voltage[k] + V[group[k]];
    store: the initial particles (x, z, vx, vy, vz host arrays, already
    started by advance_boris_init); mcc: the species' collisions or None.
    """

    def __init__(self, ctx: abi.Context, fields: abi.Fields2D, base_voltage, group, n_groups: int,
                 species: abi.Species2D, store, steps: int, mcc=None, checkpoint_every: int = 100,
                 adjoint_tol: float = 0.0):
        self.ctx, self.fields, self.species, self.mcc = ctx, fields, species, mcc
        self.base = np.ascontiguousarray(base_voltage, np.float64).ravel()
        self.group = np.ascontiguousarray(group, np.int32).ravel()
        if self.group.shape != self.base.shape:
            raise ValueError("group and base_voltage must both hold one entry per node")
        if self.group.max() >= n_groups:
            raise ValueError(f"group index {self.group.max()} out of range for {n_groups} groups")
        self.n_groups = n_groups
        self.store = tuple(np.ascontiguousarray(a, np.float32) for a in store)
        self.steps, self.checkpoint_every, self.adjoint_tol = steps, checkpoint_every, adjoint_tol
        self.info = None

    def evaluate(self, V, grad: bool = True):
        """(J, dJ/dV or None) at group voltages V (n_groups floats)."""
        V = np.asarray(V, np.float64).ravel()
        if V.shape != (self.n_groups,):
            raise ValueError(f"expected {self.n_groups} group voltages, got {V.shape}")
        volt = self.base.copy()
        on = self.group >= 0
        volt[on] += V[self.group[on]]
        self.fields.upload(abi.FIELD_VOLTAGE, volt)
        self.fields.upload(abi.FIELD_U, np.zeros_like(volt))
        parts = abi.ParticlesSoA.from_host(self.ctx, *self.store)
        J, g, self.info = abi.adjoint_energy(self.ctx, self.fields, self.species, parts, self.steps,
                                             group=self.group if grad else None, n_groups=self.n_groups,
                                             checkpoint_every=self.checkpoint_every,
                                             adjoint_tol=self.adjoint_tol, mcc=self.mcc)
        return J, g


class ElectrodeEnergy(torch.autograd.Function):
    """J(V) of an ElectrodeRun; V is a 1-d tensor of group voltages."""

    @staticmethod
    def forward(ctx, V: torch.Tensor, run: ElectrodeRun):
        J, g = run.evaluate(V.detach().to("cpu", torch.float64).numpy(), grad=V.requires_grad)
        ctx.save_for_backward(torch.as_tensor(g if g is not None else np.zeros(run.n_groups), dtype=V.dtype,
                                              device=V.device))
        return torch.tensor(J, dtype=V.dtype, device=V.device)

    @staticmethod
    def backward(ctx, gJ):
        (g,) = ctx.saved_tensors
        return gJ * g, None


def electrode_energy(V: torch.Tensor, run: ElectrodeRun) -> torch.Tensor:
    return ElectrodeEnergy.apply(V, run)
