"""Reverse-mode derivatives of a rollout: the adjoint step of the library
(sphmhd_adjoint_*) and a torch.autograd.Function over it.

``AdjointSystem`` is a ``DeviceSystem`` with one bar row per property (a
second [NFIELDS, n] storage tensor).  ``record`` runs the rollout -- the
warm start and ``nsteps`` PEC steps -- keeping the state every ``every``
steps; ``seed_totals`` sets the bars of the final state for a weighted sum of
the conserved totals; ``reverse`` sweeps back over the rollout, recomputing
each segment from its checkpoint, and leaves in the bar rows the gradient
with respect to every property of the initial state (and returns the
gradient with respect to the parameters of TPARAMS).  One reverse sweep
serves any number of inputs -- the initial field of every particle as well
as the viscosity and cleaning parameters.

``RolloutEnergy`` is the torch view: its forward records the rollout from
initial-field tensors and a parameter vector and returns the final weighted
totals (E_kin + E_mag by default); its backward runs the reverse sweep and
returns the gradient with respect to those tensors, so gradients flow on to
whatever torch computation produced them (e.g. an initial condition built
from an amplitude).  Memory: the checkpoints (one state per ``every``
steps) and one segment of ``every`` states, not the whole rollout.
"""
from __future__ import annotations

import ctypes
from typing import Mapping, Sequence

import numpy as np
import torch

from ._abi import FIELD_INDEX, NFIELDS, NTOTALS, NTPARAMS, TOTALS, TPARAMS, check
from .device import DeviceSystem

ENERGY = {"ekin": 1.0, "emag": 1.0}


def _weights(weights: Mapping[str, float]) -> np.ndarray:
    w = np.zeros(NTOTALS)
    for name, v in weights.items():
        w[TOTALS.index(name)] = float(v)
    if w[TOTALS.index("divb")] != 0.0:
        raise ValueError("the divB total has no derivative")
    return w


class AdjointSystem(DeviceSystem):
    """A device system with bar rows for the reverse sweep."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.bars = torch.zeros((NFIELDS, self.n), dtype=torch.float64, device=self.device)
        ptrs = (ctypes.c_void_p * NFIELDS)(*[self.bars[k].data_ptr() for k in range(NFIELDS)])
        check(self.lib.sphmhd_adjoint_bind(self._ctx, ptrs))

    def bar(self, name: str) -> torch.Tensor:
        """Device view of one bar row."""
        return self.bars[FIELD_INDEX[name]]

    def gradient(self, name: str) -> np.ndarray:
        torch.cuda.synchronize(self.device)
        return self.bar(name).cpu().numpy()

    def zero_bars(self) -> None:
        self.bars.zero_()

    def record(self, nsteps: int, every: int = 20, warm: bool = True) -> int:
        """The rollout from the current fields: returns the steps taken."""
        check(self.lib.sphmhd_adjoint_record(self._ctx, 1 if warm else 0, int(nsteps), int(every)))
        steps = ctypes.c_int64(0)
        check(self.lib.sphmhd_adjoint_steps(self._ctx, ctypes.byref(steps)))
        return int(steps.value)

    def seed_totals(self, weights: Mapping[str, float] = ENERGY) -> None:
        w = _weights(weights)
        check(self.lib.sphmhd_adjoint_seed_totals(self._ctx, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def reverse(self) -> np.ndarray:
        """The reverse sweep; returns the parameter gradient [ninst, NTPARAMS]."""
        out = np.zeros((self.ninst, NTPARAMS))
        check(self.lib.sphmhd_adjoint_reverse(self._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out


def with_params(params, values: Mapping[str, float]):
    """Copy of a Params with the named members of TPARAMS replaced."""
    p = type(params)()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(params), ctypes.sizeof(params))
    for name, v in values.items():
        if name not in TPARAMS:
            raise KeyError(name)
        setattr(p, name, float(v))
    return p


def energy_gradient(setup, steps: int, every: int = 20, warm: bool = True,
                    weights: Mapping[str, float] = ENERGY):
    """The weighted final totals of a single-instance rollout of ``setup`` and
    their gradient: (value, {property: d/d initial property}, {parameter:
    d/d parameter})."""
    sysm = AdjointSystem(setup.params, setup.offsets)
    try:
        sysm.set_fields(dict(setup.fields))
        sysm.record(steps, every, warm)
        tot = sysm.totals_dict(0)
        value = float(sum(w * tot[name] for name, w in weights.items()))
        sysm.zero_bars()
        sysm.seed_totals(weights)
        pbar = sysm.reverse()
        torch.cuda.synchronize(sysm.device)
        host = sysm.bars.cpu().numpy()
        grads = {name: host[FIELD_INDEX[name]].copy() for name in FIELD_INDEX}
        return value, grads, dict(zip(TPARAMS, pbar[0]))
    finally:
        sysm.close()


class RolloutEnergy(torch.autograd.Function):
    """E(fields, theta): forward records the rollout, backward reverses it.

    apply(theta, setup, names, steps, every, warm, weights, *tensors):
    theta is a 1-d tensor of the TPARAMS values (in that order) for the
    single instance of ``setup``; ``tensors`` are initial properties named by
    ``names`` (the rest of the initial state comes from ``setup.fields``).
    Returns the weighted final totals as a 0-d tensor."""

    @staticmethod
    def forward(ctx, theta, setup, names, steps, every, warm, weights, *tensors):
        params = with_params(setup.params[0], dict(zip(TPARAMS, theta.detach().cpu().tolist())))
        sysm = AdjointSystem([params], setup.offsets)
        fields = dict(setup.fields)
        for name, t in zip(names, tensors):
            fields[name] = t.detach().to(torch.float64).cpu().numpy()
        sysm.set_fields(fields)
        sysm.record(steps, every, warm)
        tot = sysm.totals_dict(0)
        value = float(sum(w * tot[name] for name, w in weights.items()))
        ctx.sysm = sysm
        ctx.names = tuple(names)
        ctx.weights = dict(weights)
        ctx.theta_meta = (theta.dtype, theta.device)
        ctx.tensor_meta = [(t.dtype, t.device) for t in tensors]
        return torch.as_tensor(value, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out):
        sysm = ctx.sysm
        try:
            g = float(grad_out)
            sysm.zero_bars()
            sysm.seed_totals({k: g * v for k, v in ctx.weights.items()})
            pbar = sysm.reverse()
            dtype, device = ctx.theta_meta
            dtheta = torch.as_tensor(pbar[0], dtype=dtype, device=device)
            dfields = [sysm.bar(name).clone().to(dtype=dt, device=dev)
                       for name, (dt, dev) in zip(ctx.names, ctx.tensor_meta)]
        finally:
            sysm.close()
            ctx.sysm = None
        return (dtheta, None, None, None, None, None, None, *dfields)


def rollout_energy(setup, fields: Mapping[str, torch.Tensor], theta: torch.Tensor, steps: int, every: int = 20,
                   warm: bool = True, weights: Mapping[str, float] = ENERGY) -> torch.Tensor:
    """Differentiable E(fields, theta) of a single-instance rollout (RolloutEnergy)."""
    names: Sequence[str] = tuple(fields)
    return RolloutEnergy.apply(theta, setup, names, steps, every, warm, dict(weights),
                               *[fields[n] for n in names])
