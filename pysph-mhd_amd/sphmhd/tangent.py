"""Derivatives of the step: the tangent-linear (forward-mode) sweep of the
library (sphmhd_tangent_*) and a torch.autograd.Function over a rollout.

``TangentSystem`` is a ``DeviceSystem`` whose every property row carries
``NTANGENTS`` directional derivatives beside its value (a second storage
tensor of shape [NTANGENTS, NFIELDS, n]).  The caller seeds the directions:
initial-state tangents through ``set_tangents`` and parameter tangents
through ``seed_params``; ``warm_start`` and ``step`` then advance values and
derivatives together, taking exactly the primal step's branches, and
``tangent_totals`` returns each conserved total with its derivatives.

For a scalar objective of up to ``NTANGENTS`` inputs (here the final-time
energy as a function of the initial field amplitude and the viscosity and
cleaning parameters) one sweep gives the whole gradient, with memory
independent of the rollout length: no tape and no checkpoints are needed.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Mapping, Sequence

import numpy as np
import torch

from ._abi import FIELD_INDEX, NFIELDS, NTANGENTS, NTOTALS, NTPARAMS, TOTALS, TPARAMS, check
from .device import DeviceSystem


class TangentSystem(DeviceSystem):
    """A device system carrying NTANGENTS tangent directions."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.tstorage = torch.zeros((NTANGENTS, NFIELDS, self.n), dtype=torch.float64, device=self.device)
        ptrs = []
        for d in range(NTANGENTS):
            for f in range(NFIELDS):
                ptrs.append(self.tstorage[d, f].data_ptr())
        arr = (ctypes.c_void_p * (NTANGENTS * NFIELDS))(*ptrs)
        check(self.lib.sphmhd_tangent_bind(self._ctx, arr))
        self.seed_params({})

    def set_tangents(self, direction: int, values: Mapping[str, np.ndarray]) -> None:
        """Initial tangent of the named properties along one direction."""
        for name, val in values.items():
            t = torch.as_tensor(np.asarray(val, dtype=np.float64), device=self.device)
            self.tstorage[direction, FIELD_INDEX[name]].copy_(t)

    def tangent(self, direction: int, name: str) -> np.ndarray:
        torch.cuda.synchronize(self.device)
        return self.tstorage[direction, FIELD_INDEX[name]].cpu().numpy()

    def seed_params(self, seeds: Mapping[tuple[int, str], float] | Sequence) -> None:
        """Parameter tangents: {(direction, parameter name): d param / d direction}
        for every instance, or an array [ninst, NTPARAMS, NTANGENTS]."""
        if isinstance(seeds, Mapping):
            arr = np.zeros((self.ninst, NTPARAMS, NTANGENTS))
            for (d, name), v in seeds.items():
                arr[:, TPARAMS.index(name), d] = v
        else:
            arr = np.ascontiguousarray(np.asarray(seeds, dtype=np.float64).reshape(self.ninst, NTPARAMS, NTANGENTS))
        check(self.lib.sphmhd_tangent_seed(self._ctx, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def warm_start(self) -> None:
        check(self.lib.sphmhd_tangent_warm_start(self._ctx))

    def step(self, nsteps: int = 1) -> int:
        return check(self.lib.sphmhd_tangent_step(self._ctx, int(nsteps)))

    def tangent_totals(self) -> dict[str, np.ndarray]:
        """{total: array [ninst, 1 + NTANGENTS]} (value, then its tangents)."""
        out = np.zeros((self.ninst, NTOTALS, 1 + NTANGENTS))
        check(self.lib.sphmhd_tangent_totals(self._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return {name: out[:, k, :] for k, name in enumerate(TOTALS)}

    def dt_tangent(self) -> np.ndarray:
        out = np.zeros((self.ninst, NTANGENTS))
        check(self.lib.sphmhd_tangent_get_dt(self._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out


def _ic_tangent(setup_fn: Callable, kw: dict, name: str) -> dict[str, np.ndarray]:
    """d(initial fields)/d(kw[name]): the builder's own exact derivative when
    it records one (``info["tangents"][name]``, e.g. alfven_2d, whose fields
    are linear in the amplitude), else a central difference of the builder
    (exact up to rounding for fields at most quadratic in the input)."""
    exact = setup_fn(**kw).info.get("tangents", {}).get(name)
    if exact is not None:
        return {k: np.asarray(v, dtype=np.float64) for k, v in exact.items()}
    x = kw[name]
    delta = 1.0e-3 * max(abs(x), 1.0)
    fp = setup_fn(**{**kw, name: x + delta}).fields
    fm = setup_fn(**{**kw, name: x - delta}).fields
    return {k: (np.asarray(fp[k]) - np.asarray(fm[k])) / (2.0 * delta) for k in fp}


def energy_rollout(setup_fn: Callable, theta: Sequence[float], names: Sequence[str], steps: int):
    """E = E_kin + E_mag after the warm start and `steps` PEC steps of the
    single-instance setup `setup_fn(**dict(zip(names, theta)))`, and dE/dtheta
    from one tangent sweep.

    An entry of ``names`` is either a parameter of TPARAMS (seeded as a
    parameter tangent) or another keyword of ``setup_fn`` that only shapes the
    initial condition, such as a field amplitude (seeded as the initial-state
    tangent).  At most NTANGENTS entries."""
    if len(theta) > NTANGENTS:
        raise ValueError(f"at most {NTANGENTS} inputs per sweep")
    kw = {nm: float(v) for nm, v in zip(names, theta)}
    setup = setup_fn(**kw)
    sysm = TangentSystem(setup.params, setup.offsets)
    sysm.set_fields(dict(setup.fields))
    seeds = {}
    for d, nm in enumerate(names):
        if nm in TPARAMS:
            seeds[(d, nm)] = 1.0
        else:
            sysm.set_tangents(d, _ic_tangent(setup_fn, kw, nm))
    sysm.seed_params(seeds)
    sysm.warm_start()
    sysm.step(steps)
    tot = sysm.tangent_totals()
    e = tot["ekin"][0] + tot["emag"][0]
    sysm.close()
    return float(e[0]), np.asarray(e[1:1 + len(names)])


def energy_primal(setup_fn: Callable, kw: Mapping[str, float], steps: int) -> float:
    """The same objective from the plain (value-only) step."""
    setup = setup_fn(**kw)
    sysm = DeviceSystem(setup.params, setup.offsets)
    sysm.set_fields(dict(setup.fields))
    sysm.warm_start()
    sysm.step(steps)
    tot = sysm.totals_dict(0)
    sysm.close()
    return float(tot["ekin"] + tot["emag"])


class EnergyRollout(torch.autograd.Function):
    """torch.autograd view of energy_rollout: forward returns E(theta) as a
    0-d tensor; backward multiplies the incoming gradient by dE/dtheta from
    the same tangent sweep."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, setup_fn, names, steps):
        e, grad = energy_rollout(setup_fn, theta.detach().cpu().tolist(), names, steps)
        ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype, device=theta.device))
        return torch.as_tensor(e, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None
