"""The tangent-linear step (sphmhd_tangent_*) on the GPU.

  * its values are those of the plain step (the dual kernels restate the
    primal bodies; only rounding may differ);
  * dE/dtheta of E = E_kin + E_mag after the warm start and a number of steps
    matches central finite differences of the plain step, for the initial
    field amplitude and two parameters, on the travelling Alfven wave;
  * torch.autograd through EnergyRollout returns the same gradient.

Central differences use eps = 1e-6 theta and are checked against eps = 1e-5
theta first, so a branch flip between theta + eps and theta - eps (a
neighbour pair crossing the support, a Newton iteration count changing)
would show as disagreement between the two rather than pass silently.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import STATE, compare

pytestmark = pytest.mark.gpu

NAMES = ("amplitude", "alpha_max", "cleaning_damping_factor")
THETA = (0.1, 1.0, 1.0)


def _alfven(amplitude=0.1, alpha_max=1.0, cleaning_damping_factor=1.0):
    from sphmhd.cases.workloads import alfven_2d

    return alfven_2d(24, 42, amplitude=amplitude, alpha_max=alpha_max,
                     cleaning_damping_factor=cleaning_damping_factor)


def test_tangent_values_equal_plain_step(gpu):
    """Floor: divBsymm, divBdiff, psi, apsi and aBevol* are measured against
    1e-3 of their natural scale (|B|/h, rho|B|/h, |B|, v_w|B|/h,
    v_w|B|/(rho h)), since they are round-off where B is uniform; every
    other field against its own norm."""
    from sphmhd.device import DeviceSystem
    from sphmhd.tangent import TangentSystem, _ic_tangent

    s = _alfven()
    plain = DeviceSystem(s.params, s.offsets)
    tan = TangentSystem(s.params, s.offsets)
    for d in (plain, tan):
        d.set_fields(dict(s.fields))
    tan.set_tangents(0, _ic_tangent(_alfven, dict(zip(NAMES, THETA)), "amplitude"))
    tan.seed_params({(1, "alpha_max"): 1.0, (2, "cleaning_damping_factor"): 1.0})
    plain.warm_start()
    tan.warm_start()
    plain.step(6)
    tan.step(6)
    # the parity standard for FP64 fields (rel-L2 <= 1e-10, psi against 1e-3
    # of its natural scale while it is pure round-off): the dual arithmetic
    # contracts differently from the primal's, and the Cullen-Dehnen limiter
    # on this exact lattice amplifies such rounding in alpha
    compare(tan, plain, STATE, 1e-10, "tangent values")
    assert np.allclose(plain.clock()[0], tan.clock()[0], rtol=1e-12, atol=0)


def test_tangent_matches_central_differences(gpu):
    from sphmhd.tangent import energy_primal, energy_rollout

    steps = 10
    e, g = energy_rollout(_alfven, THETA, NAMES, steps)
    kw = dict(zip(NAMES, THETA))
    e0 = energy_primal(_alfven, kw, steps)
    assert abs(e - e0) <= 1e-12 * abs(e0)
    for k, nm in enumerate(NAMES):
        fd = {}
        for rel in (1e-5, 1e-6):
            eps = rel * THETA[k]
            ep = energy_primal(_alfven, {**kw, nm: THETA[k] + eps}, steps)
            em = energy_primal(_alfven, {**kw, nm: THETA[k] - eps}, steps)
            fd[rel] = (ep - em) / (2.0 * eps)
        # rounding noise of a central difference of E: ~1e-16 |E| / eps
        noise = 4e-16 * abs(e0) / (1e-6 * THETA[k])
        assert abs(fd[1e-5] - fd[1e-6]) <= 1e-5 * abs(fd[1e-6]) + 2 * noise, (nm, fd)
        assert abs(g[k] - fd[1e-6]) <= 1e-5 * abs(fd[1e-6]) + noise, (nm, g[k], fd[1e-6])


def test_autograd_energy_rollout(gpu):
    import torch

    from sphmhd.tangent import EnergyRollout, energy_rollout

    theta = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    e = EnergyRollout.apply(theta, _alfven, NAMES, 4)
    (2.0 * e).backward()
    e_ref, g_ref = energy_rollout(_alfven, THETA, NAMES, 4)
    assert float(e.detach()) == pytest.approx(e_ref, rel=1e-14)
    assert np.allclose(theta.grad.numpy(), 2.0 * g_ref, rtol=1e-14, atol=0)


def test_tangent_values_equal_plain_step_split_grid(gpu):
    """The reference rotor (3-D, its dense disk splits the cells): the
    tangent step's values, whose density pass scans per-particle stencils
    as the plain pass does, are the plain step's to the parity standard.

    Floor: the cleaning fields against 1e-3 of their natural scale (|B|/h,
    rho|B|/h, |B|, v_w|B|/h, v_w|B|/(rho h)), and the fields an exact
    lattice holds at round-off against 1e-3 of a physical scale: v_w for
    velocities and dh/dt, v_w^2/h for accelerations, v_w/h for div v,
    v_w^3/h for de/dt, |B| for field components, alpha_max for the switch;
    every other field against its own norm."""
    from helpers import CASE_NX, case_setup, init_fields, physical_scales
    from sphmhd.device import DeviceSystem
    from sphmhd.tangent import TangentSystem

    s = case_setup("mhdrotor", CASE_NX["mhdrotor"])
    p = s.params[0]
    plain = DeviceSystem(s.params, s.offsets)
    tan = TangentSystem(s.params, s.offsets)
    for d in (plain, tan):
        d.set_fields(init_fields(s))
    plain.warm_start()
    tan.warm_start()
    plain.step(3)
    tan.step(3)
    compare(tan, plain, STATE, 1e-10, "tangent values, split grid",
            scales=physical_scales(plain, p.mu0, p.alpha_max, None))
