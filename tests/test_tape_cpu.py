"""The taped oracle (oracle/tape.cpp): the CPU reference gradient of a
rollout, pinned against the plain oracle.

tape.cpp compiles the oracle's own source over a reverse-mode scalar that
records every FP64 operation; one reverse sweep over the tape gives the
gradient of an objective (here E_kin + E_mag, the totals of
sphmhd_totals) with respect to every initial property and to the six
differentiated parameters.  It must be the derivative of the oracle's map:

  * the objective it returns equals the plain oracle's bit for bit (the
    arithmetic is the oracle's, value for value);
  * contracted with the exact derivative of the Alfven wave's initial fields
    with respect to the amplitude (they are linear in it), it matches central
    differences of the plain oracle in the amplitude -- a smooth direction --
    to the second-order truncation of the differences (1e-8 relative);
  * a random direction over every particle's position, and the viscosity
    parameter, match central differences likewise (1e-6: the directions
    cross a few branch-free but less smooth terms);
  * a field the objective cannot see (the initial psi: the warm start sets
    psi = 0) has gradient exactly 0.

The rollout: a jittered 10x17 Alfven wave (170 particles, Wendland C6), the
warm start and two PEC steps.
"""
from __future__ import annotations

import copy

import numpy as np
import pytest

STEPS = 2


def _setup(amplitude=0.1):
    from sphmhd.cases.workloads import alfven_2d, jittered

    return jittered(alfven_2d(10, 17, amplitude=amplitude), 0.1, seed=5)


def _energy(setup, params=None, fields=None):
    from pyoracle import OracleSystem

    o = OracleSystem(params or setup.params, setup.offsets, threads=1)
    o.set_fields(fields or setup.fields)
    o.warm_start()
    o.step(STEPS)
    t = o.totals()[0]
    o.close()
    return t[4] + t[6]


@pytest.fixture(scope="module")
def taped():
    from pyoracle import tape_rollout

    s = _setup()
    vals, grads, gpar, nodes = tape_rollout(s.params, s.offsets, s.fields, nsteps=STEPS)
    return s, vals, grads, gpar, nodes


def test_tape_value_is_the_oracles(taped):
    s, vals, _, _, nodes = taped
    assert nodes > 1000
    assert vals[0] == _energy(s)


def test_tape_amplitude_direction_matches_central_differences(taped):
    s, _, grads, _, _ = taped
    tan = s.info["tangents"]["amplitude"]
    g = sum(float(np.dot(grads[k], tan[k])) for k in tan)
    eps = 1.0e-5
    fd = (_energy(_setup(0.1 + eps)) - _energy(_setup(0.1 - eps))) / (2.0 * eps)
    assert abs(g - fd) <= 1.0e-8 * abs(fd), (g, fd)


def test_tape_position_direction_matches_central_differences(taped):
    s, _, grads, _, _ = taped
    rng = np.random.default_rng(11)
    vx, vy = rng.standard_normal(s.fields["x"].size), rng.standard_normal(s.fields["x"].size)
    g = float(np.dot(grads["x"], vx) + np.dot(grads["y"], vy))
    eps = 1.0e-6

    def moved(sign):
        f = {k: v.copy() for k, v in s.fields.items()}
        f["x"] = f["x"] + sign * eps * vx
        f["y"] = f["y"] + sign * eps * vy
        return _energy(s, fields=f)

    fd = (moved(1.0) - moved(-1.0)) / (2.0 * eps)
    assert abs(g - fd) <= 1.0e-6 * abs(fd), (g, fd)


def test_tape_viscosity_parameter_matches_central_differences(taped):
    s, _, _, gpar, _ = taped
    p = s.params[0]
    eps = 1.0e-3 * p.artificial_viscosity
    pp, pm = copy.copy(p), copy.copy(p)
    pp.artificial_viscosity += eps
    pm.artificial_viscosity -= eps
    fd = (_energy(s, params=[pp]) - _energy(s, params=[pm])) / (2.0 * eps)
    assert abs(gpar[0, 0] - fd) <= 1.0e-6 * abs(fd), (gpar[0, 0], fd)


def test_tape_gradient_of_an_unseen_field_is_zero(taped):
    _, _, grads, _, _ = taped
    assert not np.any(grads["psi"])
    assert np.any(grads["Bz"]) and np.any(grads["u"])
