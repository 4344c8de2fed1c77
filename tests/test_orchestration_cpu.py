"""Pins of the oracle's orchestration: the time-step chain and the warm-start
sequence (reference: equations.py:570-585 AdaptiveTimestep, solver.py:40,61
the step re-read after each evaluation, application.py:110-147 the warm start
and initial dt), and the solver loop's end at tf.

Each check is something the mathematics or the sequence itself fixes,
computed independently of the oracle's formulas:
  * a uniform state has no acceleration, so its step is the Courant limit
    tsf * 0.3 * h / max(v_w, c_h) in closed form;
  * the pending step is the minimum over particles, and it sits where the
    signal speed is largest;
  * a step that would pass tf is clipped to tf - t and the run stops there;
  * after the warm start the solver's re-read step (the pending minimum)
    equals the application's initial step, min of the Courant and force
    limits;
  * evaluations 2 and 3 of the warm start leave h and omega exactly as
    evaluation 1 made them (converged stays 1), and the magnetic field is
    the initial one put back and re-synced: B = (B0 / rho) * rho bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import init_fields


def _oracle(setup):
    from pyoracle import OracleSystem

    o = OracleSystem(setup.params, setup.offsets, threads=8)
    o.set_fields(init_fields(setup))
    return o


def _uniform(nx=16, vel=(0.0, 0.0, 0.0), bfield=(0.2, 0.1, 0.05), chf=1.0, adaptive=True):
    from sphmhd.cases.workloads import orszag_tang_2d

    s = orszag_tang_2d(nx, cleaning_speed_factor=chf)
    s.params[0].adaptive_timestep = 1 if adaptive else 0
    n = s.n
    for c, v in zip(("u", "v", "w"), vel):
        s.fields[c] = np.full(n, float(v))
    for c, v in zip(("Bx", "By", "Bz"), bfield):
        s.fields[c] = np.full(n, float(v))
    return s


@pytest.mark.parametrize("chf", [1.0, 2.0])
@pytest.mark.parametrize("vel", [(0.0, 0.0, 0.0), (0.3, -0.2, 0.1)])
def test_uniform_state_step_is_courant_closed_form(chf, vel):
    """Uniform state on an exact periodic lattice (at rest, or moving
    uniformly: Galilean): accelerations vanish to round-off and every signal
    speed equals v_w = sqrt(c_s^2 + B^2 / (mu0 rho)), so the pending step is
    tsf * 0.3 * h / max(v_w, c_h) with c_h = chf * v_w."""
    s = _uniform(vel=vel, chf=chf)
    p = s.params[0]
    o = _oracle(s)
    o.warm_start()
    h, rho, e = o.get("h"), o.get("rho"), o.get("e")
    b2 = sum(float(v) ** 2 for v in (0.2, 0.1, 0.05))
    pres = (p.gamma - 1.0) * rho * e
    cs2 = p.gamma * pres / rho
    vw = np.sqrt(cs2 + b2 / (p.mu0 * rho))
    ch = chf * vw
    scale = np.max(vw * vw / h)
    for c in ("au", "av", "aw"):
        assert np.max(np.abs(o.get(c))) <= 1e-12 * scale, c
    expect = np.min(p.timestep_factor * 0.3 * h / np.maximum(vw, ch))
    pend = o.pending_dt()[0]
    assert pend == pytest.approx(expect, rel=1e-13)
    t, dt = o.clock()
    assert dt[0] == pend  # solver.py:40: the re-read step replaces the initial one


def test_pending_step_is_the_minimum_over_particles():
    """Two regions: the right half ten times hotter.  The pending step is the
    smallest per-particle step, and the particle attaining it lies in the hot
    half, where the sound speed is larger."""
    from sphmhd.cases.workloads import jittered, orszag_tang_2d

    s = jittered(orszag_tang_2d(24), 0.1, seed=5)
    hot = s.fields["x"] > 0.5
    s.fields["e"] = np.where(hot, 10.0, 1.0) * s.fields["e"]
    o = _oracle(s)
    o.warm_start()
    dta = o.get("dt_adapt")
    pend = o.pending_dt()[0]
    assert pend == np.min(dta)
    assert hot[int(np.argmin(dta))]
    assert np.min(dta[hot]) < np.min(dta[~hot])
    # each particle's step is the smaller of its Courant and force limits
    p = s.params[0]
    h, cfl = o.get("h"), o.get("dt_cfl")
    a = np.sqrt(o.get("au") ** 2 + o.get("av") ** 2 + o.get("aw") ** 2)
    lim = np.minimum(p.timestep_factor * 0.3 * h / cfl,
                     np.where(a > 0, p.timestep_factor * 0.25 * np.sqrt(h / np.where(a > 0, a, 1.0)), np.inf))
    assert np.allclose(dta, lim, rtol=1e-14, atol=0)


@pytest.mark.parametrize("adaptive", [True, False])
def test_step_clipped_to_tf_and_run_stops(adaptive):
    """A run bounded by tf: the steps before it are those of an unbounded run
    (bit for bit), the step that would pass tf is cut to tf - t, the clock
    lands on tf and the run stops there instead of taking zero-length steps."""
    from sphmhd.cases.workloads import jittered, orszag_tang_2d

    base = jittered(orszag_tang_2d(16), 0.1, seed=2)
    base.params[0].adaptive_timestep = 1 if adaptive else 0
    free = _oracle(base)
    free.warm_start()
    ts = []
    for _ in range(4):
        free.step(1)
        ts.append(free.clock()[0][0])
    tf = ts[1] + 0.5 * (ts[2] - ts[1])  # inside the third step
    bounded = jittered(orszag_tang_2d(16, tf=tf), 0.1, seed=2)
    bounded.params[0].adaptive_timestep = 1 if adaptive else 0
    o = _oracle(bounded)
    o.warm_start()
    o.step(1)
    o.step(1)
    assert o.clock()[0][0] == ts[1]
    t2 = o.clock()[0][0]
    assert o.clock()[1][0] == tf - t2  # clipped
    o.step(100)
    assert o.steps_taken() == 3
    t, dt = o.clock()
    assert t[0] == tf
    assert dt[0] == 0.0


@pytest.mark.parametrize("adaptive", [True, False])
def test_initial_step_equals_application_formula(adaptive):
    """application.py:134-147 (min over particles of the Courant and force
    limits, written independently in sphmhd.application.initial_time_step)
    and the solver's re-read (solver.py:40: the pending minimum when
    adaptive, the application's step otherwise) give the same first step."""
    from sphmhd.application import initial_time_step
    from sphmhd.cases.workloads import jittered, rotor_2d

    s = jittered(rotor_2d(32), 0.1, seed=7)
    s.params[0].adaptive_timestep = 1 if adaptive else 0
    o = _oracle(s)
    o.warm_start()
    p = s.params[0]
    app_dt = initial_time_step(o.get("h"), o.get("dt_cfl"), o.get("au"), o.get("av"), o.get("aw"),
                               p.timestep_factor)
    t, dt = o.clock()
    assert dt[0] == app_dt
    assert o.pending_dt()[0] == app_dt
    # on SPH states the Courant limit binds (dt_cfl takes the largest signal
    # speed over the neighbours); the force branch is checked on its own below


def test_initial_time_step_takes_either_limit():
    """initial_time_step on arrays where each limit binds in turn, and where
    only particles with a non-zero acceleration have a force limit."""
    from sphmhd.application import initial_time_step

    h = np.array([0.01, 0.02, 0.01])
    cfl = np.array([1.0, 1.0, 1.0])
    zero = np.zeros(3)
    # no acceleration anywhere: Courant 0.3 * 0.01
    assert initial_time_step(h, cfl, zero, zero, zero, 1.0) == pytest.approx(0.003, rel=1e-15)
    # a large acceleration on particle 1: 0.25 * sqrt(0.02 / 1e4) = 3.5355e-4
    au = np.array([0.0, 6.0e3, 0.0])
    av = np.array([0.0, 8.0e3, 0.0])
    assert initial_time_step(h, cfl, au, av, zero, 1.0) == pytest.approx(0.25 * np.sqrt(0.02 / 1.0e4), rel=1e-15)
    # timestep_factor scales both limits
    assert initial_time_step(h, cfl, au, av, zero, 0.5) == pytest.approx(0.125 * np.sqrt(0.02 / 1.0e4), rel=1e-15)
    with pytest.raises(RuntimeError):
        initial_time_step(h, np.array([1.0, 0.0, 1.0]), zero, zero, zero, 1.0)


def test_warm_start_sequence_invariants():
    """Evaluation 1 from h0 = h and the synced state decides h and omega; the
    two evaluations after it run without the integrator's initialize, so
    every particle stays converged and h, omega, rho keep evaluation 1's
    bits.  The field is the initial one put back and re-synced with that rho:
    B = (B0 / rho) * rho exactly."""
    from sphmhd.cases.workloads import jittered, orszag_tang_2d

    s = jittered(orszag_tang_2d(24), 0.1, seed=9)
    one = _oracle(s)
    one.set_fields({"h0": s.fields["h"]})
    one.sync_initial_state()
    one.eval()
    full = _oracle(s)
    full.warm_start()
    assert np.all(full.get("converged") == 1.0)
    for name in ("h", "omega", "rho"):
        assert np.array_equal(full.get(name), one.get(name)), name
    rho = one.get("rho")
    for c in ("Bx", "By", "Bz"):
        b0 = s.fields[c]
        assert np.array_equal(full.get(c), (b0 / rho) * rho), c
    # evaluation 1 itself moved h off the lattice value and B off B0 (its
    # reconstruction used the lattice density), so the checks above see the
    # restore and the freeze
    assert not np.array_equal(one.get("h"), s.fields["h"])
    assert not np.allclose(one.get("Bx"), s.fields["Bx"], rtol=1e-12, atol=0)


def test_solver_stops_at_tf_with_exact_count():
    """The solver loop over an evaluator that keeps the device's contract
    (runs of up to n steps, each step the pending dt clipped to tf - t, a run
    ending after the step that lands on tf): a tf-bounded run ends at t == tf
    with one count per step taken, and dumps at the start, every pfreq steps
    and the end."""
    from sphmhd.base import PECIntegrator
    from sphmhd.solver import MHDSolver

    class Clock:
        def __init__(self, pending):
            self.pending = pending
            self.calls = []

        def initial_acceleration(self):
            pass

        def pending_dt(self):
            return np.array([self.pending])

        def run(self, n, t, dt):
            self.calls.append(n)
            taken = 0
            for _ in range(n):
                reaches = t + dt >= tf
                t = t + dt
                dt = self.pending if t + self.pending <= tf else tf - t
                taken += 1
                if reaches:
                    break
            return t, dt, taken

    tf = 1.0
    integ = PECIntegrator()
    integ.evaluator = Clock(0.07)
    solver = MHDSolver(integrator=integ, dt=1.0, tf=tf, adaptive_timestep=True, pfreq=5)
    dumps = []
    solver.dump_output = lambda: dumps.append(solver.count)
    solver.solve()
    assert solver.t == tf
    assert solver.count == int(np.ceil(tf / 0.07))  # 15 steps: 14 of 0.07 and a clipped one
    assert sum(integ.evaluator.calls) >= solver.count
    assert dumps[0] == 0 and dumps[-1] == solver.count
    assert all(c % 5 == 0 for c in dumps[1:-1])
