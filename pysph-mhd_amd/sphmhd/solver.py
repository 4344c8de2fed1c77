"""MHDSolver (reference interface: src/solver.py:15-77).

The time loop of the reference -- dump the start, make sure the
accelerations of the initial state exist, take the first step size, then
step, advance the clock, pick the next step size and dump every ``pfreq``
steps until ``tf`` or ``max_steps`` -- with the steps handed to the device in
runs that end at the next dump (``evaluator.run``): step sizes, the clock and
every field stay on the GPU, and the host re-reads the clock after each run.
A run stops early after the step that lands on ``tf`` and reports how many
steps it took, so ``count`` is exact and no zero-length step is taken once
the clock has reached ``tf``.
"""
from __future__ import annotations

import logging

from .base import EPSILON, Solver

logger = logging.getLogger(__name__)


class MHDSolver(Solver):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.initial_acceleration_is_current = False

    def _time_left(self) -> bool:
        """Steps remain while the clock is short of tf (by more than the
        end-time tolerance), the next step has a length and the step budget
        is not spent."""
        return (self.tf - self.t) > self._epsilon and self.dt > 0.0 and self.count < self.max_steps

    def _run_length(self) -> int:
        """Steps handed to the device at once: one when callbacks watch each
        step, else up to the next dump (and never past max_steps)."""
        if self.pre_step_callbacks or self.post_step_callbacks:
            return 1
        to_dump = self.pfreq - (self.count % self.pfreq) if self.pfreq > 0 else self.max_steps
        return max(1, min(to_dump, self.max_steps - self.count))

    def solve(self, show_progress=False):
        self._epsilon = EPSILON * self.tf
        self.dump_output()
        self.barrier()

        # A warm start by the application leaves the initial accelerations
        # in place; only without one are they evaluated here.
        if not self.initial_acceleration_is_current:
            self.integrator.initial_acceleration(self.t, self.dt)
        self.initial_acceleration_is_current = False

        self.dt = self._get_timestep()
        evaluator = self.integrator.evaluator
        while self._time_left():
            run = self._run_length()
            for callback in self.pre_step_callbacks:
                callback(self)
            logger.debug("Iteration=%d, time=%f, timestep=%f", self.count, self.t, self.dt)
            self.t, self.dt, taken = evaluator.run(run, self.t, self.dt)
            self.count += taken
            self._epsilon = EPSILON * self.tf * self.count
            for callback in self.post_step_callbacks:
                callback(self)
            self._dump_output_if_needed()
            if self.execute_commands is not None and self.count % self.command_interval == 0:
                self.execute_commands(self)
        self.dump_output()
