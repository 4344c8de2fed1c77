"""MHDSolver (reference: src/solver.py:1-77).

The loop keeps the reference's structure -- dump, optional startup RHS,
dt from the last evaluation, then step / advance t / new dt / dump until tf
or max_steps -- but hands the device runs of steps between dumps
(``evaluator.run``), because the time step, the clock and every field stay
on the GPU; the host clock is re-read after each run."""
from __future__ import annotations

import logging

from .base import EPSILON, Solver

logger = logging.getLogger(__name__)


class MHDSolver(Solver):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.initial_acceleration_is_current = False

    def solve(self, show_progress=False):
        self._epsilon = EPSILON * self.tf
        self.dump_output()
        self.barrier()

        # Skip the startup RHS when the application has already warmed it.
        if self.initial_acceleration_is_current:
            self.initial_acceleration_is_current = False
        else:
            self.integrator.initial_acceleration(self.t, self.dt)

        self.dt = self._get_timestep()
        ev = self.integrator.evaluator
        while (self.tf - self.t) > self._epsilon and self.count < self.max_steps:
            if self.pre_step_callbacks or self.post_step_callbacks:
                batch = 1
            else:
                to_dump = self.pfreq - (self.count % self.pfreq) if self.pfreq > 0 else self.max_steps
                batch = max(1, min(to_dump, self.max_steps - self.count))
            for callback in self.pre_step_callbacks:
                callback(self)
            logger.debug("Iteration=%d, time=%f, timestep=%f", self.count, self.t, self.dt)
            self.t, self.dt = ev.run(batch, self.t, self.dt)
            self.count += batch
            self._epsilon = EPSILON * self.tf * self.count
            for callback in self.post_step_callbacks:
                callback(self)
            self._dump_output_if_needed()
            if self.execute_commands is not None and self.count % self.command_interval == 0:
                self.execute_commands(self)
        self.dump_output()
