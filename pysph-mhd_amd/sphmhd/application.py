"""MHDApplication over the HIP backend (reference interface:
src/application.py:26-147).

Cases subclass it and override the reference's hooks -- ``create_mhd_particles``,
``bounds`` and ``refresh_initial_thermodynamics`` -- and read the settings
``initialize`` copies from the configuration.  What differs is where the
start-up runs: the particles live on the GPU, so the warm start (three force
evaluations with the magnetic field put back after the first) is one call of
the library (``sphmhd_warm_start``).  Only a case that refreshes its
thermodynamics after the first density sum (the vortex) has the host drive
the three evaluations itself, calling the hook in between on a host copy of
the summed density.  The order of operations is the reference's in both
paths (application.py:110-147); the initial time step is the smaller of the
Courant and force limits over all particles (application.py:134-147).
"""
from __future__ import annotations

import numpy as np

from .base import Application, DomainManager
from .config import SimulationConfig
from .equations import COURANT_FACTOR, FORCE_FACTOR
from .scheme import MHDScheme

# settings copied from the configuration by initialize(): attribute name ->
# (section, key) of SimulationConfig
_SETTINGS = {
    "gamma": ("physics", "gamma"),
    "mu0": ("physics", "mu0"),
    "kernel": ("numerics", "kernel"),
    "hfact": ("numerics", "hfact"),
    "density": ("numerics", "density"),
    "periodic_mode": ("numerics", "periodic_mode"),
    "tf": ("solver", "tf"),
    "pfreq": ("solver", "pfreq"),
    "adaptive_timestep": ("solver", "adaptive_timestep"),
    "timestep_factor": ("solver", "timestep_factor"),
    "nx": ("case", "nx"),
}


def initial_time_step(h, dt_cfl, au, av, aw, timestep_factor):
    """First step of a run (application.py:134-147): the smallest of the
    per-particle Courant limits tsf * 0.3 * h / dt_cfl and, over particles
    with a non-zero acceleration, the force limits tsf * 0.25 * sqrt(h/|a|).
    Raises when a signal speed is not positive or the step is not."""
    h = np.asarray(h, dtype=np.float64)
    dt_cfl = np.asarray(dt_cfl, dtype=np.float64)
    if not np.all(dt_cfl > 0.0):
        raise RuntimeError("initial dt: every dt_cfl must be positive (application.py:139)")
    limit = float(np.min(timestep_factor * COURANT_FACTOR * h / dt_cfl))
    anorm = np.sqrt(au * au + av * av + aw * aw)
    moving = anorm > 0.0
    if moving.any():
        limit = min(limit, float(np.min(timestep_factor * FORCE_FACTOR * np.sqrt(h[moving] / anorm[moving]))))
    if not limit > 0.0:
        raise RuntimeError("initial dt must be positive (application.py:147)")
    return limit


class MHDApplication(Application):
    config_type = SimulationConfig

    def __init__(self, config):
        self.config = config
        super().__init__(fname=config.case.name)

    # -- hooks of the reference -------------------------------------------------
    def initialize(self):
        for attr, (section, key) in _SETTINGS.items():
            setattr(self, attr, getattr(getattr(self.config, section), key))

    def create_mhd_particles(self):  # pragma: no cover - interface
        raise NotImplementedError

    def refresh_initial_thermodynamics(self, particle):
        """Called once, after the first density sum of the warm start; the
        base does nothing (application.py:43-44)."""

    @property
    def bounds(self):  # pragma: no cover - interface
        raise NotImplementedError

    def create_particles(self):
        arrays = self.create_mhd_particles()
        if len(arrays) != 1 or arrays[0].name != "fluid":
            raise ValueError("the MHD scheme runs exactly one particle array, named 'fluid'")
        self.scheme.setup_properties(arrays[0])
        return arrays

    def create_domain(self):
        if self.periodic_mode is None:  # open domain: no wrap, no minimum image
            return DomainManager()
        box = dict(zip(("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"), self.bounds))
        return DomainManager(periodic_in_x=True, periodic_in_y=True, periodic_in_z=True,
                             periodic_mode=self.periodic_mode, **box)

    def create_scheme(self):
        return MHDScheme(self.config)

    def configure_scheme(self):
        self.scheme.configure_solver(dt=1.0, tf=self.tf, adaptive_timestep=self.adaptive_timestep, pfreq=self.pfreq)

    def solve(self):
        self.warm_start()
        super().solve()

    # -- start-up on the device -------------------------------------------------
    def _refreshes_thermodynamics(self) -> bool:
        return type(self).refresh_initial_thermodynamics is not MHDApplication.refresh_initial_thermodynamics

    def warm_start(self):
        """Accelerations of the initial state and the first time step, then
        tell the solver not to evaluate them again (solver.py:35-38)."""
        system = self.evaluator.system
        self.evaluator.check_periodic_images()
        if self._refreshes_thermodynamics():
            self._warm_start_with_hook(system)
        else:
            system.warm_start()
        self.evaluator.check_periodic_images()  # with the summed h
        particle = self.particles[0]
        particle.gpu.pull("rho", "h", "Bx", "By", "Bz", "converged", "dt_cfl", "au", "av", "aw")
        self.solver.dt = initial_time_step(particle.h, particle.dt_cfl, particle.au, particle.av, particle.aw,
                                           self.timestep_factor)
        self.solver.initial_acceleration_is_current = True

    def _warm_start_with_hook(self, system):
        """The three evaluations run from the host so that the case's hook can
        see the first summed density: evaluation 1 from h0 = h and the synced
        predicted state, the field restored, the hook, then two evaluations in
        which h and omega stay as evaluation 1 left them (converged stays 1)."""
        field = system.field
        saved_b = [field(c).clone() for c in ("Bx", "By", "Bz")]

        def evaluate(times):
            field("h0").copy_(field("h"))
            system.sync_initial_state()
            for _ in range(times):
                system.eval()
            if not bool((field("converged") == 1.0).all()):
                raise RuntimeError("warm start: the density iteration did not converge (application.py:120,131)")

        evaluate(1)
        for dst, src in zip(("Bx", "By", "Bz"), saved_b):
            field(dst).copy_(src)
        particle = self.particles[0]
        particle.gpu.pull("rho", "h", "converged", "e")
        self.refresh_initial_thermodynamics(particle)
        particle.gpu.push("e")
        evaluate(2)
