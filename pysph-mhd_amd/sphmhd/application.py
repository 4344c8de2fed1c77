"""MHDApplication (reference: src/application.py:1-147).

Same hooks and the same start-up sequence: periodic DomainManager from the
case bounds, the three-evaluation warm start with the magnetic field restored
in between (application.py:110-132) and the initial step from dt_cfl and the
force norm (application.py:134-147).  Host arrays are numpy; the device copy
is kept in step through ``particle.gpu.push/pull`` exactly where the
reference pushes and pulls."""
from __future__ import annotations

import numpy as np

from .base import Application, DomainManager
from .config import SimulationConfig
from .equations import COURANT_FACTOR, FORCE_FACTOR
from .particles import INITIAL_PREDICTOR_PAIRS, INITIAL_STATE_GPU_PROPERTIES
from .scheme import MHDScheme


class MHDApplication(Application):
    config_type = SimulationConfig

    def __init__(self, config):
        self.config = config
        super().__init__(fname=config.case.name)

    def initialize(self):
        self.gamma = self.config.physics.gamma
        self.mu0 = self.config.physics.mu0
        self.kernel = self.config.numerics.kernel
        self.hfact = self.config.numerics.hfact
        self.density = self.config.numerics.density
        self.periodic_mode = self.config.numerics.periodic_mode
        self.tf = self.config.solver.tf
        self.pfreq = self.config.solver.pfreq
        self.adaptive_timestep = self.config.solver.adaptive_timestep
        self.timestep_factor = self.config.solver.timestep_factor
        self.nx = self.config.case.nx

    def create_mhd_particles(self):  # pragma: no cover - interface
        raise NotImplementedError

    def refresh_initial_thermodynamics(self, particle):
        pass

    @property
    def bounds(self):  # pragma: no cover - interface
        raise NotImplementedError

    def create_particles(self):
        particles = self.create_mhd_particles()
        assert len(particles) == 1
        assert particles[0].name == "fluid"
        self.scheme.setup_properties(particles[0])
        return particles

    def create_domain(self):
        if self.periodic_mode is None:
            return DomainManager()
        xmin, xmax, ymin, ymax, zmin, zmax = self.bounds
        return DomainManager(xmin=xmin, xmax=xmax, ymin=ymin, ymax=ymax, zmin=zmin, zmax=zmax, periodic_in_x=True,
                             periodic_in_y=True, periodic_in_z=True, periodic_mode=self.periodic_mode)

    def create_scheme(self):
        return MHDScheme(self.config)

    def configure_scheme(self):
        self.scheme.configure_solver(dt=1.0, tf=self.tf, adaptive_timestep=self.adaptive_timestep, pfreq=self.pfreq)

    def solve(self):
        self._warm_initial_derivatives(self.solver)
        self._compute_initial_dt(self.solver)
        super().solve()

    @staticmethod
    def _sync_initial_mhd_state(particle):
        particle.Bevolx[:] = particle.Bx / particle.rho
        particle.Bevoly[:] = particle.By / particle.rho
        particle.Bevolz[:] = particle.Bz / particle.rho
        particle.psi[:] = 0.0
        for predicted, current in INITIAL_PREDICTOR_PAIRS:
            getattr(particle, predicted)[:] = getattr(particle, current)
        if particle.gpu is not None:
            particle.gpu.push(*INITIAL_STATE_GPU_PROPERTIES)

    def _warm_initial_derivatives(self, solver):
        particle = self.particles[0]
        initial_magnetic_field = (particle.Bx.copy(), particle.By.copy(), particle.Bz.copy())
        particle.h0[:] = particle.h
        self._sync_initial_mhd_state(particle)
        particle.gpu.push("h0")
        solver.integrator.initial_acceleration(solver.t, solver.dt)
        particle.gpu.pull("rho", "h", "converged")
        if not np.all(particle.converged == 1):
            raise RuntimeError("warm start: density iteration did not converge (application.py:120)")
        particle.Bx[:], particle.By[:], particle.Bz[:] = initial_magnetic_field
        self.refresh_initial_thermodynamics(particle)
        particle.h0[:] = particle.h
        particle.gpu.push("e", "h0")
        self._sync_initial_mhd_state(particle)
        for _ in range(2):
            solver.integrator.initial_acceleration(solver.t, solver.dt)
        particle.gpu.pull("rho", "h", "Bx", "By", "Bz", "converged")
        if not np.all(particle.converged == 1):
            raise RuntimeError("warm start: density iteration did not converge (application.py:131)")
        solver.initial_acceleration_is_current = True

    def _compute_initial_dt(self, solver):
        particle = self.particles[0]
        particle.gpu.pull("h", "dt_cfl", "au", "av", "aw")
        force_norm = np.sqrt(particle.au * particle.au + particle.av * particle.av + particle.aw * particle.aw)
        if not np.all(particle.dt_cfl > 0.0):
            raise RuntimeError("initial dt: dt_cfl must be positive (application.py:139)")
        dt_courant = self.timestep_factor * COURANT_FACTOR * particle.h / particle.dt_cfl
        dt_candidates = [float(np.min(dt_courant))]
        force_mask = force_norm > 0.0
        if np.any(force_mask):
            dt_force = self.timestep_factor * FORCE_FACTOR * np.sqrt(particle.h[force_mask] / force_norm[force_mask])
            dt_candidates.append(float(np.min(dt_force)))
        solver.dt = min(dt_candidates)
        if not solver.dt > 0.0:
            raise RuntimeError("initial dt must be positive (application.py:147)")
