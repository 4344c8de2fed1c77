"""The pieces of the PySPH plug-in protocol the reference builds on, restated
for the HIP backend (reference imports: scheme.py:3-6, equations.py:4,
integrator.py:3, solver.py:4-6, application.py:4-5, cases/*.py:4).

  Equation / Group / Scheme     the group-plan vocabulary (descriptors only:
                                the arithmetic lives in libsphmhd.so)
  IntegratorStep / PECIntegrator  predictor-evaluate-corrector wiring
  CubicSpline ... SuperGaussian  kernel selectors with PySPH's class names
  ParticleArray                 host numpy properties with a device mirror
                                (``pa.gpu.push(*names)`` / ``pull``), the
                                PySPH ParticleArray + gpu model
  DomainManager                 periodic box description
  Solver / Application          time-loop and driver bases
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

import numpy as np

from ._abi import FIELD_INDEX, FIELDS

EPSILON = -1e-12  # tolerance factor of the solver's end-time test (times tf * count)


# ---------------------------------------------------------------------------
# equations and groups
# ---------------------------------------------------------------------------

class Equation:
    """An equation of the plan: destination array name, source names and its
    constructor constants.  ``initialize`` / ``loop`` / ``post_loop`` bodies
    are evaluated by the HIP kernels."""

    def __init__(self, dest, sources):
        self.dest = dest
        self.sources = sources

    def constants(self) -> dict:
        return {k: v for k, v in vars(self).items() if k not in ("dest", "sources")}

    def __repr__(self):
        return f"{type(self).__name__}(dest={self.dest!r}, sources={self.sources!r})"


class Group:
    def __init__(self, equations: Sequence[Equation], update_nnps: bool = False, iterate: bool = False,
                 max_iterations: int = 1, min_iterations: int = 0, real: bool = True):
        self.equations = list(equations)
        self.update_nnps = update_nnps
        self.iterate = iterate
        self.max_iterations = max_iterations
        self.min_iterations = min_iterations
        self.real = real


class Scheme:
    def configure_solver(self, **kw):  # pragma: no cover - interface
        raise NotImplementedError

    def setup_properties(self, particle):  # pragma: no cover - interface
        raise NotImplementedError

    def get_equations(self):  # pragma: no cover - interface
        raise NotImplementedError

    @staticmethod
    def _ensure_properties(particle, props, clean=True):
        for p in props:
            if isinstance(p, dict):
                particle.add_property(p["name"], type=p.get("type", "double"), default=p.get("default", 0))
            else:
                particle.add_property(p)


# ---------------------------------------------------------------------------
# kernels (selectors; the functions are evaluated on the device)
# ---------------------------------------------------------------------------

class _Kernel:
    kernel_id = -1
    radius_scale = 0.0

    def __init__(self, dim: int = 3):
        self.dim = dim


class CubicSpline(_Kernel):
    kernel_id, radius_scale = 0, 2.0


class QuinticSpline(_Kernel):
    kernel_id, radius_scale = 1, 3.0


class WendlandQuintic(_Kernel):
    kernel_id, radius_scale = 2, 2.0


class WendlandQuinticC4(_Kernel):
    kernel_id, radius_scale = 3, 2.0


class WendlandQuinticC6(_Kernel):
    kernel_id, radius_scale = 4, 2.0


class Gaussian(_Kernel):
    kernel_id, radius_scale = 5, 3.0


class SuperGaussian(_Kernel):
    kernel_id, radius_scale = 6, 3.0


# ---------------------------------------------------------------------------
# integrator wiring
# ---------------------------------------------------------------------------

class IntegratorStep:
    pass


class PECIntegrator:
    """initialize -> stage1 -> update_domain -> accelerations -> stage2 ->
    update_domain; ``step`` and ``initial_acceleration`` forward to the
    evaluator bound by the solver."""

    def __init__(self, **steppers):
        self.steppers = steppers
        self.evaluator = None

    def initial_acceleration(self, t, dt):
        self.evaluator.initial_acceleration()

    def step(self, t, dt):
        self.evaluator.step(1)

    def compute_time_step(self, dt, cfl):
        return self.evaluator.pending_dt()


# ---------------------------------------------------------------------------
# particles
# ---------------------------------------------------------------------------

class _GPUMirror:
    """Device copy of a ParticleArray: the bound properties live in one
    device tensor of the HIP backend; push/pull move named properties."""

    def __init__(self, pa):
        self.pa = pa
        self.system = None

    def attach(self, system):
        self.system = system

    def push(self, *names):
        if self.system is None:
            return
        names = names or tuple(n for n in self.pa.properties if n in FIELD_INDEX or n == "itype")
        self.system.set_fields({n: self.pa.properties[n] for n in names if n in FIELD_INDEX or n == "itype"})

    def pull(self, *names):
        if self.system is None:
            return
        names = names or tuple(n for n in self.pa.properties if n in FIELD_INDEX)
        state = self.system.state([n for n in names if n in FIELD_INDEX])
        for n, v in state.items():
            self.pa.properties[n][:] = v


class ParticleArray:
    """Named FP64 (and int) properties in host numpy arrays plus a device
    mirror, as PySPH's ParticleArray (``pa.x``, ``pa.add_property``,
    ``pa.gpu.push/pull``)."""

    def __init__(self, name: str, **props):
        object.__setattr__(self, "name", name)
        object.__setattr__(self, "properties", {})
        object.__setattr__(self, "output_property_arrays", [])
        n = None
        for k, v in props.items():
            arr = np.asarray(v)
            n = arr.size if n is None else n
        object.__setattr__(self, "_n", int(n or 0))
        for k, v in props.items():
            dtype = np.int32 if k in ("itype", "tag", "gid") else np.float64
            self.properties[k] = np.array(v, dtype=dtype).reshape(-1).copy()
        if "gid" not in self.properties:
            self.properties["gid"] = np.arange(self._n, dtype=np.int32)
        object.__setattr__(self, "gpu", _GPUMirror(self))

    def __len__(self):
        return self._n

    def get_number_of_particles(self):
        return self._n

    def add_property(self, name, type="double", default=0):
        if name not in self.properties:
            dtype = np.int32 if type == "int" else np.float64
            self.properties[name] = np.full(self._n, default, dtype=dtype)

    def set_output_arrays(self, names):
        object.__setattr__(self, "output_property_arrays", list(names))

    def get(self, *names):
        return [self.properties[n] for n in names] if len(names) > 1 else self.properties[names[0]]

    def __getattr__(self, name):
        props = object.__getattribute__(self, "properties")
        if name in props:
            return props[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in self.properties:
            self.properties[name][:] = value
        else:
            object.__setattr__(self, name, value)


def get_particle_array(name="fluid", **props) -> ParticleArray:
    return ParticleArray(name, **props)


# ---------------------------------------------------------------------------
# domain, solver, application bases
# ---------------------------------------------------------------------------

class DomainManager:
    def __init__(self, xmin=0.0, xmax=0.0, ymin=0.0, ymax=0.0, zmin=0.0, zmax=0.0, periodic_in_x=False,
                 periodic_in_y=False, periodic_in_z=False, periodic_mode=None, **kw):
        self.bounds = (xmin, xmax, ymin, ymax, zmin, zmax)
        self.periodic = (periodic_in_x, periodic_in_y, periodic_in_z)
        self.periodic_mode = periodic_mode

    @property
    def is_periodic(self) -> bool:
        return any(self.periodic)


class Solver:
    def __init__(self, dim=3, integrator=None, kernel=None, dt=1.0, tf=1.0, adaptive_timestep=False, pfreq=100,
                 max_steps=None, output_directory=None, **kw):
        self.dim = dim
        self.integrator = integrator
        self.kernel = kernel
        self.dt = dt
        self.tf = tf
        self.t = 0.0
        self.count = 0
        self.adaptive_timestep = adaptive_timestep
        self.pfreq = pfreq
        self.max_steps = max_steps if max_steps is not None else 2 ** 62
        self.in_parallel = False
        self.rank = 0
        self.reorder_freq = 0
        self.pre_step_callbacks = []
        self.post_step_callbacks = []
        self.execute_commands = None
        self.command_interval = 1
        self.output_directory = output_directory
        self.particles = []
        self.dumps = []

    def barrier(self):
        pass

    def _get_timestep(self):
        """Adaptive: min dt_adapt of the last evaluation (device reduction);
        the step never passes tf."""
        dt = self.dt
        if self.adaptive_timestep:
            pending = float(np.min(self.integrator.compute_time_step(self.dt, 0.3)))
            if pending > 0.0 and math.isfinite(pending):
                dt = pending
        if self.t + dt > self.tf:
            dt = self.tf - self.t
        return dt

    def update_particle_time(self):
        pass

    def reorder_particles(self):
        pass

    def dump_output(self):
        from .output import dump

        if self.output_directory:
            self.dumps.append(dump(self.output_directory, self.particles, self.t, self.dt, self.count))

    def _dump_output_if_needed(self):
        if self.pfreq > 0 and self.count % self.pfreq == 0:
            self.dump_output()


class Application:
    def __init__(self, fname=None, **kw):
        self.fname = fname
        self.args = []
        self.output_dir = None
        self.max_steps = None

    # hooks the subclasses override
    def initialize(self):
        pass

    def create_particles(self):  # pragma: no cover - interface
        raise NotImplementedError

    def create_domain(self):
        return DomainManager()

    def create_scheme(self):  # pragma: no cover - interface
        raise NotImplementedError

    def configure_scheme(self):
        pass

    def _parse(self, argv: Iterable[str] | None):
        argv = list(argv or [])
        i = 0
        while i < len(argv):
            a = argv[i]
            if a in ("-d", "--directory"):
                self.output_dir = argv[i + 1]
                i += 2
                continue
            if a == "--max-steps":
                self.max_steps = int(argv[i + 1])
                i += 2
                continue
            i += 1

    def prepare(self, argv=None):
        """Build scheme, solver, particles and domain on the host (no device
        work)."""
        self._parse(argv)
        self.initialize()
        self.scheme = self.create_scheme()
        self.configure_scheme()
        self.solver = self.scheme.solver
        if self.max_steps is not None:
            self.solver.max_steps = self.max_steps
        if self.output_dir:
            self.solver.output_directory = self.output_dir
        self.particles = self.create_particles()
        self.domain = self.create_domain()
        self.solver.particles = self.particles
        return self

    def setup(self, argv=None):
        """prepare(), then bind the particles to the device evaluator."""
        from .backend import HIPAccelerationEval

        self.prepare(argv)
        self.evaluator = HIPAccelerationEval(self.particles[0], self.scheme.get_equations(), self.solver.kernel,
                                             self.domain, self.scheme, self.solver)
        self.solver.integrator.evaluator = self.evaluator
        return self

    def solve(self):
        self.solver.solve()

    def run(self, argv=None):
        self.setup(argv)
        self.solve()
        return self
