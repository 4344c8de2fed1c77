"""Run configuration (reference: src/config.py:1-128), same TOML schema.

Frozen dataclasses loaded strictly from TOML (tomli on Python 3.10, the
reference uses tomllib) and validated with the reference's rules.  Optional
keys extend the schema without changing reference files:
  numerics.dim        2 or 3 (default 3, the reference's hard-wired value)
  numerics.alpha_min  floor of the viscosity switch (default 0.0)
  numerics.alpha_max  scale of the viscosity switch (default 1.0)
  solver.max_steps    cap on the number of steps (default: none)
execution.backend keeps the reference's values; every backend runs the HIP
kernels here ("hip" is accepted too).
"""
from __future__ import annotations

from dataclasses import dataclass, fields, is_dataclass
from pathlib import Path
from types import NoneType
from typing import get_args, get_type_hints

try:
    import tomllib as _toml  # Python 3.11+
except ModuleNotFoundError:  # pragma: no cover - 3.10
    import tomli as _toml


@dataclass(frozen=True)
class CaseConfig:
    name: str
    nx: int


@dataclass(frozen=True)
class PhysicsConfig:
    gamma: float
    mu0: float
    artificial_viscosity: float
    artificial_thermal_conductivity: float
    artificial_magnetic_dissipation: float


@dataclass(frozen=True)
class NumericsConfig:
    kernel: str
    hfact: float
    density: str
    periodic_mode: str | None
    cleaning_speed_factor: float
    cleaning_damping_factor: float
    dim: int | None = None
    alpha_min: float | None = None
    alpha_max: float | None = None


@dataclass(frozen=True)
class SolverConfig:
    tf: float
    pfreq: int
    adaptive_timestep: bool
    timestep_factor: float
    max_steps: int | None = None


@dataclass(frozen=True)
class ExecutionConfig:
    backend: str
    fused: bool
    directory: str


@dataclass(frozen=True)
class SimulationConfig:
    case: CaseConfig
    physics: PhysicsConfig
    numerics: NumericsConfig
    solver: SolverConfig
    execution: ExecutionConfig

    def validate(self):
        _validate_config(self)


def load_config(path, config_types):
    with Path(path).open("rb") as stream:
        values = _toml.load(stream)
    config = _load_dataclass(config_types[values["case"]["name"]], values)
    config.validate()
    return config


def _load_dataclass(config_type, values):
    type_hints = get_type_hints(config_type)
    config_fields = {field.name: type_hints[field.name] for field in fields(config_type)}
    optional_fields = {name for name, field_type in config_fields.items() if NoneType in get_args(field_type)}
    if not set(values) <= set(config_fields):
        raise AssertionError(f"unknown keys {sorted(set(values) - set(config_fields))}")
    if not set(config_fields) - optional_fields <= set(values):
        raise AssertionError(f"missing keys {sorted(set(config_fields) - optional_fields - set(values))}")
    arguments = {}
    for name, field_type in config_fields.items():
        if name not in values:
            continue
        if is_dataclass(field_type):
            arguments[name] = _load_dataclass(field_type, values[name])
        else:
            arguments[name] = values[name]
    arguments.update({name: None for name in optional_fields if name not in values})
    return config_type(**arguments)


def _validate_config(config):
    """The reference's rules (config.py:80-115) plus the optional keys."""
    c = config
    assert type(c.case.name) is str and c.case.name
    assert type(c.case.nx) is int and c.case.nx > 0
    assert type(c.physics.gamma) is float and c.physics.gamma > 1.0
    assert type(c.physics.mu0) is float and c.physics.mu0 > 0.0
    for v in (c.physics.artificial_viscosity, c.physics.artificial_thermal_conductivity,
              c.physics.artificial_magnetic_dissipation):
        assert type(v) is float and v >= 0.0
    assert c.numerics.kernel in {"cubic", "quintic", "wendland_c2", "wendland_c4", "wendland_c6", "gaussian",
                                 "super_gaussian"}
    assert type(c.numerics.hfact) is float and c.numerics.hfact > 0.0
    assert c.numerics.density in {"iterate", "single"}
    assert c.numerics.periodic_mode in {None, "ghost", "minimum_image"}
    assert type(c.numerics.cleaning_speed_factor) is float and c.numerics.cleaning_speed_factor >= 0.0
    assert type(c.numerics.cleaning_damping_factor) is float and c.numerics.cleaning_damping_factor >= 0.0
    assert c.numerics.dim in {None, 2, 3}
    if c.numerics.alpha_min is not None:
        assert type(c.numerics.alpha_min) is float and c.numerics.alpha_min >= 0.0
    if c.numerics.alpha_max is not None:
        assert type(c.numerics.alpha_max) is float and c.numerics.alpha_max > 0.0
    assert type(c.solver.tf) is float and c.solver.tf > 0.0
    assert type(c.solver.pfreq) is int and c.solver.pfreq > 0
    assert type(c.solver.adaptive_timestep) is bool
    assert type(c.solver.timestep_factor) is float and c.solver.timestep_factor > 0.0
    if c.solver.max_steps is not None:
        assert type(c.solver.max_steps) is int and c.solver.max_steps >= 0
    assert c.execution.backend in {"cython", "cuda", "opencl", "hip"}
    assert type(c.execution.fused) is bool
    assert not c.execution.fused or c.execution.backend in ("cuda", "hip")
    assert type(c.execution.directory) is str and c.execution.directory


def pysph_arguments(execution):
    """The reference maps the execution table to PySPH flags (config.py:118-128);
    here only the output directory matters (the backend is always HIP)."""
    return ["-d", execution.directory]
