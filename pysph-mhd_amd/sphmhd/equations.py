"""The 14 equations of the reference plan (src/equations.py), as descriptors.

Each class keeps the reference's name and constructor arguments and stores
the same constants; the plan built from them by MHDScheme.get_equations is
checked and then evaluated by the HIP kernels of libsphmhd.so:

  DensityIteration                 equations.py:14-94    k_density + k_g0_control
  IdealGasEOS                      :97-106               k_prep_force
  DensityFromSmoothingLength       :109-118              k_prep_force
  SmoothingLengthRateFromDensityRelation :121-130        (ah overwritten by G6)
  SmoothingLengthRateFromForceDivergence :133-140        k_force epilogue
  MagneticStressAccelerations      :143-195              k_force pair loop
  MHDAccelerations                 :198-280              k_force pair loop + epilogue
  CullenDehnenAlphaDiagnostics     :283-429              k_force pair loop + epilogue
  StrainDiagnostics                :432-471              (diagnostic only)
  MagneticStateReconstruction      :474-485              k_prep_force
  MagneticStateRates               :488-567              k_force pair loop + epilogue
  AdaptiveTimestep                 :570-585              k_force epilogue (+ dt min)
  EnergyLimiter                    :588-598              k_force epilogue
  DivBCorrection                   :601-624              k_force epilogue
"""
from __future__ import annotations

from .base import Equation

DENSITY_TOLERANCE = 1.0e-4
COURANT_FACTOR = 0.3
FORCE_FACTOR = 0.25
PAIR_DISTANCE_EPSILON = 1.0e-12
INTERACTION_DISTANCE_EPSILON = 1.0e-8
FLOAT_TINY = 2.2250738585072014e-308


class DensityIteration(Equation):
    def __init__(self, dest, sources, *, k, dim, iterate_only_once):
        self.density_iterations = True
        self.iterate_only_once = iterate_only_once
        self.dim = dim
        self.k = k
        self.htol = DENSITY_TOLERANCE
        self.equation_has_converged = 1
        super().__init__(dest, sources)

    def converged(self):
        """1 when every particle finalised in the last pass, else -1 (the
        backend keeps this flag on the device and reports it here)."""
        return self.equation_has_converged


class IdealGasEOS(Equation):
    def __init__(self, dest, sources, gamma):
        self.gamma = gamma
        self.gamma1 = gamma - 1.0
        super().__init__(dest, sources)


class DensityFromSmoothingLength(Equation):
    def __init__(self, dest, sources, k, dim):
        self.k = k
        self.dim = dim
        super().__init__(dest, sources)


class SmoothingLengthRateFromDensityRelation(Equation):
    def __init__(self, dest, sources, k, dim):
        self.k = k
        self.dim = dim
        super().__init__(dest, sources)


class SmoothingLengthRateFromForceDivergence(Equation):
    def __init__(self, dest, sources, dim):
        self.dim1 = 1.0 / dim
        super().__init__(dest, sources)


class MagneticStressAccelerations(Equation):
    def __init__(self, dest, sources, mu0):
        self.mu0 = mu0
        super().__init__(dest, sources)


class MHDAccelerations(Equation):
    def __init__(self, dest, sources, mu0, artificial_viscosity, artificial_thermal_conductivity):
        self.beta = 2.0
        self.mu0 = mu0
        self.artificial_viscosity = artificial_viscosity
        self.artificial_thermal_conductivity = artificial_thermal_conductivity
        super().__init__(dest, sources)


class CullenDehnenAlphaDiagnostics(Equation):
    def __init__(self, dest, sources):
        super().__init__(dest, sources)


class StrainDiagnostics(Equation):
    pass


class MagneticStateReconstruction(Equation):
    pass


class MagneticStateRates(Equation):
    def __init__(self, dest, sources, mu0, artificial_magnetic_dissipation, cleaning_speed_factor,
                 cleaning_damping_factor):
        super().__init__(dest, sources)
        self.mu0 = mu0
        self.artificial_magnetic_dissipation = artificial_magnetic_dissipation
        self.cleaning_speed_factor = cleaning_speed_factor
        self.cleaning_damping_factor = cleaning_damping_factor
        self.cleaning_active = float(cleaning_speed_factor > 0.0)


class AdaptiveTimestep(Equation):
    def __init__(self, dest, sources, timestep_factor):
        self.c_cour = timestep_factor * COURANT_FACTOR
        self.c_force = timestep_factor * FORCE_FACTOR
        self.bignumber = 1.0e29
        super().__init__(dest, sources)


class EnergyLimiter(Equation):
    def __init__(self, dest, sources, timestep_factor):
        self.c_cour = timestep_factor * COURANT_FACTOR
        super().__init__(dest, sources)


class DivBCorrection(Equation):
    def __init__(self, dest, sources, mu0):
        self.mu0 = mu0
        super().__init__(dest, sources)
