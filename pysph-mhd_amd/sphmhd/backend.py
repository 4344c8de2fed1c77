"""The equation-group evaluation on the GPU (replaces PySPH's AccelerationEval
and its compiled group kernels for the plan of scheme.py:67-81).

HIPAccelerationEval checks that the plan handed to it is the reference's
10-group plan (same equation classes, order, iteration settings and mutually
consistent constants) and rejects anything else, then builds the library's
parameter block from the plan constants, the domain and the solver, and
binds the particle array's device mirror to a DeviceSystem.
"""
from __future__ import annotations

import numpy as np

from . import equations as E
from ._abi import FIELDS, make_params

PLAN = [
    [E.DensityIteration],
    [E.SmoothingLengthRateFromDensityRelation, E.DensityFromSmoothingLength],
    [E.IdealGasEOS],
    [E.CullenDehnenAlphaDiagnostics],
    [E.StrainDiagnostics, E.MagneticStateReconstruction],
    [E.MHDAccelerations],
    [E.SmoothingLengthRateFromForceDivergence],
    [E.MagneticStressAccelerations],
    [E.MagneticStateRates],
    [E.EnergyLimiter, E.DivBCorrection, E.AdaptiveTimestep],
]


class PlanError(ValueError):
    pass


def _same(values, what):
    vals = [v for v in values if v is not None]
    if vals and any(abs(v - vals[0]) > 1e-15 * max(1.0, abs(vals[0])) for v in vals):
        raise PlanError(f"inconsistent {what} across the plan: {vals}")
    return vals[0] if vals else None


def read_plan(groups) -> dict:
    """Validate the group plan and return its constants."""
    if len(groups) != len(PLAN):
        raise PlanError(f"expected {len(PLAN)} groups, got {len(groups)}")
    for gi, (grp, want) in enumerate(zip(groups, PLAN)):
        got = [type(e) for e in grp.equations]
        if got != want:
            raise PlanError(f"group {gi}: expected {[w.__name__ for w in want]}, got {[g.__name__ for g in got]}")
        for e in grp.equations:
            if e.dest != "fluid":
                raise PlanError("only the single 'fluid' array of the reference is supported")
    g0 = groups[0]
    di = g0.equations[0]
    if not g0.update_nnps:
        raise PlanError("group 0 must update the neighbour search")
    iterate = bool(g0.iterate)
    if iterate == bool(di.iterate_only_once):
        raise PlanError("DensityIteration.iterate_only_once must be the negation of the group's iterate flag")
    eqs = {type(e): e for grp in groups for e in grp.equations}
    k = _same([di.k, eqs[E.DensityFromSmoothingLength].k, eqs[E.SmoothingLengthRateFromDensityRelation].k], "hfact")
    dim = int(_same([di.dim, eqs[E.DensityFromSmoothingLength].dim, eqs[E.SmoothingLengthRateFromDensityRelation].dim,
                     1.0 / eqs[E.SmoothingLengthRateFromForceDivergence].dim1], "dim"))
    mu0 = _same([eqs[E.MHDAccelerations].mu0, eqs[E.MagneticStressAccelerations].mu0,
                 eqs[E.MagneticStateRates].mu0, eqs[E.DivBCorrection].mu0], "mu0")
    tsf_a = eqs[E.AdaptiveTimestep].c_cour / E.COURANT_FACTOR
    tsf_e = eqs[E.EnergyLimiter].c_cour / E.COURANT_FACTOR
    tsf = _same([tsf_a, tsf_e], "timestep_factor")
    mr = eqs[E.MagneticStateRates]
    return dict(
        iterate=iterate,
        min_iterations=int(g0.min_iterations),
        max_iterations=int(g0.max_iterations),
        hfact=float(k),
        dim=dim,
        gamma=float(eqs[E.IdealGasEOS].gamma),
        mu0=float(mu0),
        artificial_viscosity=float(eqs[E.MHDAccelerations].artificial_viscosity),
        artificial_thermal_conductivity=float(eqs[E.MHDAccelerations].artificial_thermal_conductivity),
        artificial_magnetic_dissipation=float(mr.artificial_magnetic_dissipation),
        cleaning_speed_factor=float(mr.cleaning_speed_factor),
        cleaning_damping_factor=float(mr.cleaning_damping_factor),
        timestep_factor=float(tsf),
    )


def domain_box(pa, domain):
    """(lo, hi, periodic) of the cell grid for a particle array and a domain.

    A periodic domain gives its bounds.  An open domain (the reference's
    DomainManager() without bounds) gives the particles' starting box plus a
    margin of its size on each side: a particle that later leaves it is
    clamped into an edge cell, which only enlarges candidate sets (the radius
    test decides), so the box sets the speed of the search, never its result.
    """
    if domain is not None and getattr(domain, "is_periodic", False):
        b = domain.bounds
        return (b[0], b[2], b[4]), (b[1], b[3], b[5]), True
    lo, hi = [], []
    for name in ("x", "y", "z"):
        c = np.asarray(pa.properties[name], dtype=np.float64) if name in pa.properties else np.zeros(1)
        c0, c1 = (float(c.min()), float(c.max())) if len(c) else (0.0, 0.0)
        pad = max(0.5 * (c1 - c0), 1.0e-12 * max(abs(c0), abs(c1), 1.0))
        lo.append(c0 - pad)
        hi.append(c1 + pad)
    return tuple(lo), tuple(hi), False


def plan_params(pa, groups, kernel, domain, scheme=None, solver=None):
    """Validate the plan and return (Params, plan constants) for the library:
    constants from the equations, the grid box from the domain, the
    dissipation switch range from the scheme and tf / adaptivity from the
    solver."""
    plan = read_plan(groups)
    dim = plan["dim"]
    if kernel is not None and kernel.dim != dim:
        raise PlanError("kernel dimension differs from the plan's")
    lo, hi, periodic = domain_box(pa, domain)
    p = make_params(
        dim=dim, kernel=kernel.kernel_id, lo=lo, hi=hi, gamma=plan["gamma"], mu0=plan["mu0"], hfact=plan["hfact"],
        artificial_viscosity=plan["artificial_viscosity"],
        artificial_thermal_conductivity=plan["artificial_thermal_conductivity"],
        artificial_magnetic_dissipation=plan["artificial_magnetic_dissipation"],
        cleaning_speed_factor=plan["cleaning_speed_factor"], cleaning_damping_factor=plan["cleaning_damping_factor"],
        timestep_factor=plan["timestep_factor"],
        alpha_min=getattr(scheme, "alpha_min", 0.0), alpha_max=getattr(scheme, "alpha_max", 1.0),
        tf=float(getattr(solver, "tf", 1.0e30)), density="iterate" if plan["iterate"] else "single",
        periodic=periodic, adaptive_timestep=bool(getattr(solver, "adaptive_timestep", True)),
        max_iterations=plan["max_iterations"])
    p.min_iterations = plan["min_iterations"]
    return p, plan


class GhostImagesError(RuntimeError):
    pass


class HIPAccelerationEval:
    def __init__(self, pa, groups, kernel, domain, scheme=None, solver=None):
        from .device import DeviceSystem

        self.pa = pa
        self.groups = groups
        self.params, self.plan = plan_params(pa, groups, kernel, domain, scheme, solver)
        # periodic_mode "ghost" (config.py:99; application.py:63-78 passes it
        # to the DomainManager) runs as the minimum image: the two give the
        # same neighbours, with XIJ differing in rounding only, while every
        # periodic length exceeds twice the support radius.  Below that a
        # ghost layer holds several images of one particle, which this build
        # does not enumerate, so such a state is refused.
        self.ghost = bool(domain is not None and getattr(domain, "periodic_mode", None) == "ghost")
        self.radius = kernel.radius_scale if kernel is not None else 2.0
        self.density_eq = groups[0].equations[0]
        self.system = DeviceSystem(self.params, n=len(pa))
        pa.gpu.attach(self.system)
        pa.gpu.push(*[n for n in FIELDS if n in pa.properties], "itype")

    def check_periodic_images(self) -> None:
        """Raise GhostImagesError when ghost periodicity would need more than
        one image of a particle (a periodic length <= 2 R h_max)."""
        if not self.ghost:
            return
        hmax = float(self.system.field("h").max())
        p = self.params
        for a in range(p.dim):
            length = p.hi[a] - p.lo[a]
            if not length > 2.0 * self.radius * hmax:
                raise GhostImagesError(
                    f"periodic_mode 'ghost': axis {a} is {length:g} long, not more than twice the support "
                    f"radius {self.radius * hmax:g}, so a ghost layer would hold several images of a particle; "
                    "this build runs ghost periodicity as the minimum image, which equals it only below that")

    def initial_acceleration(self) -> int:
        self.check_periodic_images()
        rc = self.system.eval()
        conv = self.system.get("converged")
        self.density_eq.equation_has_converged = 1 if np.all(conv == 1.0) else -1
        return rc

    def step(self, nsteps: int = 1) -> int:
        return self.system.step(nsteps)

    def run(self, nsteps: int, t: float, dt: float):
        """Up to nsteps PEC steps starting from the host clock (t, dt); the
        device stops after the step that reaches tf.  Returns the new (t, dt)
        from the device clock and the number of steps taken."""
        self.check_periodic_images()
        self.system.set_clock(t, dt)
        before = self.system.steps_taken()
        self.system.step(nsteps)
        taken = self.system.steps_taken() - before
        tt, dtt = self.system.clock()
        return float(tt[0]), float(dtt[0]), int(taken)

    def pending_dt(self):
        return self.system.pending_dt()
