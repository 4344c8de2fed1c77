"""sphmhd -- MI355X-native SPH-MHD step with the reference's plug-in API.

The modules mirror the reference package (src/ of xsjk/pysph-mhd):
equations, integrator, scheme, solver, application, particles, config, run
and cases.  Equation classes are descriptors of the 10-group plan; the plan
is evaluated by hand-written HIP kernels in libsphmhd.so (see device.py).
"""
__all__ = ["MHDApplication", "MHDScheme"]


def __getattr__(name):
    if name == "MHDApplication":
        from .application import MHDApplication
        return MHDApplication
    if name == "MHDScheme":
        from .scheme import MHDScheme
        return MHDScheme
    raise AttributeError(name)
