"""The reference's cases, under the reference's case names, on its
close-packed lattices (reference: src/cases/*.py, src/run.py:16-29).

Each class keeps the reference's ``bounds`` and initial state: the same
lattice (``close_packed_lattice``, bit for bit), the same fields from the
same expressions (so a reference TOML run here starts from the reference's
state), and the same hooks (``refresh_initial_thermodynamics`` for the
vortex, frozen ``itype`` 3 walls for the shock tube).  The jets run in an
open domain (their TOMLs set no periodic mode), the other cases in the
reference's minimum-image box.
"""
from __future__ import annotations

from dataclasses import dataclass
from math import cos, pi, sin

import numpy as np

from ..application import MHDApplication
from ..base import get_particle_array
from ..config import SimulationConfig
from .lattice import close_packed_lattice


def _fluid(count, **fields):
    zeros = np.zeros(count)
    base = dict(u=zeros, v=zeros, w=zeros, Bx=zeros, By=zeros, Bz=zeros, alpha1=zeros)
    base.update(fields)
    return [get_particle_array(name="fluid", **base)]


def _slab(nx, xlo, xhi, ylo, yhi, k=4.0):
    """Bounds of the quasi-2-D cases: a slab k sqrt(6)/nx thick on each side of z = 0."""
    zmax = k * np.sqrt(6.0) / nx
    return xlo, xhi, ylo, yhi, -zmax, zmax


class OrszagTang(MHDApplication):
    """cases/orstang.py: vortex in [-0.5, 0.5)^2, B0 = 1/sqrt(4 pi), beta = 10/3."""

    @property
    def bounds(self):
        return _slab(self.nx, -0.5, 0.5, -0.5, 0.5)

    def create_mhd_particles(self):
        xmin, xmax, ymin, ymax, zmin, zmax = self.bounds
        x, y, z = close_packed_lattice(self.bounds, (xmax - xmin) / self.nx)
        n = len(x)
        b0 = 1.0 / np.sqrt(4.0 * np.pi)
        pressure = 0.5 * b0 * b0 * (10.0 / 3.0)
        rho = self.gamma * pressure
        mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
        return _fluid(n, x=x, y=y, z=z, u=-np.sin(2.0 * np.pi * (y - ymin)), v=np.sin(2.0 * np.pi * (x - xmin)),
                      rho=np.full(n, rho), h=np.full(n, self.hfact * (mass / rho) ** (1.0 / 3.0)),
                      m=np.full(n, mass), e=np.full(n, pressure / ((self.gamma - 1.0) * rho)),
                      Bx=-b0 * np.sin(2.0 * np.pi * (y - ymin)), By=b0 * np.sin(4.0 * np.pi * (x - xmin)))


class MHDRotor(MHDApplication):
    """cases/mhdrotor.py: a disk of radius 0.1 ten times denser (a lattice
    10^(1/3) finer, equal masses) spinning at 2/0.1, B = 5/sqrt(4 pi) x."""

    @property
    def bounds(self):
        return _slab(self.nx, -0.5, 0.5, -0.5, 0.5, k=2.0)

    def create_mhd_particles(self):
        xmin, xmax, ymin, ymax, zmin, zmax = self.bounds
        spacing = (xmax - xmin) / self.nx
        r0 = 0.1
        ox, oy, oz = close_packed_lattice(self.bounds, spacing)
        keep_out = ox * ox + oy * oy >= r0**2
        ix, iy, iz = close_packed_lattice(self.bounds, spacing * 10.0 ** (-1.0 / 3.0))
        keep_in = ix * ix + iy * iy <= r0**2
        x = np.concatenate((ox[keep_out], ix[keep_in]))
        y = np.concatenate((oy[keep_out], iy[keep_in]))
        z = np.concatenate((oz[keep_out], iz[keep_in]))
        n = len(x)
        inside = np.sqrt(x * x + y * y) <= r0
        rho = np.where(inside, 10.0, 1.0)
        mass = ((xmax - xmin) * (ymax - ymin) + 10.0 * np.pi * r0**2) * (zmax - zmin) / n
        omega = np.where(inside, 2.0 / r0, 0.0)
        return _fluid(n, x=x, y=y, z=z, u=-omega * y, v=omega * x, rho=rho, h=self.hfact * (mass / rho) ** (1.0 / 3.0),
                      m=np.full(n, mass), e=1.0 / ((self.gamma - 1.0) * rho),
                      Bx=np.full(n, 5.0 / np.sqrt(4.0 * np.pi)))


class MHDBlast(MHDApplication):
    """cases/mhdblast.py: p = 100 inside r < 0.125 of the unit cube, B = 10/sqrt(2) (x + z)."""

    @property
    def bounds(self):
        return -0.5, 0.5, -0.5, 0.5, -0.5, 0.5

    def create_mhd_particles(self):
        xmin, xmax, ymin, ymax, zmin, zmax = self.bounds
        x, y, z = close_packed_lattice(self.bounds, (xmax - xmin) / self.nx)
        n = len(x)
        rho = 1.0
        mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
        pressure = np.where(x * x + y * y + z * z < 0.125**2, 100.0, 1.0)
        b = 10.0 / np.sqrt(2.0)
        return _fluid(n, x=x, y=y, z=z, rho=np.full(n, rho), h=np.full(n, self.hfact * (mass / rho) ** (1.0 / 3.0)),
                      m=np.full(n, mass), e=pressure / (rho * (self.gamma - 1.0)), Bx=np.full(n, b),
                      Bz=np.full(n, b))


class AlfvenWave(MHDApplication):
    """cases/alfven.py: circularly polarised wave along x in 3 x 1.5 x 1.5,
    positions and mass rounded through float32 as the reference does."""

    @property
    def bounds(self):
        return -1.5, 1.5, -0.75, 0.75, -0.75, 0.75

    @property
    def dx(self):
        return 3.0 / self.nx

    def create_mhd_particles(self):
        x, y, z = (c.astype(np.float32).astype(float) for c in close_packed_lattice(self.bounds, self.dx))
        phase = 2.0 * np.pi * (x + 1.5)
        by = 0.1 * np.sin(phase)
        bz = 0.1 * np.cos(phase)
        n = len(x)
        mass = float(np.float32(6.75 / n))
        return _fluid(n, x=x, y=y, z=z, v=by, w=bz, rho=np.ones(n), h=np.full(n, self.hfact * mass ** (1.0 / 3.0)),
                      m=np.full(n, mass), e=np.full(n, 0.15), Bx=np.ones(n), By=by, Bz=bz)


class CurrentLoopAdvection(MHDApplication):
    """cases/jadvect.py: a weak field loop of radius 0.3 advected at (2, 1, 0.1 sqrt 5)."""

    @property
    def bounds(self):
        return _slab(self.nx, -1.0, 1.0, -0.5, 0.5)

    def create_mhd_particles(self):
        xmin, xmax, ymin, ymax, zmin, zmax = self.bounds
        x, y, z = close_packed_lattice(self.bounds, (xmax - xmin) / self.nx)
        n = len(x)
        rho = 1.0
        mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
        r = np.sqrt(x * x + y * y)
        inside = r < 0.3
        bx = np.zeros(n)
        by = np.zeros(n)
        bx[inside] = -0.001 * y[inside] / r[inside]
        by[inside] = 0.001 * x[inside] / r[inside]
        return _fluid(n, x=x, y=y, z=z, u=np.full(n, 2.0), v=np.full(n, 1.0), w=np.full(n, 0.1 * np.sqrt(5.0)),
                      rho=np.full(n, rho), h=np.full(n, self.hfact * (mass / rho) ** (1.0 / 3.0)),
                      m=np.full(n, mass), e=np.full(n, 1.5), Bx=bx, By=by)


class MHDShock(MHDApplication):
    """cases/mhdshock.py: Brio-Wu tube, left lattice twice as fine, equal
    masses, six frozen (itype 3) layers at each end; x padded by 1000 cells
    so the minimum image never wraps."""

    @property
    def bounds(self):
        dxl = 0.5 / self.nx
        dxr = 2.0 * dxl
        ymax = 6.0 * dxr * np.sqrt(3.0 / 4.0)
        zmax = 6.0 * dxr * np.sqrt(6.0) / 3.0
        return -0.5 - 1000.0 * dxl, 0.5 + 1000.0 * dxr, -ymax, ymax, -zmax, zmax

    def create_mhd_particles(self):
        assert self.nx % 2 == 0
        _, _, ymin, ymax, zmin, zmax = self.bounds
        dxl = 0.5 / self.nx
        dxr = 2.0 * dxl
        left = close_packed_lattice((-0.5, 0.0, ymin, ymax, zmin, zmax), dxl)
        right = close_packed_lattice((0.0, 0.5, ymin, ymax, zmin, zmax), dxr)
        x, y, z = (np.concatenate((a, b)) for a, b in zip(left, right))
        n = len(x)
        is_left = x <= 0.0
        rho = np.where(is_left, 1.0, 0.125)
        pressure = np.where(is_left, 1.0, 0.1)
        mass = 0.5 * (ymax - ymin) * (zmax - zmin) / len(left[0])
        wall = (x < -0.5 + 6.0 * dxl) | (x > 0.5 - 6.0 * dxr)
        return _fluid(n, x=x, y=y, z=z, rho=rho, h=self.hfact * (mass / rho) ** (1.0 / 3.0), m=np.full(n, mass),
                      e=pressure / ((self.gamma - 1.0) * rho), Bx=np.full(n, 0.75),
                      By=np.where(is_left, 1.0, -1.0), itype=np.where(wall, 3, 1).astype(np.int32))


class MHDSine(MHDApplication):
    """cases/mhdsine.py: static gas with B_y = sin(2 pi (x - xmin))."""

    @property
    def bounds(self):
        return _slab(self.nx, -0.5, 0.5, -0.5, 0.5)

    def create_mhd_particles(self):
        xmin, xmax, ymin, ymax, zmin, zmax = self.bounds
        x, y, z = close_packed_lattice(self.bounds, (xmax - xmin) / self.nx)
        n = len(x)
        rho = 2.0
        mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
        return _fluid(n, x=x, y=y, z=z, rho=np.full(n, rho), h=np.full(n, self.hfact * (mass / rho) ** (1.0 / 3.0)),
                      m=np.full(n, mass), e=np.full(n, 7.5), By=np.sin(2.0 * np.pi * (x - xmin)))


def mhdvortex_pressure(x, y):
    """Pressure of the vortex in equilibrium with its field (cases/mhdvortex.py:10-12)."""
    s = 1.0 - x * x - y * y
    return 1.0 + np.exp(s) * (s - 1.0) / (8.0 * np.pi**2)


class MHDVortex(MHDApplication):
    """cases/mhdvortex.py: advected MHD vortex in [-5, 5)^2; e is refreshed
    from the pressure profile after the density is first summed."""

    @property
    def bounds(self):
        return _slab(self.nx, -5.0, 5.0, -5.0, 5.0)

    def refresh_initial_thermodynamics(self, particle):
        particle.e[:] = mhdvortex_pressure(particle.x, particle.y) / ((self.gamma - 1.0) * particle.rho)

    def create_mhd_particles(self):
        xmin, xmax, ymin, ymax, zmin, zmax = self.bounds
        x, y, z = close_packed_lattice(self.bounds, (xmax - xmin) / self.nx)
        n = len(x)
        rho = 1.0
        mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
        f = np.exp(0.5 * (1.0 - x * x - y * y)) / (2.0 * np.pi)
        return _fluid(n, x=x, y=y, z=z, u=1.0 - y * f, v=1.0 + x * f, w=np.ones(n), rho=np.full(n, rho),
                      h=np.full(n, self.hfact * (mass / rho) ** (1.0 / 3.0)), m=np.full(n, mass),
                      e=mhdvortex_pressure(x, y) / (rho * (self.gamma - 1.0)), Bx=-y * f, By=x * f)


class MHDWave(MHDApplication):
    """cases/mhdwave.py: a Gaussian velocity pulse 3h wide along B = sqrt(2/3) x."""

    @property
    def bounds(self):
        return _slab(self.nx, -2.0, 2.0, -1.0, 1.0)

    def create_mhd_particles(self):
        xmin, xmax, ymin, ymax, zmin, zmax = self.bounds
        x, y, z = close_packed_lattice(self.bounds, (xmax - xmin) / self.nx)
        n = len(x)
        rho = 2.0
        mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
        h = self.hfact * (mass / rho) ** (1.0 / 3.0)
        return _fluid(n, x=x, y=y, z=z, u=0.01 * np.exp(-((x / (3.0 * h)) ** 2)), rho=np.full(n, rho),
                      h=np.full(n, h), m=np.full(n, mass), e=np.full(n, 0.75), Bx=np.full(n, np.sqrt(2.0 / 3.0)))


# ---- jets (cases/jet.py): SI units, argon, open domain ----------------------

ARGON_MASS_NUMBER = 39.948
ARGON_MOLAR_MASS = 0.039948
ATOMIC_MASS = 1.6726231e-27
EV_TO_K = 11604.505
JET_ANGLES = {"1jet": (0.0,), "3jet": (0.0, 30.0, -30.0), "12jet": tuple(30.0 * k for k in range(12))}


@dataclass(frozen=True)
class JetConfig:
    inner_radius: float
    outer_radius: float
    radius: float
    end_taper_fraction: float
    axial_speed: float
    azimuthal_speed: float
    temperature_ev: float
    number_density: float
    axial_field: float
    toroidal_field: float


@dataclass(frozen=True)
class JetSimulationConfig(SimulationConfig):
    jet: JetConfig

    def validate(self):
        super().validate()
        j = self.jet
        assert self.case.name in JET_ANGLES
        assert type(j.inner_radius) is float and j.inner_radius >= 0.0
        assert type(j.outer_radius) is float and j.outer_radius > j.inner_radius
        assert type(j.radius) is float and j.radius > 0.0
        assert type(j.end_taper_fraction) is float and 0.0 < j.end_taper_fraction <= 0.5
        assert type(j.axial_speed) is float and j.axial_speed > 0.0
        assert type(j.azimuthal_speed) is float and j.azimuthal_speed >= 0.0
        assert type(j.temperature_ev) is float and j.temperature_ev > 0.0
        assert type(j.number_density) is float and j.number_density > 0.0
        assert type(j.axial_field) is float and j.axial_field >= 0.0
        assert type(j.toroidal_field) is float and j.toroidal_field >= 0.0


def _jet_profiles(axial, cy, cz, jet):
    """Radial and axial shapes shared by the jet's field and velocity
    (jet.py:95-135): r, 1/r (0 on the axis), r/R, 1 - (r/R)^2, the tanh end
    tapers and their axial derivative."""
    r = np.sqrt(cy * cy + cz * cz)
    inv_r = np.divide(1.0, r, out=np.zeros_like(r), where=r > 0.0)
    ratio = r / jet.radius
    rem = 1.0 - ratio * ratio
    length = jet.outer_radius - jet.inner_radius
    taper = jet.end_taper_fraction * length
    local = 0.5 * (jet.inner_radius + jet.outer_radius) - axial
    ap = (local + 0.5 * length) / taper
    am = (local - 0.5 * length) / taper
    envelope = 0.5 * (np.tanh(ap) - np.tanh(am))
    denv = 0.5 * (1.0 / np.cosh(ap) ** 2 - 1.0 / np.cosh(am) ** 2) / taper
    return r, inv_r, ratio, rem, envelope, denv


TOROIDAL_PEAK = 216.0 / (343.0 * np.sqrt(7.0))  # max of s (1 - s^2)^3 over [0, 1]


class JMXJets(MHDApplication):
    """cases/jet.py: one, three or twelve argon plasma jets converging on the
    origin, each a column of radius R between inner and outer radius along
    its axis, with a divergence-free helical field and a rotating flow."""

    config_type = JetSimulationConfig

    @property
    def bounds(self):
        return -2.0, 2.0, -2.0, 2.0, -2.0, 2.0

    @property
    def dx(self):
        return 2.0 * self.config.jet.radius / self.nx

    def _single_jet(self):
        j = self.config.jet
        a, cy, cz = close_packed_lattice((j.inner_radius, j.outer_radius, -j.radius, j.radius, -j.radius,
                                          j.radius), self.dx)
        inside = cy * cy + cz * cz < j.radius**2
        return a[inside], cy[inside], cz[inside]

    @staticmethod
    def _field(axial, cy, cz, c, s, jet):
        r, inv_r, ratio, rem, envelope, denv = _jet_profiles(axial, cy, cz, jet)
        bax = jet.axial_field * (rem**3 * (1.0 - 5.0 * ratio**2)) * envelope
        brad = -0.5 * jet.axial_field * r * rem**4 * denv
        btor = jet.toroidal_field * ratio * rem**3 / TOROIDAL_PEAK * envelope
        by = (brad * cy + btor * cz) * inv_r
        bz = (brad * cz - btor * cy) * inv_r
        return -bax * c - bz * s, by, -bax * s + bz * c

    @staticmethod
    def _velocity(axial, cy, cz, c, s, jet):
        r, inv_r, ratio, rem, envelope, _ = _jet_profiles(axial, cy, cz, jet)
        vphi = jet.azimuthal_speed * ratio * rem**3 / TOROIDAL_PEAK * envelope
        vy = vphi * cz * inv_r
        vz = -vphi * cy * inv_r
        return -jet.axial_speed * c - vz * s, vy, -jet.axial_speed * s + vz * c

    def create_mhd_particles(self):
        axial, cy, cz = self._single_jet()
        j = self.config.jet
        pos, vel, mag = [], [], []
        for deg in JET_ANGLES[self.config.case.name]:
            ang = deg * pi / 180.0
            c, s = cos(ang), sin(ang)
            pos.append((axial * c - cz * s, cy, axial * s + cz * c))
            vel.append(self._velocity(axial, cy, cz, c, s, j))
            mag.append(self._field(axial, cy, cz, c, s, j))
        x, y, z = (np.concatenate([p[k] for p in pos]) for k in range(3))
        per_jet = len(axial)
        n = len(x)
        rho = j.number_density * ARGON_MASS_NUMBER * ATOMIC_MASS
        pressure = rho * (j.temperature_ev * EV_TO_K) * 8.31 / ARGON_MOLAR_MASS
        volume = pi * j.radius**2 * (j.outer_radius - j.inner_radius) / per_jet
        u, v, w = (np.concatenate([q[k] for q in vel]) for k in range(3))
        bx, by, bz = (np.concatenate([q[k] for q in mag]) for k in range(3))
        return _fluid(n, x=x, y=y, z=z, u=u, v=v, w=w, rho=np.full(n, rho),
                      h=np.full(n, self.hfact * volume ** (1.0 / 3.0)), m=np.full(n, rho * volume),
                      e=np.full(n, pressure / ((self.gamma - 1.0) * rho)), gid=np.arange(n, dtype=np.uint32),
                      Bx=bx, By=by, Bz=bz)


CASES = {
    "1jet": JMXJets,
    "3jet": JMXJets,
    "12jet": JMXJets,
    "alfven": AlfvenWave,
    "jadvect": CurrentLoopAdvection,
    "mhdblast": MHDBlast,
    "mhdrotor": MHDRotor,
    "mhdshock": MHDShock,
    "mhdsine": MHDSine,
    "mhdvortex": MHDVortex,
    "mhdwave": MHDWave,
    "orstang": OrszagTang,
}


# ---- the same initial states generated on the GPU ---------------------------
# Each function returns the case id of sphmhd_case_fields, its constants (the
# particle-independent values, from the host method's own expressions), the
# device positions the kernel reads and, for the jets, the single jet's rows.
# Where the host method raises an array to a power (h of the two densities of
# the rotor and the shock tube), the constants come from the same numpy array
# operation, so that they round as the host arrays do.

def _orstang_device(app, lattice):
    xmin, xmax, ymin, ymax, zmin, zmax = app.bounds
    pos = lattice(app.bounds, (xmax - xmin) / app.nx)
    n = len(pos[0])
    b0 = 1.0 / np.sqrt(4.0 * np.pi)
    pressure = 0.5 * b0 * b0 * (10.0 / 3.0)
    rho = app.gamma * pressure
    mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
    return 5, [xmin, ymin, rho, mass, app.hfact * (mass / rho) ** (1.0 / 3.0),
               pressure / ((app.gamma - 1.0) * rho), b0], pos, None


def _mhdrotor_device(app, lattice):
    import torch

    xmin, xmax, ymin, ymax, zmin, zmax = app.bounds
    spacing = (xmax - xmin) / app.nx
    r0 = 0.1
    ox, oy, oz = lattice(app.bounds, spacing)
    keep_out = ox * ox + oy * oy >= r0**2
    ix, iy, iz = lattice(app.bounds, spacing * 10.0 ** (-1.0 / 3.0))
    keep_in = ix * ix + iy * iy <= r0**2
    pos = tuple(torch.cat((o[keep_out], i[keep_in])) for o, i in ((ox, ix), (oy, iy), (oz, iz)))
    n = len(pos[0])
    mass = ((xmax - xmin) * (ymax - ymin) + 10.0 * np.pi * r0**2) * (zmax - zmin) / n
    h_in, h_out = app.hfact * (mass / np.array([10.0, 1.0])) ** (1.0 / 3.0)
    return 6, [r0, 2.0 / r0, mass, h_in, h_out, app.gamma - 1.0, 5.0 / np.sqrt(4.0 * np.pi)], pos, None


def _mhdblast_device(app, lattice):
    xmin, xmax, ymin, ymax, zmin, zmax = app.bounds
    pos = lattice(app.bounds, (xmax - xmin) / app.nx)
    n = len(pos[0])
    rho = 1.0
    mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
    return 7, [rho, mass, app.hfact * (mass / rho) ** (1.0 / 3.0), app.gamma - 1.0, 10.0 / np.sqrt(2.0),
               0.125**2], pos, None


def _alfven_device(app, lattice):
    pos = lattice(app.bounds, app.dx)  # rounded through float32 by the kernel
    mass = float(np.float32(6.75 / len(pos[0])))
    return 8, [mass, app.hfact * mass ** (1.0 / 3.0)], pos, None


def _jadvect_device(app, lattice):
    xmin, xmax, ymin, ymax, zmin, zmax = app.bounds
    pos = lattice(app.bounds, (xmax - xmin) / app.nx)
    n = len(pos[0])
    rho = 1.0
    mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
    return 9, [rho, mass, app.hfact * (mass / rho) ** (1.0 / 3.0), 0.1 * np.sqrt(5.0)], pos, None


def _mhdshock_device(app, lattice):
    import torch

    assert app.nx % 2 == 0
    _, _, ymin, ymax, zmin, zmax = app.bounds
    dxl = 0.5 / app.nx
    dxr = 2.0 * dxl
    left = lattice((-0.5, 0.0, ymin, ymax, zmin, zmax), dxl)
    right = lattice((0.0, 0.5, ymin, ymax, zmin, zmax), dxr)
    pos = tuple(torch.cat((a, b)) for a, b in zip(left, right))
    mass = 0.5 * (ymax - ymin) * (zmax - zmin) / len(left[0])
    h_left, h_right = app.hfact * (mass / np.array([1.0, 0.125])) ** (1.0 / 3.0)
    return 10, [mass, h_left, h_right, app.gamma - 1.0, -0.5 + 6.0 * dxl, 0.5 - 6.0 * dxr], pos, None


def _mhdsine_device(app, lattice):
    xmin, xmax, ymin, ymax, zmin, zmax = app.bounds
    pos = lattice(app.bounds, (xmax - xmin) / app.nx)
    n = len(pos[0])
    rho = 2.0
    mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
    return 11, [xmin, rho, mass, app.hfact * (mass / rho) ** (1.0 / 3.0)], pos, None


def _mhdvortex_device(app, lattice):
    xmin, xmax, ymin, ymax, zmin, zmax = app.bounds
    pos = lattice(app.bounds, (xmax - xmin) / app.nx)
    n = len(pos[0])
    rho = 1.0
    mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
    return 12, [rho, mass, app.hfact * (mass / rho) ** (1.0 / 3.0), app.gamma - 1.0, 8.0 * np.pi**2], pos, None


def _mhdwave_device(app, lattice):
    xmin, xmax, ymin, ymax, zmin, zmax = app.bounds
    pos = lattice(app.bounds, (xmax - xmin) / app.nx)
    n = len(pos[0])
    rho = 2.0
    mass = rho * (xmax - xmin) * (ymax - ymin) * (zmax - zmin) / n
    return 13, [rho, mass, app.hfact * (mass / rho) ** (1.0 / 3.0), np.sqrt(2.0 / 3.0)], pos, None


def _jets_device(app, lattice):
    import torch

    j = app.config.jet
    a, cy, cz = lattice((j.inner_radius, j.outer_radius, -j.radius, j.radius, -j.radius, j.radius), app.dx)
    inside = cy * cy + cz * cz < j.radius**2
    aux = torch.stack((a[inside], cy[inside], cz[inside])).contiguous()
    per_jet = aux.shape[1]
    rho = j.number_density * ARGON_MASS_NUMBER * ATOMIC_MASS
    pressure = rho * (j.temperature_ev * EV_TO_K) * 8.31 / ARGON_MOLAR_MASS
    volume = pi * j.radius**2 * (j.outer_radius - j.inner_radius) / per_jet
    consts = [rho, rho * volume, app.hfact * volume ** (1.0 / 3.0), pressure / ((app.gamma - 1.0) * rho), j.radius,
              j.inner_radius, j.outer_radius, j.end_taper_fraction * (j.outer_radius - j.inner_radius),
              j.axial_speed, j.azimuthal_speed, j.axial_field, j.toroidal_field, TOROIDAL_PEAK]
    for deg in JET_ANGLES[app.config.case.name]:
        ang = deg * pi / 180.0
        consts += [cos(ang), sin(ang)]
    return 14, consts, None, aux


_DEVICE = {
    OrszagTang: _orstang_device, MHDRotor: _mhdrotor_device, MHDBlast: _mhdblast_device,
    AlfvenWave: _alfven_device, CurrentLoopAdvection: _jadvect_device, MHDShock: _mhdshock_device,
    MHDSine: _mhdsine_device, MHDVortex: _mhdvortex_device, MHDWave: _mhdwave_device, JMXJets: _jets_device,
}


def device_particles(app, device="cuda"):
    """GPU twin of ``app.create_mhd_particles()``: the lattice from
    sphmhd_hcp_lattice, masked and joined as the case does with torch, then
    the fields from sphmhd_case_fields.  Returns (rows, itype): an FP64 device
    tensor [NFIELDS, n] in the bound layout (rows the case does not set are
    zero) and an int32 device tensor of n particle types (1, or 3 for the
    shock tube's walls)."""
    import torch

    from .device import new_rows, set_positions, write_case
    from .lattice import close_packed_lattice_device

    def lattice(bounds, spacing):
        return close_packed_lattice_device(bounds, spacing, device=device)

    maker = _DEVICE.get(type(app))
    if maker is None:
        raise TypeError(f"no device generator for {type(app).__name__}")
    with torch.cuda.device(torch.device(device)):
        case_id, consts, pos, aux = maker(app, lattice)
        n = aux.shape[1] * (len(consts) - 13) // 2 if aux is not None else len(pos[0])
        rows = new_rows(n, device=device)
        if pos is not None:
            set_positions(rows, *pos)
        itype = torch.ones(n, dtype=torch.int32, device=device)
        write_case(case_id, consts, rows, itype=itype, aux=aux)
    return rows, itype
