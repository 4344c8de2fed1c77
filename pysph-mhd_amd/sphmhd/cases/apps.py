"""Case applications of the benchmark workloads, written against the
reference's case interface (cases/*.py: ``bounds`` + ``create_mhd_particles``
returning one ``fluid`` ParticleArray).  Their TOMLs are in configs/.

  OrszagTang2D   orszag_tang_2d  (2-D version of cases/orstang.py)
  MHDRotor2D     rotor_2d        (2-D version of cases/mhdrotor.py)
  MHDBlast3D     blast_3d        (cases/mhdblast.py on a cubic lattice)
  AlfvenWave2D   alfven_2d       (2-D tilted version of cases/alfven.py)
"""
from __future__ import annotations

from ..application import MHDApplication
from ..base import get_particle_array
from . import workloads as B


def _pa(fields):
    keep = ("x", "y", "z", "u", "v", "w", "rho", "h", "m", "e", "Bx", "By", "Bz", "alpha1")
    return [get_particle_array(name="fluid", **{k: fields[k] for k in keep})]


class OrszagTang2D(MHDApplication):
    @property
    def bounds(self):
        return 0.0, 1.0, 0.0, 1.0, 0.0, 0.0

    def create_mhd_particles(self):
        alpha0 = self.config.numerics.alpha_min or 0.0
        return _pa(B.orszag_tang_fields(self.nx, 1.0, self.gamma, self.hfact, alpha0=alpha0))


class MHDRotor2D(MHDApplication):
    @property
    def bounds(self):
        return 0.0, 1.0, 0.0, 1.0, 0.0, 0.0

    def create_mhd_particles(self):
        return _pa(B.rotor_2d(self.nx, kernel=self.kernel, hfact=self.hfact).fields)


class MHDBlast3D(MHDApplication):
    @property
    def bounds(self):
        return 0.0, 1.0, 0.0, 1.0, 0.0, 1.0

    def create_mhd_particles(self):
        return _pa(B.blast_3d(self.nx, kernel=self.kernel, hfact=self.hfact).fields)


class AlfvenWave2D(MHDApplication):
    ny_ratio = 168.0 / 97.0

    @property
    def bounds(self):
        s = B.alfven_2d(self.nx, int(round(self.nx * self.ny_ratio)))
        p = s.params[0]
        return p.lo[0], p.hi[0], p.lo[1], p.hi[1], 0.0, 0.0

    def create_mhd_particles(self):
        s = B.alfven_2d(self.nx, int(round(self.nx * self.ny_ratio)), kernel=self.kernel, hfact=self.hfact)
        return _pa(s.fields)
