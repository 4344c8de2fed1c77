"""Initial conditions of the benchmark workloads (host numpy, FP64).

Each builder returns a ``Setup``: per-instance parameters, CSR instance
offsets and the initial properties (the ones an Application's
create_particles fills: x, y, z, u, v, w, rho, h, m, e, Bx, By, Bz, alpha1).
Particles sit at lattice cell centres (i + 1/2) dx of periodic boxes.
Every lattice setup also carries ``device``: per instance, the case id and
constants with which sphmhd_case_fields writes the same fields on the GPU
(``cases.device.load_setup``); ``host_fields=False`` skips the host arrays.

  orszag_tang_2d   2-D Orszag-Tang vortex (the reference case orstang.py:18-44
                   on a square lattice: quintic kernel, h = 1.2 dx,
                   alpha in [0.1, 1])
  rotor_2d         2-D MHD rotor (mhdrotor.py) with the linear density taper
                   over 0.1 < r < 0.115, variable mass on a uniform lattice
  blast_3d         3-D MHD blast (mhdblast.py) on a cubic lattice, quintic kernel
                   with about 60 neighbours
  ensemble_ot      64 Orszag-Tang instances of uneven size with jittered
                   field amplitude and dissipation/cleaning parameters
  ensemble_of      a batch of independent single-instance setups
  alfven_2d        circularly polarised Alfven wave tilted by 30 degrees
                   (alfven.py), with viscosity and cleaning switched on
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .._abi import Params, make_params


@dataclass
class Setup:
    name: str
    params: list[Params]
    offsets: np.ndarray
    fields: dict[str, np.ndarray]
    itype: np.ndarray | None = None
    info: dict = field(default_factory=dict)
    device: list | None = None  # per instance: (case id, constants), or None

    @property
    def n(self) -> int:
        return int(self.offsets[-1])


def _lattice_2d(nx: int, ny: int, lx: float, ly: float, x0: float = 0.0, y0: float = 0.0):
    dx, dy = lx / nx, ly / ny
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    x = x0 + (i.ravel() + 0.5) * dx
    y = y0 + (j.ravel() + 0.5) * dy
    return x, y, dx, dy


def _lattice_3d(n: int, l: float = 1.0):
    dx = l / n
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # x fastest, then y, then z
    x = (k.ravel() + 0.5) * dx
    y = (j.ravel() + 0.5) * dx
    z = (i.ravel() + 0.5) * dx
    return x, y, z, dx


# case ids of sphmhd_case_fields (include/sphmhd.h)
CASE_ORSZAG_TANG_2D, CASE_ROTOR_2D, CASE_BLAST_3D, CASE_ALFVEN_2D = 1, 2, 3, 4


def orszag_tang_constants(nx: int, b0_factor: float = 1.0, gamma: float = 5.0 / 3.0, hfact: float = 1.2,
                          alpha0: float = 0.1) -> list[float]:
    """nx, dx, rho0, m, h, e, B0, alpha0 of the vortex on an nx^2 lattice."""
    dx = 1.0 / nx
    rho0 = 25.0 / (36.0 * np.pi)
    p0 = 5.0 / (12.0 * np.pi)
    b0 = b0_factor / np.sqrt(4.0 * np.pi)
    return [nx, dx, rho0, rho0 * dx * dx, hfact * dx, p0 / ((gamma - 1.0) * rho0), b0, alpha0]


def orszag_tang_fields(nx: int, b0_factor: float = 1.0, gamma: float = 5.0 / 3.0, hfact: float = 1.2,
                       alpha0: float = 0.1):
    x, y, dx, _ = _lattice_2d(nx, nx, 1.0, 1.0)
    n = x.size
    _, _, rho0, m, h, e, b0, _ = orszag_tang_constants(nx, b0_factor, gamma, hfact, alpha0)
    two_pi = 2.0 * np.pi
    f = {
        "x": x, "y": y, "z": np.zeros(n),
        "u": -np.sin(two_pi * y), "v": np.sin(two_pi * x), "w": np.zeros(n),
        "rho": np.full(n, rho0),
        "m": np.full(n, m),
        "h": np.full(n, h),
        "e": np.full(n, e),
        "Bx": -b0 * np.sin(two_pi * y), "By": b0 * np.sin(2.0 * two_pi * x), "Bz": np.zeros(n),
        "alpha1": np.full(n, alpha0),
    }
    return f


def orszag_tang_2d(nx: int = 128, *, kernel: str = "quintic", hfact: float = 1.2, b0_factor: float = 1.0,
                   artificial_viscosity: float = 1.0, cleaning_speed_factor: float = 1.0,
                   cleaning_damping_factor: float = 1.0, tf: float = 1.0e30, host_fields: bool = True) -> Setup:
    gamma = 5.0 / 3.0
    p = make_params(dim=2, kernel=kernel, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 0.0), gamma=gamma, hfact=hfact,
                    artificial_viscosity=artificial_viscosity, artificial_thermal_conductivity=1.0,
                    artificial_magnetic_dissipation=1.0, cleaning_speed_factor=cleaning_speed_factor,
                    cleaning_damping_factor=cleaning_damping_factor, timestep_factor=1.0,
                    alpha_min=0.1, alpha_max=1.0, tf=tf)
    f = orszag_tang_fields(nx, b0_factor, gamma, hfact, alpha0=0.1) if host_fields else {}
    dev = [(CASE_ORSZAG_TANG_2D, orszag_tang_constants(nx, b0_factor, gamma, hfact, alpha0=0.1))]
    return Setup(f"orszag_tang_2d_{nx}", [p], np.array([0, nx * nx], dtype=np.int64), f,
                 info={"nx": nx, "kernel": kernel}, device=dev)


def rotor_2d(nx: int = 512, *, kernel: str = "cubic", hfact: float = 1.2, tf: float = 1.0e30,
             host_fields: bool = True) -> Setup:
    gamma = 1.4
    b = 5.0 / np.sqrt(4.0 * np.pi)
    p = make_params(dim=2, kernel=kernel, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 0.0), gamma=gamma, hfact=hfact,
                    alpha_min=0.0, alpha_max=1.0, tf=tf)
    dev = [(CASE_ROTOR_2D, [nx, 1.0 / nx, hfact * (1.0 / nx), gamma, b])]
    info = {"nx": nx, "kernel": kernel}
    if not host_fields:
        return Setup(f"rotor_2d_{nx}", [p], np.array([0, nx * nx], dtype=np.int64), {}, info=info, device=dev)
    x, y, dx, _ = _lattice_2d(nx, nx, 1.0, 1.0)
    n = x.size
    r0, r1, omega = 0.1, 0.115, 20.0
    rx, ry = x - 0.5, y - 0.5
    r = np.sqrt(rx * rx + ry * ry)
    f_taper = np.clip((r1 - r) / (r1 - r0), 0.0, 1.0)
    inside = r < r0
    rho = np.where(inside, 10.0, 1.0 + 9.0 * f_taper)
    # rigid rotation inside r0, the same angular momentum density tapered to 0 at r1
    safe_r = np.where(r > 0.0, r, 1.0)
    vfac = np.where(inside, omega, f_taper * omega * r0 / safe_r)
    u = -vfac * ry
    v = vfac * rx
    f = {
        "x": x, "y": y, "z": np.zeros(n),
        "u": u, "v": v, "w": np.zeros(n),
        "rho": rho, "m": rho * dx * dx, "h": np.full(n, hfact * dx),
        "e": 1.0 / ((gamma - 1.0) * rho),
        "Bx": np.full(n, b), "By": np.zeros(n), "Bz": np.zeros(n),
        "alpha1": np.zeros(n),
    }
    return Setup(f"rotor_2d_{nx}", [p], np.array([0, n], dtype=np.int64), f, info=info, device=dev)


def blast_3d(nx: int = 128, *, kernel: str = "quintic", hfact: float = 0.81, tf: float = 1.0e30,
             host_fields: bool = True) -> Setup:
    gamma = 5.0 / 3.0
    b = 1.0 / np.sqrt(2.0)
    p = make_params(dim=3, kernel=kernel, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), gamma=gamma, hfact=hfact,
                    alpha_min=0.0, alpha_max=1.0, tf=tf)
    dx = 1.0 / nx
    dev = [(CASE_BLAST_3D, [nx, dx, dx ** 3, hfact * dx, gamma, b, 0.1 ** 2])]
    info = {"nx": nx, "kernel": kernel}
    if not host_fields:
        return Setup(f"blast_3d_{nx}", [p], np.array([0, nx ** 3], dtype=np.int64), {}, info=info, device=dev)
    x, y, z, dx = _lattice_3d(nx)
    n = x.size
    r2 = (x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2
    pres = np.where(r2 < 0.1 ** 2, 100.0, 1.0)
    f = {
        "x": x, "y": y, "z": z,
        "u": np.zeros(n), "v": np.zeros(n), "w": np.zeros(n),
        "rho": np.ones(n), "m": np.full(n, dx ** 3), "h": np.full(n, hfact * dx),
        "e": pres / (gamma - 1.0),
        "Bx": np.full(n, b), "By": np.full(n, b), "Bz": np.zeros(n),
        "alpha1": np.zeros(n),
    }
    return Setup(f"blast_3d_{nx}", [p], np.array([0, n], dtype=np.int64), f, info=info, device=dev)


def ensemble_table(count: int = 64, seed: int = 0):
    """Seeded per-instance table: lattice size, B0 factor and the
    (artificial_viscosity, cleaning_speed_factor, cleaning_damping_factor)
    jitter, drawn in this order."""
    rng = np.random.default_rng(seed)
    n = rng.integers(64, 257, count)
    b0f = rng.uniform(0.5, 1.5, count)
    jit = rng.uniform(0.8, 1.2, (count, 3))
    return n, b0f, jit


def ensemble_ot(count: int = 64, seed: int = 0, *, sizes=None, kernel: str = "quintic", hfact: float = 1.2,
                tf: float = 1.0e30, host_fields: bool = True) -> Setup:
    n_tab, b0f, jit = ensemble_table(count, seed)
    if sizes is not None:
        n_tab = np.asarray(sizes, dtype=np.int64)[:count]
    params, parts, dev = [], [], []
    offsets = [0]
    for k in range(count):
        s = orszag_tang_2d(int(n_tab[k]), kernel=kernel, hfact=hfact, b0_factor=float(b0f[k]),
                           artificial_viscosity=float(jit[k, 0]), cleaning_speed_factor=float(jit[k, 1]),
                           cleaning_damping_factor=float(jit[k, 2]), tf=tf, host_fields=host_fields)
        params.append(s.params[0])
        parts.append(s.fields)
        dev += s.device
        offsets.append(offsets[-1] + s.n)
    fields = {name: np.concatenate([p[name] for p in parts]) for name in parts[0]}
    return Setup(f"ensemble_ot_{count}", params, np.asarray(offsets, dtype=np.int64), fields,
                 info={"sizes": n_tab.tolist(), "b0_factor": b0f.tolist(), "jitter": jit.tolist()}, device=dev)


def ensemble_of(setups) -> Setup:
    """One batch of independent single-instance setups, in the given order."""
    params, offsets = [], [0]
    for s in setups:
        if len(s.params) != 1:
            raise ValueError("ensemble_of takes single-instance setups")
        params.append(s.params[0])
        offsets.append(offsets[-1] + s.n)
    fields = {name: np.concatenate([s.fields[name] for s in setups]) for name in setups[0].fields}
    dev = [d for s in setups for d in s.device] if all(s.device for s in setups) else None
    return Setup(f"ensemble_of_{len(setups)}", params, np.asarray(offsets, dtype=np.int64), fields,
                 info={"members": [s.name for s in setups]}, device=dev)


def alfven_2d(nx: int = 97, ny: int = 168, *, amplitude: float = 0.1, kernel: str = "wendland_c6",
              hfact: float = 1.9, alpha_max: float = 1.0, cleaning_damping_factor: float = 1.0,
              tf: float = 1.0e30, host_fields: bool = True) -> Setup:
    """Travelling circularly polarised Alfven wave along (cos 30, sin 30) in the
    periodic box 1/cos30 x 1/sin30 (one wavelength along each diagonal)."""
    gamma = 5.0 / 3.0
    lx, ly = 1.0 / np.cos(np.pi / 6.0), 1.0 / np.sin(np.pi / 6.0)
    ca, sa = np.cos(np.pi / 6.0), np.sin(np.pi / 6.0)
    p = make_params(dim=2, kernel=kernel, lo=(0.0, 0.0, 0.0), hi=(lx, ly, 0.0), gamma=gamma, hfact=hfact,
                    artificial_viscosity=1.0, artificial_thermal_conductivity=0.0,
                    artificial_magnetic_dissipation=0.0, cleaning_speed_factor=1.0,
                    cleaning_damping_factor=cleaning_damping_factor, timestep_factor=0.5, alpha_min=0.0,
                    alpha_max=alpha_max, tf=tf)
    dx, dy = lx / nx, ly / ny
    rho, pres = 1.0, 0.1
    dev = [(CASE_ALFVEN_2D, [nx, dx, dy, amplitude, ca, sa, rho * dx * dy, hfact * np.sqrt(dx * dy),
                             pres / ((gamma - 1.0) * rho)])]
    info = {"nx": nx, "ny": ny, "amplitude": amplitude}
    if not host_fields:
        return Setup(f"alfven_2d_{nx}x{ny}", [p], np.array([0, nx * ny], dtype=np.int64), {}, info=info,
                     device=dev)
    x, y, dx, dy = _lattice_2d(nx, ny, lx, ly)
    n = x.size
    xpar = x * ca + y * sa
    ph = 2.0 * np.pi * xpar
    bperp = amplitude * np.sin(ph)
    bz = amplitude * np.cos(ph)
    # parallel field 1, perpendicular component along (-sin, cos)
    bx = 1.0 * ca - bperp * sa
    by = 1.0 * sa + bperp * ca
    vx = -bperp * sa
    vy = bperp * ca
    f = {
        "x": x, "y": y, "z": np.zeros(n),
        "u": vx, "v": vy, "w": bz.copy(),
        "rho": np.full(n, rho), "m": np.full(n, rho * dx * dy),
        "h": np.full(n, hfact * np.sqrt(dx * dy)),
        "e": np.full(n, pres / ((gamma - 1.0) * rho)),
        "Bx": bx, "By": by, "Bz": bz,
        "alpha1": np.zeros(n),
    }
    # the fields are linear in the amplitude: their exact derivative, the
    # initial tangent of a rollout differentiated with respect to it
    s_ph, c_ph = np.sin(ph), np.cos(ph)
    d_amp = {"u": -s_ph * sa, "v": s_ph * ca, "w": c_ph.copy(), "Bx": -s_ph * sa, "By": s_ph * ca, "Bz": c_ph.copy()}
    info["tangents"] = {"amplitude": d_amp}
    return Setup(f"alfven_2d_{nx}x{ny}", [p], np.array([0, n], dtype=np.int64), f, info=info, device=dev)


def jittered(setup: Setup, frac: float = 0.1, seed: int = 1) -> Setup:
    """Copy of a lattice setup with positions jittered by +-frac*dx (uniform),
    to avoid exact neighbour-radius ties in tests (host fields only: the
    jittered positions are no lattice the device generator knows)."""
    if not setup.fields:
        raise ValueError("jittered needs a setup built with host fields")
    rng = np.random.default_rng(seed)
    f = {k: v.copy() for k, v in setup.fields.items()}
    p0 = setup.params[0]
    for k in range(len(setup.offsets) - 1):
        a, b = int(setup.offsets[k]), int(setup.offsets[k + 1])
        p = setup.params[k]
        nloc = b - a
        dx = (p.hi[0] - p.lo[0]) / max(1.0, round(nloc ** (1.0 / p.dim)))
        for ax, name in enumerate(("x", "y", "z")[: p.dim]):
            f[name][a:b] += rng.uniform(-frac, frac, nloc) * dx
            L = p.hi[ax] - p.lo[ax]
            f[name][a:b] = p.lo[ax] + np.mod(f[name][a:b] - p.lo[ax], L)
    del p0
    return Setup(setup.name + "_jit", setup.params, setup.offsets, f, setup.itype, dict(setup.info))
