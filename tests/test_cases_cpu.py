"""The reference's cases on the host: the close-packed lattice against the
reference's own positions (golden fixture) and against the geometry of a
close packing, each case's initial state against invariants its
construction fixes, the TOML configs of all twelve cases, the parameters
the plan derives from them, and the open-domain neighbour search of the
oracle against brute force.  No GPU."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from helpers import CASE_NX, case_app, case_setup

GOLDEN = Path(__file__).resolve().parent / "golden" / "lattice_golden.npz"
CONFIGS = Path(__file__).resolve().parents[1] / "pysph-mhd_amd" / "configs"


def _golden():
    z = np.load(GOLDEN, allow_pickle=False)
    return [(z["bounds"][c], float(z["spacing"][c]), z["x"][a:b], z["y"][a:b], z["z"][a:b])
            for c, (a, b) in enumerate(zip(z["offsets"][:-1], z["offsets"][1:]))]


@pytest.mark.parametrize("case", range(11))
def test_lattice_matches_reference_positions(case):
    """Host lattice == the reference's close_packed_lattice, bit for bit, on
    the bounds and spacings of its cases (thin slabs, the blast cube, the
    rotor's finer disk lattice, both halves of the shock tube, a jet)."""
    from sphmhd.cases.lattice import close_packed_lattice, lattice_shape

    bounds, spacing, gx, gy, gz = _golden()[case]
    x, y, z = close_packed_lattice(bounds, spacing)
    nx, ny, nz, *_ = lattice_shape(bounds, spacing)
    assert len(x) == nx * ny * nz == len(gx)
    assert ny % 2 == 0 and nz % 3 == 0
    assert np.array_equal(x, gx) and np.array_equal(y, gy) and np.array_equal(z, gz)


@pytest.mark.parametrize("cells", [(6, 4, 3), (5, 6, 6), (8, 4, 9)])
def test_lattice_is_close_packed(cells):
    """In a periodic box of exactly (cx s, cy s sqrt(3/4), cz s sqrt(6)/3)
    every site has 12 nearest neighbours at distance s and none closer: the
    coordination of a close packing.  Dropping a row or layer offset, or
    using the wrong layer height, breaks it."""
    from sphmhd.cases.lattice import close_packed_lattice

    s = 0.1
    cx, cy, cz = cells
    lx, ly, lz = cx * s, cy * s * np.sqrt(0.75), cz * s * np.sqrt(6.0) / 3.0
    x, y, z = close_packed_lattice((0.0, lx, 0.0, ly, 0.0, lz), s)
    assert len(x) == cx * cy * cz
    p = np.stack([x, y, z], axis=1)
    d = p[:, None, :] - p[None, :, :]
    box = np.array([lx, ly, lz])
    d -= box * np.round(d / box)
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, np.inf)
    assert r.min() == pytest.approx(s, rel=1e-12)
    near = (r < s * (1 + 1e-9)).sum(1)
    assert np.all(near == 12)
    # next shell of a close packing sits at sqrt(2) s
    assert np.all(r[r > s * (1 + 1e-9)] >= np.sqrt(2.0) * s * (1 - 1e-9))


def _uniform_mass_check(pa, rho, volume):
    assert np.sum(pa.m) == pytest.approx(rho * volume, rel=1e-12)
    assert np.all(pa.m == pa.m[0])


def _box(app):
    b = app.bounds
    return (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4])


@pytest.mark.parametrize("name", sorted(CASE_NX))
def test_reference_case_builds(name):
    """Each case at a small size: a finite initial state with the shape its
    construction fixes."""
    from sphmhd.cases.lattice import lattice_shape

    app = case_app(name, CASE_NX[name])
    pa = app.particles[0]
    n = len(pa)
    for k in ("x", "y", "z", "u", "v", "w", "rho", "h", "m", "e", "Bx", "By", "Bz"):
        assert np.all(np.isfinite(pa.properties[k])), k
    assert np.all(pa.h > 0) and np.all(pa.m > 0) and np.all(pa.e > 0) and np.all(pa.rho > 0)
    nx = CASE_NX[name]
    b = app.bounds
    if name in ("orstang", "mhdsine", "mhdvortex", "mhdwave", "jadvect", "mhdblast"):
        lx, ly, lz = lattice_shape(b, (b[1] - b[0]) / nx)[:3]
        assert n == lx * ly * lz
        rho = {"orstang": app.gamma * 0.5 / (4.0 * np.pi) * (10.0 / 3.0), "mhdsine": 2.0, "mhdwave": 2.0}.get(name, 1.0)
        _uniform_mass_check(pa, rho, _box(app))
        # every site inside the box
        for a, c in enumerate(("x", "y", "z")):
            assert np.all(pa.properties[c] > b[2 * a]) and np.all(pa.properties[c] < b[2 * a + 1])
    if name == "mhdrotor":
        r = np.hypot(pa.x, pa.y)
        lz = b[5] - b[4]
        inside = r <= 0.1
        # equal masses on a lattice 10^(1/3) finer: density 10 in the disk, 1 outside
        dens_in = inside.sum() * pa.m[0] / (np.pi * 0.01 * lz)
        dens_out = (~inside).sum() * pa.m[0] / ((1.0 - np.pi * 0.01) * lz)
        assert dens_in == pytest.approx(10.0, rel=0.1) and dens_out == pytest.approx(1.0, rel=0.05)
        assert np.allclose(pa.u[inside], -20.0 * pa.y[inside]) and np.all(pa.u[~inside] == 0.0)
    if name == "mhdshock":
        left = pa.x <= 0.0
        # the right lattice is twice as coarse in all three directions: 1/8 the sites
        assert left.sum() == 8 * (~left).sum()
        ymax, zmax = b[3], b[5]
        assert left.sum() * pa.m[0] == pytest.approx(0.5 * (2 * ymax) * (2 * zmax), rel=1e-12)
        walls = pa.itype == 3
        nl = left.sum() // lattice_shape((-0.5, 0.0, -ymax, ymax, -zmax, zmax), 0.5 / nx)[0]
        nr = (~left).sum() // lattice_shape((0.0, 0.5, -ymax, ymax, -zmax, zmax), 1.0 / nx)[0]
        assert walls.sum() == 6 * nl + 6 * nr  # six layers frozen at each end
        assert set(np.unique(pa.itype)) == {1, 3}
        # the tube is padded by 1000 cells so the minimum image never joins its ends
        assert b[0] < -0.5 - 900 * 0.5 / nx and b[1] > 0.5 + 900 / nx
    if name == "alfven":
        for c in ("x", "y", "z"):
            v = pa.properties[c]
            assert np.array_equal(v, v.astype(np.float32).astype(np.float64)), c
        assert pa.m[0] == float(np.float32(6.75 / n))
        bperp = np.hypot(pa.By, pa.Bz)
        assert np.allclose(bperp, 0.1, rtol=1e-12)
        assert np.array_equal(pa.v, pa.By) and np.array_equal(pa.w, pa.Bz)
    if name == "mhdvortex":
        # far from the centre the vortex decays to the uniform state
        far = np.hypot(pa.x, pa.y) > 4.5
        assert np.allclose(pa.u[far], 1.0, atol=1e-3) and np.allclose(pa.Bx[far], 0.0, atol=1e-3)
    if name.endswith("jet"):
        from sphmhd.cases.reference import JET_ANGLES

        jet = app.config.jet
        angles = JET_ANGLES[name]
        per = n // len(angles)
        assert per * len(angles) == n
        for k, deg in enumerate(angles):
            sl = slice(k * per, (k + 1) * per)
            c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
            axial = pa.x[sl] * c + pa.z[sl] * s
            cross = np.hypot(pa.y[sl], -pa.x[sl] * s + pa.z[sl] * c)
            assert np.all(axial >= jet.inner_radius) and np.all(axial <= jet.outer_radius)
            assert np.all(cross < jet.radius)
            # the flow runs along the axis towards the origin
            vax = pa.u[sl] * c + pa.w[sl] * s
            assert np.allclose(vax, -jet.axial_speed, rtol=1e-12)
        vol = np.pi * jet.radius ** 2 * (jet.outer_radius - jet.inner_radius)
        rho0 = jet.number_density * 39.948 * 1.6726231e-27
        assert np.sum(pa.m[:per]) == pytest.approx(rho0 * vol, rel=1e-12)


def test_case_configs_load_and_validate():
    """The twelve TOMLs (the reference's schema and values) load and pass
    validation; a jet TOML without its [jet] table, or with a taper fraction
    out of range, is rejected."""
    from sphmhd.config import load_config
    from sphmhd.run import CASES

    types = {n: a.config_type for n, a in CASES.items()}
    names = sorted(p.stem for p in CONFIGS.glob("*.toml"))
    assert set(CASE_NX) <= set(names)
    for name in CASE_NX:
        cfg = load_config(CONFIGS / f"{name}.toml", types)
        assert cfg.case.name == name
        assert (cfg.numerics.periodic_mode is None) == name.endswith("jet")
    text = (CONFIGS / "1jet.toml").read_text()
    start, end = text.index("[jet]"), text.index("[physics]")
    import tempfile

    with tempfile.TemporaryDirectory() as tmp:
        p = Path(tmp) / "nojet.toml"
        p.write_text(text[:start] + text[end:])
        with pytest.raises(Exception):
            load_config(p, types)
        p.write_text(text.replace("end_taper_fraction = 0.1", "end_taper_fraction = 0.7"))
        with pytest.raises(AssertionError):
            load_config(p, types)


KERNEL_IDS = {"cubic": 0, "quintic": 1, "wendland_c4": 3, "wendland_c6": 4}


@pytest.mark.parametrize("name", sorted(CASE_NX))
def test_case_parameters(name):
    """The plan of each case yields the library parameters its TOML names:
    kernel, hfact, gamma, mu0, dissipation and cleaning factors; periodic
    minimum-image boxes for all but the jets, whose open grid box holds
    every particle with a margin."""
    s = case_setup(name, CASE_NX[name])
    p = s.params[0]
    cfg = s.info["app"].config
    assert p.kernel == KERNEL_IDS[cfg.numerics.kernel]
    assert p.hfact == cfg.numerics.hfact and p.gamma == cfg.physics.gamma and p.mu0 == cfg.physics.mu0
    assert p.artificial_viscosity == cfg.physics.artificial_viscosity
    assert p.cleaning_damping_factor == cfg.numerics.cleaning_damping_factor
    assert p.timestep_factor == pytest.approx(cfg.solver.timestep_factor, rel=1e-15)
    assert p.dim == 3
    if name.endswith("jet"):
        assert p.periodic == 0
        for a, c in enumerate(("x", "y", "z")):
            v = s.fields[c]
            span = v.max() - v.min()
            assert p.lo[a] <= v.min() - 0.49 * span and p.hi[a] >= v.max() + 0.49 * span
    else:
        assert p.periodic == 1
        b = s.info["app"].bounds
        assert tuple(p.lo) == (b[0], b[2], b[4]) and tuple(p.hi) == (b[1], b[3], b[5])


@pytest.mark.parametrize("name", ["1jet", "3jet"])
def test_open_domain_neighbours_match_brute_force(name):
    """Oracle cell search in an open domain == all-pairs search, also with
    particles moved far outside the grid box (clamped into edge cells)."""
    from pyoracle import OracleSystem

    s = case_setup(name, CASE_NX[name])
    f = dict(s.fields)
    x = f["x"].copy()
    x[:40] += 5.0  # beyond the box on the right, as a group
    x[40:45] -= 3.0
    f["x"] = x
    o = OracleSystem(s.params, s.offsets)
    o.set_fields(f)
    oc, ic = o.neighbours()
    ob, ib = o.neighbours_brute()
    assert np.array_equal(oc, ob)
    for i in range(len(x)):
        assert np.array_equal(np.sort(ic[oc[i]:oc[i + 1]]), np.sort(ib[ob[i]:ob[i + 1]])), i
    assert oc[-1] > 20 * len(x)


@pytest.mark.parametrize("name,nx,hbig", [("mhdrotor", 16, None), ("mhdrotor", 16, 0.2), ("mhdshock", 8, None)])
def test_split_cells_neighbours_match_brute_force(name, nx, hbig):
    """Where h varies by 2x or more the cells are split below R h_max (the
    rotor's disk has h 10^(-1/3) of the rest, the shock tube's left half
    half the right's; one particle with a huge h drives the split to its
    cap of 4): the cell search still equals the all-pairs search."""
    from pyoracle import OracleSystem

    s = case_setup(name, nx)
    f = dict(s.fields)
    if hbig is not None:
        f["h"] = f["h"].copy()
        f["h"][3] = hbig
    o = OracleSystem(s.params, s.offsets)
    o.set_fields(f)
    if s.itype is not None:
        o.set_fields({"itype": s.itype})
    oc, ic = o.neighbours()
    ob, ib = o.neighbours_brute()
    assert np.array_equal(oc, ob)
    for i in range(len(f["x"])):
        assert np.array_equal(np.sort(ic[oc[i]:oc[i + 1]]), np.sort(ib[ob[i]:ob[i + 1]])), i
