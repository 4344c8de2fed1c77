"""periodic_mode = "ghost" (reference: config.py:99, application.py:63-78)
against the minimum image this build runs it as.

A ghost layer copies every particle within the support radius of a periodic
face to the far side (its position shifted by +-L, so rounded), and the
neighbour search then runs without any wrap.  While every periodic length
exceeds twice the support radius R h_max, each particle has at most one
image within reach of any other, so the ghost search finds exactly the
minimum-image neighbour sets; XIJ differs only by the rounding of the shifted
copy.  Here the ghost search is emulated independently (explicit copies,
plain differences, a KD-tree over real and ghost particles) and compared,
set for set, with the oracle's minimum-image lists on the reference cases
that can set the mode, at sizes where the condition holds; the -m gpu test
does the same with the library's lists after a warm start and two steps.
Below the condition (the reference's vortex slab, 1.22 thick against a
support of ~5) a ghost layer holds several images of one particle, which
the library refuses (backend.GhostImagesError) rather than answer
differently; that refusal is tested too.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import case_setup, init_fields

# the periodic reference cases (the jets are open), at sizes where every
# periodic length exceeds 2 R h_max after the warm start
GHOST_CASES = {"orstang": 16, "mhdrotor": 16, "mhdblast": 10, "alfven": 24, "jadvect": 16, "mhdshock": 16,
               "mhdsine": 16, "mhdwave": 16}


def ghost_neighbours(x, h, lo, hi, dim, radius_scale):
    """Neighbour index sets of a ghost-layer search: copies of the particles
    within R h_max of each periodic face, shifted by +-L per axis (and their
    combinations across edges and corners), then the symmetric radius test
    r^2 < (R max(h_i, h_j))^2 over real and ghost particles with plain
    differences.  Returns (sets, multi): multi counts particles that see two
    images of one neighbour."""
    from scipy.spatial import cKDTree

    n = x.shape[0]
    L = np.array([hi[a] - lo[a] for a in range(dim)])
    reach = radius_scale * float(h.max())
    pos, ids = [x[:, :dim]], [np.arange(n)]
    shifts = np.array(np.meshgrid(*[[-1, 0, 1]] * dim, indexing="ij")).reshape(dim, -1).T
    for sh in shifts:
        if not sh.any():
            continue
        sel = np.ones(n, dtype=bool)
        for a in range(dim):
            if sh[a] == 1:    # copy of particles near the low face, placed beyond the high face
                sel &= x[:, a] < lo[a] + reach
            elif sh[a] == -1:
                sel &= x[:, a] >= hi[a] - reach
        pos.append(x[sel, :dim] + sh * L)
        ids.append(np.nonzero(sel)[0])
    allpos = np.concatenate(pos)
    allid = np.concatenate(ids)
    allh = h[allid]
    tree = cKDTree(allpos)
    sets, multi = [], 0
    for i in range(n):
        cand = np.asarray(tree.query_ball_point(x[i, :dim], reach * (1 + 1e-12)), dtype=np.int64)
        d = x[i, :dim] - allpos[cand]
        r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
        if dim == 3:
            r2 = r2 + d[:, 2] * d[:, 2]
        rad = radius_scale * np.maximum(h[i], allh[cand])
        got = allid[cand[r2 < rad * rad]]
        if len(np.unique(got)) != len(got):
            multi += 1
        sets.append(np.unique(got))
    return sets, multi


def _positions(sysm, dim):
    cols = [sysm.get(c) for c in ("x", "y", "z")[:dim]]
    if dim == 2:
        cols.append(np.zeros_like(cols[0]))
    return np.stack(cols, axis=1)


def _compare(sets, csr):
    offs, idx = csr
    for i, want in enumerate(sets):
        got = idx[offs[i]: offs[i + 1]]
        assert np.array_equal(got, want), (i, got, want)


@pytest.mark.parametrize("name", sorted(GHOST_CASES))
def test_ghost_layer_equals_minimum_image(name):
    from pyoracle import OracleSystem, kernel_radius

    s = case_setup(name, GHOST_CASES[name])
    p = s.params[0]
    o = OracleSystem(s.params, s.offsets, threads=8)
    o.set_fields(init_fields(s))
    o.warm_start()
    h = o.get("h")
    R = kernel_radius(p.kernel)
    assert all(p.hi[a] - p.lo[a] > 2 * R * h.max() for a in range(p.dim))
    sets, multi = ghost_neighbours(_positions(o, p.dim), h, p.lo, p.hi, p.dim, R)
    assert multi == 0
    _compare(sets, o.neighbours())


def test_ghost_with_several_images_is_refused():
    """The reference's vortex slab: a periodic length below 2 R h_max, where
    a ghost layer would hold several images of one particle; the ghost
    search finds them, and the library's evaluator refuses the mode."""
    from pyoracle import OracleSystem, kernel_radius

    s = case_setup("mhdvortex", 16)
    p = s.params[0]
    o = OracleSystem(s.params, s.offsets, threads=8)
    o.set_fields(init_fields(s))
    h = o.get("h")
    R = kernel_radius(p.kernel)
    assert min(p.hi[a] - p.lo[a] for a in range(p.dim)) <= 2 * R * h.max()
    _, multi = ghost_neighbours(_positions(o, p.dim), h, p.lo, p.hi, p.dim, R)
    assert multi > 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GHOST_CASES))
def test_ghost_layer_equals_library_lists(gpu, name):
    """The library's lists after the warm start and two steps (particles
    moved, h summed) equal the ghost search's sets."""
    from pyoracle import kernel_radius
    from sphmhd.device import DeviceSystem

    s = case_setup(name, GHOST_CASES[name])
    p = s.params[0]
    dev = DeviceSystem(s.params, s.offsets)
    dev.set_fields(init_fields(s))
    dev.warm_start()
    dev.step(2)
    h = dev.get("h")
    R = kernel_radius(p.kernel)
    sets, multi = ghost_neighbours(_positions(dev, p.dim), h, p.lo, p.hi, p.dim, R)
    assert multi == 0
    _compare(sets, dev.neighbours())


@pytest.mark.gpu
def test_ghost_mode_refused_on_thin_slab(gpu, tmp_path):
    from helpers import case_app
    from sphmhd.backend import GhostImagesError

    app = case_app("mhdvortex", 16, directory=str(tmp_path / "out"))
    app.periodic_mode = "ghost"
    app.domain = app.create_domain()
    from sphmhd.backend import HIPAccelerationEval

    app.evaluator = HIPAccelerationEval(app.particles[0], app.scheme.get_equations(), app.solver.kernel,
                                        app.domain, app.scheme, app.solver)
    app.solver.integrator.evaluator = app.evaluator
    with pytest.raises(GhostImagesError):
        app.warm_start()
