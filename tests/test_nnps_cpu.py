"""The oracle's neighbour search against brute force and its own definition.

  * neighbour sets of the cell-list search equal an O(N^2) enumeration of the
    same radius test r^2 < (R max(h_i, h_j))^2 with the minimum image
    (periodic 2-D/3-D, jittered lattices, variable h, boxes so small that the
    stencil wraps onto itself, non-periodic boxes, ensembles);
  * cell keys and the sorted permutation equal a direct numpy evaluation of
    key = floor((x - lo)/c_eff) linearised x-fastest, stably sorted.
"""
from __future__ import annotations

import numpy as np
import pytest


def _setup(kind, seed=0):
    from sphmhd.cases.workloads import blast_3d, ensemble_ot, jittered, orszag_tang_2d, rotor_2d

    if kind == "ot2d":
        return jittered(orszag_tang_2d(24), 0.2, seed=seed)
    if kind == "ot2d_tiny":
        return jittered(orszag_tang_2d(6), 0.2, seed=seed)
    if kind == "rotor":
        s = jittered(rotor_2d(24), 0.2, seed=seed)
        rng = np.random.default_rng(seed)
        s.fields["h"] = s.fields["h"] * rng.uniform(0.7, 1.4, s.n)
        return s
    if kind == "blast3d":
        return jittered(blast_3d(9), 0.2, seed=seed)
    if kind == "ensemble":
        return jittered(ensemble_ot(3, sizes=[8, 13, 10]), 0.2, seed=seed)
    raise ValueError(kind)


def _oracle(s):
    from pyoracle import OracleSystem

    o = OracleSystem(s.params, s.offsets)
    o.set_fields({k: v for k, v in s.fields.items()})
    return o


@pytest.mark.parametrize("kind", ["ot2d", "ot2d_tiny", "rotor", "blast3d", "ensemble"])
def test_cell_list_equals_brute_force(kind):
    s = _setup(kind)
    o = _oracle(s)
    oc, ic = o.neighbours()
    ob, ib = o.neighbours_brute()
    assert np.array_equal(oc, ob)
    assert np.array_equal(ic, ib)
    counts = np.diff(oc)
    assert counts.min() >= 1  # self pair
    assert np.all(ic[oc[:-1]] <= np.arange(s.n))  # sorted lists contain i itself


def test_non_periodic_box():
    from sphmhd.cases.workloads import jittered, orszag_tang_2d

    s = jittered(orszag_tang_2d(16), 0.2, seed=9)
    s.params[0].periodic = 0
    o = _oracle(s)
    oc, ic = o.neighbours()
    ob, ib = o.neighbours_brute()
    assert np.array_equal(oc, ob) and np.array_equal(ic, ib)


def test_minimum_image_pairs_across_the_boundary():
    from sphmhd.cases.workloads import orszag_tang_2d

    s = orszag_tang_2d(16)
    o = _oracle(s)
    oc, ic = o.neighbours()
    # particle 0 sits in the corner cell; its neighbours include the far side
    nb0 = ic[oc[0]: oc[1]]
    x = s.fields["x"]
    assert np.any(x[nb0] > 0.8) and np.any(x[nb0] < 0.2)


def _direct_keys(s, hmax_by_inst):
    from pyoracle import kernel_radius

    keys = np.zeros(s.n, dtype=np.int64)
    base = 0
    for k, p in enumerate(s.params):
        a, b = int(s.offsets[k]), int(s.offsets[k + 1])
        nk = b - a
        cap = max(1, 2 * nk)
        # cells of R h_max, halved in 3-D (uniform h: no further split)
        c = kernel_radius(p.kernel) * hmax_by_inst[k] / (2.0 if p.dim == 3 else 1.0)
        while True:
            n = [max(1, int(np.floor((p.hi[ax] - p.lo[ax]) / c))) if ax < p.dim else 1 for ax in range(3)]
            if n[0] * n[1] * n[2] <= cap:
                break
            c = 2.0 * c
        ceff = [(p.hi[ax] - p.lo[ax]) / n[ax] if ax < p.dim else 1.0 for ax in range(3)]
        coords = []
        for ax, name in enumerate(("x", "y", "z")):
            if ax < p.dim:
                f = np.floor((s.fields[name][a:b] - p.lo[ax]) / ceff[ax])
                coords.append(np.clip(f, 0, n[ax] - 1).astype(np.int64))
            else:
                coords.append(np.zeros(nk, dtype=np.int64))
        keys[a:b] = base + coords[0] + n[0] * (coords[1] + n[1] * coords[2])
        base += n[0] * n[1] * n[2]
    return keys


@pytest.mark.parametrize("kind", ["ot2d", "rotor", "blast3d", "ensemble"])
def test_cell_keys_and_stable_sort(kind):
    s = _setup(kind)
    o = _oracle(s)
    keys, perm = o.cells()
    hmax = [float(s.fields["h"][int(s.offsets[k]): int(s.offsets[k + 1])].max()) for k in range(len(s.params))]
    direct = _direct_keys(s, hmax)
    want_perm = np.argsort(direct, kind="stable")
    assert np.array_equal(perm, want_perm)
    assert np.array_equal(keys.astype(np.int64), direct[want_perm])


@pytest.mark.parametrize("bulk_frac", [1.0, 0.97])
def test_3d_cells_follow_the_bulk_h(bulk_frac):
    """A single 3-D instance sizes its cells from its h histogram: where
    most particles share one h the cells are half of R times that h (to the
    1/16-octave level edge), whatever a few larger h are (here a ball of
    particles at 2.5 times the bulk h, which the h_max rule would size the
    cells for: c = R h_max / 4), and the search still equals the all-pairs
    one."""
    from pyoracle import OracleSystem
    from sphmhd.cases.workloads import blast_3d

    nx = 20
    s = blast_3d(nx)
    f = {k: v.copy() for k, v in s.fields.items()}
    hb = f["h"][0]
    if bulk_frac < 1.0:
        r = np.sqrt((f["x"] - 0.5) ** 2 + (f["y"] - 0.5) ** 2 + (f["z"] - 0.5) ** 2)
        ball = r < np.quantile(r, 1.0 - bulk_frac)
        f["h"][ball] = 2.5 * hb
    o = OracleSystem(s.params, s.offsets)
    o.set_fields(f)
    keys, _ = o.cells()
    n = int(round((int(keys.max()) + 1) ** (1.0 / 3.0)))
    assert n ** 3 == int(keys.max()) + 1  # the corner cell holds a lattice site
    R = 3.0  # quintic
    c = 1.0 / n
    assert 0.5 * R * hb * (1.0 - 1e-12) <= c <= 0.5 * R * hb * 2.0 ** (1.0 / 16.0) * 1.07, (n, c / (R * hb))
    oc, ic = o.neighbours()
    ob, ib = o.neighbours_brute()
    assert np.array_equal(oc, ob)
    for i in range(0, len(f["x"]), 7):
        assert np.array_equal(np.sort(ic[oc[i]:oc[i + 1]]), np.sort(ib[ob[i]:ob[i + 1]])), i
    o.close()
