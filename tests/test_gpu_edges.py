"""Edge cases of the HIP path against the oracle: no particles, one
particle (the density iteration cannot converge: the warm start must fail
the way the reference's assert does, application.py:120), an ensemble with
an empty instance between two others, and particles placed exactly on the
periodic box faces."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import RATES, STATE, compare, init_fields

pytestmark = pytest.mark.gpu

TOL = 1.0e-10


def _ot(nx, seed):
    from sphmhd.cases.workloads import jittered, orszag_tang_2d

    return jittered(orszag_tang_2d(nx), 0.1, seed=seed)


def test_no_particles_is_rejected(gpu):
    """A context needs at least one particle (an empty instance inside an
    ensemble is fine, below): creating one with none fails with EINVAL."""
    from sphmhd._abi import SPHMHDError
    from sphmhd.device import DeviceSystem

    p = _ot(16, 1).params[0]
    with pytest.raises(SPHMHDError) as err:
        DeviceSystem([p], np.array([0, 0]))
    assert err.value.code == -1


def test_one_particle_does_not_converge(gpu):
    """A lone particle's density sum m W(0, h) scales like its h-relation
    rho = m (hfact / h)^d, so the iteration has no fixed point: the warm
    start ends with SPHMHD_ENOTCONV on both sides."""
    from pyoracle import OracleSystem
    from sphmhd._abi import SPHMHDError
    from sphmhd.device import DeviceSystem

    s = _ot(16, 1)
    one = {k: v[:1].copy() for k, v in init_fields(s).items()}
    dev = DeviceSystem(s.params, np.array([0, 1]))
    dev.set_fields(one)
    orc = OracleSystem(s.params, np.array([0, 1]))
    orc.set_fields(one)
    with pytest.raises(RuntimeError, match="-5"):
        orc.warm_start()
    with pytest.raises(SPHMHDError) as err:
        dev.warm_start()
    assert err.value.code == -5


def test_ragged_ensemble_with_empty_instance(gpu):
    """Instances of 576, 0 and 1024 particles (in that order) in one context:
    each non-empty one matches the oracle run of the same ensemble, and the
    empty one keeps the clock the oracle gives it.

    Floor: divBsymm, divBdiff, psi, apsi and aBevol* are measured against
    1e-3 of their natural scale (|B|/h, rho|B|/h, |B|, v_w|B|/h,
    v_w|B|/(rho h)), since they are round-off where B is uniform; every
    other field against its own norm."""
    from pyoracle import OracleSystem
    from sphmhd.device import DeviceSystem

    a, b = _ot(24, 2), _ot(32, 3)
    params = [a.params[0], a.params[0], b.params[0]]
    offsets = np.array([0, a.n, a.n, a.n + b.n])
    fields = {k: np.concatenate([init_fields(a)[k], init_fields(b)[k]]) for k in init_fields(a)}
    dev = DeviceSystem(params, offsets)
    orc = OracleSystem(params, offsets, threads=16)
    dev.set_fields(fields)
    orc.set_fields(fields)
    dev.warm_start()
    orc.warm_start()
    compare(dev, orc, STATE + RATES, TOL, "warm start")
    for k in range(3):
        dev.step(1)
        orc.step(1)
        compare(dev, orc, STATE + RATES, TOL, f"step {k + 1}")
    td, dtd = dev.clock()
    to, dto = orc.clock()
    assert td[1] == to[1]  # the empty instance's clock, exactly
    # the others sum steps whose dt follows h, which the density sums give
    # to the last place of their own summation order
    assert np.allclose(td, to, rtol=1e-12, atol=0)
    assert np.allclose(dtd, dto, rtol=1e-12, atol=0)


def test_particles_on_box_faces(gpu):
    """Particles exactly at lo, at hi (outside the half-open box, binned into
    the last cell and wrapped by the first stage) and one ulp below hi: cells
    and neighbour sets bit-exact, then three steps to 1e-10.

    Floor: divBsymm, divBdiff, psi, apsi and aBevol* are measured against
    1e-3 of their natural scale (|B|/h, rho|B|/h, |B|, v_w|B|/h,
    v_w|B|/(rho h)), since they are round-off where B is uniform; every
    other field against its own norm."""
    from pyoracle import OracleSystem
    from sphmhd.device import DeviceSystem

    s = _ot(32, 4)
    f = init_fields(s)
    p = s.params[0]
    x, y = f["x"].copy(), f["y"].copy()
    x[:5] = p.lo[0]
    x[5:10] = p.hi[0]
    y[10:15] = np.nextafter(p.hi[1], -np.inf)
    y[15:20] = p.hi[1]
    f["x"], f["y"] = x, y
    dev = DeviceSystem(s.params, s.offsets)
    orc = OracleSystem(s.params, s.offsets, threads=16)
    dev.set_fields(f)
    orc.set_fields(f)
    kd, pd = dev.cells()
    ko, po = orc.cells()
    assert np.array_equal(kd, ko) and np.array_equal(pd, po)
    od, idd = dev.neighbours()
    oo, io = orc.neighbours()
    assert np.array_equal(od, oo) and np.array_equal(idd, io)
    dev.warm_start()
    orc.warm_start()
    for k in range(3):
        dev.step(1)
        orc.step(1)
        compare(dev, orc, STATE + RATES, TOL, f"step {k + 1}")
    assert np.all(dev.get("x") < p.hi[0]) and np.all(dev.get("x") >= p.lo[0])
