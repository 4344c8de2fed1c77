"""Launch shapes of the density, list and force kernels: the results of a
step do not depend on the workgroup size (each particle's sums run in the
same order whatever the shape), so every shape, and the autotuned choice,
reproduce the default bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import RATES, STATE, init_fields

pytestmark = pytest.mark.gpu

SHAPES = [(256, 256, 256), (64, 64, 64), (128, 128, 128), (64, 256, 128), (256, 64, 64)]


def _case(which):
    from sphmhd.cases.workloads import blast_3d, ensemble_ot, jittered, orszag_tang_2d, rotor_2d

    if which == "ot2d":
        return jittered(orszag_tang_2d(40), 0.1, seed=40)
    if which == "rotor":
        return jittered(rotor_2d(48), 0.1, seed=4)
    if which == "blast3d":
        return jittered(blast_3d(14), 0.05, seed=2)
    return jittered(ensemble_ot(4, sizes=[24, 40, 32, 28]), 0.1, seed=9)


def _run(setup, blocks, steps, modes=(0, 0, 0)):
    from sphmhd.device import DeviceSystem

    d = DeviceSystem(setup.params, setup.offsets)
    d.set_launch(blocks)
    d.set_persistent(modes)
    d.set_fields(init_fields(setup))
    d.warm_start()
    d.step(steps)
    out = {nm: d.get(nm) for nm in STATE + RATES}
    clock = d.clock()
    st = d.stats()
    d.close()
    return out, clock, st


@pytest.mark.parametrize("which", ["ot2d", "rotor", "blast3d", "ensemble"])
def test_every_shape_gives_the_same_steps(gpu, which):
    s = _case(which)
    ref, ref_clock, _ = _run(s, SHAPES[0], 4)
    for blocks in SHAPES[1:]:
        got, clock, st = _run(s, blocks, 4)
        assert (st["block_density"], st["block_lists"], st["block_force"]) == blocks
        for nm in STATE + RATES:
            assert np.array_equal(got[nm], ref[nm]), (blocks, nm)
        assert np.array_equal(clock[0], ref_clock[0]) and np.array_equal(clock[1], ref_clock[1])


@pytest.mark.parametrize("which", ["ot2d", "blast3d", "ensemble"])
def test_persistent_launches_give_the_same_steps(gpu, which):
    """Persistent launches (workgroups pulling tiles of any instance from
    per-XCD work queues, the queues reset by each launch's last workgroup)
    against plain launches, bit for bit, in several mixes of modes and
    shapes; the modes reported are those set."""
    s = _case(which)
    ref, ref_clock, _ = _run(s, (256, 256, 256), 4)
    for blocks, modes in (((256, 256, 256), (1, 1, 1)), ((64, 128, 256), (1, 0, 1)), ((128, 64, 64), (0, 1, 1))):
        got, clock, st = _run(s, blocks, 4, modes)
        assert (st["persist_density"], st["persist_lists"], st["persist_force"]) == modes
        for nm in STATE + RATES:
            assert np.array_equal(got[nm], ref[nm]), (blocks, modes, nm)
        assert np.array_equal(clock[0], ref_clock[0]) and np.array_equal(clock[1], ref_clock[1])


def test_autotuned_run_equals_fixed_run(gpu, monkeypatch):
    """The first steps of a context rotate the configurations (shape and
    plain or persistent launch, timed with events) and keep the fastest; the
    trajectory is that of any fixed configuration."""
    import os

    from sphmhd.device import DeviceSystem

    if os.environ.get("SPHMHD_PERSIST") is not None or os.environ.get("SPHMHD_TUNE") == "0":
        pytest.skip("launch tuning turned off by the environment (SPHMHD_PERSIST / SPHMHD_TUNE=0)")
    monkeypatch.delenv("SPHMHD_SPEC", raising=False)  # one force launch per step below
    s = _case("ot2d")
    ref, ref_clock, _ = _run(s, (256, 256, 256), 15)
    d = DeviceSystem(s.params, s.offsets)
    d.set_launch(None)
    d.set_fields(init_fields(s))
    d.warm_start()
    d.set_timing(True)
    d.step(5)  # the 12 tuning steps span the two calls
    d.step(10)
    blocks, tuned = d.launch()
    assert tuned and all(b in (64, 128, 256) for b in blocks)
    st = d.stats()
    assert st["launch_tuned"] == 1
    for nm in STATE + RATES:
        assert np.array_equal(d.get(nm), ref[nm]), nm
    assert np.array_equal(d.clock()[0], ref_clock[0])
    # the caller's timing survives the tuning steps: only the 3 steps after
    # tuning are in the sums (one force launch each)
    assert d.timing()["force_launches"] == 3
    d.close()


def test_set_launch_rejects_bad_shapes(gpu):
    from sphmhd._abi import SPHMHDError
    from sphmhd.device import DeviceSystem

    s = _case("ot2d")
    d = DeviceSystem(s.params, s.offsets)
    with pytest.raises(SPHMHDError):
        d.set_launch((32, 256, 256))
    d.close()
