"""sphmhd_totals_noalloc (the allocation-free conserved totals) against
sphmhd_totals through the C ABI: two calls in a row on one GPU, the second
on a smaller context (fewer instances and particles), each to 1e-13 of the
totals' own scale (momentum against sqrt(2 M E_kin), whose total is zero up
to round-off): the two entry points add the same terms in different orders.
The second call also checks that nothing of the first call's larger grid
leaks into its result, since both share the library's static scratch."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import init_fields

pytestmark = pytest.mark.gpu


def _check(a, b, label):
    for k in range(a.shape[0]):
        mass, ekin = b[k, 0], b[k, 4]
        pscale = np.sqrt(2.0 * mass * max(ekin, 1e-300))
        for q in range(a.shape[1]):
            scale = pscale if q in (1, 2, 3) else max(abs(b[k, q]), 1e-300)
            assert abs(a[k, q] - b[k, q]) <= 1e-13 * scale, (label, k, q, a[k, q], b[k, q])


def _system(setup):
    from sphmhd.device import DeviceSystem

    dev = DeviceSystem(setup.params, setup.offsets)
    dev.set_fields(init_fields(setup))
    dev.warm_start()
    dev.step(2)
    return dev


def test_totals_noalloc_matches_totals_twice(gpu):
    import torch

    from sphmhd.cases.workloads import ensemble_ot, jittered, orszag_tang_2d

    big = _system(jittered(ensemble_ot(6, sizes=[96, 40, 131, 64, 77, 150]), 0.1, seed=3))
    small = _system(jittered(orszag_tang_2d(24), 0.1, seed=4))
    for label, dev in (("large ensemble", big), ("small instance", small)):
        ref = dev.totals()
        got = dev.totals_noalloc()
        torch.cuda.synchronize()
        _check(got.cpu().numpy(), ref, label)
        assert got.shape == ref.shape


def test_totals_noalloc_rejects_bad_output(gpu):
    from sphmhd._abi import SPHMHDError
    from sphmhd.cases.workloads import jittered, orszag_tang_2d

    dev = _system(jittered(orszag_tang_2d(16), 0.1, seed=5))
    with pytest.raises(SPHMHDError):
        from sphmhd._abi import check

        check(dev.lib.sphmhd_totals_noalloc(dev._ctx, None))
