"""Initial conditions generated on the GPU (sphmhd_case_fields) against the
host builders: the benchmark lattices through cases.device.load_setup and
the reference's twelve cases through cases.reference.device_particles.

Fields computed with + - * / and sqrt only must agree bit for bit (the
kernel keeps the host's order of operations with contraction off, and takes
every particle-independent constant from the host's own expression); fields
that pass through sin, cos, exp, tanh, cosh or pow may differ by the last
places of the two math libraries, bounded here by 1e-14 of the field's
largest magnitude.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import CASE_NX, case_app

pytestmark = pytest.mark.gpu

NEAR_TOL = 1.0e-14
PROPS = ("x", "y", "z", "u", "v", "w", "rho", "h", "m", "e", "Bx", "By", "Bz", "alpha1")

# fields that go through a transcendental function, per case
WORKLOAD_NEAR = {
    "orszag_tang_2d": ("u", "v", "Bx", "By"),
    "rotor_2d": (),
    "blast_3d": (),
    "alfven_2d": ("u", "v", "w", "Bx", "By", "Bz"),
}
CASE_NEAR = {
    "orstang": ("u", "v", "Bx", "By"),
    "mhdrotor": (), "mhdblast": (), "jadvect": (), "mhdshock": (),
    "alfven": ("v", "w", "By", "Bz"),
    "mhdsine": ("By",),
    "mhdvortex": ("u", "v", "e", "Bx", "By"),
    "mhdwave": ("u",),
    "1jet": ("u", "v", "w", "Bx", "By", "Bz"),
    "3jet": ("u", "v", "w", "Bx", "By", "Bz"),
    "12jet": ("u", "v", "w", "Bx", "By", "Bz"),
}


def _check(label, host: dict, dev: dict, near):
    for name in PROPS:
        h, d = np.asarray(host[name], dtype=np.float64), dev[name]
        assert h.shape == d.shape, (label, name)
        if name in near:
            scale = max(float(np.abs(h).max()), 1e-300)
            err = float(np.abs(d - h).max()) / scale
            assert err <= NEAR_TOL, f"{label}.{name}: max diff {err:.3e} of max |field|"
        else:
            bad = np.flatnonzero(d != h)
            assert bad.size == 0, f"{label}.{name}: {bad.size} entries differ, first {bad[:3]} {d[bad[:3]]} {h[bad[:3]]}"


@pytest.mark.parametrize("name,args", [
    ("orszag_tang_2d", dict(nx=48)),
    ("rotor_2d", dict(nx=96)),
    ("blast_3d", dict(nx=20)),
    ("alfven_2d", dict(nx=17, ny=30)),
])
def test_workload_fields_on_device(gpu, name, args):
    import torch

    from sphmhd.cases import workloads as B
    from sphmhd.cases.device import load_setup
    from sphmhd.device import DeviceSystem

    host = getattr(B, name)(**args)
    lean = getattr(B, name)(**args, host_fields=False)
    assert not lean.fields and lean.n == host.n
    dev = DeviceSystem(lean.params, lean.offsets, device=gpu)
    load_setup(dev, lean)
    torch.cuda.synchronize()
    _check(name, host.fields, dev.state(PROPS), WORKLOAD_NEAR[name])
    dev.close()


def test_ensemble_fields_on_device(gpu):
    """Uneven instances, each generated into its slice of the storage rows."""
    import torch

    from sphmhd.cases import workloads as B
    from sphmhd.cases.device import load_setup
    from sphmhd.device import DeviceSystem

    sizes = [16, 33, 24, 7]
    host = B.ensemble_ot(4, seed=3, sizes=sizes)
    lean = B.ensemble_ot(4, seed=3, sizes=sizes, host_fields=False)
    dev = DeviceSystem(lean.params, lean.offsets, device=gpu)
    load_setup(dev, lean)
    torch.cuda.synchronize()
    _check("ensemble_ot", host.fields, dev.state(PROPS), WORKLOAD_NEAR["orszag_tang_2d"])
    dev.close()


@pytest.mark.parametrize("case", sorted(CASE_NEAR))
def test_reference_case_fields_on_device(gpu, case):
    import torch

    from sphmhd._abi import FIELD_INDEX
    from sphmhd.cases.reference import device_particles

    app = case_app(case, CASE_NX[case])
    pa = app.create_mhd_particles()[0]
    host = {k: np.asarray(pa.properties[k], dtype=np.float64) for k in PROPS}
    rows, itype = device_particles(app, device=gpu)
    torch.cuda.synchronize()
    got = rows.cpu().numpy()
    dev = {k: got[FIELD_INDEX[k]] for k in PROPS}
    _check(case, host, dev, CASE_NEAR[case])
    want_itype = np.asarray(pa.properties["itype"], dtype=np.int32) if "itype" in pa.properties else None
    if want_itype is not None:
        assert np.array_equal(itype.cpu().numpy(), want_itype)
    else:
        assert bool((itype == 1).all())


def test_case_fields_rejects_bad_arguments(gpu):
    import ctypes

    import torch

    from sphmhd._abi import NFIELDS, load_library

    lib = load_library()
    rows = torch.zeros((NFIELDS, 8), dtype=torch.float64, device=gpu)
    ptrs = (ctypes.c_void_p * NFIELDS)(*[rows[k].data_ptr() for k in range(NFIELDS)])
    consts = (ctypes.c_double * 8)(*([4.0] + [0.25] * 7))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.sphmhd_case_fields(99, consts, 8, ptrs, None, 8, 0, None, stream) != 0
    assert lib.sphmhd_case_fields(1, consts, 49, ptrs, None, 8, 0, None, stream) != 0
    assert lib.sphmhd_case_fields(14, consts, 8, ptrs, None, 8, 0, None, stream) != 0  # jets without rows
    torch.cuda.synchronize()
    assert bool((rows == 0).all())
    assert lib.sphmhd_case_fields(1, consts, 8, ptrs, None, 8, 0, None, stream) == 0
    torch.cuda.synchronize()
    assert float(rows[6].min()) == 0.25  # rho of the vortex constants above
