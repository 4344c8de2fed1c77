"""Initial conditions generated on the GPU (sphmhd_case_fields).

``load_setup`` fills a DeviceSystem from a lattice Setup without host
arrays: each instance's fields (and, for the benchmark lattices, positions)
are computed by one kernel launch into the instance's slice of the storage
rows.  ``write_case`` is the thin wrapper of the C entry point; the reference
cases (``cases.reference``) call it through ``device_particles``.

The kernel follows the host builders' expressions in their order with
contraction off, and every particle-independent constant comes from the same
host expression as in the host builder, so the device fields equal the host
fields bit for bit except through sin, cos, exp, tanh, cosh and pow.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import torch

from .._abi import FIELD_INDEX, NFIELDS, check, load_library

MAX_CONSTANTS = 48


def write_case(case_id: int, constants: Sequence[float], rows: torch.Tensor, itype: torch.Tensor | None = None,
               aux: torch.Tensor | None = None) -> None:
    """Enqueue sphmhd_case_fields on torch's current stream.

    rows: FP64 device tensor [NFIELDS, n] whose rows are contiguous (a
    DeviceSystem's storage, or a column slice of it for one instance);
    itype: int32 device tensor of n entries (written by the shock tube);
    aux: for the jets, the single jet's (axial, y, z) as an FP64 [3, n_aux]
    device tensor, with n a multiple of n_aux."""
    if rows.dtype != torch.float64 or rows.dim() != 2 or rows.shape[0] != NFIELDS or rows.stride(1) != 1:
        raise ValueError("rows must be an FP64 tensor [NFIELDS, n] with contiguous rows")
    if len(constants) > MAX_CONSTANTS:
        raise ValueError(f"at most {MAX_CONSTANTS} constants")
    n = int(rows.shape[1])
    if itype is not None and (itype.dtype != torch.int32 or itype.numel() != n or not itype.is_contiguous()):
        raise ValueError("itype must be a contiguous int32 tensor of n entries")
    n_aux, aux_ptr = 0, None
    if aux is not None:
        if aux.dtype != torch.float64 or aux.dim() != 2 or aux.shape[0] != 3 or not aux.is_contiguous():
            raise ValueError("aux must be a contiguous FP64 tensor [3, n_aux]")
        n_aux, aux_ptr = int(aux.shape[1]), ctypes.c_void_p(aux.data_ptr())
    lib = load_library()
    ptrs = (ctypes.c_void_p * NFIELDS)(*[rows[k].data_ptr() for k in range(NFIELDS)])
    consts = (ctypes.c_double * max(len(constants), 1))(*[float(c) for c in constants])
    stream = torch.cuda.current_stream(rows.device).cuda_stream
    check(lib.sphmhd_case_fields(int(case_id), consts, len(constants), ptrs,
                                 None if itype is None else ctypes.c_void_p(itype.data_ptr()), n, n_aux, aux_ptr,
                                 ctypes.c_void_p(stream)))


def load_setup(system, setup) -> None:
    """Initial state of ``setup`` into the DeviceSystem ``system``: its host
    arrays when it has them, else every instance generated on the GPU."""
    if setup.fields:
        f = dict(setup.fields)
        if setup.itype is not None:
            f["itype"] = setup.itype
        system.set_fields(f)
        return
    if not setup.device or len(setup.device) != system.ninst:
        raise ValueError(f"setup {setup.name} has neither host fields nor a device generator per instance")
    with torch.cuda.stream(system.stream):
        for k, (case_id, consts) in enumerate(setup.device):
            a, b = int(system.offsets[k]), int(system.offsets[k + 1])
            write_case(case_id, consts, system.storage[:, a:b])


def new_rows(n: int, device="cuda") -> torch.Tensor:
    """Zeroed FP64 storage [NFIELDS, n] for a generated state."""
    return torch.zeros((NFIELDS, n), dtype=torch.float64, device=device)


def set_positions(rows: torch.Tensor, x: torch.Tensor, y: torch.Tensor, z: torch.Tensor) -> None:
    rows[FIELD_INDEX["x"]].copy_(x)
    rows[FIELD_INDEX["y"]].copy_(y)
    rows[FIELD_INDEX["z"]].copy_(z)
