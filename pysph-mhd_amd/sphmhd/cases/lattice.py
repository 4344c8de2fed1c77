"""Close-packed (HCP, ABC-stacked) lattice of the reference's cases
(reference: src/cases/lattice.py:4-34), on the host and on the GPU.

Layer counts: n_x = ceil(L_x / s), n_y = ceil(L_y / (s sqrt(3/4))) rounded
down to even, n_z = ceil(L_z / (s sqrt(6)/3)) rounded down to a multiple of
3 (each quotient first stepped one ulp towards -inf, so an exact multiple is
not rounded up); the spacings are then stretched to fill the box exactly.
Rows alternate by half a spacing in x with the row parity and the layer
index mod 3, and the three layers of a stack sit at y offsets 0, 1/3 and 2/3
of a row.  The arithmetic, and its order, is the reference's so that the
positions agree bit for bit (the reference cases start from them).
"""
from __future__ import annotations

import ctypes

import numpy as np


def lattice_shape(bounds, spacing):
    """(n_x, n_y, n_z, dx, dy, dz) of the lattice (lattice.py:6-14)."""
    xmin, xmax, ymin, ymax, zmin, zmax = (float(b) for b in bounds)
    nx = int(np.ceil(np.nextafter((xmax - xmin) / spacing, -np.inf)))
    ny = int(np.ceil(np.nextafter((ymax - ymin) / (spacing * np.sqrt(3.0 / 4.0)), -np.inf)))
    nz = int(np.ceil(np.nextafter((zmax - zmin) / (spacing * np.sqrt(6.0) / 3.0), -np.inf)))
    ny = 2 * (ny // 2)
    nz = 3 * (nz // 3)
    return nx, ny, nz, (xmax - xmin) / nx, (ymax - ymin) / ny, (zmax - zmin) / nz


def close_packed_lattice(bounds, spacing):
    """Host version: x, y, z arrays in (i_x, i_y, i_z) C order."""
    xmin, _, ymin, _, zmin, _ = (float(b) for b in bounds)
    nx, ny, nz, dx, dy, dz = lattice_shape(bounds, spacing)
    k = np.arange(nx, dtype=np.int64)[:, None, None]
    ly = np.arange(1, ny + 1, dtype=np.int64)[None, :, None]
    lz = np.arange(1, nz + 1, dtype=np.int64)[None, None, :]
    k, ly, lz = np.broadcast_arrays(k, ly, lz)
    ypar = ly % 2
    zmod = lz % 3
    shifted = ((zmod == 0) & (ypar == 0)) | ((zmod == 2) & (ypar == 1)) | ((zmod == 1) & (ypar == 0))
    x_off = np.where(shifted, 0.5, 0.0)
    y_off = np.where(zmod == 0, 2.0 / 3.0, np.where(zmod == 2, 1.0 / 3.0, 0.0))
    x = xmin + 0.25 * spacing + (k + x_off) * dx
    y = ymin + (ly - 1.0 + 1.0 / 6.0 + y_off) * dy
    z = zmin + (lz - 0.5) * dz
    return x.ravel(), y.ravel(), z.ravel()


def close_packed_lattice_device(bounds, spacing, device="cuda"):
    """GPU version (sphmhd_hcp_lattice): the same positions, generated on
    the device into three new FP64 tensors."""
    import torch

    from .._abi import check, load_library

    lib = load_library()
    nx, ny, nz, *_ = lattice_shape(bounds, spacing)
    count = nx * ny * nz
    out = torch.empty((3, max(count, 1)), dtype=torch.float64, device=device)
    b = (ctypes.c_double * 6)(*[float(v) for v in bounds])
    got = ctypes.c_int64(0)
    stream = torch.cuda.current_stream(out.device).cuda_stream
    check(lib.sphmhd_hcp_lattice(b, ctypes.c_double(float(spacing)), ctypes.c_void_p(out[0].data_ptr()),
                                 ctypes.c_void_p(out[1].data_ptr()), ctypes.c_void_p(out[2].data_ptr()), count,
                                 ctypes.byref(got), ctypes.c_void_p(stream)))
    assert got.value == count
    return out[0, :count], out[1, :count], out[2, :count]
