"""Write golden positions of the reference's close-packed lattice.

Runs in the build container only (the reference is not available anywhere
else): loads src/cases/lattice.py of the reference from /root/reference by
file path (it imports numpy only) and evaluates close_packed_lattice on the
bounds and spacings the reference's cases use, at small sizes: the thin
slabs of orstang/mhdsine/mhdvortex/jadvect/mhdwave/mhdrotor (z half-width
k sqrt(6)/nx), the blast cube, the rotor's denser inner lattice (spacing
10^(-1/3) smaller), the two halves of the shock tube and a jet shell.

Output: tests/golden/lattice_golden.npz, arrays ``bounds`` [C, 6],
``spacing`` [C], ``offsets`` [C + 1] and ``x``, ``y``, ``z`` (the C cases'
positions concatenated).
"""
from __future__ import annotations

import importlib.util
from pathlib import Path

import numpy as np

REFERENCE = Path("/root/reference/src/cases/lattice.py")
OUT = Path(__file__).resolve().parent / "lattice_golden.npz"


def cases():
    out = []
    for nx, k, (x0, x1, y0, y1) in ((12, 4.0, (-0.5, 0.5, -0.5, 0.5)), (10, 2.0, (-0.5, 0.5, -0.5, 0.5)),
                                     (16, 4.0, (-1.0, 1.0, -0.5, 0.5)), (14, 4.0, (-2.0, 2.0, -1.0, 1.0)),
                                     (9, 4.0, (-5.0, 5.0, -5.0, 5.0))):
        zmax = k * np.sqrt(6.0) / nx
        b = (x0, x1, y0, y1, -zmax, zmax)
        out.append((b, (x1 - x0) / nx))
    out.append(((-0.5, 0.5, -0.5, 0.5, -0.5, 0.5), 1.0 / 7))
    zmax = 2.0 * np.sqrt(6.0) / 10
    out.append(((-0.5, 0.5, -0.5, 0.5, -zmax, zmax), (1.0 / 10) * 10.0 ** (-1.0 / 3.0)))
    nx = 16
    dxl = 0.5 / nx
    dxr = 2.0 * dxl
    ymax = 6.0 * dxr * np.sqrt(3.0 / 4.0)
    zmax = 6.0 * dxr * np.sqrt(6.0) / 3.0
    out.append(((-0.5, 0.0, -ymax, ymax, -zmax, zmax), dxl))
    out.append(((0.0, 0.5, -ymax, ymax, -zmax, zmax), dxr))
    out.append(((0.810, 0.900, -0.040, 0.040, -0.040, 0.040), 2.0 * 0.040 / 8))
    out.append(((-1.5, 1.5, -0.75, 0.75, -0.75, 0.75), 3.0 / 12))
    return out


def main():
    spec = importlib.util.spec_from_file_location("reference_lattice", REFERENCE)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    bounds, spacing, offsets, xs, ys, zs = [], [], [0], [], [], []
    for b, s in cases():
        x, y, z = mod.close_packed_lattice(b, s)
        bounds.append(b)
        spacing.append(s)
        xs.append(x)
        ys.append(y)
        zs.append(z)
        offsets.append(offsets[-1] + len(x))
    np.savez_compressed(OUT, bounds=np.array(bounds, dtype=np.float64), spacing=np.array(spacing),
                        offsets=np.array(offsets, dtype=np.int64), x=np.concatenate(xs), y=np.concatenate(ys),
                        z=np.concatenate(zs))
    print(f"{OUT}: {len(bounds)} cases, {offsets[-1]} sites")


if __name__ == "__main__":
    main()
