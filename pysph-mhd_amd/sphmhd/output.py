"""Output dumps and full-state checkpoints.

``dump`` writes the reference's OUTPUT_PROPERTIES (particles.py:53-71) of
every particle array plus, under ``state/``, every device-bound property, so
a dump is also an exact restart point (the reference's dumps lack psi, B/rho,
the *0 copies and alpha10).  ``restore`` loads such a file back into a
particle array and its device mirror.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from ._abi import FIELDS
from .particles import OUTPUT_PROPERTIES


def dump(directory, particles, t, dt, count) -> Path:
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    out = {"t": np.array(t), "dt": np.array(dt), "count": np.array(count)}
    for pa in particles:
        names = [n for n in FIELDS if n in pa.properties]
        pa.gpu.pull(*names)
        for name in OUTPUT_PROPERTIES:
            if name in pa.properties:
                out[f"{pa.name}/{name}"] = pa.properties[name]
        for name in names + ["itype"]:
            out[f"state/{pa.name}/{name}"] = pa.properties[name]
    path = d / f"{particles[0].name}_{int(count):07d}.npz"
    np.savez(path, **out)
    return path


def restore(path, particles):
    """Load a dump into the particle arrays (host and device); returns
    (t, dt, count)."""
    z = np.load(path, allow_pickle=False)
    for pa in particles:
        names = []
        for key in z.files:
            pre = f"state/{pa.name}/"
            if key.startswith(pre):
                name = key[len(pre):]
                if name in pa.properties:
                    pa.properties[name][:] = z[key]
                    names.append(name)
        pa.gpu.push(*names)
    return float(z["t"]), float(z["dt"]), int(z["count"])
