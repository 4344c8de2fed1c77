"""Shared helpers for the parity tests (CPU oracle vs HIP path)."""
from __future__ import annotations

import numpy as np


# properties compared after an evaluation or a step (live outputs of the
# group plan and the integrator state)
STATE = ("x", "y", "u", "v", "w", "e", "h", "rho", "p", "cs", "Bx", "By", "Bz", "Bevolx", "Bevoly",
         "Bevolz", "psi", "alpha1", "omega")
RATES = ("au", "av", "aw", "ae", "ah", "aBevolx", "aBevoly", "aBevolz", "apsi", "alphaloc", "dt_cfl",
         "dt_adapt", "div", "divBsymm", "divBdiff")
INIT = ("x", "y", "z", "u", "v", "w", "rho", "h", "m", "e", "Bx", "By", "Bz", "alpha1")


def init_fields(setup):
    f = {k: v for k, v in setup.fields.items() if k in INIT}
    if setup.itype is not None:
        f["itype"] = setup.itype
    return f


def natural_scales(orc) -> dict[str, float]:
    """Norms of the magnitude each divergence-cleaning field would have if B
    varied on the scale h.  Fields that are analytically zero for a given
    state (div B of a uniform field, psi before any divergence has grown) hold
    only round-off, so their relative error is measured against 1e-3 of this
    scale instead of their own norm."""
    bx, by, bz = orc.get("Bx"), orc.get("By"), orc.get("Bz")
    rho, h, cs = orc.get("rho"), orc.get("h"), orc.get("cs")
    b = np.sqrt(bx * bx + by * by + bz * bz)
    vw = np.sqrt(cs * cs + b * b / rho)
    sc = {
        "divBsymm": np.linalg.norm(b / h),
        "divBdiff": np.linalg.norm(rho * b / h),
        "psi": np.linalg.norm(b),
        "apsi": np.linalg.norm(vw * b / h),
    }
    for c in ("x", "y", "z"):
        sc[f"aBevol{c}"] = np.linalg.norm(vw * b / (rho * h))
    return sc


def compare(dev, orc, names, tol, label=""):
    worst = {}
    scales = natural_scales(orc)
    for nm in names:
        a = dev.get(nm)
        b = orc.get(nm)
        den = max(np.linalg.norm(b), 1e-3 * scales.get(nm, 0.0))
        err = float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
        worst[nm] = err
    bad = {k: v for k, v in worst.items() if not (v <= tol)}
    assert not bad, f"{label} rel-L2 above {tol}: {bad}"
    return worst
