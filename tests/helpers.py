"""Shared helpers for the parity tests (CPU oracle vs HIP path)."""
from __future__ import annotations

import numpy as np

from conftest import rel_l2

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


def compare(dev, orc, names, tol, label=""):
    worst = {}
    for nm in names:
        a = dev.get(nm)
        b = orc.get(nm)
        err = rel_l2(a, b)
        worst[nm] = err
    bad = {k: v for k, v in worst.items() if not (v <= tol)}
    assert not bad, f"{label} rel-L2 above {tol}: {bad}"
    return worst
