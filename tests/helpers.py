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


def natural_scales(orc, mask=None) -> dict[str, float]:
    """Norms of the magnitude each divergence-cleaning field would have if B
    varied on the scale h.  Fields that are analytically zero for a given
    state (div B of a uniform field, psi before any divergence has grown) hold
    only round-off, so their relative error is measured against 1e-3 of this
    scale instead of their own norm.  ``mask`` selects the particles (one
    instance of an ensemble), as in compare()."""
    sel = (lambda v: v) if mask is None else (lambda v: v[mask])
    bx, by, bz = sel(orc.get("Bx")), sel(orc.get("By")), sel(orc.get("Bz"))
    rho, h, cs = sel(orc.get("rho")), sel(orc.get("h")), sel(orc.get("cs"))
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


def compare(dev, orc, names, tol, label="", scales=None, mask=None):
    """Relative L2 error of each named property, with the norm floored at
    1e-3 of its natural scale; ``tol`` is a number or a {name: tolerance}
    map, ``mask`` selects the particles compared."""
    worst = {}
    scales = natural_scales(orc, mask) if scales is None else scales
    for nm in names:
        a = dev.get(nm)
        b = orc.get(nm)
        if mask is not None:
            a, b = a[mask], b[mask]
        den = max(np.linalg.norm(b), 1e-3 * scales.get(nm, 0.0))
        err = float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
        worst[nm] = err
    bad = {k: v for k, v in worst.items() if not (v <= (tol.get(k, 0.0) if isinstance(tol, dict) else tol))}
    assert not bad, f"{label} rel-L2 above {tol}: {bad}"
    return worst


# ---- the reference's cases ---------------------------------------------------

# small sizes of the reference's twelve cases (particle counts 432 to 10,368)
CASE_NX = {"orstang": 16, "mhdrotor": 16, "mhdblast": 10, "alfven": 12, "jadvect": 16, "mhdshock": 16,
           "mhdsine": 16, "mhdvortex": 16, "mhdwave": 16, "1jet": 8, "3jet": 6, "12jet": 4}


def case_app(name, nx=None, directory="/tmp/sphmhd_case"):
    """Host-side application of a reference case from its TOML in
    pysph-mhd_amd/configs, at size nx, prepared (scheme, solver, particles,
    domain) without touching a device."""
    import dataclasses
    from pathlib import Path

    from sphmhd.config import load_config
    from sphmhd.run import CASES

    path = Path(__file__).resolve().parents[1] / "pysph-mhd_amd" / "configs" / f"{name}.toml"
    cfg = load_config(path, {n: a.config_type for n, a in CASES.items()})
    if nx is not None:
        cfg = dataclasses.replace(cfg, case=dataclasses.replace(cfg.case, nx=int(nx)))
    return CASES[name](cfg).prepare(["-d", directory])


def case_setup(name, nx=None):
    """Setup (parameters from the application's plan, domain and solver;
    initial properties from its particle array) of a reference case."""
    from sphmhd.backend import plan_params
    from sphmhd.cases.workloads import Setup

    app = case_app(name, nx)
    pa = app.particles[0]
    params, _ = plan_params(pa, app.scheme.get_equations(), app.solver.kernel, app.domain, app.scheme, app.solver)
    fields = {k: np.array(pa.properties[k], dtype=np.float64) for k in INIT if k in pa.properties}
    itype = np.asarray(pa.properties["itype"], dtype=np.int32).copy() if "itype" in pa.properties else None
    return Setup(name, [params], np.array([0, len(pa)], dtype=np.int64), fields, itype, info={"app": app})


def physical_scales(orc, mu0=1.0, alpha_max=1.0, mask=None) -> dict[str, float]:
    """natural_scales plus floors for the fields that a case can hold at
    round-off level (a uniform flow has div v = 0, a slab symmetric in z has
    w = 0, a static gas has no acceleration but the round-off of its pressure
    sums): velocities against the magnetosonic speed v_w = sqrt(c_s^2 +
    B^2/(mu0 rho)), accelerations against v_w^2/h, div v against v_w/h,
    dh/dt against v_w, de/dt against v_w^3/h, field components against |B|
    and the dissipation switch against alpha_max.  With mu0 in v_w the floors hold in SI units
    too.  ``mask`` selects the particles, as in compare()."""
    sel = (lambda v: v) if mask is None else (lambda v: v[mask])
    bx, by, bz = (sel(orc.get(c)) for c in ("Bx", "By", "Bz"))
    rho, h, cs = (sel(orc.get(c)) for c in ("rho", "h", "cs"))
    b2 = bx * bx + by * by + bz * bz
    vw = np.sqrt(cs * cs + b2 / (mu0 * rho))
    b = np.sqrt(b2)
    sc = {
        "divBsymm": np.linalg.norm(b / h),
        "divBdiff": np.linalg.norm(rho * b / h),
        "psi": np.linalg.norm(b / np.sqrt(mu0)),
        "apsi": np.linalg.norm(vw * b / (np.sqrt(mu0) * h)),
        "div": np.linalg.norm(vw / h),
        "ah": np.linalg.norm(vw),
        "ae": np.linalg.norm(vw ** 3 / h),
        "alphaloc": alpha_max * np.sqrt(len(h)),
        "alpha1": alpha_max * np.sqrt(len(h)),
    }
    for c in ("x", "y", "z"):
        sc[f"aBevol{c}"] = np.linalg.norm(vw * b / (rho * h))
        sc[f"Bevol{c}"] = np.linalg.norm(b / rho)
        sc[f"B{c}"] = np.linalg.norm(b)
    for c in ("u", "v", "w"):
        sc[c] = np.linalg.norm(vw)
        sc["a" + c] = np.linalg.norm(vw * vw / h)
    return sc


def nudged(fields, seed=1):
    """Copy of initial fields whose x and y positions are moved by one ulp
    each, up or down at random per particle: the rounding-level perturbation
    that measures how much the reference algorithm itself amplifies rounding
    for a given state (a uniform nudge of x alone keeps lattice symmetries
    and can miss it)."""
    rng = np.random.default_rng(seed)
    out = dict(fields)
    for c in ("x", "y"):
        v = np.asarray(fields[c], dtype=np.float64)
        out[c] = np.nextafter(v, rng.choice([-np.inf, np.inf], v.size))
    return out
