"""Write golden input/output vectors of the reference's equation bodies.

Runs in the build container only (the reference is not available anywhere
else): loads src/equations.py and src/integrator.py of the reference from
/root/reference with two stand-ins -- a base class for
``pysph.sph.equation.Equation`` / ``pysph.sph.integrator_step.IntegratorStep``
(the bodies never call into them) and ``typing.override`` (Python 3.12+) as
the identity -- and calls every ``initialize`` / ``loop`` / ``post_loop`` /
``stage`` method on a two-particle property set (index 0 = destination,
1 = source) with seeded random inputs that cover each branch.  The kernel
values WI, DWI, DWJ, GHI and the pair symbols are inputs here, so the
vectors pin the bodies' arithmetic independently of any kernel.

Output: tests/golden/reference_bodies.npz with two arrays: ``values`` (all
numbers, FP64, flat) and ``manifest`` (JSON).  Each manifest record holds the
body name, the offsets of its params (k, dim, iterate_only_once, gamma, mu0,
av, ac, ar, chf, sigma, tsf), its 15 pair symbols (XIJ[3], RIJ, R2IJ, RHOIJ,
WI, DWI[3], DWJ[3], GHI, dt), the returned flag and, per property, the
offsets of its two input and two output values.
"""
from __future__ import annotations

import importlib.util
import inspect
import sys
import types
import typing
from pathlib import Path

import numpy as np

REF = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent / "reference_bodies.npz"


def load_reference():
    if not hasattr(typing, "override"):
        typing.override = lambda f: f  # identity decorator of Python 3.12
    pysph = types.ModuleType("pysph")
    sph = types.ModuleType("pysph.sph")
    eqm = types.ModuleType("pysph.sph.equation")
    ism = types.ModuleType("pysph.sph.integrator_step")

    class Equation:
        def __init__(self, dest, sources):
            self.dest = dest
            self.sources = sources

    class IntegratorStep:
        pass

    eqm.Equation = Equation
    ism.IntegratorStep = IntegratorStep
    sys.modules.update({"pysph": pysph, "pysph.sph": sph, "pysph.sph.equation": eqm,
                        "pysph.sph.integrator_step": ism})
    mods = {}
    for name in ("equations", "integrator"):
        spec = importlib.util.spec_from_file_location(f"ref_{name}", REF / f"{name}.py")
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mods[name] = mod
    return mods["equations"], mods["integrator"]


SYM_NAMES = ("XIJ", "RIJ", "R2IJ", "RHOIJ", "WI", "DWI", "DWJ", "GHI")
POSITIVE = {"rho", "h", "h0", "m", "cs", "p", "e", "epred", "e0", "omega", "dt_cfl"}


def random_props(rng, names, scale=1.0):
    vals = {}
    for nm in names:
        if nm in POSITIVE:
            vals[nm] = rng.uniform(0.5, 2.0, 2)
        else:
            vals[nm] = rng.normal(0.0, scale, 2)
    return vals


def call(method, props, sym, dt=None):
    sig = inspect.signature(method)
    kwargs = {}
    for pname in sig.parameters:
        if pname == "d_idx":
            kwargs[pname] = 0
        elif pname == "s_idx":
            kwargs[pname] = 1
        elif pname.startswith("d_") or pname.startswith("s_"):
            kwargs[pname] = props[pname[2:]]
        elif pname == "dt":
            kwargs[pname] = dt
        elif pname in ("XIJ", "DWI", "DWJ"):
            kwargs[pname] = sym[pname]
        else:
            kwargs[pname] = sym[pname]
    return method(**kwargs)


def prop_names_of(method):
    return sorted({p[2:] for p in inspect.signature(method).parameters
                   if p[:2] in ("d_", "s_") and p not in ("d_idx", "s_idx")})


def make_sym(rng, tiny=None):
    xij = rng.normal(0.0, 0.05, 3)
    if tiny is not None:
        xij = np.array([tiny, 0.0, 0.0])
    rij = float(np.sqrt(xij @ xij))
    return {
        "XIJ": xij.tolist(), "RIJ": rij, "R2IJ": rij * rij, "RHOIJ": float(rng.uniform(0.5, 2.0)),
        "WI": float(rng.uniform(0.0, 5.0)), "DWI": rng.normal(0, 20, 3).tolist(),
        "DWJ": rng.normal(0, 20, 3).tolist(), "GHI": float(rng.normal(0, 30)),
    }


def sym_vector(sym, dt):
    return np.array(list(sym["XIJ"]) + [sym["RIJ"], sym["R2IJ"], sym["RHOIJ"], sym["WI"]] + list(sym["DWI"])
                    + list(sym["DWJ"]) + [sym["GHI"], dt if dt is not None else 0.0])


def main():
    eq, integ = load_reference()
    rng = np.random.default_rng(20261203)
    values: list[float] = []
    manifest: list[dict] = []

    def put(arr) -> int:
        off = len(values)
        values.extend(float(x) for x in np.asarray(arr, dtype=np.float64).ravel())
        return off

    def record(body, params, method, props, sym, dt=None, flag_attr=None, obj=None):
        inp = {k: v.copy() for k, v in props.items()}
        ret = call(method, props, sym, dt)
        flag = getattr(obj, flag_attr) if flag_attr else 0
        rec = {"body": body, "params": put(params), "sym": put(sym_vector(sym, dt)),
               "flag": float(flag if flag is not None else 0.0),
               "props": {k: [put(inp[k]), put(props[k])] for k in props}}
        manifest.append(rec)
        return ret

    def P(k=1.2, dim=3, once=False, gamma=5.0 / 3.0, mu0=1.0, av=1.0, ac=1.0, ar=1.0, chf=1.0, sigma=1.0,
          tsf=1.0):
        return [k, dim, float(once), gamma, mu0, av, ac, ar, chf, sigma, tsf]

    for case in range(16):
        dim = 3 if case % 2 == 0 else 2
        mu0 = 1.0 if case % 3 else 1.25663706212e-6 * 1e6
        prm = dict(k=float(rng.uniform(1.0, 1.5)), dim=dim, gamma=float(rng.uniform(1.2, 2.0)), mu0=mu0,
                   av=float(rng.uniform(0.5, 1.5)), ac=float(rng.uniform(0.0, 1.5)), ar=float(rng.uniform(0.0, 1.5)),
                   chf=float(rng.uniform(0.0, 1.5)) if case % 5 else 0.0, sigma=float(rng.uniform(0.0, 1.5)),
                   tsf=float(rng.uniform(0.3, 1.0)))
        prv = P(**prm)
        tiny = [None, None, None, 0.0, 5e-9, 5e-13][case % 6]
        sym = make_sym(rng, tiny)

        # DensityIteration (equations.py:14-94)
        for once in (False, True):
            d = eq.DensityIteration("fluid", ["fluid"], k=prm["k"], dim=dim, iterate_only_once=once)
            pv = P(**{**prm, "dim": dim}); pv[2] = float(once)
            props = random_props(rng, set(prop_names_of(d.initialize)) | set(prop_names_of(d.loop))
                                 | set(prop_names_of(d.post_loop)))
            record("DENSITY_INIT", pv, d.initialize, props, sym)
            record("DENSITY_LOOP", pv, d.loop, props, sym)
            # post_loop: unconverged particle, rho sum near or far from m/(h/k)^d
            props["converged"][:] = [0.0 if case % 4 else 1.0, 0.0]
            hi, mi = props["h"][0], props["m"][0]
            target = mi / (hi / prm["k"]) ** dim
            props["rho"][0] = target * (1.0 + (1e-7 if case % 3 == 0 else rng.uniform(-0.3, 0.3)))
            props["h0"][0] = hi * rng.uniform(0.9, 1.1)
            props["dwdh"][0] = rng.normal(0.0, 1.0) * target / hi
            d.equation_has_converged = 1
            record("DENSITY_POST", pv, d.post_loop, props, sym, flag_attr="equation_has_converged", obj=d)

        e1 = eq.IdealGasEOS("fluid", None, gamma=prm["gamma"])
        props = random_props(rng, prop_names_of(e1.loop))
        record("EOS", prv, e1.loop, props, sym)
        e2 = eq.DensityFromSmoothingLength("fluid", None, k=prm["k"], dim=dim)
        props = random_props(rng, prop_names_of(e2.initialize))
        record("RHO_FROM_H", prv, e2.initialize, props, sym)
        e3 = eq.SmoothingLengthRateFromDensityRelation("fluid", None, k=prm["k"], dim=dim)
        props = random_props(rng, prop_names_of(e3.initialize))
        record("AH_FROM_RHO", prv, e3.initialize, props, sym)
        e4 = eq.SmoothingLengthRateFromForceDivergence("fluid", None, dim=dim)
        props = random_props(rng, prop_names_of(e4.initialize))
        record("AH_FROM_DIV", prv, e4.initialize, props, sym)

        e5 = eq.MagneticStressAccelerations("fluid", ["fluid"], mu0=mu0)
        props = random_props(rng, prop_names_of(e5.loop))
        record("MAGSTRESS_LOOP", prv, e5.loop, props, sym)

        e6 = eq.MHDAccelerations("fluid", ["fluid"], mu0=mu0, artificial_viscosity=prm["av"],
                                 artificial_thermal_conductivity=prm["ac"])
        names = set(prop_names_of(e6.initialize)) | set(prop_names_of(e6.loop)) | set(prop_names_of(e6.post_loop))
        props = random_props(rng, names)
        record("MHDACC_INIT", prv, e6.initialize, props, sym)
        props = random_props(rng, names)
        if case % 2:  # approaching pair (dot < 0 branch)
            props["upred"][0], props["upred"][1] = -abs(props["upred"][0]) * np.sign(sym["XIJ"][0] or 1.0), 0.0
        record("MHDACC_LOOP", prv, e6.loop, props, sym)
        record("MHDACC_POST", prv, e6.post_loop, props, sym)

        e7 = eq.CullenDehnenAlphaDiagnostics("fluid", ["fluid"])
        names = set(prop_names_of(e7.initialize)) | set(prop_names_of(e7.loop)) | set(prop_names_of(e7.post_loop))
        props = random_props(rng, names)
        record("CD_INIT", prv, e7.initialize, props, sym)
        props = random_props(rng, names)
        record("CD_LOOP", prv, e7.loop, props, sym)
        if case % 3 == 1:  # singular R matrix -> fallback branch (equations.py:400-410)
            for q in ("cd_rxz", "cd_ryz", "cd_rzz"):
                props[q][0] = 0.0
        record("CD_POST", prv, e7.post_loop, props, sym)

        e8 = eq.StrainDiagnostics("fluid", None)
        props = random_props(rng, prop_names_of(e8.initialize))
        if case % 3 == 1:
            for q in ("cd_rxz", "cd_ryz", "cd_rzz"):
                props[q][0] = 0.0
        record("STRAIN", prv, e8.initialize, props, sym)
        e9 = eq.MagneticStateReconstruction("fluid", None)
        props = random_props(rng, prop_names_of(e9.initialize))
        record("BRECON", prv, e9.initialize, props, sym)

        e10 = eq.MagneticStateRates("fluid", ["fluid"], mu0=mu0, artificial_magnetic_dissipation=prm["ar"],
                                    cleaning_speed_factor=prm["chf"], cleaning_damping_factor=prm["sigma"])
        names = set(prop_names_of(e10.initialize)) | set(prop_names_of(e10.loop)) | set(prop_names_of(e10.post_loop))
        props = random_props(rng, names)
        record("MAGRATES_INIT", prv, e10.initialize, props, sym)
        props = random_props(rng, names)
        record("MAGRATES_LOOP", prv, e10.loop, props, sym)
        record("MAGRATES_POST", prv, e10.post_loop, props, sym)

        e11 = eq.AdaptiveTimestep("fluid", None, timestep_factor=prm["tsf"])
        props = random_props(rng, prop_names_of(e11.initialize))
        if case % 4 == 1:
            props["dt_cfl"][0] = 0.0
        if case % 4 == 2:
            for q in ("au", "av", "aw"):
                props[q][0] = 0.0
        record("ADAPTIVE_DT", prv, e11.initialize, props, sym)

        e12 = eq.EnergyLimiter("fluid", None, timestep_factor=prm["tsf"])
        props = random_props(rng, prop_names_of(e12.initialize))
        if case % 2:  # strong cooling: triggers the limiter branch
            props["ae"][0] = -10.0 * props["epred"][0] * props["dt_cfl"][0] / (0.3 * prm["tsf"] * props["h"][0])
        record("ENERGY_LIMITER", prv, e12.initialize, props, sym)

        e13 = eq.DivBCorrection("fluid", None, mu0=mu0)
        props = random_props(rng, prop_names_of(e13.initialize))
        if case % 4 == 3:
            for q in ("Bxpred", "Bypred", "Bzpred"):
                props[q][0] = 0.0
        if case % 4 == 1:  # low plasma beta: full correction
            props["p"][0] = 1e-3
        record("DIVB_CORRECTION", prv, e13.initialize, props, sym)

        st = integ.MHDStep(prm["gamma"])
        names = set(prop_names_of(st.initialize)) | set(prop_names_of(st.stage1)) | set(prop_names_of(st.stage2))
        props = random_props(rng, names)
        props["itype"] = np.array([1.0 if case % 5 else 3.0, 1.0])
        props["converged"][:] = 1.0
        dt = float(rng.uniform(1e-4, 1e-2))
        record("STEP_INITIALIZE", prv, st.initialize, props, sym, dt)
        props = random_props(rng, names)
        props["itype"] = np.array([1.0 if case % 5 else 3.0, 1.0])
        if case % 2:
            props["alphaloc"][0] = props["alpha10"][0] + 0.5
        record("STEP_STAGE1", prv, st.stage1, props, sym, dt)
        props = random_props(rng, names)
        props["itype"] = np.array([1.0 if case % 5 else 3.0, 1.0])
        record("STEP_STAGE2", prv, st.stage2, props, sym, dt)

    import json

    np.savez_compressed(OUT, values=np.asarray(values, dtype=np.float64),
                        manifest=np.frombuffer(json.dumps(manifest).encode(), dtype=np.uint8))
    print(f"wrote {len(manifest)} records, {len(values)} values to {OUT}")


if __name__ == "__main__":
    main()
