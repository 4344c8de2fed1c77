"""Throughput of the SPH-MHD step on one MI355X.

One step = one full predictor-evaluate-corrector step (integrator.py:12-82
around the group plan of scheme.py:70-81, all density iterations included)
over every particle of the workload.  Inputs are device-resident before the
timed region; the timed region is exactly --steps steps between two stream
synchronisations, after the warm start (application.py:110-147) and
--warmup untimed steps.

Headline workload: the 3-D MHD blast on a 128^3 lattice (2,097,152
particles, quintic kernel, FP64), the largest single-GPU configuration of
the benchmark set.  The JSON line also carries the other single-GPU
workloads (the 512^2 rotor, the 1M-particle Orszag-Tang vortex and the
64-instance uneven ensemble), the roofline of the dominant kernel (the fused
force pass) with the density passes' beside it, and the CPU oracle timed on
a bounded sample of the headline workload on this host.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "pysph-mhd_amd"))

METRIC = "particle-steps/s & HBM roofline % (plain vs batched vs fwd+adjoint), 1 GPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64, vector and matrix alike

# Algorithmic bytes of one force launch per particle (compulsory traffic: each
# live input read once, each output written once, FP64): reads x, y, h, m,
# rho, p, cs, epred, upred(3), Bpred(3), psipred, omega, alpha1, a_prev(3)
# (20 fields, 160 B in 2-D; 21 in 3-D) and writes a(3), ae, ah, aBevol(3),
# apsi, alphaloc, dt_adapt (11 fields, 88 B).
FORCE_BYTES_PER_PARTICLE = {2: 248.0, 3: 256.0}

# The first steps of a context time each workgroup size and launch mode
# (plain or persistent) of the density, list and force launches and keep the
# fastest (results do not depend on it); the bench runs them before the
# warm-up steps, outside the timed region.
TUNE_STEPS = 12

# Measured HBM traffic of the force kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE
# per dispatch, scaled by a plain copy of known size), kept in the profile
# summaries as "traffic_bytes_per_launch[<workload>]: <bytes>".
TRAFFIC_FILE = ROOT / "profiles" / "k_force_hbm_traffic.md"


# FP64 operations the force kernel executes per neighbour entry, counted by
# rocprofv3 (SQ_INSTS_VALU_ADD_F64, _MUL_F64, _FMA_F64, _TRANS_F64 per
# dispatch: 64 lanes x (add + mul + trans + 2 fma) / pairs), kept in the
# profile summaries as "fp64_flops_per_pair[<workload>]: <flops>".
FLOPS_FILE = ROOT / "profiles" / "k_force_fp64_ops.md"

# Algorithmic FP64 work per neighbour entry of the force pass and of a
# density pass: the oracle's own arithmetic counted over an operation-counting
# scalar (oracle/flopcount.cpp, tests/test_flops_cpu.py), kept as
# "alg_flops_per_pair[<workload>]" and "alg_density_flops_per_pair[<workload>]".
# The Orszag-Tang variants share the 1M vortex's figures.
ALG_FLOPS_FILE = ROOT / "profiles" / "k_force_alg_flops.md"
ALG_FLOPS_KEY = {"ot128": "ot1m", "ot128x64": "ot1m", "ensemble64": "ot1m"}

# Used only for a workload with no counted figure: the operation count of the
# fused reference bodies written out by hand (add/sub/mul/div/sqrt 1, fma 2).
FLOPS_PER_PAIR_ESTIMATE = {2: 330.0, 3: 390.0}

# Density pass: compulsory bytes per particle of one full pass (reads x, y,
# (z), h, m, h0, converged; writes rho, h, omega, converged) and, where no
# counted figure exists, FP64 operations per neighbour entry (distance, kernel
# value and h-derivative, two sums), add/sub/mul/div/sqrt 1, fma 2.
DENSITY_BYTES_PER_PARTICLE = {2: 80.0, 3: 88.0}
DENSITY_FLOPS_PER_PAIR = 30.0


def _profile_value(path: Path, key: str):
    """The number after "<key>:" on a line of a committed profile summary,
    or None."""
    try:
        for line in path.read_text().splitlines():
            if line.strip().startswith(key + ":"):
                return float(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def measured_traffic(workload: str):
    """HBM bytes per force launch for this workload from the committed
    profile summary, or None when it has not been profiled."""
    return _profile_value(TRAFFIC_FILE, f"traffic_bytes_per_launch[{workload}]")


def flops_per_pair(workload: str, dim: int):
    """(algorithmic FP64 flops per neighbour entry of the force pass, how it
    was obtained): the oracle's operation count, else the estimate."""
    key = ALG_FLOPS_KEY.get(workload, workload)
    v = _profile_value(ALG_FLOPS_FILE, f"alg_flops_per_pair[{key}]")
    if v is not None:
        return v, "algorithmic: the oracle's reference bodies over an operation-counting scalar"
    return FLOPS_PER_PAIR_ESTIMATE[dim], "estimate (hand count of the reference bodies)"


def executed_flops_per_pair(workload: str):
    """FP64 operations the GPU force kernel executes per neighbour entry
    (rocprofv3 VALU instruction counters), or None when not profiled."""
    return _profile_value(FLOPS_FILE, f"fp64_flops_per_pair[{workload}]")


def density_flops_per_pair(workload: str):
    key = ALG_FLOPS_KEY.get(workload, workload)
    v = _profile_value(ALG_FLOPS_FILE, f"alg_density_flops_per_pair[{key}]")
    return v if v is not None else DENSITY_FLOPS_PER_PAIR


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="blast128",
                    choices=["rotor512", "ot1m", "blast128", "ensemble64", "ot128", "ot128x64"])
    ap.add_argument("--no-extras", action="store_true", help="only the headline workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the C5 gradient by the adjoint step against the plain rollout (one JSON line)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 512 MiB FP64 buffer 5 times (scales HBM byte counters)")
    return ap.parse_args()


def calibrate_copy():
    """A plain device copy of known size: the FETCH_SIZE / WRITE_SIZE counters
    of its dispatches give the factor between counted and moved bytes."""
    import torch

    n = 512 * 1024 * 1024 // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                      "copy_GBps": round(2 * n * 8 / dt / 1e9, 1)}))


def make_setup(name: str, host_fields: bool = True):
    """The workload's setup; host_fields=False leaves the initial fields to
    the GPU generator (cases.device.load_setup)."""
    from sphmhd.cases import workloads as B

    hf = host_fields
    if name == "rotor512":
        return B.rotor_2d(512, host_fields=hf), "rotor_2d 512x512 (262,144 particles), cubic kernel, h=1.2dx"
    if name == "ot1m":
        return B.orszag_tang_2d(1024, host_fields=hf), "orszag_tang_2d 1024x1024 (1,048,576 particles), quintic, h=1.2dx"
    if name == "ot128":
        return B.orszag_tang_2d(128, host_fields=hf), "orszag_tang_2d 128x128 (16,384 particles), quintic, h=1.2dx, one instance"
    if name == "ot128x64":
        s = B.ensemble_of([B.orszag_tang_2d(128, host_fields=hf) for _ in range(64)])
        return s, "64 identical orszag_tang_2d 128x128 instances in one batch (1,048,576 particles)"
    if name == "blast128":
        return B.blast_3d(128, host_fields=hf), "blast_3d 128^3 (2,097,152 particles), quintic, h=0.81dx"
    if name == "ensemble64":
        s = B.ensemble_ot(64, host_fields=hf)
        return s, f"ensemble_ot 64 instances, N_k = n_k^2 with n_k~U{{64..256}} seed 0 ({s.n:,} particles)"
    raise ValueError(name)


def run_gpu(name: str, steps: int, warmup: int, timing_steps: int = 10):
    """Timed region: `steps` steps in one sphmhd_step call (steps after the
    first replay the captured hipGraph), between two stream synchronisations,
    without per-phase events (each event record drains the queue; on the
    512^2 rotor they cost ~10% of a step).  Then `timing_steps` steps
    with HIP events on the library's stream give the per-kernel durations the
    roofline uses (one sphmhd_step call: all but its last step skip the
    diagnostic rows, as in the timed region)."""
    import torch

    from sphmhd.cases.device import load_setup
    from sphmhd.device import DeviceSystem

    setup, desc = make_setup(name, host_fields=False)
    dev = DeviceSystem(setup.params, setup.offsets, device="cuda:0")
    load_setup(dev, setup)  # initial fields generated on the GPU
    dev.warm_start()
    dev.step(TUNE_STEPS)
    if warmup:
        dev.step(warmup)
    dev.set_timing(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.step(steps)
    dev.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = dev.stats()
    dev.set_timing(True)
    dev.step(timing_steps)  # one call, as in the timed region (per-phase events, no enqueueing ahead)
    dev.synchronize()
    tm = dev.timing()
    pairs = dev.stats()["pairs_force"]
    res = {
        "workload": name,
        "desc": desc,
        "particles": setup.n,
        "instances": len(setup.params),
        "dim": int(setup.params[0].dim),
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "particle_steps_per_s": setup.n * steps / wall,
        "timing": tm,
        "timing_steps": timing_steps,
        "stats": st,
        "launch": dev.launch()[0],
        "persistent": [st["persist_density"], st["persist_lists"], st["persist_force"]],
        "pairs_timing": pairs,
    }
    dev.close()
    del dev
    torch.cuda.empty_cache()
    return res


def run_tangent(steps: int, warmup: int = 2):
    """Forward step with derivatives against the plain step on the travelling
    Alfven wave (97x168, Wendland C6, h=1.9dx): the tangent-linear step
    advances the values and NTANGENTS directional derivatives (here d/dA of
    the initial amplitude, d/d alpha_max, d/d sigma), so one sweep gives the
    gradient of any final-time total in those three inputs."""
    import torch

    from sphmhd.cases.workloads import alfven_2d
    from sphmhd.device import DeviceSystem
    from sphmhd.tangent import TangentSystem, _ic_tangent

    names = ("amplitude", "alpha_max", "cleaning_damping_factor")
    theta = dict(zip(names, (0.1, 1.0, 1.0)))
    setup = alfven_2d(97, 168)
    out = {"desc": "alfven_2d 97x168 (16,296 particles), wendland_c6, h=1.9dx; derivatives along "
                   "d/d amplitude, d/d alpha_max, d/d cleaning_damping_factor",
           "particles": setup.n, "steps": steps}
    for label, cls in (("plain", DeviceSystem), ("tangent", TangentSystem)):
        d = cls(setup.params, setup.offsets, device="cuda:0")
        d.set_fields(dict(setup.fields))
        if cls is TangentSystem:
            d.set_tangents(0, _ic_tangent(lambda **kw: alfven_2d(97, 168, **kw), theta, "amplitude"))
            d.seed_params({(1, "alpha_max"): 1.0, (2, "cleaning_damping_factor"): 1.0})
        d.warm_start()
        d.step(warmup + TUNE_STEPS)
        d.set_timing(False)
        d.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.step(steps)
        d.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out[label + "_ms_per_step"] = round(1e3 * wall / steps, 4)
        out[label + "_particle_steps_per_s"] = round(setup.n * steps / wall, 1)
        d.close()
        del d
    torch.cuda.empty_cache()
    out["tangent_over_plain"] = round(out["tangent_ms_per_step"] / out["plain_ms_per_step"], 3)
    return out


def run_adjoint(steps: int = 400, every: int = 20):
    """The gradient of the final E_kin + E_mag of the C5 rollout (the
    travelling Alfven wave, 97x168, warm start + `steps` steps) with respect
    to every initial property and the six parameters, by the adjoint step:
    the forward rollout recorded with checkpoints every `every` steps, then
    one reverse sweep (each segment recomputed from its checkpoint, each
    step's evaluation recomputed before it is reversed).  Beside it, the plain
    rollout (warm start + `steps` steps) on the same context type."""
    import torch

    from sphmhd.adjoint import ENERGY, AdjointSystem
    from sphmhd.cases.workloads import alfven_2d
    from sphmhd.device import DeviceSystem

    setup = alfven_2d(97, 168)
    out = {"desc": "alfven_2d 97x168 (16,296 particles), wendland_c6, h=1.9dx: gradient of the final E_kin + E_mag "
                   "with respect to every initial property and the six parameters",
           "particles": setup.n, "steps": steps, "checkpoint_every": every}
    d = DeviceSystem(setup.params, setup.offsets, device="cuda:0")
    d.set_fields(dict(setup.fields))
    d.warm_start()
    d.step(TUNE_STEPS + 2)
    d.set_fields(dict(setup.fields))
    d.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.warm_start()
    d.step(steps)
    d.synchronize()
    fwd = time.perf_counter() - t0
    d.close()
    a = AdjointSystem(setup.params, setup.offsets, device="cuda:0")
    a.set_fields(dict(setup.fields))
    a.record(TUNE_STEPS + 2, every)  # launch-shape tuning and first-call costs, untimed
    a.zero_bars()
    a.seed_totals(ENERGY)
    a.reverse()
    a.set_fields(dict(setup.fields))
    a.zero_bars()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    a.record(steps, every)
    torch.cuda.synchronize()
    rec = time.perf_counter() - t0
    a.seed_totals(ENERGY)
    t0 = time.perf_counter()
    a.reverse()
    torch.cuda.synchronize()
    rev = time.perf_counter() - t0
    a.close()
    torch.cuda.empty_cache()
    out["forward_ms"] = round(1e3 * fwd, 2)
    out["forward_ms_per_step"] = round(1e3 * fwd / steps, 4)
    out["record_ms"] = round(1e3 * rec, 2)
    out["reverse_ms"] = round(1e3 * rev, 2)
    out["reverse_ms_per_step"] = round(1e3 * rev / steps, 4)
    out["gradient_over_forward"] = round((rec + rev) / fwd, 3)
    out["reverse_over_forward"] = round(rev / fwd, 3)
    out["note"] = ("gradient_over_forward: (record + reverse sweep) / plain rollout, warm start included in both; "
                   "the reverse sweep recomputes each segment and each step's evaluation from the checkpoints")
    return out


def roofline(res):
    """Roofline of the dominant kernel (the fused force pass).  Its binding
    roof is the FP64 vector rate (no step is a dense contraction, so no MFMA
    runs): per launch it does flops_per_pair x pairs of algorithmic FP64 work
    (the reference bodies' operation count) on bytes_alg of compulsory HBM
    traffic, about 20-30 flop/B against a ridge of 9.8 flop/B.  The HBM
    fraction of the same launch, the FP64 rate the kernel's executed
    instructions give, and the density passes' figures are reported beside
    it."""
    tm = res["timing"]
    n = res["particles"]
    dim = res["dim"]
    launches = max(1, tm["force_launches"])
    t_force = tm["force_ms"] / launches * 1e-3
    bytes_alg = FORCE_BYTES_PER_PARTICLE[dim] * n
    pairs = res["pairs_timing"]
    fpp, fpp_how = flops_per_pair(res["workload"], dim)
    flops = fpp * pairs
    gbs = bytes_alg / t_force / 1e9
    tflops = flops / t_force / 1e12
    t_attain = max(bytes_alg / (HBM_PEAK_GBS * 1e9), flops / (FP64_PEAK_TFLOPS * 1e12))
    # density passes: per step, one full pass's compulsory bytes and flops
    # (later passes only revisit unconverged particles)
    steps = max(1, res["timing_steps"])
    t_dens = tm["density_ms"] / steps * 1e-3
    dens = None
    if t_dens > 0:
        dbytes = DENSITY_BYTES_PER_PARTICLE[dim] * n
        dflops = density_flops_per_pair(res["workload"]) * pairs
        dens = {
            "kernel": "k_density (G0 summation density + Newton step for h), all passes of a step",
            "ms_per_step": round(t_dens * 1e3, 4),
            "bytes_alg_per_step": dbytes,
            "flops_alg_per_step": dflops,
            "hbm_GBps": round(dbytes / t_dens / 1e9, 2),
            "hbm_frac": round(dbytes / t_dens / 1e9 / HBM_PEAK_GBS, 5),
            "fp64_TFLOPs": round(dflops / t_dens / 1e12, 3),
            "fp64_frac": round(dflops / t_dens / 1e12 / FP64_PEAK_TFLOPS, 4),
        }
    return {
        "bound": "fp64",
        "bound_note": "FP64 vector ALU rate (78.6 TFLOP/s on MI355X); no MFMA: no step is a dense contraction",
        "kernel": "k_force (fused G3+G5+G7+G8 pair loop, G3/G5/G6/G8/G9 epilogue)",
        "achieved": round(tflops, 3),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
        "traffic": measured_traffic(res["workload"]),
        "avg_launch_ms": round(tm["force_ms"] / launches, 5),
        "launches": tm["force_launches"],
        "flops_per_pair": fpp,
        "flops_per_pair_source": fpp_how,
        "fp64_executed_per_pair": executed_flops_per_pair(res["workload"]),
        "fp64_executed_TFLOPs": (round(executed_flops_per_pair(res["workload"]) * pairs / t_force / 1e12, 3)
                                 if executed_flops_per_pair(res["workload"]) else None),
        "pairs_per_launch": pairs,
        "bytes_alg_per_launch": bytes_alg,
        "hbm": {
            "achieved": round(gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 5),
        },
        "attainable_frac": round(t_attain / t_force, 4),
        "density": dens,
    }


def cpu_baseline(name: str, steps: int):
    sys.path.insert(0, str(ROOT / "oracle"))
    from pyoracle import OracleSystem

    cores = len(os.sched_getaffinity(0))
    threads = min(cores, 16)
    setup, desc = make_setup(name)
    orc = OracleSystem(setup.params, setup.offsets, threads=threads)
    orc.set_fields({k: v for k, v in setup.fields.items()})
    orc.warm_start()
    t0 = time.perf_counter()
    orc.step(steps)
    wall = time.perf_counter() - t0
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    orc.close()
    return {
        "value": round(setup.n * steps / wall, 1),
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{desc}: warm start untimed, then {steps} PEC steps timed ({wall:.1f} s); "
                  f"FP64 C oracle, OpenMP {threads} threads on {model or platform.processor()}",
    }


def cpu_ot128(steps: int = 100, steps_one_thread: int = 20):
    """The oracle on the 2-D Orszag-Tang vortex at 128^2 particles (quintic,
    PEC steps after the warm start) at one thread (a bounded sample of
    steps_one_thread steps, about 10 s) and at
    every host core (the full 100 steps), with the conserved totals at steps
    0 and 100 of the latter: particle-steps/s, the all-core/1-thread ratio
    and the relative drift of mass, momentum and total energy."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from pyoracle import OracleSystem

    from sphmhd.cases.workloads import orszag_tang_2d

    cores = len(os.sched_getaffinity(0))
    s = orszag_tang_2d(128)
    out = {"config": "Orszag-Tang 2-D, 128^2 = 16,384 particles, quintic, h = 1.2 dx, FP64",
           "cores_available": cores}
    for label, threads, nst in (("one_thread", 1, steps_one_thread), ("all_cores", min(cores, 16), steps)):
        orc = OracleSystem(s.params, s.offsets, threads=threads)
        orc.set_fields(dict(s.fields))
        orc.warm_start()
        t0 = orc.totals()[0]
        tic = time.perf_counter()
        orc.step(nst)
        wall = time.perf_counter() - tic
        t1 = orc.totals()[0]
        orc.close()
        out[label] = {"threads": threads, "steps": nst, "particle_steps_per_s": round(s.n * nst / wall, 1),
                      "seconds": round(wall, 2)}
    e0 = t0[4] + t0[5] + t0[6]
    e1 = t1[4] + t1[5] + t1[6]
    out["speedup_all_cores"] = round(out["all_cores"]["particle_steps_per_s"] / out["one_thread"]["particle_steps_per_s"], 2)
    out["totals_drift"] = {  # over the all-core run's steps
        "mass": float(abs(t1[0] - t0[0]) / abs(t0[0])),
        "momentum": float(np.hypot(t1[1] - t0[1], t1[2] - t0[2]) / np.sqrt(2.0 * t0[0] * t0[4])),
        "energy": float(abs(e1 - e0) / abs(e0)),
        "divb_start": float(t0[7]), "divb_end": float(t1[7]),
        "note": "momentum against sqrt(2 M E_kin); energy = kinetic + thermal + magnetic",
    }
    return out


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.adjoint_only:
        print(json.dumps(run_adjoint(400, 20)))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs one GPU per context; sharding a workload over several GPUs "
                                   "is not implemented, so --gpus must be 1"}))
        sys.exit(2)
    head = run_gpu(args.workload, args.steps, args.warmup)
    extras = {}
    if not args.no_extras:
        for name, steps in (("rotor512", 20), ("ot1m", 20), ("blast128", 10), ("ensemble64", 20), ("ot128", 50),
                            ("ot128x64", 20)):
            if name == args.workload:
                continue
            r = run_gpu(name, steps, 2)
            extras[name] = {
                "desc": r["desc"],
                "particles": r["particles"],
                "instances": r["instances"],
                "steps": steps,
                "particle_steps_per_s": round(r["particle_steps_per_s"], 1),
                "ms_per_step": round(r["ms_per_step"], 4),
                "pair_interactions_per_s": round(r["stats"]["pairs_force"] * steps / r["wall_s"], 1),
                "launch_blocks": list(r["launch"]),
                "persistent": r["persistent"],
                "roofline": roofline(r),
            }
        if "ensemble64" in extras and "ot1m" in extras:
            extras["ensemble_vs_plain_equal_n"] = {
                "ratio": round(extras["ensemble64"]["particle_steps_per_s"] / extras["ot1m"]["particle_steps_per_s"], 3),
                "what": "particle-steps/s of the 64-instance uneven ensemble (1.89M particles) over one plain "
                        "Orszag-Tang instance of 1M particles",
            }
        if "ot128" in extras and "ot128x64" in extras:
            extras["batched_vs_plain"] = {
                "ratio": round(extras["ot128x64"]["particle_steps_per_s"] / extras["ot128"]["particle_steps_per_s"], 2),
                "what": "particle-steps/s of 64 OT 128x128 instances in one batch over one OT 128x128 instance alone",
            }
        extras["fwd_tangent"] = run_tangent(100)
        extras["fwd_adjoint"] = run_adjoint(400, 20)
    line = {
        "metric": METRIC,
        "value": round(head["particle_steps_per_s"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (lattice initial conditions of the named workload)",
        "config": {"workload": head["desc"], "particles": head["particles"], "steps_timed": args.steps,
                   "step": "PEC step incl. all density iterations", "parallelism": "single GPU"},
        "pair_interactions_per_s": round(head["stats"]["pairs_force"] * args.steps / head["wall_s"], 1),
        "timing_ms": {k: round(v, 4) if isinstance(v, float) else v for k, v in head["timing"].items()},
        "timing_note": f"HIP-event sums over one call of {head['timing_steps']} steps run after the timed region",
        "graph_replays": head["stats"]["graph_replays"],
        "launch_blocks": {"density": head["launch"][0], "lists": head["launch"][1], "force": head["launch"][2],
                          "persistent": head["persistent"],
                          "note": "workgroup sizes chosen by timing the first steps"},
        "roofline": roofline(head),
        "extra_workloads": extras,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_steps)
        if not args.no_extras:
            line["cpu_ot128"] = cpu_ot128()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
