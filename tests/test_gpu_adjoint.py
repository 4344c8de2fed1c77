"""The adjoint (reverse-mode) step with checkpointed recomputation on the GPU
(sphmhd_adjoint_*, sphmhd/adjoint.py).

  * the reverse sweep's gradient with respect to every initial property and
    the six parameters equals the taped oracle's (oracle/tape.cpp: the
    oracle's own arithmetic recorded and reversed, the CPU reference
    gradient) on a jittered Alfven wave through the warm start and three
    steps, at the derivative standard of 1e-8 (two exact derivatives of maps
    that agree to rounding);
  * along the directions of the tangent-linear step (the initial amplitude's
    exact field derivative, alpha_max, the cleaning damping) the adjoint
    gradient equals the tangent sweep's derivatives: <grad, v> == J v;
  * the checkpoint interval does not change the gradient (every 1, 3, 20 and
    one segment for the whole rollout give the same bits);
  * the full C5 rollout (97 x 168, 400 steps, checkpoints every 20 steps):
    dE/d amplitude, dE/d alpha_max and dE/d cleaning damping against central
    differences of the plain step at eps = 1e-6 and 1e-5 relative, to 1e-5;
  * torch.autograd through RolloutEnergy: the gradient reaches the amplitude
    the initial fields were built from in torch;
  * 3-D (a jittered 6^3 blast: the Cullen-Dehnen velocity-gradient inversion,
    z components) and frozen particles (itype != 1) against the taped
    oracle; an uneven two-instance ensemble equals each instance alone.

Fields compared by relative L2 over the particles; a field whose reference
gradient is zero must be zero to 1e-12 of the largest gradient's scale.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# initial properties the rollout reads (the rest is overwritten before it is
# read, or constant): their gradients are compared
INPUTS = ("x", "y", "u", "v", "w", "rho", "h", "e", "Bx", "By", "Bz", "alpha1", "au", "av", "aw")
TPARAMS = ("artificial_viscosity", "artificial_thermal_conductivity", "artificial_magnetic_dissipation",
           "cleaning_speed_factor", "cleaning_damping_factor", "alpha_max")


def _small(amplitude=0.1, **kw):
    from sphmhd.cases.workloads import alfven_2d, jittered

    return jittered(alfven_2d(10, 17, amplitude=amplitude, **kw), 0.1, seed=5)


def _print(*a):
    print(*a, flush=True)


def test_adjoint_matches_taped_oracle(gpu):
    from pyoracle import tape_rollout
    from sphmhd.adjoint import energy_gradient

    s = _small()
    # nonzero initial accelerations so that their (Cullen-Dehnen) gradient is exercised
    rng = np.random.default_rng(2)
    f = dict(s.fields)
    for k in ("au", "av"):
        f[k] = 0.05 * rng.standard_normal(f["x"].size)
    s = type(s)(s.name, s.params, s.offsets, f, info=s.info)
    steps = 3
    _print("adjoint vs taped oracle: alfven 10x17 jittered, warm start + 3 steps")
    value, grads, gpar = energy_gradient(s, steps, every=2)
    tvals, tgrads, tgpar, nodes = tape_rollout(s.params, s.offsets, s.fields, nsteps=steps)
    assert value == pytest.approx(tvals[0], rel=1e-12)
    scale = max(float(np.abs(tgrads[k]).max()) for k in INPUTS)
    worst = 0.0
    for k in INPUTS:
        a, b = grads[k], tgrads[k]
        nb = float(np.linalg.norm(b))
        if nb <= 1e-12 * scale:
            assert float(np.abs(a).max()) <= 1e-12 * scale, (k, float(np.abs(a).max()))
            continue
        e = float(np.linalg.norm(a - b)) / nb
        worst = max(worst, e)
        _print(f"  d/d {k}: rel-L2 {e:.2e}")
        assert e <= 1e-8, (k, e)
    for p, name in enumerate(TPARAMS):
        ref = tgpar[0, p]
        _print(f"  d/d {name}: adjoint {gpar[name]:.12e} tape {ref:.12e}")
        assert abs(gpar[name] - ref) <= 1e-8 * abs(ref) + 1e-14 * abs(value), name
    _print(f"adjoint vs taped oracle: worst field rel-L2 {worst:.2e} ({nodes} tape nodes)")


def test_adjoint_equals_tangent_directions(gpu):
    from sphmhd.adjoint import energy_gradient
    from sphmhd.cases.workloads import alfven_2d
    from sphmhd.tangent import energy_rollout

    def setup_fn(amplitude=0.1, alpha_max=1.0, cleaning_damping_factor=1.0):
        return alfven_2d(24, 42, amplitude=amplitude, alpha_max=alpha_max,
                         cleaning_damping_factor=cleaning_damping_factor)

    names = ("amplitude", "alpha_max", "cleaning_damping_factor")
    theta = (0.1, 1.0, 1.0)
    steps = 10
    _print("adjoint vs tangent: alfven 24x42, 10 steps")
    e_t, g_t = energy_rollout(setup_fn, theta, names, steps)
    s = setup_fn()
    value, grads, gpar = energy_gradient(s, steps, every=4)
    assert value == pytest.approx(e_t, rel=1e-12)
    tan = s.info["tangents"]["amplitude"]
    g_amp = sum(float(np.dot(grads[k], tan[k])) for k in tan)
    got = (g_amp, gpar["alpha_max"], gpar["cleaning_damping_factor"])
    for k, nm in enumerate(names):
        _print(f"  dE/d {nm}: adjoint {got[k]:.12e} tangent {g_t[k]:.12e}")
        assert abs(got[k] - g_t[k]) <= 1e-9 * abs(g_t[k]) + 1e-14 * abs(value), (nm, got[k], g_t[k])


def test_adjoint_checkpoint_interval_does_not_change_the_gradient(gpu):
    from sphmhd.adjoint import energy_gradient

    s = _small()
    _print("adjoint checkpoint intervals: alfven 10x17, 6 steps")
    ref = None
    for every in (1, 3, 20):
        value, grads, gpar = energy_gradient(s, 6, every=every)
        got = (value, np.concatenate([grads[k] for k in INPUTS]), np.array(list(gpar.values())))
        if ref is None:
            ref = got
            continue
        assert got[0] == ref[0]
        assert np.array_equal(got[1], ref[1]), every
        assert np.array_equal(got[2], ref[2]), every


def test_adjoint_recomputation_redone_with_waits_equals_default(gpu, monkeypatch):
    """The reverse sweep recomputes each segment without a host wait per
    evaluation; an evaluation that needs more density passes than were
    launched raises the abort flag and the segment is recomputed with a wait
    after each evaluation.  With one speculative pass (SPHMHD_SPEC=1) every
    evaluation aborts, so every segment takes that second path: the value and
    the gradient equal the default run's bit for bit.  The density passes run
    as launches of their own here (SPHMHD_DENSITY_FUSED=0): fused into the
    first launch, one pass never falls short."""
    from sphmhd.adjoint import energy_gradient

    monkeypatch.setenv("SPHMHD_DENSITY_FUSED", "0")
    s = _small()
    _print("adjoint recomputation with waits: alfven 10x17, 6 steps, segments of 3")
    value, grads, gpar = energy_gradient(s, 6, every=3)
    monkeypatch.setenv("SPHMHD_SPEC", "1")
    v1, g1, p1 = energy_gradient(s, 6, every=3)
    assert v1 == value
    for k in INPUTS:
        assert np.array_equal(g1[k], grads[k]), k
    assert list(p1.values()) == list(gpar.values())


def test_c5_adjoint_400_steps_vs_central_differences(gpu):
    """C5: the 97 x 168 Alfven wave, 400 steps, checkpoints every 20 steps;
    the three gradients of the tangent sweep's test, from one reverse sweep."""
    from sphmhd.adjoint import energy_gradient
    from sphmhd.cases.workloads import alfven_2d
    from sphmhd.tangent import energy_primal

    steps = 400
    names = ("amplitude", "alpha_max", "cleaning_damping_factor")
    theta = (0.1, 1.0, 1.0)
    kw = dict(zip(names, theta))
    _print("C5 adjoint: alfven 97x168, 400 steps, every 20")
    s = alfven_2d(**kw)
    e, grads, gpar = energy_gradient(s, steps, every=20)
    tan = s.info["tangents"]["amplitude"]
    g = (sum(float(np.dot(grads[k], tan[k])) for k in tan), gpar["alpha_max"], gpar["cleaning_damping_factor"])
    e0 = energy_primal(alfven_2d, kw, steps)
    assert abs(e - e0) <= 1e-12 * abs(e0)
    for k, nm in enumerate(names):
        fd = {}
        for rel in (1e-5, 1e-6):
            eps = rel * theta[k]
            ep = energy_primal(alfven_2d, {**kw, nm: theta[k] + eps}, steps)
            em = energy_primal(alfven_2d, {**kw, nm: theta[k] - eps}, steps)
            fd[rel] = (ep - em) / (2.0 * eps)
        noise = 4e-16 * abs(e0) / (1e-6 * theta[k])
        _print(f"  dE/d {nm}: adjoint {g[k]:.10e} FD(1e-6) {fd[1e-6]:.10e} FD(1e-5) {fd[1e-5]:.10e}")
        assert abs(fd[1e-5] - fd[1e-6]) <= 1e-5 * abs(fd[1e-6]) + 2 * noise, (nm, fd)
        assert abs(g[k] - fd[1e-6]) <= 1e-5 * abs(fd[1e-6]) + noise, (nm, g[k], fd[1e-6])


def test_rollout_energy_autograd_reaches_the_amplitude(gpu):
    import torch

    from sphmhd.adjoint import energy_gradient, rollout_energy

    s = _small()
    tan = s.info["tangents"]["amplitude"]
    base = {k: np.asarray(s.fields[k]) - 0.1 * np.asarray(tan[k]) for k in tan}  # the fields are linear in A
    amp = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    fields = {k: torch.as_tensor(base[k]) + amp * torch.as_tensor(np.asarray(tan[k])) for k in tan}
    p = s.params[0]
    theta = torch.tensor([getattr(p, n) for n in TPARAMS], dtype=torch.float64, requires_grad=True)
    e = rollout_energy(s, fields, theta, steps=4, every=2)
    (2.0 * e).backward()
    value, grads, gpar = energy_gradient(s, 4, every=2)
    assert float(e.detach()) == pytest.approx(value, rel=1e-12)
    g_amp = sum(float(np.dot(grads[k], tan[k])) for k in tan)
    assert float(amp.grad) == pytest.approx(2.0 * g_amp, rel=1e-10)
    assert np.allclose(theta.grad.numpy(), 2.0 * np.array([gpar[n] for n in TPARAMS]), rtol=1e-12, atol=0)


def _compare_fields(grads, ref, label, tol):
    scale = max(float(np.abs(ref[k]).max()) for k in INPUTS)
    worst = 0.0
    for k in INPUTS:
        a, b = grads[k], ref[k]
        nb = float(np.linalg.norm(b))
        if nb <= 1e-12 * scale:
            assert float(np.abs(a).max()) <= 1e-12 * scale, (label, k, float(np.abs(a).max()))
            continue
        e = float(np.linalg.norm(a - b)) / nb
        worst = max(worst, e)
        assert e <= tol, (label, k, e)
    return worst


def test_adjoint_3d_matches_taped_oracle(gpu):
    """3-D: a jittered 6^3 blast (quintic), warm start and one step: the
    velocity-gradient inversion of the Cullen-Dehnen switch and the z
    components."""
    from pyoracle import tape_rollout
    from sphmhd.adjoint import energy_gradient
    from sphmhd.cases.workloads import blast_3d, jittered

    s = jittered(blast_3d(6), 0.1, seed=4)
    _print("adjoint vs taped oracle: blast 6^3 jittered, warm start + 1 step")
    value, grads, gpar = energy_gradient(s, 1, every=1)
    tvals, tgrads, tgpar, nodes = tape_rollout(s.params, s.offsets, s.fields, nsteps=1)
    assert value == pytest.approx(tvals[0], rel=1e-12)
    inputs = INPUTS + ("z",)
    worst = _compare_fields(grads, tgrads, "3-D", 1e-8)
    e = float(np.linalg.norm(grads["z"] - tgrads["z"]) / np.linalg.norm(tgrads["z"]))
    assert e <= 1e-8, e
    for p, name in enumerate(TPARAMS):
        ref = tgpar[0, p]
        assert abs(gpar[name] - ref) <= 1e-8 * abs(ref) + 1e-14 * abs(value), (name, gpar[name], ref)
    _print(f"  3-D worst field rel-L2 {max(worst, e):.2e} over {inputs}")


def test_adjoint_frozen_particles_match_taped_oracle(gpu):
    """Particles with itype != 1 keep their state through the stages
    (integrator.py:38, 67): their bars pass through them."""
    from pyoracle import tape_rollout
    from sphmhd.adjoint import energy_gradient

    s = _small()
    f = dict(s.fields)
    it = np.ones(f["x"].size, dtype=np.int32)
    it[::7] = 3
    f["itype"] = it
    s = type(s)(s.name, s.params, s.offsets, f, info=s.info)
    value, grads, gpar = energy_gradient(s, 3, every=2)
    tvals, tgrads, tgpar, _ = tape_rollout(s.params, s.offsets, s.fields, nsteps=3)
    assert value == pytest.approx(tvals[0], rel=1e-12)
    worst = _compare_fields(grads, tgrads, "frozen", 1e-8)
    for p, name in enumerate(TPARAMS):
        ref = tgpar[0, p]
        assert abs(gpar[name] - ref) <= 1e-8 * abs(ref) + 1e-14 * abs(value), (name, gpar[name], ref)
    _print(f"  frozen particles: worst field rel-L2 {worst:.2e}")


def test_adjoint_ensemble_instances_equal_alone(gpu):
    """Two uneven Orszag-Tang instances in one context: each instance's
    gradient (fields and parameters) equals the instance run alone."""
    import torch

    from sphmhd.adjoint import AdjointSystem
    from sphmhd.cases.workloads import Setup, ensemble_ot, jittered

    ens = jittered(ensemble_ot(2, seed=3, sizes=[14, 18]), 0.1, seed=9)

    def grad_of(setup):
        sysm = AdjointSystem(setup.params, setup.offsets)
        sysm.set_fields(dict(setup.fields))
        sysm.record(3, 2)
        sysm.zero_bars()
        sysm.seed_totals({"ekin": 1.0, "emag": 1.0})
        pb = sysm.reverse()
        torch.cuda.synchronize()
        g = {k: sysm.gradient(k) for k in INPUTS}
        sysm.close()
        return g, pb

    g_all, pb_all = grad_of(ens)
    for k in range(2):
        a, b = int(ens.offsets[k]), int(ens.offsets[k + 1])
        alone = Setup(f"inst{k}", [ens.params[k]], np.array([0, b - a], dtype=np.int64),
                      {name: np.asarray(v)[a:b].copy() for name, v in ens.fields.items()})
        g1, pb1 = grad_of(alone)
        for name in INPUTS:
            ref = g1[name]
            got = g_all[name][a:b]
            assert np.allclose(got, ref, rtol=1e-11, atol=1e-13 * float(np.abs(ref).max() + 1e-300)), (k, name)
        assert np.allclose(pb_all[k], pb1[0], rtol=1e-11, atol=1e-18), (k, pb_all[k], pb1[0])
