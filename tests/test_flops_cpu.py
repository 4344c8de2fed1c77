"""Algorithmic FP64 work per pair, counted by the oracle over an
operation-counting scalar (oracle/flopcount.cpp: the oracle's own source
with every double a Counted; + - * / sqrt cbrt exp count one each, while a
pair is evaluated).

The figures bench.py divides by the force pass's time are committed in
profiles/k_force_alg_flops.md; this test recomputes them on each workload at
a reduced size (warm start, two steps, then one counted evaluation, one
thread) and holds the committed ones to 0.1%:

  force   the symbols the four force passes read (r, rho_ij, grad W_i,
          grad W_j), once per pair as the fused pass evaluates them, plus the
          loop bodies of the Cullen-Dehnen, MHD acceleration, magnetic stress
          and magnetic state-rate passes (equations.py:143-567);
  density the symbols the density loop's live sums read (r, W_i, dW_i/dh)
          plus its two live sums, rho += m_j W_i and dwdh += m_j dW_i/dh
          (two multiplies, two adds; equations.py:45,50), per pass.

The pair separation (XIJ, r^2: 8 operations in 3-D) is computed in the
neighbour search and not counted.  That the counting build computes exactly
what the oracle computes is checked field by field, bit for bit.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

PROFILE = Path(__file__).resolve().parents[1] / "profiles" / "k_force_alg_flops.md"

STATE = ("x", "y", "z", "u", "v", "w", "e", "h", "rho", "omega", "au", "av", "ae", "ah", "alphaloc", "Bevolx",
         "Bevoly", "psi", "alpha1", "aBevolx", "apsi", "div", "dt_cfl", "dt_adapt", "divBsymm", "divBdiff")


def _setup(workload):
    from sphmhd.cases import workloads as B

    return {
        "blast128": lambda: B.blast_3d(24),
        "rotor512": lambda: B.rotor_2d(64),
        "ot1m": lambda: B.orszag_tang_2d(64),
    }[workload]()


def _counts(setup):
    from pyoracle import OracleSystem

    o = OracleSystem(setup.params, setup.offsets, threads=1, counted=True)
    o.set_fields(dict(setup.fields))
    o.warm_start()
    o.step(2)
    o.flop_reset()
    o.eval()
    body, sym, pairs = o.flop_counts()
    o.close()
    return body, sym, pairs


def _figures(body, sym, pairs):
    per_sym = sym / pairs.sum()  # every pass evaluates every symbol once per pair
    per_body = body / np.maximum(pairs, 1)
    force = per_sym[0] + per_sym[1] + per_sym[3] + per_sym[4] + per_body[1:].sum()
    density = per_sym[0] + per_sym[2] + per_sym[5] + 4.0
    return force, density


def _committed(key):
    for line in PROFILE.read_text().splitlines():
        if line.strip().startswith(key + ":"):
            return float(line.split(":", 1)[1])
    raise KeyError(key)


@pytest.mark.parametrize("workload", ["blast128", "rotor512", "ot1m"])
def test_committed_figures(workload):
    body, sym, pairs = _counts(_setup(workload))
    # one pass per force group, two density iterations (the reference's minimum)
    assert pairs[1] == pairs[2] == pairs[3] == pairs[4] > 0
    assert pairs[0] == 2 * pairs[1]
    force, density = _figures(body, sym, pairs)
    print(f"{workload}: force {force:.2f}, density {density:.2f} flop per pair")
    assert force == pytest.approx(_committed(f"alg_flops_per_pair[{workload}]"), rel=1e-3)
    assert density == pytest.approx(_committed(f"alg_density_flops_per_pair[{workload}]"), rel=1e-3)


@pytest.mark.parametrize("which", ["ot", "blast"])
def test_counting_build_computes_the_oracle(which):
    """The counted library is the oracle: a warm start and three steps give
    every field bit for bit."""
    from pyoracle import OracleSystem
    from sphmhd.cases import workloads as B

    s = B.jittered(B.orszag_tang_2d(24)) if which == "ot" else B.jittered(B.blast_3d(10))
    out = []
    for counted in (False, True):
        o = OracleSystem(s.params, s.offsets, threads=1, counted=counted)
        o.set_fields(dict(s.fields))
        o.warm_start()
        o.step(3)
        out.append({k: o.get(k).copy() for k in STATE})
        o.close()
    for k in STATE:
        assert np.array_equal(out[0][k], out[1][k]), k
