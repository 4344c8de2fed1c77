"""The hand-written reverse bodies of the adjoint step (csrc/sph_adjoint.h),
checked on the host against forward-mode duals.

tests/adjoint_selfcheck.hip draws random pair and particle states (every
branch of the bodies taken in some trials) and checks the dot-product
identity <g, v> == <ybar, f'(x) v> between the reverse functions (g, the
gradient of <ybar, f>) and a dual evaluation of the forward functions along a
random direction v: the reverse of the fused force pair body (2-D cubic and
Wendland C6, 3-D quintic and Gaussian) and of the force epilogue (2-D and
3-D, with the velocity-gradient inversion, the energy limiter, the divB
correction's clamp and the alpha_loc floor).  Agreement to rounding (1e-11
relative to the terms' magnitudes); a transposition mistake shows as O(1).
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


def _hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    return None


@pytest.fixture(scope="module")
def selfcheck(tmp_path_factory):
    cc = _hipcc()
    if cc is None:
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("adj") / "selfcheck"
    cmd = [cc, "-std=c++17", "-O1", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", str(HERE / "adjoint_selfcheck.hip"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = {}
    for line in out.splitlines():
        parts = line.split()
        if parts and not parts[0].endswith("_branches"):
            rows[parts[0]] = (int(parts[1]), float(parts[2]))
        elif parts:
            rows[parts[0]] = {parts[k]: int(parts[k + 1]) for k in range(1, len(parts) - 1, 2)}
    return rows


@pytest.mark.parametrize("name", ["pair_cubic_2d", "pair_wendland_c6_2d", "pair_quintic_3d", "pair_gaussian_3d"])
def test_pair_reverse_matches_duals(selfcheck, name):
    trials, worst = selfcheck[name]
    assert trials >= 200
    assert worst <= 1e-11, f"{name}: reverse and dual disagree by {worst:.3e}"


@pytest.mark.parametrize("name", ["epi_2d", "epi_3d"])
def test_epilogue_reverse_matches_duals(selfcheck, name):
    trials, worst = selfcheck[name]
    assert trials >= 500
    assert worst <= 1e-11, f"{name}: reverse and dual disagree by {worst:.3e}"
    br = selfcheck[name + "_branches"]
    # every branch family is exercised by a good share of the trials
    assert br["limiter"] >= 50 and br["alpha"] >= 50 and br["divb_clamp_open"] >= 50, br
