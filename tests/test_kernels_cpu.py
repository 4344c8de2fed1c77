"""The oracle's SPH kernels against what the mathematics fixes.

PySPH's kernel code is not in this container, so the kernels (W, grad W,
dW/dh; scheme.py:48-57 selects them) are pinned by closed forms:
  * normalisation: the integral of W over its support is 1 in 2-D and 3-D
    (the Gaussians are truncated at q = 3, which removes a known tail);
  * grad W and dW/dh agree with central finite differences of W;
  * W vanishes at and beyond the support radius R h and is continuous at
    the piece boundaries.
A wrong normalisation constant, sign or derivative fails one of these.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import integrate, special

KERNELS = {"cubic": 0, "quintic": 1, "wendland_c2": 2, "wendland_c4": 3, "wendland_c6": 4, "gaussian": 5,
           "super_gaussian": 6}


def _truncated_mass(name, dim):
    """Integral of the untruncated-normalised kernel over q < 3."""
    if name == "gaussian":
        if dim == 2:
            return 1.0 - math.exp(-9.0)
        return special.erf(3.0) - 6.0 / math.sqrt(math.pi) * math.exp(-9.0)
    if name == "super_gaussian":
        # exp(-q^2)(d/2 + 1 - q^2) / pi^(d/2): radial mass up to 3
        f = (lambda q: 2 * q * math.exp(-q * q) * (2.0 - q * q)) if dim == 2 else \
            (lambda q: 4 * q * q * math.exp(-q * q) * (2.5 - q * q) / math.sqrt(math.pi))
        return integrate.quad(f, 0.0, 3.0, epsabs=1e-13, epsrel=1e-13, limit=200)[0]
    return 1.0


@pytest.mark.parametrize("name", list(KERNELS))
@pytest.mark.parametrize("dim", [2, 3])
def test_normalisation(name, dim):
    import pyoracle as po

    k = KERNELS[name]
    h = 0.37
    R = po.kernel_radius(k) * h
    area = (lambda r: 2.0 * math.pi * r) if dim == 2 else (lambda r: 4.0 * math.pi * r * r)
    pts = [h * q for q in (1.0, 2.0) if h * q < R]
    mass = integrate.quad(lambda r: po.kernel_w(k, dim, r, h) * area(r), 0.0, R, points=pts or None,
                          epsabs=1e-13, epsrel=1e-13, limit=200)[0]
    assert mass == pytest.approx(_truncated_mass(name, dim), rel=2e-10, abs=2e-10)


@pytest.mark.parametrize("name", list(KERNELS))
@pytest.mark.parametrize("dim", [2, 3])
def test_gradient_matches_finite_difference(name, dim):
    import pyoracle as po

    k = KERNELS[name]
    h = 0.8
    rng = np.random.default_rng(k * 10 + dim)
    R = po.kernel_radius(k) * h
    for _ in range(40):
        x = rng.normal(0.0, 1.0, 3)
        if dim == 2:
            x[2] = 0.0
        x *= rng.uniform(0.05, 0.98) * R / np.linalg.norm(x)
        r = float(np.linalg.norm(x))
        g = po.kernel_grad(k, dim, x, r, h)
        for a in range(dim):
            e = 1e-6 * h
            xp, xm = x.copy(), x.copy()
            xp[a] += e
            xm[a] -= e
            fd = (po.kernel_w(k, dim, float(np.linalg.norm(xp)), h) - po.kernel_w(k, dim, float(np.linalg.norm(xm)), h)) / (2 * e)
            assert g[a] == pytest.approx(fd, rel=2e-6, abs=1e-7 * abs(po.kernel_w(k, dim, 0.0, h)))
        # dW/dh
        e = 1e-6 * h
        fdh = (po.kernel_w(k, dim, r, h + e) - po.kernel_w(k, dim, r, h - e)) / (2 * e)
        assert po.kernel_gradh(k, dim, r, h) == pytest.approx(fdh, rel=2e-6, abs=1e-7 * abs(po.kernel_w(k, dim, 0.0, h)) / h)


@pytest.mark.parametrize("name", list(KERNELS))
def test_compact_support_and_continuity(name):
    import pyoracle as po

    k = KERNELS[name]
    h = 1.0
    R = po.kernel_radius(k)
    for dim in (2, 3):
        assert po.kernel_w(k, dim, R, h) == 0.0
        assert po.kernel_w(k, dim, R * (1 + 1e-9), h) == 0.0
        assert po.kernel_dwdq(k, dim, R * (1 + 1e-9), h) == 0.0
        assert po.kernel_w(k, dim, 0.0, h) > 0.0
        for q in (1.0, 2.0):
            if q < R:
                a = po.kernel_w(k, dim, q * (1 - 1e-12), h)
                b = po.kernel_w(k, dim, q * (1 + 1e-12), h)
                assert a == pytest.approx(b, rel=1e-9, abs=1e-12)
                da = po.kernel_dwdq(k, dim, q * (1 - 1e-12), h)
                db = po.kernel_dwdq(k, dim, q * (1 + 1e-12), h)
                assert da == pytest.approx(db, rel=1e-9, abs=1e-12)


def test_self_pair_gradient_is_zero():
    """grad W of a self pair (r = 0) is 0 (PySPH guards r <= 1e-12)."""
    import pyoracle as po

    for k in KERNELS.values():
        g = po.kernel_grad(k, 3, [0.0, 0.0, 0.0], 0.0, 1.0)
        assert np.all(g == 0.0)


def test_quintic_matches_published_normalisation():
    """sigma_2 = 7/(478 pi), sigma_3 = 1/(120 pi) for the quintic spline
    (Morris 1996) and 10/(7 pi), 1/pi for the cubic (Monaghan 1992)."""
    import pyoracle as po

    assert po.kernel_w(1, 2, 0.0, 1.0) == pytest.approx(7.0 / (478.0 * math.pi) * 66.0, rel=1e-15)
    assert po.kernel_w(1, 3, 0.0, 1.0) == pytest.approx(1.0 / (120.0 * math.pi) * 66.0, rel=1e-15)
    assert po.kernel_w(0, 2, 0.0, 1.0) == pytest.approx(10.0 / (7.0 * math.pi), rel=1e-15)
    assert po.kernel_w(0, 3, 0.0, 1.0) == pytest.approx(1.0 / math.pi, rel=1e-15)
