"""Pins of the oracle's Phi (the 3-D Fourier transforms of the bases).

Each check is independent of the closed forms being tested:
  * numerical quadrature of Phi(k) = (2/k) * int_0^R rho phi(rho) sin(2 pi k rho) drho
    for the basis functions themselves (M4 spline with support 2, Wendland C2 on
    the unit ball, exp(-pi rho^2));
  * special values Phi(0) = int phi dV (pi, 2pi/21, 1);
  * continuity of the reference's quadratic fits with its closed forms at the
    switch radii (a wrong fit coefficient breaks it), and the fits' bounded
    deviation from the exact transform below the switch.
"""
import math
import warnings

import numpy as np
import pytest
from scipy import integrate

import oracle


def m4(q):
    if q < 1.0:
        return 1.0 - 1.5 * q * q + 0.75 * q ** 3
    if q < 2.0:
        return 0.25 * (2.0 - q) ** 3
    return 0.0


def wendland(q):
    return (1.0 - q) ** 4 * (4.0 * q + 1.0) if q < 1.0 else 0.0


def gaussian(q):
    return math.exp(-math.pi * q * q)


BASES = {"SPH": (m4, 2.0), "WENDLAND_D3_C2": (wendland, 1.0), "GAUSSIAN": (gaussian, 7.0)}


def quad_phi(basis, k):
    phi, R = BASES[basis]
    if k == 0.0:
        val, _ = integrate.quad(lambda r: 4 * math.pi * r * r * phi(r), 0, R, limit=400, epsabs=1e-14, epsrel=1e-13,
                                points=[1.0] if basis == "SPH" else None)
        return val
    f = lambda r: r * phi(r)
    # Where the transform is far below epsabs (the Gaussian at large k) QUADPACK reports roundoff
    # before reaching it; the comparisons below bound the error absolutely, so that is harmless.
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", integrate.IntegrationWarning)
        val, _ = integrate.quad(f, 0, R, weight="sin", wvar=2 * math.pi * k, limit=800, epsabs=1e-15, epsrel=1e-13)
    return 2.0 / k * val


R_GRID = np.concatenate([[0.0, 1e-3, 0.01, 0.05], np.linspace(0.06, 6.0, 37)])


@pytest.mark.parametrize("basis", ["SPH", "WENDLAND_D3_C2", "GAUSSIAN"])
def test_analytic_phi_matches_quadrature(basis):
    got = oracle.phi64(basis, "analytic", R_GRID)
    scale = quad_phi(basis, 0.0)
    for r, g in zip(R_GRID, got):
        q = quad_phi(basis, float(r))
        assert abs(g - q) <= 1e-9 * scale + 1e-8 * abs(q), (basis, r, g, q)


def test_phi_at_zero():
    assert oracle.phi64("SPH", "analytic", [0.0])[0] == pytest.approx(math.pi, rel=1e-14)
    assert oracle.phi64("WENDLAND_D3_C2", "analytic", [0.0])[0] == pytest.approx(2 * math.pi / 21, rel=1e-14)
    assert oracle.phi64("GAUSSIAN", "analytic", [0.0])[0] == 1.0


def test_reference_gaussian_has_the_reference_sign():
    # fourier_transform_cpu.cpp:51 returns exp(+pi r^2); the transform of exp(-pi x^2) is exp(-pi r^2).
    assert oracle.phi64("GAUSSIAN", "reference", [1.0])[0] == pytest.approx(math.exp(math.pi), rel=1e-14)
    assert oracle.phi64("GAUSSIAN", "analytic", [1.0])[0] == pytest.approx(math.exp(-math.pi), rel=1e-14)


@pytest.mark.parametrize("basis,switch", [("SPH", 0.06), ("WENDLAND_D3_C2", 0.24)])
def test_reference_fits(basis, switch):
    scale = oracle.phi64(basis, "analytic", [0.0])[0]
    above = np.linspace(switch, 5.0, 200)
    ref = oracle.phi64(basis, "reference", above)
    ana = oracle.phi64(basis, "analytic", above)
    # Above the switch the reference closed form is the exact transform.
    assert np.max(np.abs(ref - ana)) <= 1e-10 * scale
    # The fit meets the closed form at the switch ...
    below_edge = np.nextafter(switch, 0.0)
    fit_edge = oracle.phi64(basis, "reference", [below_edge])[0]
    exact_edge = oracle.phi64(basis, "analytic", [switch])[0]
    assert abs(fit_edge - exact_edge) <= 1e-6 * scale
    # ... and deviates from the exact transform below it by at most 3e-4 of Phi(0), not zero.
    below = np.linspace(0.0, below_edge, 200)
    dev = np.abs(oracle.phi64(basis, "reference", below) - oracle.phi64(basis, "analytic", below)) / scale
    assert dev.max() <= 3e-4
    assert dev.max() >= 1e-7


@pytest.mark.parametrize("basis", ["SPH", "WENDLAND_D3_C2", "GAUSSIAN"])
def test_port_phi32_follows_reference_formula(basis):
    r = np.linspace(0.0, 3.0, 301).astype(np.float32)
    p32 = oracle.phi32_port(basis, r).astype(np.float64)
    p64 = oracle.phi64(basis, "reference", r.astype(np.float64))
    scale = np.maximum(np.abs(p64), oracle.phi64(basis, "reference", [0.0])[0] * 1e-3)
    # FP32 cancellation in the reference closed forms near the switch radii is ~1e-4.
    assert np.max(np.abs(p32 - p64) / scale) <= 3e-4
