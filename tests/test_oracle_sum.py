"""Pins of the oracle's direct sum (spectrum) and of its FP32 reference port.

Known answers and invariants, each chosen so that a plausible slip fails:
  * one particle at (1/4, 0, 0): F at x=1 is -i*s*w*Phi(1) (phase sign, conjugate, scale);
  * one particle at the origin: F[y][x0] = s*w*Phi(r) with the reference's index map
    (x0 > cx -> x = x0 - 2cx) and unequal step sizes (axis swaps, index slips);
  * translation by d multiplies F by exp(-2 pi i k.d); linearity in the weights;
    y=0 row Hermitian; group additivity;
  * a pure-Python brute force with Phi from quadrature on a tiny case;
  * the FP32 port against the FP64 oracle on the 4096-particle config.
"""
import cmath
import math

import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import workloads as W

from test_oracle_phi import quad_phi


def xy_of(cx, cy):
    x0 = np.arange(2 * cx)
    x = np.where(x0 > cx, x0 - 2 * cx, x0)
    y = np.arange(cy)
    return np.broadcast_to(x[None, :], (cy, 2 * cx)), np.broadcast_to(y[:, None], (cy, 2 * cx))


def test_single_particle_quarter_phase():
    pos = np.array([[0.25, 0.0, 0.0]], np.float32)
    w = np.array([2.0], np.float32)
    F = oracle.spectrum64([(pos, w, None, "GAUSSIAN")], (4, 4), (1.0, 1.0), (1, 0, 0), (0, 1, 0), "analytic")
    expect = -1j * 2.0 * math.exp(-math.pi)
    assert abs(F[0, 1] - expect) <= 1e-14
    # x = -1 (x0 = 7) is the conjugate phase
    assert abs(F[0, 7] - np.conj(expect)) <= 1e-14


@pytest.mark.parametrize("mode", ["analytic", "reference"])
def test_index_map_particle_at_origin(mode):
    cx, cy, sx, sy = 5, 3, 0.3, 0.5
    pos = np.zeros((1, 3), np.float32)
    F = oracle.spectrum64([(pos, np.ones(1, np.float32), None, "GAUSSIAN")], (cx, cy), (sx, sy), (1, 0, 0), (0, 1, 0),
                          mode)
    x, y = xy_of(cx, cy)
    r2 = (np.float32(sx) * x.astype(np.float64)) ** 2 + (np.float32(sy) * y.astype(np.float64)) ** 2
    sign = -1.0 if mode == "analytic" else 1.0
    s = float(np.float32(sx)) * float(np.float32(sy))
    expect = s * np.exp(sign * math.pi * r2)
    assert np.max(np.abs(F - expect)) <= 1e-13 * np.abs(expect).max()


def _cloud(rng, n, basis="WENDLAND_D3_C2"):
    return (rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32), rng.uniform(-1, 1, n).astype(np.float32),
            rng.uniform(0.02, 0.2, n).astype(np.float32), basis)


VIEW = dict(cutoff=(6, 5), step=(0.8, 1.1), u=(0.6, 0.8, 0.0), v=(-0.48, 0.36, 0.8))


def test_translation_is_a_phase_ramp():
    rng = np.random.default_rng(1)
    _, w, h, b = _cloud(rng, 50)
    pos = (rng.integers(-512, 512, (50, 3)) / 1024.0).astype(np.float32)  # dyadic: pos + d is exact in FP32
    d = np.array([0.125, -0.25, 0.5], np.float32)
    F0 = oracle.spectrum64([(pos, w, h, b)], **VIEW, phi_mode="analytic")
    F1 = oracle.spectrum64([(pos + d, w, h, b)], **VIEW, phi_mode="analytic")
    x, y = xy_of(*VIEW["cutoff"])
    u = np.array(VIEW["u"], np.float32).astype(np.float64)
    v = np.array(VIEW["v"], np.float32).astype(np.float64)
    sx, sy = (float(np.float32(t)) for t in VIEW["step"])
    k = (sx * x)[..., None] * u + (sy * y)[..., None] * v
    ramp = np.exp(-2j * math.pi * (k @ d.astype(np.float64)))
    assert oracle.rel_l2(F1, F0 * ramp) <= 1e-12


def test_linearity_and_group_additivity():
    rng = np.random.default_rng(2)
    pos, w1, h, b = _cloud(rng, 40)
    w2 = rng.uniform(-1, 1, 40).astype(np.float32)
    F1 = oracle.spectrum64([(pos, w1, h, b)], **VIEW)
    F2 = oracle.spectrum64([(pos, w2, h, b)], **VIEW)
    F12 = oracle.spectrum64([(pos, w1, h, b), (pos, w2, h, b)], **VIEW)
    assert oracle.rel_l2(F12, F1 + F2) <= 1e-13
    w3 = (w1 * np.float32(2.0)).astype(np.float32)
    F3 = oracle.spectrum64([(pos, w3, h, b)], **VIEW)
    assert oracle.rel_l2(F3, 2 * F1) <= 1e-13


def test_hermitian_y0_row():
    rng = np.random.default_rng(3)
    pos, w, h, b = _cloud(rng, 30, "SPH")
    cx = VIEW["cutoff"][0]
    F = oracle.spectrum64([(pos, w, h, b)], **VIEW)
    for x in range(1, cx):
        assert abs(F[0, 2 * cx - x] - np.conj(F[0, x])) <= 1e-13 * abs(F).max()


def test_brute_force_tiny():
    rng = np.random.default_rng(4)
    n = 5
    pos, w, h, b = _cloud(rng, n)
    cx, cy = 3, 2
    sx, sy = np.float32(0.9), np.float32(0.7)
    u = np.array([0.0, 0.6, 0.8], np.float32)
    v = np.array([1.0, 0.0, 0.0], np.float32)
    F = oracle.spectrum64([(pos, w, h, b)], (cx, cy), (sx, sy), u, v, "analytic")
    for y in range(cy):
        for x0 in range(2 * cx):
            x = x0 - 2 * cx if x0 > cx else x0
            fu, fv = float(sx) * x, float(sy) * y
            k = [fu * float(u[i]) + fv * float(v[i]) for i in range(3)]
            r = math.hypot(fu, fv)
            acc = 0j
            for i in range(n):
                th = 2 * math.pi * sum(k[j] * float(pos[i, j]) for j in range(3))
                acc += float(w[i]) * quad_phi("WENDLAND_D3_C2", r * float(h[i])) * cmath.exp(-1j * th)
            acc *= float(sx) * float(sy)
            assert abs(F[y, x0] - acc) <= 1e-9, (x0, y, F[y, x0], acc)


def test_port_against_fp64_on_uniform_gaussian_4096():
    wl = W.uniform_gaussian_4096()
    vw = wl.views[0]
    F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode)
    F32 = oracle.port_spectrum32([wl.group()], wl.cutoff, wl.step, vw.u, vw.v)
    e = oracle.rel_l2(F32, F64)
    # The reference's own FP32 error on this config is a few 1e-6 (serial FP32 sums, FP32 phase).
    assert 1e-7 <= e <= 2e-5


def test_port_row_subset_leaves_other_rows():
    rng = np.random.default_rng(5)
    g = _cloud(rng, 64, "GAUSSIAN")
    full = oracle.port_spectrum32([g], (4, 8), (1, 1), (1, 0, 0), (0, 1, 0))
    part = oracle.port_spectrum32([g], (4, 8), (1, 1), (1, 0, 0), (0, 1, 0), row_stride=4)
    assert np.array_equal(part[::4], full[::4])
    assert not np.any(part[1::4])


def test_port_analytic_phi_mode():
    """The port with the exact transforms in place of the reference's fits: within the FP32 port's
    own error of the FP64 analytic oracle, and measurably apart from the reference-formula port
    (the Wendland fit below r = 0.24 deviates from the exact transform by up to 2.9e-4 of Phi(0))."""
    rng = np.random.default_rng(8)
    g = _cloud(rng, 3000, "WENDLAND_D3_C2")
    u, v = (0.6, 0.8, 0.0), (-0.48, 0.36, 0.8)
    cutoff, step = (16, 12), (0.5, 0.5)
    F64 = oracle.spectrum64([g], cutoff, step, u, v, "analytic")
    Fan = oracle.port_spectrum32([g], cutoff, step, u, v, phi_mode="analytic")
    Fref = oracle.port_spectrum32([g], cutoff, step, u, v, phi_mode="reference")
    e = oracle.rel_l2(Fan, F64)
    assert 1e-8 <= e <= 2e-5, e
    assert oracle.rel_l2(Fref, F64) > 10 * e


def test_port_at_indices_equals_whole_plane():
    """Frequencies computed one by one (idx) are the whole-plane port's values bit for bit,
    for a first group and an accumulated second group."""
    rng = np.random.default_rng(9)
    groups = [_cloud(rng, 300, "GAUSSIAN"), _cloud(rng, 200, "SPH")]
    u, v = (1, 0, 0), (0, 0.6, 0.8)
    full = oracle.port_spectrum32(groups, (6, 5), (0.7, 0.9), u, v).reshape(-1)
    idx = rng.choice(full.size, 17, replace=False)
    part = oracle.port_spectrum32(groups, (6, 5), (0.7, 0.9), u, v, idx=idx)
    assert np.array_equal(part, full[idx])
