"""Seeded random configurations, each against the FP64 oracle on the whole half plane.

Every case draws, from its own seed: 1-3 groups (basis, particle count 1-3000, radii log-uniform
over up to two decades, sometimes one shared radius, sometimes some radii 0, weights of one sign
or both), a random view, step, cutoff (1-64 per axis) and image size (a multiple of 2 cx, 2 cy of
any factorisation), the form, the nodes per bin, the output tile, the partial-spectrum budget
(which may split the render into passes) and whether fully aliased terms are culled (the oracle
then drops the same terms).  The spectrum must be within the form's bar of the FP64 oracle, and
the image within 3e-6 of numpy's irfft2 of the arranged spectrum.  Bars: plain FP32 and split
FP16 1e-5, split BF16 5e-5, single FP16 5e-3 (measured at most 4.1e-6, 8.2e-6, 4.3e-7 and 2.7e-4;
images 0.8-2.0e-7)."""
import math

import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import abi, host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu

BASES = ["SPH", "GAUSSIAN", "WENDLAND_D3_C2"]
FORMS = ["plain", "mfma_split_f16", "mfma_split_f16", "mfma_split_bf16", "mfma_f16"]
BAR = {"plain": 1e-5, "mfma_split_f16": 1e-5, "mfma_split_bf16": 5e-5, "mfma_f16": 5e-3}
CULL_K = {"SPH": 2.0, "WENDLAND_D3_C2": 1.0, "GAUSSIAN": 3.0}


def draw(seed, bases=BASES):
    rng = np.random.default_rng(seed)
    groups = []
    for _ in range(int(rng.integers(1, 4))):
        n = int(rng.integers(1, 3001))
        basis = bases[int(rng.integers(0, 3)) % len(bases)]
        pos = rng.uniform(-0.6, 0.6, (n, 3)).astype(np.float32)
        signed = rng.random() < 0.5
        w = rng.uniform(-1 if signed else 0.1, 1, n).astype(np.float32)
        kind = rng.random()
        if kind < 0.15:
            h = np.full(n, rng.uniform(0.005, 0.05), np.float32)  # one shared radius
        else:
            lo = float(np.exp(rng.uniform(np.log(0.002), np.log(0.02))))
            hi = lo * float(np.exp(rng.uniform(0.0, np.log(100.0))))
            h = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.float32)
            if kind > 0.9:
                h[rng.random(n) < 0.1] = 0.0
        groups.append((pos, w, h, basis))
    # Block length 64 needs 2 cx cy to be a multiple of 64 (vis_config_check).
    cx = int(rng.integers(1, 65))
    g = 32 // math.gcd(cx, 32)
    cy = g * int(rng.integers(1, max(1, 64 // g) + 1))
    nx = 2 * cx * int(rng.integers(1, 3)) + int(rng.integers(0, 7))
    ny = 2 * cy * int(rng.integers(1, 3)) + int(rng.integers(0, 7))
    step = (float(rng.uniform(0.3, 1.0)), float(rng.uniform(0.3, 1.0)))
    form = FORMS[int(rng.integers(0, len(FORMS)))]
    nodes = [0, 0, 2, 3, 4][int(rng.integers(0, 5))]
    tile = ["default", "half_x", "half_y", "quarter"][int(rng.integers(0, 4))]
    budget = [0, 0, 1 << 16, 1][int(rng.integers(0, 4))]
    cull = rng.random() < 0.25
    vw = W.random_view(rng)
    return groups, (cx, cy), (nx, ny), step, form, nodes, tile, budget, cull, vw


def culled(groups, step, vw):
    """The oracle's copy of the groups with the culled terms' weights set to 0 (the library's
    rule: the square of half-side k r0 around the projected position lies outside the window)."""
    out = []
    # The config holds the step and the axes in FP32; the test is evaluated in FP64 from those.
    sx, sy = (float(np.float32(x)) for x in step)
    u = np.asarray(vw.u, np.float32).astype(np.float64)
    v = np.asarray(vw.v, np.float32).astype(np.float64)
    hu, hv = 0.5 / abs(sx), 0.5 / abs(sy)
    for pos, w, h, basis in groups:
        p = pos.astype(np.float64)
        du, dv = p @ u, p @ v
        r = CULL_K[basis] * h.astype(np.float64)
        drop = (du + r < -hu) | (du - r > hu) | (dv + r < -hv) | (dv - r > hv)
        out.append((pos, np.where(drop, 0.0, w).astype(np.float32), h, basis))
    return out


@pytest.mark.parametrize("seed,phi", [(s, "analytic") for s in range(48)] + [(s, "reference") for s in range(200, 216)])
def test_random_config(seed, phi):
    """phi: the exact transforms, or (SPH and Wendland groups only: the reference's Gaussian
    exp(+pi r^2) overflows at these radii) the reference's formulas with their quadratic fits."""
    bases = BASES if phi == "analytic" else ["SPH", "WENDLAND_D3_C2"]
    groups, cutoff, samples, step, form, nodes, tile, budget, cull, vw = draw(seed, bases)
    desc = (f"seed {seed} {phi}: {[(g[3], g[0].shape[0]) for g in groups]} cutoff {cutoff} samples {samples} "
            f"form {form} nodes {nodes} tile {tile} budget {budget} cull {cull}")
    print(f"[case] {desc}", flush=True)
    ds = host.Dataset(groups)
    try:
        abi.lib().vis_set_partial_budget(budget)
        with host.Config(cutoff=cutoff, samples=samples, step=step, u=vw.u, v=vw.v, form=form,
                         phi_mode=phi, nodes=nodes, tile=tile, block_length=64, cull=cull) as cfg:
            cfg.register(ds)
            try:
                cfg.render(ds)
                F, img = cfg.spectrum(), cfg.image()
            finally:
                cfg.unregister(ds)
    finally:
        abi.lib().vis_set_partial_budget(0)
    ref_groups = culled(groups, step, vw) if cull else groups
    F64 = oracle.spectrum64(ref_groups, cutoff, step, vw.u, vw.v, phi)
    e = oracle.rel_l2(F, F64)
    ref = np.fft.irfft2(oracle.arrange64(F.astype(np.complex128), *cutoff, *samples), s=samples, norm="forward")
    e_img = oracle.rel_l2(img, ref)
    print(f"{desc}: rel-L2 vs FP64 {e:.2e}, image vs irfft2 {e_img:.2e}", flush=True)
    assert e <= BAR[form]
    assert e_img <= 3e-6


@pytest.mark.parametrize("seed", range(100, 112))
def test_random_batch(seed):
    """A batch of 2-5 renders (random datasets and views) sharing one random grid and MFMA setting:
    the first call (no graph), the capture and a replay give the same images bit for bit, each
    within 1e-6 of its one-at-a-time render and its spectrum within the form's bar of FP64."""
    rng = np.random.default_rng(seed)
    _, cutoff, samples, step, form, nodes, tile, budget, _, _ = draw(seed)
    if form == "plain":
        form = "mfma_split_f16"
    k = int(rng.integers(2, 6))
    cases = [draw(seed * 1000 + i) for i in range(k)]
    sets = [host.Dataset(c[0]) for c in cases]
    vws = [c[9] for c in cases]
    desc = f"seed {seed}: {k} renders, cutoff {cutoff} samples {samples} {form} nodes {nodes} tile {tile} budget {budget}"
    print(f"[case] {desc}", flush=True)
    cfgs = [host.Config(cutoff=cutoff, samples=samples, step=step, u=vw.u, v=vw.v, form=form, phi_mode="analytic",
                        nodes=nodes, tile=tile, block_length=64) for vw in vws]
    try:
        abi.lib().vis_set_partial_budget(budget)
        for ds in sets:
            cfgs[0].register(ds)
        single = []
        for ds, cfg in zip(sets, cfgs):
            cfg.render(ds)
            single.append((cfg.image(), cfg.spectrum()))
        first = None
        for rep in range(3):
            host.render_batched(sets, cfgs)
            imgs = [cfg.image() for cfg in cfgs]
            if first is None:
                first = imgs
                spectra = [cfg.spectrum() for cfg in cfgs]
            else:
                for a, b in zip(imgs, first):
                    assert np.array_equal(a, b), (desc, rep)
    finally:
        abi.lib().vis_set_partial_budget(0)
        for ds in sets:
            cfgs[0].unregister(ds)
        for cfg in cfgs:
            cfg.close()
    worst = 0.0
    for (img1, _), img, F, c in zip(single, first, spectra, cases):
        assert oracle.rel_l2(img, img1) <= 1e-6, desc
        F64 = oracle.spectrum64(c[0], cutoff, step, c[9].u, c[9].v, "analytic")
        e = oracle.rel_l2(F, F64)
        worst = max(worst, e)
        assert e <= BAR[form], desc
    print(f"{desc}: batched vs single <= 1e-6, worst rel-L2 vs FP64 {worst:.2e}", flush=True)
