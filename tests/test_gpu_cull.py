"""cull_fully_aliased_terms (the reference's dead culling, meshless_vis.cu:39-127, made
to work): a term whose projected footprint lies wholly outside the image's period
window is dropped from the sum.  Pinned by construction, not by restating the rule:
particles placed well inside the window, well outside it, and straddling its edge.
  * data inside the window: the culled render equals the unculled one bit for bit;
  * mixed data: the culled render equals the render of the data without the outside
    particles (relative L2 <= 1e-6: the sums run in another order) and the FP64 oracle
    of that data (<= 1e-5); straddling particles are kept;
  * the batched graph is keyed on the flag (toggling it re-captures).  Its culled renders
    are held to <= 2e-6 of the kept-only render: the two bin different particle sets
    (the outside particles keep their h-bin slots with weight 0), so their h-interpolation
    errors, each about 1-2e-6 against FP64, differ; and to <= 1e-5 of the FP64 oracle."""
import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import host

pytestmark = pytest.mark.gpu

CUTOFF, SAMPLES, STEP = (32, 32), (64, 64), (1.0, 1.0)  # period window |u|, |v| <= 0.5
U, V = (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)


def cloud(rng, n, xlo, xhi, r):
    pos = np.empty((n, 3), np.float32)
    pos[:, 0] = rng.uniform(xlo, xhi, n)
    pos[:, 1] = rng.uniform(-0.3, 0.3, n)
    pos[:, 2] = rng.uniform(-2.0, 2.0, n)  # depth does not matter for the projection
    w = rng.uniform(0.1, 1.0, n).astype(np.float32)
    h = rng.uniform(0.8 * r, 1.2 * r, n).astype(np.float32)  # per-particle radii: binned MFMA kinds
    return pos, w, h


def parts(seed):
    rng = np.random.default_rng(seed)
    inside = cloud(rng, 3000, -0.3, 0.3, 0.03)
    outside = cloud(rng, 1500, 0.7, 0.9, 0.03)  # footprint <= 3 * 0.036 from the centre: wholly outside
    straddle = cloud(rng, 200, 0.49, 0.51, 0.03)  # crosses |u| = 0.5: kept
    return inside, outside, straddle


def cat(*ps):
    return tuple(np.concatenate([p[i] for p in ps]) for i in range(3))


def render(groups, form, phi_mode, cull):
    ds = host.Dataset(groups)
    with host.Config(cutoff=CUTOFF, samples=SAMPLES, step=STEP, u=U, v=V, form=form, phi_mode=phi_mode,
                     cull=cull) as cfg:
        cfg.register(ds)
        try:
            cfg.render(ds)
            F, img = cfg.spectrum(), cfg.image()
        finally:
            cfg.unregister(ds)
    return F, img


CASES = [("plain", "GAUSSIAN", "analytic"), ("plain", "SPH", "reference"),
         ("mfma_split_f16", "WENDLAND_D3_C2", "analytic"), ("mfma_split_f16", "GAUSSIAN", "analytic")]


@pytest.mark.parametrize("form,basis,phi_mode", CASES)
def test_cull_inside_data_unchanged(form, basis, phi_mode):
    inside, _, _ = parts(1)
    g = [(*inside, basis)]
    print(f"[case] cull inside {form} {basis}", flush=True)
    F0, i0 = render(g, form, phi_mode, False)
    F1, i1 = render(g, form, phi_mode, True)
    np.testing.assert_array_equal(F0, F1)
    np.testing.assert_array_equal(i0, i1)


@pytest.mark.parametrize("form,basis,phi_mode", CASES)
def test_cull_drops_outside_terms_only(form, basis, phi_mode):
    inside, outside, straddle = parts(2)
    everything = [(*cat(inside, outside, straddle), basis)]
    kept = [(*cat(inside, straddle), basis)]
    print(f"[case] cull mixed {form} {basis}", flush=True)
    F_cull, img_cull = render(everything, form, phi_mode, True)
    F_kept, img_kept = render(kept, form, phi_mode, False)
    F_all, _ = render(everything, form, phi_mode, False)
    e = oracle.rel_l2(F_cull, F_kept)
    F64 = oracle.spectrum64(kept, CUTOFF, STEP, U, V, phi_mode)
    e64 = oracle.rel_l2(F_cull, F64)
    print(f"culled vs kept-only render {e:.2e}, vs FP64 oracle {e64:.2e}, "
          f"culled vs unculled {oracle.rel_l2(F_cull, F_all):.2e}")
    assert e <= 1e-6
    assert e64 <= 1e-5
    assert oracle.rel_l2(img_cull, img_kept) <= 1e-6
    assert oracle.rel_l2(F_cull, F_all) > 1e-3  # the outside particles did contribute


def test_cull_batched_graph_follows_the_flag():
    inside, outside, straddle = parts(3)
    groups = [(*cat(inside, outside, straddle), "WENDLAND_D3_C2")]
    kept = [(*cat(inside, straddle), "WENDLAND_D3_C2")]
    ds = host.Dataset(groups)
    cfgs = [host.Config(cutoff=CUTOFF, samples=SAMPLES, step=STEP, u=U, v=V, form="mfma_split_f16",
                        phi_mode="analytic") for _ in range(2)]
    try:
        cfgs[0].register(ds)
        out = {}
        for cull in (True, False, True):
            for c in cfgs:
                c.cs.cull_fully_aliased_terms = cull
            host.render_batched([ds, ds], cfgs)
            out.setdefault(cull, []).append([c.spectrum() for c in cfgs])
        cfgs[0].unregister(ds)
    finally:
        for c in cfgs:
            c.close()
    F_kept, _ = render(kept, "mfma_split_f16", "analytic", False)
    F_all, _ = render(groups, "mfma_split_f16", "analytic", False)
    F64 = oracle.spectrum64(kept, CUTOFF, STEP, U, V, "analytic")
    for F in out[True][0] + out[True][1]:
        print(f"culled batch vs kept-only render {oracle.rel_l2(F, F_kept):.2e}, vs FP64 {oracle.rel_l2(F, F64):.2e}")
        assert oracle.rel_l2(F, F_kept) <= 2e-6
        assert oracle.rel_l2(F, F64) <= 1e-5
    for F in out[False][0]:
        assert oracle.rel_l2(F, F_all) <= 1e-6
    np.testing.assert_array_equal(out[True][0][0], out[True][1][0])
