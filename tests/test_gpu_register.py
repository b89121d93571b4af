"""Asynchronous registration (vis_register_meshless_dataset_async): uploads on the library's copy
stream, ordered before later work on the config's stream.  Pinned by image equality with the
synchronous path (meshless_vis.cu:264-290 semantics), including an upload of timestep t+1 that
overlaps the render of timestep t."""
import numpy as np
import pytest

from libmeshlessvis_amd import host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["mfma_split_f16", "plain"])
def test_async_registration_matches_sync(form):
    steps = W.drifting_clusters(steps=4, views_per_step=1, n_lo=2000, n_hi=40000, seed=7)
    sets = [host.Dataset([wl.group()]) for wl in steps]
    wl0 = steps[0]
    cfg = host.Config(cutoff=wl0.cutoff, samples=wl0.samples, step=wl0.step, form=form, phi_mode=wl0.phi_mode)
    ref = []
    for ds, wl in zip(sets, steps):
        cfg.set_view(wl.views[0].u, wl.views[0].v)
        cfg.register(ds)
        cfg.render(ds)
        ref.append(cfg.image())
        cfg.unregister(ds)
    # Pipeline: the upload of t+1 is enqueued while the render of t is in flight.
    cfg.register_async(sets[0])
    for t, (ds, wl) in enumerate(zip(sets, steps)):
        cfg.set_view(wl.views[0].u, wl.views[0].v)
        cfg.render(ds)
        if t + 1 < len(sets):
            cfg.register_async(sets[t + 1])
        img = cfg.image()
        print(f"[case] {form} t={t} N={wl.n}", flush=True)
        assert np.array_equal(img, ref[t]), t
        cfg.unregister(ds)
    cfg.close()


def test_async_registration_then_immediate_unregister():
    """Unregistering right after an asynchronous registration waits for the copies."""
    rng = np.random.default_rng(2)
    ds = host.Dataset([(rng.uniform(-0.5, 0.5, (5000, 3)).astype(np.float32),
                        rng.uniform(0, 1, 5000).astype(np.float32), np.full(5000, 0.02, np.float32), "GAUSSIAN")])
    cfg = host.Config(cutoff=(32, 32), samples=(64, 64), step=(0.5, 0.5), phi_mode="analytic")
    for _ in range(3):
        cfg.register_async(ds)
        cfg.unregister(ds)
    cfg.register_async(ds)
    cfg.render(ds)
    a = cfg.image()
    cfg.unregister(ds)
    cfg.register(ds)
    cfg.render(ds)
    assert np.isfinite(a).all()
    assert np.array_equal(a, cfg.image())
    cfg.unregister(ds)
    cfg.close()


@pytest.mark.parametrize("basis", ["WENDLAND_D3_C2", "GAUSSIAN"])
def test_bins_from_registration_match_bins_from_first_render(basis):
    """A registration through an MFMA config bins the groups from the host arrays; one through a
    plain config leaves the binning to the first MFMA render (from a device copy).  Both give the
    same image bit for bit."""
    rng = np.random.default_rng(9)
    n = 20000
    ds = host.Dataset([(rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32), rng.uniform(0, 1, n).astype(np.float32),
                        rng.uniform(0.005, 0.05, n).astype(np.float32), basis)])
    geo = dict(cutoff=(64, 48), samples=(128, 128), step=(0.5, 0.5), u=(0.6, 0.8, 0.0), v=(-0.48, 0.36, 0.8),
               phi_mode="analytic")
    mf = host.Config(form="mfma_split_f16", **geo)
    pl = host.Config(form="plain", **geo)
    mf.register(ds)  # bins built here
    mf.render(ds)
    a = mf.image()
    mf.unregister(ds)
    pl.register(ds)  # no bins here: the MFMA render below builds them
    mf.render(ds)
    b = mf.image()
    pl.unregister(ds)
    mf.close()
    pl.close()
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)
