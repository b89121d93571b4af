"""Batched rendering (one persistent MFMA launch per kernel variant inside a captured HIP graph)
against one-at-a-time renders, and against the FP64 oracle on a batch of drifting clusters.

The graph is replayed several times with eager renders and synchronisations between the
replays.  A batch's first call runs its launches without a graph and the next call captures
and replays them: every replay must equal that first call bit for bit (same kernels, same work
split, same ordered reductions).  Against one-at-a-time renders the images agree to FP32
rounding (<= 1e-6 relative): a batch cuts its jobs into work items for the whole batch, a
single render for itself, so the FP32 partial sums end at other particles.
"""
import numpy as np
import pytest

import oracle
from libmeshlessvis_amd import host
from libmeshlessvis_amd import workloads as W

pytestmark = pytest.mark.gpu


def views(rng, k):
    return [W.random_view(rng) for _ in range(k)]


def make_sets(rng):
    sets = []
    for n, basis in ((3000, "GAUSSIAN"), (5000, "WENDLAND_D3_C2"), (1200, "WENDLAND_D3_C2")):
        pos = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
        h = np.exp(rng.uniform(np.log(0.005), np.log(0.08), n)).astype(np.float32)
        w = (h.astype(np.float64) ** 3 * rng.uniform(0.5, 2, n)).astype(np.float32)
        sets.append(host.Dataset([(pos, w, h, basis)]))
    return sets


@pytest.mark.parametrize("tile", ["default", "quarter"])
def test_batched_graph_replays_match_eager(tile):
    print(f"[case] tile {tile}", flush=True)
    rng = np.random.default_rng(21)
    sets = make_sets(rng)
    vws = views(rng, 2)
    pairs, cfgs = [], []
    for ds in sets:
        for vw in vws:
            cfg = host.Config(cutoff=(64, 32), samples=(128, 96), step=(0.5, 0.5), u=vw.u, v=vw.v,
                              form="mfma_split_f16", phi_mode="analytic", tile=tile)
            pairs.append(ds)
            cfgs.append(cfg)
    for ds in sets:
        cfgs[0].register(ds)
    eager = []
    for ds, cfg in zip(pairs, cfgs):
        cfg.render(ds)
        eager.append(cfg.image())
    other = host.Config(cutoff=(32, 32), samples=(64, 64), step=(0.5, 0.5), form="plain", phi_mode="analytic")
    first = None
    for rep in range(4):  # rep 0 without a graph, rep 1 captures, reps 2, 3 replay
        host.render_batched(pairs, cfgs)
        imgs = [cfg.image() for cfg in cfgs]
        spectra = [cfg.spectrum() for cfg in cfgs]
        if first is None:
            first = (imgs, spectra)
        for a, b in zip(imgs, eager):
            assert oracle.rel_l2(a, b) <= 1e-6, rep
        for a, b in zip(imgs + spectra, first[0] + first[1]):
            assert np.array_equal(a, b), rep
        other.render(sets[2])  # eager work between replays
        other.synchronize()
    # A changed view is a new batch signature: the graph is rebuilt and stays right.
    cfgs[1].set_view(vws[0].u, vws[0].v)
    host.render_batched(pairs, cfgs)
    cfgs[0].synchronize()
    assert oracle.rel_l2(cfgs[1].image(), eager[0]) <= 1e-6
    for ds in sets:
        cfgs[0].unregister(ds)
    for cfg in cfgs + [other]:
        cfg.close()


def test_batched_drifting_clusters_against_oracle():
    steps = W.drifting_clusters(steps=3, views_per_step=3, n_lo=2000, n_hi=20000, seed=4)
    sets = [host.Dataset([wl.group()]) for wl in steps]
    pairs, cfgs, meta = [], [], []
    for ds, wl in zip(sets, steps):
        for vw in wl.views:
            cfgs.append(host.Config(cutoff=wl.cutoff, samples=wl.samples, step=wl.step, u=vw.u, v=vw.v,
                                    form="mfma_split_f16", phi_mode=wl.phi_mode))
            pairs.append(ds)
            meta.append((wl, vw))
    for ds in sets:
        cfgs[0].register(ds)
    host.render_batched(pairs, cfgs)
    rng = np.random.default_rng(3)
    for cfg, (wl, vw) in zip(cfgs, meta):
        F = cfg.spectrum().reshape(-1)
        idx = np.concatenate([np.arange(0, 64), rng.choice(F.size, 192, replace=False)])
        F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode, idx=idx)
        e = oracle.rel_l2(F[idx], F64)
        print(f"{wl.name} N={wl.n}: rel-L2 {e:.3e}")
        assert e <= 1e-5
    for ds in sets:
        cfgs[0].unregister(ds)
    for cfg in cfgs:
        cfg.close()


def test_mixed_forms_fall_back_to_sequential():
    rng = np.random.default_rng(5)
    sets = make_sets(rng)[:2]
    a = host.Config(cutoff=(32, 16), samples=(64, 64), step=(0.5, 0.5), form="plain", phi_mode="analytic")
    b = host.Config(cutoff=(32, 16), samples=(64, 64), step=(0.5, 0.5), form="mfma_split_f16", phi_mode="analytic")
    for ds in sets:
        a.register(ds)
    host.render_batched(sets, [a, b])
    ia, ib = a.image(), b.image()
    a.render(sets[0])
    b.render(sets[1])
    assert np.array_equal(ia, a.image())
    assert oracle.rel_l2(ib, b.image()) <= 1e-6
    for ds in sets:
        a.unregister(ds)
    a.close()
    b.close()


def test_batched_graph_after_reregistration():
    """A dataset unregistered and registered again gets new bins (its device arrays may come back
    at the same addresses): the batch graph captured for the old registration must not be replayed."""
    rng = np.random.default_rng(31)
    ds = make_sets(rng)[1]  # Wendland with radii: binned kernel, per-registration bins
    vws = views(rng, 2)
    cfgs = [host.Config(cutoff=(64, 32), samples=(128, 96), step=(0.5, 0.5), u=vw.u, v=vw.v,
                        form="mfma_split_f16", phi_mode="analytic") for vw in vws]
    eager = []
    for rep in range(3):
        cfgs[0].register(ds)
        host.render_batched([ds, ds], cfgs)
        imgs = [c.image() for c in cfgs]
        if not eager:
            for c in cfgs:
                c.render(ds)
                eager.append(c.image())
        for a, b in zip(imgs, eager):
            assert oracle.rel_l2(a, b) <= 1e-6, rep
        cfgs[0].unregister(ds)
    for c in cfgs:
        c.close()


def test_second_call_captures_third_replays():
    """Configs whose first render is the batched call itself (their first render allocates
    buffers the batch key names): the first call runs without a graph, the second captures
    and launches it, the third replays it; a new view is captured anew."""
    import ctypes as C
    from libmeshlessvis_amd import abi

    print("[case] batch graph counts", flush=True)
    rng = np.random.default_rng(23)
    sets = make_sets(rng)[:2]
    vws = views(rng, 2)
    pairs, cfgs = [], []
    for ds in sets:
        for vw in vws:
            pairs.append(ds)
            cfgs.append(host.Config(cutoff=(64, 32), samples=(128, 96), step=(0.5, 0.5), u=vw.u, v=vw.v,
                                    form="mfma_split_f16", phi_mode="analytic"))
    for ds in sets:
        cfgs[0].register(ds)

    def counts():
        cap, launches = C.c_int(-1), C.c_longlong(-1)
        abi.lib().vis_debug_batch_graph_counts(cfgs[0].p, C.byref(cap), C.byref(launches))
        abi.check("vis_debug_batch_graph_counts")
        return cap.value, launches.value

    expected = [(0, 0), (1, 1), (1, 2), (1, 3)]
    images = []
    for call, want in enumerate(expected):
        host.render_batched(pairs, cfgs)
        images.append([cfg.image() for cfg in cfgs])
        assert counts() == want, (call, counts())
    for imgs in images[1:]:
        for a, b in zip(imgs, images[0]):
            assert np.array_equal(a, b)
    cfgs[1].set_view(vws[0].u, vws[0].v)  # a new batch: its first call runs without a graph
    host.render_batched(pairs, cfgs)
    assert counts() == (1, 3)
    host.render_batched(pairs, cfgs)
    assert counts() == (2, 4)
    for ds in sets:
        cfgs[0].unregister(ds)
    for c in cfgs:
        c.close()
